/* Oracle (test infrastructure only).  Nothing in the product links this code:
 * only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark load it, as the checker.
 *
 * BLS12-381 scalar field Fr, base field Fq and the G1 group law in textbook
 * Jacobian coordinates.  This is a plain CPU restatement of what the reference
 * CPU backend obtains from arkworks 0.6 (ark-ff / ark-ec / ark-bls12-381):
 *   - field arithmetic used by backend/src/cpu/ (all .rs files);
 *   - E::G1::msm_unchecked(..).into_affine() used by
 *     backend/src/cpu/multilinear_kzg.rs:55,131;
 *   - the canonical encodings of backend/src/ascend910b3/encoding.rs:12-108
 *     (Fr LE32, Fq LE48, G1 packet = tag | x LE48 | y LE48).
 */
#ifndef HP_ORACLE_CORE_H
#define HP_ORACLE_CORE_H

#include <stdint.h>
#include <string.h>

#define FN 4
#define FP fr
#include "field_impl.h"
#undef FN
#undef FP

#define FN 6
#define FP fq
#include "field_impl.h"
#undef FN
#undef FP

/* Affine point: inf != 0 means the identity (coordinates ignored). */
typedef struct { fq_t x, y; int inf; } g1_aff;
/* Jacobian point (X/Z^2, Y/Z^3); Z == 0 means the identity. */
typedef struct { fq_t X, Y, Z; } g1_jac;

void oracle_init(void);

static inline void g1_jac_set_inf(g1_jac *r) {
  fq_one(&r->X);
  fq_one(&r->Y);
  fq_zero(&r->Z);
}

static inline int g1_jac_is_inf(const g1_jac *p) { return fq_is_zero(&p->Z); }

static inline void g1_jac_from_aff(g1_jac *r, const g1_aff *a) {
  if (a->inf) {
    g1_jac_set_inf(r);
    return;
  }
  r->X = a->x;
  r->Y = a->y;
  fq_one(&r->Z);
}

/* Doubling on y^2 = x^3 + b (a = 0):
 *   S = 4 X Y^2, M = 3 X^2, X' = M^2 - 2 S, Y' = M (S - X') - 8 Y^4, Z' = 2 Y Z. */
static inline void g1_jac_double(g1_jac *r, const g1_jac *p) {
  if (g1_jac_is_inf(p) || fq_is_zero(&p->Y)) {
    g1_jac_set_inf(r);
    return;
  }
  fq_t yy, s, m, t, x3, y3, z3, yyyy;
  fq_sqr(&yy, &p->Y);
  fq_mul(&s, &p->X, &yy);
  fq_add(&s, &s, &s);
  fq_add(&s, &s, &s);
  fq_sqr(&t, &p->X);
  fq_add(&m, &t, &t);
  fq_add(&m, &m, &t);
  fq_sqr(&x3, &m);
  fq_sub(&x3, &x3, &s);
  fq_sub(&x3, &x3, &s);
  fq_sqr(&yyyy, &yy);
  fq_add(&yyyy, &yyyy, &yyyy);
  fq_add(&yyyy, &yyyy, &yyyy);
  fq_add(&yyyy, &yyyy, &yyyy);
  fq_sub(&t, &s, &x3);
  fq_mul(&y3, &m, &t);
  fq_sub(&y3, &y3, &yyyy);
  fq_mul(&z3, &p->Y, &p->Z);
  fq_add(&z3, &z3, &z3);
  r->X = x3;
  r->Y = y3;
  r->Z = z3;
}

/* General addition: U1 = X1 Z2^2, U2 = X2 Z1^2, S1 = Y1 Z2^3, S2 = Y2 Z1^3,
 * H = U2 - U1, R = S2 - S1; H = 0 means equal x: double or cancel. */
static inline void g1_jac_add(g1_jac *r, const g1_jac *p, const g1_jac *q) {
  if (g1_jac_is_inf(p)) {
    *r = *q;
    return;
  }
  if (g1_jac_is_inf(q)) {
    *r = *p;
    return;
  }
  fq_t z1z1, z2z2, u1, u2, s1, s2, h, rr, t, hh, hhh, v, x3, y3, z3;
  fq_sqr(&z1z1, &p->Z);
  fq_sqr(&z2z2, &q->Z);
  fq_mul(&u1, &p->X, &z2z2);
  fq_mul(&u2, &q->X, &z1z1);
  fq_mul(&t, &q->Z, &z2z2);
  fq_mul(&s1, &p->Y, &t);
  fq_mul(&t, &p->Z, &z1z1);
  fq_mul(&s2, &q->Y, &t);
  fq_sub(&h, &u2, &u1);
  fq_sub(&rr, &s2, &s1);
  if (fq_is_zero(&h)) {
    if (fq_is_zero(&rr)) {
      g1_jac_double(r, p);
    } else {
      g1_jac_set_inf(r);
    }
    return;
  }
  fq_sqr(&hh, &h);
  fq_mul(&hhh, &hh, &h);
  fq_mul(&v, &u1, &hh);
  fq_sqr(&x3, &rr);
  fq_sub(&x3, &x3, &hhh);
  fq_sub(&x3, &x3, &v);
  fq_sub(&x3, &x3, &v);
  fq_sub(&t, &v, &x3);
  fq_mul(&y3, &rr, &t);
  fq_mul(&t, &s1, &hhh);
  fq_sub(&y3, &y3, &t);
  fq_mul(&z3, &p->Z, &q->Z);
  fq_mul(&z3, &z3, &h);
  r->X = x3;
  r->Y = y3;
  r->Z = z3;
}

/* Mixed addition P + Q with Q affine (Z2 = 1): U2 = x2 Z1^2, S2 = y2 Z1^3,
 * H = U2 - X1, R = S2 - Y1; X3 = R^2 - H^3 - 2 X1 H^2,
 * Y3 = R (X1 H^2 - X3) - Y1 H^3, Z3 = Z1 H (8 products and 3 squarings,
 * against 12 and 4 for the general formula with Z2 = 1).  H = 0 means equal
 * x: double or cancel. */
static inline void g1_jac_add_aff(g1_jac *r, const g1_jac *p, const g1_aff *q) {
  if (q->inf) {
    *r = *p;
    return;
  }
  if (g1_jac_is_inf(p)) {
    g1_jac_from_aff(r, q);
    return;
  }
  fq_t z1z1, u2, s2, h, rr, hh, hhh, v, x3, y3, z3, t;
  fq_sqr(&z1z1, &p->Z);
  fq_mul(&u2, &q->x, &z1z1);
  fq_mul(&t, &p->Z, &z1z1);
  fq_mul(&s2, &q->y, &t);
  fq_sub(&h, &u2, &p->X);
  fq_sub(&rr, &s2, &p->Y);
  if (fq_is_zero(&h)) {
    if (fq_is_zero(&rr)) {
      g1_jac_double(r, p);
    } else {
      g1_jac_set_inf(r);
    }
    return;
  }
  fq_sqr(&hh, &h);
  fq_mul(&hhh, &hh, &h);
  fq_mul(&v, &p->X, &hh);
  fq_sqr(&x3, &rr);
  fq_sub(&x3, &x3, &hhh);
  fq_sub(&x3, &x3, &v);
  fq_sub(&x3, &x3, &v);
  fq_sub(&t, &v, &x3);
  fq_mul(&y3, &rr, &t);
  fq_mul(&t, &p->Y, &hhh);
  fq_sub(&y3, &y3, &t);
  fq_mul(&z3, &p->Z, &h);
  r->X = x3;
  r->Y = y3;
  r->Z = z3;
}

static inline void g1_jac_to_aff(g1_aff *r, const g1_jac *p) {
  if (g1_jac_is_inf(p)) {
    memset(r, 0, sizeof *r);
    r->inf = 1;
    return;
  }
  fq_t zi, zi2, zi3;
  fq_inv(&zi, &p->Z);
  fq_sqr(&zi2, &zi);
  fq_mul(&zi3, &zi2, &zi);
  fq_mul(&r->x, &p->X, &zi2);
  fq_mul(&r->y, &p->Y, &zi3);
  r->inf = 0;
}

static inline void g1_aff_neg(g1_aff *r, const g1_aff *a) {
  *r = *a;
  if (!a->inf) fq_neg(&r->y, &a->y);
}

/* y^2 == x^3 + 4 */
static inline int g1_aff_on_curve(const g1_aff *a) {
  if (a->inf) return 1;
  fq_t lhs, rhs, four;
  fq_sqr(&lhs, &a->y);
  fq_sqr(&rhs, &a->x);
  fq_mul(&rhs, &rhs, &a->x);
  fq_from_u64(&four, 4);
  fq_add(&rhs, &rhs, &four);
  return fq_eq(&lhs, &rhs);
}

/* Double-and-add over canonical little-endian scalar limbs (MSB first). */
static inline void g1_jac_mul_limbs(g1_jac *r, const g1_jac *p, const uint64_t *k, int nlimbs) {
  g1_jac acc;
  g1_jac_set_inf(&acc);
  for (int i = nlimbs - 1; i >= 0; --i) {
    for (int b = 63; b >= 0; --b) {
      g1_jac_double(&acc, &acc);
      if ((k[i] >> b) & 1) g1_jac_add(&acc, &acc, p);
    }
  }
  *r = acc;
}

/* 97-byte packet codec: tag (0 finite, 1 identity with zero coordinates). */
int g1_from_packet(g1_aff *r, const uint8_t *packet);
void g1_to_packet(uint8_t *packet, const g1_aff *a);

#endif
