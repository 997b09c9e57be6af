/* Oracle (test infrastructure only): field parameters, encodings and batch
 * field operations.  Moduli are the published BLS12-381 constants (the
 * reference restates them at backend/src/ascend910b3/native/field_params.h:9-38). */
#include <stdlib.h>
#include "oracle.h"
#include "oracle_core.h"

fr_params_t fr_P;
fq_params_t fq_P;

static const uint64_t FR_MODULUS[4] = {
    0xffffffff00000001ULL, 0x53bda402fffe5bfeULL, 0x3339d80809a1d805ULL, 0x73eda753299d7d48ULL};
static const uint64_t FQ_MODULUS[6] = {
    0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL, 0x6730d2a0f6b0f624ULL,
    0x64774b84f38512bfULL, 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};

static int g_initialized = 0;

void oracle_init(void) {
  if (g_initialized) return;
  fr_init_params(FR_MODULUS);
  fq_init_params(FQ_MODULUS);
  g_initialized = 1;
}

int g1_from_packet(g1_aff *r, const uint8_t *packet) {
  if (packet[0] == 1) {
    for (int i = 1; i < 97; ++i)
      if (packet[i]) return 0;
    memset(r, 0, sizeof *r);
    r->inf = 1;
    return 1;
  }
  if (packet[0] != 0) return 0;
  if (!fq_from_bytes(&r->x, packet + 1)) return 0;
  if (!fq_from_bytes(&r->y, packet + 49)) return 0;
  r->inf = 0;
  if (fq_is_zero(&r->x) && fq_is_zero(&r->y)) return 0;
  return g1_aff_on_curve(r);
}

void g1_to_packet(uint8_t *packet, const g1_aff *a) {
  memset(packet, 0, 97);
  if (a->inf) {
    packet[0] = 1;
    return;
  }
  fq_to_bytes(packet + 1, &a->x);
  fq_to_bytes(packet + 49, &a->y);
}

#define DEFINE_OP(PFX, BYTES)                                                          \
  int oracle_##PFX##_op(int op, const uint8_t *a, const uint8_t *b, uint8_t *out,      \
                        size_t n) {                                                    \
    oracle_init();                                                                     \
    for (size_t i = 0; i < n; ++i) {                                                   \
      PFX##_t x, y, z;                                                                 \
      if (!PFX##_from_bytes(&x, a + BYTES * i)) return ORACLE_INVALID_PARAMETERS;      \
      if (op <= 2 && !PFX##_from_bytes(&y, b + BYTES * i)) return ORACLE_INVALID_PARAMETERS; \
      switch (op) {                                                                    \
        case 0: PFX##_add(&z, &x, &y); break;                                          \
        case 1: PFX##_sub(&z, &x, &y); break;                                          \
        case 2: PFX##_mul(&z, &x, &y); break;                                          \
        case 3: PFX##_neg(&z, &x); break;                                              \
        case 4: PFX##_sqr(&z, &x); break;                                              \
        case 5:                                                                        \
          if (PFX##_is_zero(&x)) return ORACLE_INVALID_PARAMETERS;                     \
          PFX##_inv(&z, &x);                                                           \
          break;                                                                       \
        default: return ORACLE_INVALID_PARAMETERS;                                     \
      }                                                                                \
      PFX##_to_bytes(out + BYTES * i, &z);                                             \
    }                                                                                  \
    return ORACLE_OK;                                                                  \
  }

DEFINE_OP(fr, 32)
DEFINE_OP(fq, 48)

/* The standard BLS12-381 G1 generator (published decimal coordinates). */
static const uint8_t G1_GEN_X[48] = {
    0xbb, 0xc6, 0x22, 0xdb, 0x0a, 0xf0, 0x3a, 0xfb, 0xef, 0x1a, 0x7a, 0xf9, 0x3f, 0xe8, 0x55, 0x6c,
    0x58, 0xac, 0x1b, 0x17, 0x3f, 0x3a, 0x4e, 0xa1, 0x05, 0xb9, 0x74, 0x97, 0x4f, 0x8c, 0x68, 0xc3,
    0x0f, 0xac, 0xa9, 0x4f, 0x8c, 0x63, 0x95, 0x26, 0x94, 0xd7, 0x97, 0x31, 0xa7, 0xd3, 0xf1, 0x17};
static const uint8_t G1_GEN_Y[48] = {
    0xe1, 0xe7, 0xc5, 0x46, 0x29, 0x23, 0xaa, 0x0c, 0xe4, 0x8a, 0x88, 0xa2, 0x44, 0xc7, 0x3c, 0xd0,
    0xed, 0xb3, 0x04, 0x2c, 0xcb, 0x18, 0xdb, 0x00, 0xf6, 0x0a, 0xd0, 0xd5, 0x95, 0xe0, 0xf5, 0xfc,
    0xe4, 0x8a, 0x1d, 0x74, 0xed, 0x30, 0x9e, 0xa0, 0xf1, 0xa0, 0xaa, 0xe3, 0x81, 0xf4, 0xb3, 0x08};

void oracle_generator_aff(g1_aff *g) {
  oracle_init();
  fq_from_bytes(&g->x, G1_GEN_X);
  fq_from_bytes(&g->y, G1_GEN_Y);
  g->inf = 0;
}

int oracle_g1_generator(uint8_t *out97) {
  g1_aff g;
  oracle_generator_aff(&g);
  g1_to_packet(out97, &g);
  return ORACLE_OK;
}

int oracle_g1_on_curve(const uint8_t *p97) {
  oracle_init();
  g1_aff a;
  return g1_from_packet(&a, p97) ? 1 : 0;
}

int oracle_g1_add(const uint8_t *p97, const uint8_t *q97, uint8_t *out97) {
  oracle_init();
  g1_aff a, b, c;
  if (!g1_from_packet(&a, p97) || !g1_from_packet(&b, q97)) return ORACLE_INVALID_PARAMETERS;
  g1_jac ja, r;
  g1_jac_from_aff(&ja, &a);
  g1_jac_add_aff(&r, &ja, &b);
  g1_jac_to_aff(&c, &r);
  g1_to_packet(out97, &c);
  return ORACLE_OK;
}

int oracle_g1_mul(const uint8_t *p97, const uint8_t *k32, uint8_t *out97) {
  oracle_init();
  g1_aff a, c;
  if (!g1_from_packet(&a, p97)) return ORACLE_INVALID_PARAMETERS;
  uint64_t k[4];
  for (int i = 0; i < 4; ++i) {
    uint64_t w = 0;
    for (int j = 7; j >= 0; --j) w = (w << 8) | k32[8 * i + j];
    k[i] = w;
  }
  g1_jac ja, r;
  g1_jac_from_aff(&ja, &a);
  g1_jac_mul_limbs(&r, &ja, k, 4);
  g1_jac_to_aff(&c, &r);
  g1_to_packet(out97, &c);
  return ORACLE_OK;
}
