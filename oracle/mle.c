/* Oracle (test infrastructure only): multilinear-extension operations, the
 * eq table, multilinear KZG opening and SRS layers.
 *
 * Each function follows the reference CPU backend line by line in meaning:
 *   oracle_mle_evaluate        backend/src/cpu/multilinear_polynomial.rs:84-100
 *                              (ark-poly fix_variables, point[0] = lowest bit)
 *   oracle_mle_fold            backend/src/cpu/sum_check.rs:103-118
 *   oracle_mle_linear_combination  backend/src/cpu/multilinear_polynomial.rs:50-82
 *   oracle_eq_x_r              backend/src/common/eq.rs:38-95
 *   oracle_eq_eval             backend/src/common/eq.rs:15-29
 *   oracle_frac                backend/src/cpu/multilinear_polynomial.rs:110-140
 *   oracle_product             backend/src/cpu/multilinear_polynomial.rs:142-181
 *   oracle_nums_denoms         backend/src/cpu/multilinear_polynomial.rs:183-217
 *   oracle_product_factors     backend/src/cpu/multilinear_polynomial.rs:219-253
 *   oracle_kzg_open            backend/src/cpu/multilinear_kzg.rs:79-137
 *   oracle_srs_layer           subroutines/src/pcs/multilinear_kzg/srs.rs:158-220
 */
#include <pthread.h>
#include <stdlib.h>
#include "oracle.h"
#include "oracle_core.h"

static fr_t *load_table(const uint8_t *bytes, size_t n, int *ok) {
  fr_t *t = (fr_t *)malloc(sizeof(fr_t) * (n ? n : 1));
  *ok = 1;
  for (size_t i = 0; i < n; ++i)
    if (!fr_from_bytes(&t[i], bytes + 32 * i)) *ok = 0;
  if (!*ok) {
    free(t);
    return NULL;
  }
  return t;
}

static void store_table(uint8_t *bytes, const fr_t *t, size_t n) {
  for (size_t i = 0; i < n; ++i) fr_to_bytes(bytes + 32 * i, &t[i]);
}

/* In place: t[b] = t[2b] + r (t[2b+1] - t[2b]) for b < half. */
static void fold_in_place(fr_t *t, size_t half, const fr_t *r) {
  for (size_t b = 0; b < half; ++b) {
    fr_t lo = t[2 * b], d;
    fr_sub(&d, &t[2 * b + 1], &lo);
    fr_mul(&d, &d, r);
    fr_add(&t[b], &lo, &d);
  }
}

int oracle_mle_evaluate(const uint8_t *table, size_t nv, const uint8_t *point, uint8_t *out) {
  oracle_init();
  if (nv >= 40) return ORACLE_INVALID_PARAMETERS;
  int ok;
  size_t n = (size_t)1 << nv;
  fr_t *t = load_table(table, n, &ok);
  if (!ok) return ORACLE_INVALID_PARAMETERS;
  for (size_t i = 0; i < nv; ++i) {
    fr_t r;
    if (!fr_from_bytes(&r, point + 32 * i)) {
      free(t);
      return ORACLE_INVALID_PARAMETERS;
    }
    fold_in_place(t, n >> (i + 1), &r);
  }
  fr_to_bytes(out, &t[0]);
  free(t);
  return ORACLE_OK;
}

int oracle_mle_fold(const uint8_t *table, size_t nv, const uint8_t *r, uint8_t *out) {
  oracle_init();
  if (nv == 0 || nv >= 40) return ORACLE_INVALID_PARAMETERS;
  int ok;
  size_t n = (size_t)1 << nv;
  fr_t *t = load_table(table, n, &ok);
  fr_t rr;
  if (!ok) return ORACLE_INVALID_PARAMETERS;
  if (!fr_from_bytes(&rr, r)) {
    free(t);
    return ORACLE_INVALID_PARAMETERS;
  }
  fold_in_place(t, n / 2, &rr);
  store_table(out, t, n / 2);
  free(t);
  return ORACLE_OK;
}

int oracle_mle_linear_combination(const uint8_t *tables, const uint8_t *coeffs, size_t count,
                                  size_t nv, uint8_t *out) {
  oracle_init();
  if (count == 0 || nv >= 40) return ORACLE_INVALID_PARAMETERS;
  size_t n = (size_t)1 << nv;
  fr_t *acc = (fr_t *)calloc(n, sizeof(fr_t));
  for (size_t k = 0; k < count; ++k) {
    fr_t c;
    if (!fr_from_bytes(&c, coeffs + 32 * k)) {
      free(acc);
      return ORACLE_INVALID_PARAMETERS;
    }
    if (fr_is_zero(&c)) continue;
    for (size_t j = 0; j < n; ++j) {
      fr_t v;
      if (!fr_from_bytes(&v, tables + 32 * (k * n + j))) {
        free(acc);
        return ORACLE_INVALID_PARAMETERS;
      }
      fr_mul(&v, &v, &c);
      fr_add(&acc[j], &acc[j], &v);
    }
  }
  store_table(out, acc, n);
  free(acc);
  return ORACLE_OK;
}

/* Recursive construction of eq.rs:67-95: start from [1 - r_last, r_last] and
 * prepend one coordinate per level (even child b - r b, odd child r b). */
int oracle_eq_x_r(const uint8_t *r, size_t nv, uint8_t *out) {
  oracle_init();
  if (nv == 0 || nv >= 40) return ORACLE_INVALID_PARAMETERS;
  fr_t *rs = (fr_t *)malloc(sizeof(fr_t) * nv);
  for (size_t i = 0; i < nv; ++i)
    if (!fr_from_bytes(&rs[i], r + 32 * i)) {
      free(rs);
      return ORACLE_INVALID_PARAMETERS;
    }
  size_t n = (size_t)1 << nv;
  fr_t *buf = (fr_t *)malloc(sizeof(fr_t) * n);
  fr_t *res = (fr_t *)malloc(sizeof(fr_t) * n);
  fr_t one;
  fr_one(&one);
  fr_sub(&buf[0], &one, &rs[nv - 1]);
  buf[1] = rs[nv - 1];
  size_t len = 2;
  for (size_t k = nv - 1; k-- > 0;) {
    for (size_t i = 0; i < 2 * len; ++i) {
      fr_t bi = buf[i >> 1], tmp;
      fr_mul(&tmp, &rs[k], &bi);
      if ((i & 1) == 0)
        fr_sub(&res[i], &bi, &tmp);
      else
        res[i] = tmp;
    }
    len *= 2;
    fr_t *sw = buf;
    buf = res;
    res = sw;
  }
  store_table(out, buf, n);
  free(rs);
  free(buf);
  free(res);
  return ORACLE_OK;
}

int oracle_eq_eval(const uint8_t *x, const uint8_t *y, size_t n, uint8_t *out) {
  oracle_init();
  fr_t res, one;
  fr_one(&one);
  res = one;
  for (size_t i = 0; i < n; ++i) {
    fr_t xi, yi, xy, t;
    if (!fr_from_bytes(&xi, x + 32 * i) || !fr_from_bytes(&yi, y + 32 * i))
      return ORACLE_INVALID_PARAMETERS;
    fr_mul(&xy, &xi, &yi);
    fr_add(&t, &xy, &xy);
    fr_sub(&t, &t, &xi);
    fr_sub(&t, &t, &yi);
    fr_add(&t, &t, &one);
    fr_mul(&res, &res, &t);
  }
  fr_to_bytes(out, &res);
  return ORACLE_OK;
}

int oracle_frac(const uint8_t *nums, const uint8_t *dens, size_t count, size_t nv, uint8_t *out) {
  oracle_init();
  if (count == 0 || nv >= 40) return ORACLE_INVALID_PARAMETERS;
  size_t n = (size_t)1 << nv;
  int ok;
  fr_t *num = load_table(nums, n, &ok);
  if (!ok) return ORACLE_INVALID_PARAMETERS;
  fr_t *den = load_table(dens, n, &ok);
  if (!ok) {
    free(num);
    return ORACLE_INVALID_PARAMETERS;
  }
  int status = ORACLE_OK;
  for (size_t k = 1; k < count && status == ORACLE_OK; ++k) {
    for (size_t j = 0; j < n; ++j) {
      fr_t f, g;
      if (!fr_from_bytes(&f, nums + 32 * (k * n + j)) || !fr_from_bytes(&g, dens + 32 * (k * n + j))) {
        status = ORACLE_INVALID_PARAMETERS;
        break;
      }
      fr_mul(&num[j], &num[j], &f);
      fr_mul(&den[j], &den[j], &g);
    }
  }
  for (size_t j = 0; j < n && status == ORACLE_OK; ++j)
    if (fr_is_zero(&den[j])) status = ORACLE_INVALID_PARAMETERS;
  if (status == ORACLE_OK) {
    /* Montgomery batch inversion (ark-ff batch_inversion semantics). */
    fr_t *prefix = (fr_t *)malloc(sizeof(fr_t) * n);
    fr_t acc;
    fr_one(&acc);
    for (size_t j = 0; j < n; ++j) {
      prefix[j] = acc;
      fr_mul(&acc, &acc, &den[j]);
    }
    fr_t inv;
    fr_inv(&inv, &acc);
    for (size_t j = n; j-- > 0;) {
      fr_t dj;
      fr_mul(&dj, &inv, &prefix[j]);
      fr_mul(&inv, &inv, &den[j]);
      fr_mul(&num[j], &num[j], &dj);
    }
    free(prefix);
    store_table(out, num, n);
  }
  free(num);
  free(den);
  return status;
}

int oracle_product(const uint8_t *frac, size_t nv, uint8_t *out) {
  oracle_init();
  if (nv == 0 || nv >= 40) return ORACLE_INVALID_PARAMETERS;
  size_t n = (size_t)1 << nv;
  int ok;
  fr_t *f = load_table(frac, n, &ok);
  if (!ok) return ORACLE_INVALID_PARAMETERS;
  fr_t *e = (fr_t *)malloc(sizeof(fr_t) * n);
  size_t len = 0;
  for (size_t i = 0; i + 1 < n; i += 2) fr_mul(&e[len++], &f[i], &f[i + 1]);
  for (size_t child = 0; child + 2 < n; child += 2) {
    fr_mul(&e[len], &e[child], &e[child + 1]);
    ++len;
  }
  fr_zero(&e[len++]);
  store_table(out, e, n);
  free(f);
  free(e);
  return ORACLE_OK;
}

int oracle_nums_denoms(const uint8_t *beta, const uint8_t *gamma, const uint8_t *fxs,
                       const uint8_t *gxs, const uint8_t *perms, size_t count, size_t nv,
                       uint8_t *out_nums, uint8_t *out_dens) {
  oracle_init();
  if (count == 0 || nv >= 40) return ORACLE_INVALID_PARAMETERS;
  fr_t b, g;
  if (!fr_from_bytes(&b, beta) || !fr_from_bytes(&g, gamma)) return ORACLE_INVALID_PARAMETERS;
  size_t n = (size_t)1 << nv;
  for (size_t c = 0; c < count; ++c) {
    fr_t id;
    fr_from_u64(&id, (uint64_t)(c * n));
    fr_mul(&id, &id, &b);
    for (size_t j = 0; j < n; ++j) {
      size_t at = c * n + j;
      fr_t f, gx, p, num, den, t;
      if (!fr_from_bytes(&f, fxs + 32 * at) || !fr_from_bytes(&gx, gxs + 32 * at) ||
          !fr_from_bytes(&p, perms + 32 * at))
        return ORACLE_INVALID_PARAMETERS;
      fr_add(&num, &f, &id);
      fr_add(&num, &num, &g);
      fr_mul(&t, &b, &p);
      fr_add(&den, &gx, &t);
      fr_add(&den, &den, &g);
      fr_to_bytes(out_nums + 32 * at, &num);
      fr_to_bytes(out_dens + 32 * at, &den);
      fr_add(&id, &id, &b);
    }
  }
  return ORACLE_OK;
}

int oracle_product_factors(const uint8_t *frac, const uint8_t *prod, size_t nv, uint8_t *p1,
                           uint8_t *p2) {
  oracle_init();
  if (nv == 0 || nv >= 40) return ORACLE_INVALID_PARAMETERS;
  size_t n = (size_t)1 << nv, half = n / 2;
  for (size_t i = 0; i < half; ++i) {
    memcpy(p1 + 32 * i, frac + 32 * (2 * i), 32);
    memcpy(p2 + 32 * i, frac + 32 * (2 * i + 1), 32);
    memcpy(p1 + 32 * (half + i), prod + 32 * (2 * i), 32);
    memcpy(p2 + 32 * (half + i), prod + 32 * (2 * i + 1), 32);
  }
  return ORACLE_OK;
}

int oracle_kzg_open(const uint8_t *layers97, const uint8_t *table, size_t nv, const uint8_t *point,
                    int threads, uint8_t *out_quotients97, uint8_t *out_eval) {
  oracle_init();
  if (nv >= 40) return ORACLE_INVALID_PARAMETERS;
  size_t n = (size_t)1 << nv;
  int ok;
  fr_t *f = load_table(table, n, &ok);
  if (!ok) return ORACLE_INVALID_PARAMETERS;
  if (nv == 0) {
    fr_to_bytes(out_eval, &f[0]);
    free(f);
    return ORACLE_OK;
  }
  fr_t *q = (fr_t *)malloc(sizeof(fr_t) * (n / 2));
  uint8_t *qbytes = (uint8_t *)malloc(32 * (n / 2));
  size_t layer_offset = 0;
  for (size_t j = 0; j < nv; ++j) {
    fr_t x;
    if (!fr_from_bytes(&x, point + 32 * j)) {
      free(f);
      free(q);
      free(qbytes);
      return ORACLE_INVALID_PARAMETERS;
    }
    size_t cur = n >> (j + 1);
    for (size_t b = 0; b < cur; ++b) {
      fr_t lo = f[2 * b], t;
      fr_sub(&q[b], &f[2 * b + 1], &lo);
      fr_mul(&t, &q[b], &x);
      fr_add(&f[b], &lo, &t);
    }
    store_table(qbytes, q, cur);
    int st = oracle_msm_pippenger(layers97 + 97 * layer_offset, qbytes, cur, threads,
                                  out_quotients97 + 97 * j);
    if (st != ORACLE_OK) {
      free(f);
      free(q);
      free(qbytes);
      return st;
    }
    layer_offset += cur;
  }
  fr_to_bytes(out_eval, &f[0]);
  free(f);
  free(q);
  free(qbytes);
  return ORACLE_OK;
}

/* ---- SRS layer: g * eq(t, x) for x in {0,1}^len (bit 0 <-> t[0]). ---- */
void oracle_generator_aff(g1_aff *g);

typedef struct {
  const g1_aff *table; /* 32 windows x 256 digits, affine */
  const fr_t *scalars;
  uint8_t *out97;
  size_t begin, end;
} srs_job;

static void *srs_worker(void *arg) {
  srs_job *job = (srs_job *)arg;
  size_t cnt = job->end - job->begin;
  g1_jac *acc = (g1_jac *)malloc(sizeof(g1_jac) * (cnt ? cnt : 1));
  fr_t *pref = (fr_t *)malloc(sizeof(fq_t) * (cnt ? cnt : 1) + 64);
  fq_t *prefix = (fq_t *)pref;
  for (size_t i = 0; i < cnt; ++i) {
    uint64_t k[4];
    fr_to_canonical_limbs(k, &job->scalars[job->begin + i]);
    g1_jac_set_inf(&acc[i]);
    for (int w = 0; w < 32; ++w) {
      unsigned d = (unsigned)((k[w / 8] >> (8 * (w % 8))) & 0xff);
      if (d) g1_jac_add_aff(&acc[i], &acc[i], &job->table[w * 256 + d]);
    }
  }
  /* batch inversion of the Z coordinates (identity points skipped) */
  fq_t run;
  fq_one(&run);
  for (size_t i = 0; i < cnt; ++i) {
    prefix[i] = run;
    if (!g1_jac_is_inf(&acc[i])) fq_mul(&run, &run, &acc[i].Z);
  }
  fq_t inv;
  fq_inv(&inv, &run);
  for (size_t i = cnt; i-- > 0;) {
    g1_aff a;
    if (g1_jac_is_inf(&acc[i])) {
      a.inf = 1;
      g1_to_packet(job->out97 + 97 * (job->begin + i), &a);
      continue;
    }
    fq_t zi, zi2, zi3;
    fq_mul(&zi, &inv, &prefix[i]);
    fq_mul(&inv, &inv, &acc[i].Z);
    fq_sqr(&zi2, &zi);
    fq_mul(&zi3, &zi2, &zi);
    fq_mul(&a.x, &acc[i].X, &zi2);
    fq_mul(&a.y, &acc[i].Y, &zi3);
    a.inf = 0;
    g1_to_packet(job->out97 + 97 * (job->begin + i), &a);
  }
  free(acc);
  free(pref);
  return NULL;
}

int oracle_srs_layer(const uint8_t *t, size_t len, int threads, uint8_t *out97) {
  oracle_init();
  if (len >= 40) return ORACLE_INVALID_PARAMETERS;
  size_t n = (size_t)1 << len;
  fr_t *eq = (fr_t *)malloc(sizeof(fr_t) * n);
  if (len == 0) {
    fr_one(&eq[0]);
  } else {
    uint8_t *tmp = (uint8_t *)malloc(32 * n);
    int st = oracle_eq_x_r(t, len, tmp);
    if (st != ORACLE_OK) {
      free(tmp);
      free(eq);
      return st;
    }
    for (size_t i = 0; i < n; ++i) fr_from_bytes(&eq[i], tmp + 32 * i);
    free(tmp);
  }
  /* fixed-base table: table[w][d] = d * 2^(8w) * g */
  g1_aff *table = (g1_aff *)malloc(sizeof(g1_aff) * 32 * 256);
  g1_aff g;
  oracle_generator_aff(&g);
  g1_jac base;
  g1_jac_from_aff(&base, &g);
  for (int w = 0; w < 32; ++w) {
    g1_jac cur;
    g1_jac_set_inf(&cur);
    table[w * 256].inf = 1;
    for (int d = 1; d < 256; ++d) {
      g1_jac_add(&cur, &cur, &base);
      g1_jac_to_aff(&table[w * 256 + d], &cur);
    }
    for (int s = 0; s < 8; ++s) g1_jac_double(&base, &base);
  }
  if (threads < 1) threads = 1;
  pthread_t *tids = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
  srs_job *jobs = (srs_job *)calloc((size_t)threads, sizeof(srs_job));
  size_t slice = (n + threads - 1) / threads;
  for (int i = 0; i < threads; ++i) {
    jobs[i].table = table;
    jobs[i].scalars = eq;
    jobs[i].out97 = out97;
    jobs[i].begin = slice * i < n ? slice * i : n;
    jobs[i].end = slice * (i + 1) < n ? slice * (i + 1) : n;
    pthread_create(&tids[i], NULL, srs_worker, &jobs[i]);
  }
  for (int i = 0; i < threads; ++i) pthread_join(tids[i], NULL);
  free(tids);
  free(jobs);
  free(table);
  free(eq);
  return ORACLE_OK;
}
