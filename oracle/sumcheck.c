/* Oracle (test infrastructure only): SumCheck prover state.
 *
 * Follows backend/src/cpu/sum_check.rs:35-217 in meaning:
 *   - create validates n >= 1 and d >= 1 and keeps (coefficient, factor
 *     indices) products with repeats, over copies of the distinct factor
 *     tables (inputs are never mutated);
 *   - round r > 0 first folds every table with the previous challenge
 *     (fix_variables, lowest variable first, :103-118);
 *   - each product is evaluated at t = 0..arity by stepping eval += step
 *     (:129-163), scaled by its coefficient, and extrapolated to t = arity+1..d
 *     with barycentric weights over the points 0..arity (:165-217).
 * The device path evaluates every t directly instead; both are the same
 * polynomial values, so this route is an independent check of it.
 * Validation order and status kinds follow the reference bridge contract
 * (backend/src/ascend910b3/native/bridge_sumcheck.cpp:226-331,519-550). */
#include <stdlib.h>
#include "oracle.h"
#include "oracle_core.h"

typedef struct {
  size_t nv, degree, round;
  size_t factor_count, product_count;
  fr_t *tables; /* factor_count x 2^nv, folded in place */
  fr_t *coeffs;
  size_t *offsets, *indices;
  fr_t **weights; /* weights[k] for points 0..k, k = 1..degree-1 */
} sc_state;

static void batch_invert(fr_t *v, size_t n) {
  fr_t *prefix = (fr_t *)malloc(sizeof(fr_t) * (n ? n : 1));
  fr_t acc;
  fr_one(&acc);
  for (size_t i = 0; i < n; ++i) {
    prefix[i] = acc;
    if (!fr_is_zero(&v[i])) fr_mul(&acc, &acc, &v[i]);
  }
  fr_t inv;
  fr_inv(&inv, &acc);
  for (size_t i = n; i-- > 0;) {
    if (fr_is_zero(&v[i])) continue;
    fr_t t;
    fr_mul(&t, &inv, &prefix[i]);
    fr_mul(&inv, &inv, &v[i]);
    v[i] = t;
  }
  free(prefix);
}

void *oracle_sumcheck_create(size_t nv, size_t max_degree, const uint8_t *tables,
                             size_t factor_count, const uint8_t *coeffs, const size_t *offsets,
                             size_t product_count, const size_t *indices, size_t index_count,
                             int *status) {
  oracle_init();
  *status = ORACLE_INVALID_PARAMETERS;
  if (nv == 0 || max_degree == 0 || nv >= 40 || max_degree > 4096) return NULL;
  if (offsets == NULL || offsets[0] != 0) return NULL;
  for (size_t p = 0; p < product_count; ++p) {
    if (offsets[p + 1] < offsets[p]) return NULL;
    size_t arity = offsets[p + 1] - offsets[p];
    if (arity == 0 || arity > max_degree) return NULL;
  }
  if (offsets[product_count] != index_count) return NULL;
  for (size_t i = 0; i < index_count; ++i)
    if (indices[i] >= factor_count) return NULL;
  sc_state *s = (sc_state *)calloc(1, sizeof(sc_state));
  s->nv = nv;
  s->degree = max_degree;
  s->factor_count = factor_count;
  s->product_count = product_count;
  size_t n = (size_t)1 << nv;
  s->tables = (fr_t *)malloc(sizeof(fr_t) * (factor_count * n + 1));
  s->coeffs = (fr_t *)malloc(sizeof(fr_t) * (product_count + 1));
  s->offsets = (size_t *)malloc(sizeof(size_t) * (product_count + 1));
  s->indices = (size_t *)malloc(sizeof(size_t) * (index_count + 1));
  int ok = 1;
  for (size_t i = 0; i < factor_count * n; ++i)
    if (!fr_from_bytes(&s->tables[i], tables + 32 * i)) ok = 0;
  for (size_t p = 0; p < product_count; ++p)
    if (!fr_from_bytes(&s->coeffs[p], coeffs + 32 * p)) ok = 0;
  memcpy(s->offsets, offsets, sizeof(size_t) * (product_count + 1));
  if (index_count) memcpy(s->indices, indices, sizeof(size_t) * index_count);
  s->weights = (fr_t **)calloc(max_degree + 1, sizeof(fr_t *));
  for (size_t k = 1; k < max_degree; ++k) {
    s->weights[k] = (fr_t *)malloc(sizeof(fr_t) * (k + 1));
    for (size_t j = 0; j <= k; ++j) {
      fr_t w;
      fr_one(&w);
      for (size_t i = 0; i <= k; ++i) {
        if (i == j) continue;
        fr_t pj, pi, d;
        fr_from_u64(&pj, j);
        fr_from_u64(&pi, i);
        fr_sub(&d, &pj, &pi);
        fr_mul(&w, &w, &d);
      }
      s->weights[k][j] = w;
    }
    batch_invert(s->weights[k], k + 1);
  }
  if (!ok) {
    oracle_sumcheck_destroy(s);
    return NULL;
  }
  *status = ORACLE_OK;
  return s;
}

/* extrapolate(points 0..k, weights, evals, at) of cpu/sum_check.rs:202-217 */
static void extrapolate(fr_t *out, const fr_t *weights, const fr_t *evals, size_t k, size_t at) {
  fr_t *coeffs = (fr_t *)malloc(sizeof(fr_t) * (k + 1));
  fr_t a;
  fr_from_u64(&a, at);
  for (size_t i = 0; i <= k; ++i) {
    fr_t p;
    fr_from_u64(&p, i);
    fr_sub(&coeffs[i], &a, &p);
  }
  batch_invert(coeffs, k + 1);
  fr_t sum;
  fr_zero(&sum);
  for (size_t i = 0; i <= k; ++i) {
    fr_mul(&coeffs[i], &coeffs[i], &weights[i]);
    fr_add(&sum, &sum, &coeffs[i]);
  }
  fr_t sum_inv;
  if (fr_is_zero(&sum))
    fr_zero(&sum_inv);
  else
    fr_inv(&sum_inv, &sum);
  fr_t acc;
  fr_zero(&acc);
  for (size_t i = 0; i <= k; ++i) {
    fr_t t;
    fr_mul(&t, &coeffs[i], &evals[i]);
    fr_add(&acc, &acc, &t);
  }
  fr_mul(out, &acc, &sum_inv);
  free(coeffs);
}

int oracle_sumcheck_round(void *state, const uint8_t *challenge_or_null, uint8_t *out,
                          size_t count) {
  sc_state *s = (sc_state *)state;
  if (s == NULL) return ORACLE_INVALID_PARAMETERS;
  if (out == NULL) return ORACLE_INVALID_PARAMETERS;
  if (count != s->degree + 1) return ORACLE_INVALID_PARAMETERS;
  if (s->round >= s->nv) return ORACLE_INVALID_STATE;
  size_t n = (size_t)1 << s->nv;
  if (s->round == 0) {
    if (challenge_or_null != NULL) return ORACLE_INVALID_STATE;
  } else {
    if (challenge_or_null == NULL) return ORACLE_INVALID_STATE;
    fr_t c;
    if (!fr_from_bytes(&c, challenge_or_null)) return ORACLE_INVALID_PARAMETERS;
    size_t cur = n >> (s->round - 1);
    for (size_t f = 0; f < s->factor_count; ++f) {
      fr_t *t = s->tables + f * n;
      for (size_t b = 0; b < cur / 2; ++b) {
        fr_t lo = t[2 * b], d;
        fr_sub(&d, &t[2 * b + 1], &lo);
        fr_mul(&d, &d, &c);
        fr_add(&t[b], &lo, &d);
      }
    }
  }
  s->round += 1;
  size_t pairs = n >> s->round;
  fr_t *total = (fr_t *)calloc(s->degree + 1, sizeof(fr_t));
  fr_t *eval = (fr_t *)malloc(sizeof(fr_t) * (s->degree + 1));
  fr_t *step = (fr_t *)malloc(sizeof(fr_t) * (s->degree + 1));
  fr_t *sum = (fr_t *)malloc(sizeof(fr_t) * (s->degree + 2));
  for (size_t p = 0; p < s->product_count; ++p) {
    size_t begin = s->offsets[p], arity = s->offsets[p + 1] - begin;
    for (size_t t = 0; t <= arity; ++t) fr_zero(&sum[t]);
    for (size_t b = 0; b < pairs; ++b) {
      for (size_t k = 0; k < arity; ++k) {
        const fr_t *tab = s->tables + s->indices[begin + k] * n;
        eval[k] = tab[2 * b];
        fr_sub(&step[k], &tab[2 * b + 1], &tab[2 * b]);
      }
      fr_t prod = eval[0];
      for (size_t k = 1; k < arity; ++k) fr_mul(&prod, &prod, &eval[k]);
      fr_add(&sum[0], &sum[0], &prod);
      for (size_t t = 1; t <= arity; ++t) {
        for (size_t k = 0; k < arity; ++k) fr_add(&eval[k], &eval[k], &step[k]);
        prod = eval[0];
        for (size_t k = 1; k < arity; ++k) fr_mul(&prod, &prod, &eval[k]);
        fr_add(&sum[t], &sum[t], &prod);
      }
    }
    for (size_t t = 0; t <= arity; ++t) {
      fr_mul(&sum[t], &sum[t], &s->coeffs[p]);
      fr_add(&total[t], &total[t], &sum[t]);
    }
    for (size_t i = 0; i + arity < s->degree; ++i) {
      fr_t e;
      extrapolate(&e, s->weights[arity], sum, arity, arity + 1 + i);
      fr_add(&total[arity + 1 + i], &total[arity + 1 + i], &e);
    }
  }
  for (size_t t = 0; t <= s->degree; ++t) fr_to_bytes(out + 32 * t, &total[t]);
  free(total);
  free(eval);
  free(step);
  free(sum);
  return ORACLE_OK;
}

void oracle_sumcheck_destroy(void *state) {
  sc_state *s = (sc_state *)state;
  if (!s) return;
  if (s->weights) {
    for (size_t k = 0; k <= s->degree; ++k) free(s->weights[k]);
    free(s->weights);
  }
  free(s->tables);
  free(s->coeffs);
  free(s->offsets);
  free(s->indices);
  free(s);
}
