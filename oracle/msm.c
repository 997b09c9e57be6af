/* Oracle (test infrastructure only): variable-base MSM over G1.
 *
 * Restates E::G1::msm_unchecked (ark-ec 0.6.0, called at
 * backend/src/cpu/multilinear_kzg.rs:55,131): sum_i s_i * P_i, result as the
 * unique affine point.  Two independent routes:
 *   - oracle_msm_naive: per-point double-and-add, summed (small inputs);
 *   - oracle_msm_pippenger: unsigned c-bit windows, Jacobian buckets filled by
 *     mixed (Jacobian + affine) additions, running sums, threads over disjoint
 *     slices of the points (the CPU baseline).
 * Both are checked against Python-integer fixtures of the closed form
 * (sum a_i s_i) * G for bases a_i * G. */
#include <pthread.h>
#include <stdlib.h>
#include "oracle.h"
#include "oracle_core.h"

static void scalar_limbs(uint64_t k[4], const uint8_t *s) {
  for (int i = 0; i < 4; ++i) {
    uint64_t w = 0;
    for (int j = 7; j >= 0; --j) w = (w << 8) | s[8 * i + j];
    k[i] = w;
  }
}

static int scalar_canonical(const uint8_t *s) {
  fr_t x;
  return fr_from_bytes(&x, s);
}

int oracle_msm_naive(const uint8_t *points97, const uint8_t *scalars32, size_t n, uint8_t *out97) {
  oracle_init();
  g1_jac acc;
  g1_jac_set_inf(&acc);
  for (size_t i = 0; i < n; ++i) {
    g1_aff a;
    if (!g1_from_packet(&a, points97 + 97 * i)) return ORACLE_INVALID_PARAMETERS;
    if (!scalar_canonical(scalars32 + 32 * i)) return ORACLE_INVALID_PARAMETERS;
    uint64_t k[4];
    scalar_limbs(k, scalars32 + 32 * i);
    g1_jac p, t;
    g1_jac_from_aff(&p, &a);
    g1_jac_mul_limbs(&t, &p, k, 4);
    g1_jac_add(&acc, &acc, &t);
  }
  g1_aff r;
  g1_jac_to_aff(&r, &acc);
  g1_to_packet(out97, &r);
  return ORACLE_OK;
}

typedef struct {
  const g1_aff *points;
  const uint64_t *scalars; /* 4 limbs each */
  size_t begin, end;
  int c;
  g1_jac result;
} pip_job;

static unsigned window_digit(const uint64_t k[4], int bit, int c) {
  unsigned v = 0;
  for (int b = 0; b < c; ++b) {
    int pos = bit + b;
    if (pos >= 256) break;
    v |= (unsigned)((k[pos >> 6] >> (pos & 63)) & 1) << b;
  }
  return v;
}

static void *pip_worker(void *arg) {
  pip_job *job = (pip_job *)arg;
  const int c = job->c;
  const int windows = (255 + c - 1) / c;
  const size_t nb = (size_t)1 << c;
  g1_jac *buckets = (g1_jac *)malloc(sizeof(g1_jac) * nb);
  g1_jac total;
  g1_jac_set_inf(&total);
  for (int w = windows - 1; w >= 0; --w) {
    for (int d = 0; d < c; ++d) g1_jac_double(&total, &total);
    for (size_t b = 0; b < nb; ++b) g1_jac_set_inf(&buckets[b]);
    for (size_t i = job->begin; i < job->end; ++i) {
      unsigned d = window_digit(job->scalars + 4 * i, w * c, c);
      if (d) g1_jac_add_aff(&buckets[d], &buckets[d], &job->points[i]);
    }
    g1_jac running, sum;
    g1_jac_set_inf(&running);
    g1_jac_set_inf(&sum);
    for (size_t b = nb - 1; b >= 1; --b) {
      g1_jac_add(&running, &running, &buckets[b]);
      g1_jac_add(&sum, &sum, &running);
    }
    g1_jac_add(&total, &total, &sum);
  }
  free(buckets);
  job->result = total;
  return NULL;
}

int oracle_msm_pippenger(const uint8_t *points97, const uint8_t *scalars32, size_t n, int threads,
                         uint8_t *out97) {
  oracle_init();
  if (threads < 1) threads = 1;
  g1_aff *pts = (g1_aff *)malloc(sizeof(g1_aff) * (n ? n : 1));
  uint64_t *ks = (uint64_t *)malloc(sizeof(uint64_t) * 4 * (n ? n : 1));
  for (size_t i = 0; i < n; ++i) {
    if (!g1_from_packet(&pts[i], points97 + 97 * i) || !scalar_canonical(scalars32 + 32 * i)) {
      free(pts);
      free(ks);
      return ORACLE_INVALID_PARAMETERS;
    }
    scalar_limbs(ks + 4 * i, scalars32 + 32 * i);
  }
  if ((size_t)threads > n / 64 + 1) threads = (int)(n / 64 + 1);
  size_t slice = (n + threads - 1) / threads;
  int c = 1;
  while (((size_t)1 << (c + 1)) * 4 < slice && c < 16) ++c; /* ~ log2(slice) - 2 */
  pip_job *jobs = (pip_job *)calloc((size_t)threads, sizeof(pip_job));
  pthread_t *tids = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 0; t < threads; ++t) {
    jobs[t].points = pts;
    jobs[t].scalars = ks;
    jobs[t].begin = slice * t < n ? slice * t : n;
    jobs[t].end = slice * (t + 1) < n ? slice * (t + 1) : n;
    jobs[t].c = c;
    pthread_create(&tids[t], NULL, pip_worker, &jobs[t]);
  }
  g1_jac acc;
  g1_jac_set_inf(&acc);
  for (int t = 0; t < threads; ++t) {
    pthread_join(tids[t], NULL);
    g1_jac_add(&acc, &acc, &jobs[t].result);
  }
  g1_aff r;
  g1_jac_to_aff(&r, &acc);
  g1_to_packet(out97, &r);
  free(jobs);
  free(tids);
  free(pts);
  free(ks);
  return ORACLE_OK;
}
