/* Oracle (test infrastructure only): exported C entry points of the CPU
 * restatement.  Every value crosses this interface in the reference's canonical
 * host encodings (backend/src/ascend910b3/encoding.rs:12-108): Fr = 32-byte
 * little-endian canonical integer, Fq = 48-byte LE, G1 = 97-byte packet
 * (tag 0 finite / 1 identity, x LE48, y LE48).  Status: 0 success,
 * 1 invalid parameters, 2 invalid state (backend/src/errors.rs:16-21). */
#ifndef HP_ORACLE_H
#define HP_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#define ORACLE_OK 0
#define ORACLE_INVALID_PARAMETERS 1
#define ORACLE_INVALID_STATE 2

void oracle_init(void);

/* op: 0 add, 1 sub, 2 mul, 3 neg, 4 square, 5 inverse (zero input -> 1). */
int oracle_fr_op(int op, const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n);
int oracle_fq_op(int op, const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n);

int oracle_g1_generator(uint8_t *out97);
int oracle_g1_add(const uint8_t *p97, const uint8_t *q97, uint8_t *out97);
int oracle_g1_mul(const uint8_t *p97, const uint8_t *k32, uint8_t *out97);
int oracle_g1_on_curve(const uint8_t *p97);

int oracle_msm_naive(const uint8_t *points97, const uint8_t *scalars32, size_t n, uint8_t *out97);
int oracle_msm_pippenger(const uint8_t *points97, const uint8_t *scalars32, size_t n, int threads,
                         uint8_t *out97);

int oracle_mle_evaluate(const uint8_t *table, size_t nv, const uint8_t *point, uint8_t *out);
int oracle_mle_fold(const uint8_t *table, size_t nv, const uint8_t *r, uint8_t *out);
int oracle_mle_linear_combination(const uint8_t *tables, const uint8_t *coeffs, size_t count,
                                  size_t nv, uint8_t *out);
int oracle_eq_x_r(const uint8_t *r, size_t nv, uint8_t *out);
int oracle_eq_eval(const uint8_t *x, const uint8_t *y, size_t n, uint8_t *out);
int oracle_frac(const uint8_t *nums, const uint8_t *dens, size_t count, size_t nv, uint8_t *out);
int oracle_product(const uint8_t *frac, size_t nv, uint8_t *out);
int oracle_nums_denoms(const uint8_t *beta, const uint8_t *gamma, const uint8_t *fxs,
                       const uint8_t *gxs, const uint8_t *perms, size_t count, size_t nv,
                       uint8_t *out_nums, uint8_t *out_dens);
int oracle_product_factors(const uint8_t *frac, const uint8_t *prod, size_t nv, uint8_t *p1,
                           uint8_t *p2);

void *oracle_sumcheck_create(size_t nv, size_t max_degree, const uint8_t *tables,
                             size_t factor_count, const uint8_t *coeffs, const size_t *offsets,
                             size_t product_count, const size_t *indices, size_t index_count,
                             int *status);
int oracle_sumcheck_round(void *state, const uint8_t *challenge_or_null, uint8_t *out,
                          size_t count);
void oracle_sumcheck_destroy(void *state);

int oracle_kzg_open(const uint8_t *layers97, const uint8_t *table, size_t nv,
                    const uint8_t *point, int threads, uint8_t *out_quotients97,
                    uint8_t *out_eval);
int oracle_srs_layer(const uint8_t *t, size_t len, int threads, uint8_t *out97);
#endif
