/* A C program linked against the static archive libhp_mi355x.a (as the
 * reference's Rust glue links its bridge, backend/src/ascend910b3/ffi.rs:34)
 * that runs on the GPU: the known answers of evaluate_mle
 * (backend/src/cpu/multilinear_polynomial.rs:281-284: [2,3,5,7] at (2,3) is
 * 19) and of two sumcheck rounds (backend/src/cpu/sum_check.rs:225-253:
 * 2 f + 3 f^2 over f = [2,3,5,7] gives [101,194,317], then [56,261,616] at
 * challenge 2). */
#include <stdio.h>
#include <string.h>

#include "../../include/hp_mi355x.h"

static void fr(uint8_t out[32], uint64_t v) {
  memset(out, 0, 32);
  for (int i = 0; i < 8; ++i) out[i] = (uint8_t)(v >> (8 * i));
}

static int is_small(const uint8_t in[32], uint64_t v) {
  uint8_t want[32];
  fr(want, v);
  return memcmp(in, want, 32) == 0;
}

int main(void) {
  HpMi355xContext *ctx = NULL;
  if (hp_mi355x_v1_context_create(0, &ctx).kind != 0) return 1;
  uint8_t table[4 * 32];
  const uint64_t vals[4] = {2, 3, 5, 7};
  for (int i = 0; i < 4; ++i) fr(table + 32 * i, vals[i]);
  HpMi355xDeviceAddress f = NULL;
  if (hp_mi355x_v1_mle_upload(ctx, 2, table, 4, &f).kind != 0) return 2;
  uint8_t point[64], value[32];
  fr(point, 2);
  fr(point + 32, 3);
  if (hp_mi355x_v1_mle_evaluate(ctx, f, 2, point, value).kind != 0 || !is_small(value, 19)) return 3;
  HpMi355xDeviceAddress factors[1] = {f};
  uint8_t coeffs[64];
  fr(coeffs, 2);
  fr(coeffs + 32, 3);
  const size_t offsets[3] = {0, 1, 3};
  const size_t indices[3] = {0, 0, 0};
  HpMi355xSumCheckState *st = NULL;
  if (hp_mi355x_v1_sumcheck_prover_state_create(ctx, 2, 2, factors, 1, coeffs, offsets, 2, indices, 3,
                                                &st).kind != 0)
    return 4;
  uint8_t msg[96], ch[32];
  if (hp_mi355x_v1_sumcheck_prover_state_round(ctx, st, NULL, msg, 3).kind != 0) return 5;
  if (!is_small(msg, 101) || !is_small(msg + 32, 194) || !is_small(msg + 64, 317)) return 6;
  fr(ch, 2);
  if (hp_mi355x_v1_sumcheck_prover_state_round(ctx, st, ch, msg, 3).kind != 0) return 7;
  if (!is_small(msg, 56) || !is_small(msg + 32, 261) || !is_small(msg + 64, 616)) return 8;
  if (hp_mi355x_v1_sumcheck_prover_state_destroy(ctx, st).kind != 0) return 9;
  if (hp_mi355x_v1_buffer_retire(ctx, f).kind != 0) return 10;
  if (hp_mi355x_v1_context_destroy(ctx).kind != 0) return 11;
  printf("STATIC_GPU PASS\n");
  return 0;
}
