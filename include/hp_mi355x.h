/* C ABI of the MI355X HyperPlonk prover backend (hp_mi355x_v1_*).
 *
 * It is the drop-in replacement for the reference's accelerator bridge
 * (backend/src/ascend910b3/native/bridge.h:29-62, bound from Rust in
 * backend/src/ascend910b3/ffi.rs:34-173): the same 20 entry points with the
 * same argument meaning, ownership, validation order and status kinds, so the
 * Rust trait glue of that backend (MleBackend, MultilinearKzgBackend,
 * SumCheckProver) binds to these symbols unchanged apart from the prefix and
 * the archive's name: both link a static archive (libhp_mi355x.a here, plus
 * -lamdhip64 and the C++ runtime; the shared libhp_mi355x.so is built from the
 * same objects for ctypes and the C++ host programs).
 *
 * Host exchange formats (not device layouts): Fr = canonical 32-byte
 * little-endian integer; G1 = 97 bytes: tag (0 finite, 1 identity with zero
 * coordinates) | x LE48 | y LE48.  Byte widths: evaluations 32, commitments
 * 97, point/r 32*num_vars, SRS layers 97*point_count, open quotients
 * 97*num_vars, rounds 32*evaluation_count.  Zero-count arrays may be null,
 * except product_offsets, which always holds product_count+1 entries
 * starting at 0.  Host buffers are borrowed only until return; host outputs
 * are fully written on success.  Resource outputs are unique and non-null on
 * success; on failure output slots are cleared.  Inputs are immutable.  A
 * context is used by one host thread at a time.  MLE handles are device base
 * addresses of Montgomery-form tables owned by the context; retire hands a
 * handle back, and its memory is reused only after every prover state that
 * still reads it is destroyed.  Zero denominators (including 0/0) are
 * InvalidParameters.  The first sumcheck round requires a null challenge,
 * later rounds a non-null one; ordering errors and exhaustion return
 * InvalidState without advancing the state.  A failed device execution
 * poisons the prover state it was advancing and the context; later calls on
 * them return InvalidState with the first error's code, while buffer_retire
 * and the params / state destroys still release their handle (the owner's
 * drop path), and context_destroy releases everything and returns that
 * first error.  Handles are checked against the context's live set before
 * anything is read through them.  No C++ exception crosses this interface.
 */
#ifndef HP_MI355X_H
#define HP_MI355X_H
#include <stddef.h>
#include <stdint.h>

#define HP_MI355X_STATUS_SUCCESS 0
#define HP_MI355X_STATUS_INVALID_PARAMETERS 1
#define HP_MI355X_STATUS_INVALID_STATE 2
#define HP_MI355X_STATUS_NOT_IMPLEMENTED 3

/* bridge.h:34 Ascend910B3Status: {kind, native_code}; only {0, 0} is success. */
typedef struct HpMi355xStatus {
  int32_t kind;
  int32_t native_code;
} HpMi355xStatus;

typedef struct HpMi355xContext HpMi355xContext;
typedef struct HpMi355xKzgParams HpMi355xKzgParams;
typedef struct HpMi355xSumCheckState HpMi355xSumCheckState;
typedef void *HpMi355xDeviceAddress;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime (bridge.h:43-45) ------------------------------------------ */
/* replaces hp_ascend910b3_v1_context_create (bridge.h:43) */
HpMi355xStatus hp_mi355x_v1_context_create(int32_t device_id, HpMi355xContext **out);
/* replaces hp_ascend910b3_v1_context_destroy (bridge.h:44) */
HpMi355xStatus hp_mi355x_v1_context_destroy(HpMi355xContext *ctx);
/* replaces hp_ascend910b3_v1_buffer_retire (bridge.h:45) */
HpMi355xStatus hp_mi355x_v1_buffer_retire(HpMi355xContext *ctx, HpMi355xDeviceAddress address);

/* ---- MLE (bridge.h:46-53; MleBackend, backend/src/multilinear_polynomial.rs:33-121) */
/* mle_from_evaluations: replaces hp_ascend910b3_v1_mle_upload (bridge.h:46) */
HpMi355xStatus hp_mi355x_v1_mle_upload(HpMi355xContext *ctx, size_t num_vars,
                                       const uint8_t *evaluations, size_t evaluation_count,
                                       HpMi355xDeviceAddress *out);
/* linear_combination: replaces bridge.h:47 */
HpMi355xStatus hp_mi355x_v1_mle_linear_combination(HpMi355xContext *ctx, size_t num_vars,
                                                   const HpMi355xDeviceAddress *polynomials,
                                                   const uint8_t *coefficients, size_t count,
                                                   HpMi355xDeviceAddress *out);
/* evaluate_mle: replaces bridge.h:48 */
HpMi355xStatus hp_mi355x_v1_mle_evaluate(HpMi355xContext *ctx, HpMi355xDeviceAddress polynomial,
                                         size_t num_vars, const uint8_t *point,
                                         uint8_t *out_evaluation);
/* build_eq_x_r: replaces bridge.h:49 */
HpMi355xStatus hp_mi355x_v1_mle_build_eq_x_r(HpMi355xContext *ctx, const uint8_t *r,
                                             size_t num_vars, HpMi355xDeviceAddress *out);
/* compute_frac_poly: replaces bridge.h:50 */
HpMi355xStatus hp_mi355x_v1_mle_compute_frac_poly(HpMi355xContext *ctx, size_t num_vars,
                                                  const HpMi355xDeviceAddress *numerators,
                                                  const HpMi355xDeviceAddress *denominators,
                                                  size_t count, HpMi355xDeviceAddress *out);
/* compute_product_poly: replaces bridge.h:51 */
HpMi355xStatus hp_mi355x_v1_mle_compute_product_poly(HpMi355xContext *ctx,
                                                     HpMi355xDeviceAddress frac, size_t num_vars,
                                                     HpMi355xDeviceAddress *out);
/* compute_nums_and_denoms: replaces bridge.h:52 */
HpMi355xStatus hp_mi355x_v1_mle_compute_nums_and_denoms(
    HpMi355xContext *ctx, size_t num_vars, const uint8_t *beta, const uint8_t *gamma,
    const HpMi355xDeviceAddress *fxs, const HpMi355xDeviceAddress *gxs,
    const HpMi355xDeviceAddress *permutations, size_t count,
    HpMi355xDeviceAddress *out_numerators, HpMi355xDeviceAddress *out_denominators);
/* compute_product_factors: replaces bridge.h:53 */
HpMi355xStatus hp_mi355x_v1_mle_compute_product_factors(HpMi355xContext *ctx,
                                                        HpMi355xDeviceAddress frac,
                                                        HpMi355xDeviceAddress product,
                                                        size_t num_vars,
                                                        HpMi355xDeviceAddress *out_p1,
                                                        HpMi355xDeviceAddress *out_p2);

/* ---- multilinear KZG (bridge.h:54-59; backend/src/multilinear_kzg.rs:20-55) */
/* prepare_prover_param, step 1: replaces bridge.h:54 */
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_create(HpMi355xContext *ctx, size_t num_vars,
                                                          HpMi355xKzgParams **out);
/* prepare_prover_param, step 2 (per layer): replaces bridge.h:55 */
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_upload_layer(HpMi355xContext *ctx,
                                                                HpMi355xKzgParams *params,
                                                                size_t layer_index,
                                                                const uint8_t *points,
                                                                size_t point_count);
/* replaces bridge.h:56 */
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_destroy(HpMi355xContext *ctx,
                                                           HpMi355xKzgParams *params);
/* commit: replaces bridge.h:57 */
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_commit(HpMi355xContext *ctx,
                                                   const HpMi355xKzgParams *params,
                                                   HpMi355xDeviceAddress polynomial,
                                                   size_t num_vars, uint8_t *out_commitment);
/* multi_commit: replaces bridge.h:58 */
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_multi_commit(HpMi355xContext *ctx,
                                                         const HpMi355xKzgParams *params,
                                                         const HpMi355xDeviceAddress *polynomials,
                                                         const size_t *num_vars, size_t count,
                                                         uint8_t *out_commitments);
/* open: replaces bridge.h:59 */
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_open(HpMi355xContext *ctx,
                                                 const HpMi355xKzgParams *params,
                                                 HpMi355xDeviceAddress polynomial, size_t num_vars,
                                                 const uint8_t *point, uint8_t *out_quotients,
                                                 uint8_t *out_evaluation);

/* ---- SumCheck (bridge.h:60-62; backend/src/sum_check.rs:17-38) --------- */
/* prover_init: replaces bridge.h:60 */
HpMi355xStatus hp_mi355x_v1_sumcheck_prover_state_create(
    HpMi355xContext *ctx, size_t num_vars, size_t max_degree,
    const HpMi355xDeviceAddress *factors, size_t factor_count, const uint8_t *coefficients,
    const size_t *product_offsets, size_t product_count, const size_t *factor_indices,
    size_t index_count, HpMi355xSumCheckState **out);
/* prove_round_and_update_state: replaces bridge.h:61 */
HpMi355xStatus hp_mi355x_v1_sumcheck_prover_state_round(HpMi355xContext *ctx,
                                                        HpMi355xSumCheckState *state,
                                                        const uint8_t *challenge_or_null,
                                                        uint8_t *out_evaluations,
                                                        size_t evaluation_count);
/* replaces bridge.h:62 */
HpMi355xStatus hp_mi355x_v1_sumcheck_prover_state_destroy(HpMi355xContext *ctx,
                                                          HpMi355xSumCheckState *state);

/* ---- additive extensions (no counterpart in the reference bridge) ------ */
/* Copy a resident table back as canonical LE32 values (tests, diagnostics). */
HpMi355xStatus hp_mi355x_v1_mle_download(HpMi355xContext *ctx, HpMi355xDeviceAddress polynomial,
                                         size_t num_vars, uint8_t *out_evaluations,
                                         size_t evaluation_count);
/* Fill every layer of params on the device from a trapdoor t (num_vars
 * canonical Fr): layer i = { g * eq(t[i..], x) }, last layer = [g]; the
 * construction of subroutines/src/pcs/multilinear_kzg/srs.rs:158-220 with
 * g the standard generator. */
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_generate(HpMi355xContext *ctx,
                                                            HpMi355xKzgParams *params,
                                                            const uint8_t *t);
/* Copy one prepared layer back as 97-byte packets (tests, SRS files). */
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_download_layer(
    HpMi355xContext *ctx, const HpMi355xKzgParams *params, size_t layer_index, uint8_t *points,
    size_t point_count);
/* Run one commit-sized MSM (layer N - num_vars) with HIP events between its
 * stages on the context stream.  stage_ms (>= 7 floats): digits, sort, tasks,
 * accumulate, bucket sums, bucket reduction, whole device part; plan_info
 * (4 u32): window bits c, windows W, buckets B, non-zero digits. */
HpMi355xStatus hp_mi355x_v1_msm_profile(HpMi355xContext *ctx, const HpMi355xKzgParams *params,
                                        HpMi355xDeviceAddress polynomial, size_t num_vars,
                                        float *stage_ms, size_t stage_count,
                                        uint32_t *plan_info);
/* One round of each of `count` distinct sumcheck states in shared launches:
 * per state i the arguments and checks of sumcheck_prover_state_round
 * (challenges_or_null may be null when every state is at round zero).  All
 * states are checked before any work; on a device failure every state of the
 * batch is poisoned.  Used to batch the rounds of independent proofs. */
HpMi355xStatus hp_mi355x_v1_sumcheck_prover_state_round_batch(
    HpMi355xContext *ctx, HpMi355xSumCheckState *const *states,
    const uint8_t *const *challenges_or_null, uint8_t *const *out_evaluations,
    const size_t *evaluation_counts, size_t count);
/* Evaluate `count` resident tables, table i (num_vars[i] variables) at
 * points[i] (num_vars[i] canonical Fr, low variable first), into
 * out_evaluations (32 * count bytes): evaluate_mle for many tables in one
 * pass; all arguments are checked before any work. */
HpMi355xStatus hp_mi355x_v1_mle_evaluate_batch(HpMi355xContext *ctx,
                                               const HpMi355xDeviceAddress *polynomials,
                                               const size_t *num_vars,
                                               const uint8_t *const *points,
                                               uint8_t *out_evaluations, size_t count);
/* Time `reps` folds of a resident table by its lowest variable (value r,
 * one canonical Fr) with HIP events on the context stream; *ms_per_fold gets
 * the mean (benchmark of the fold's HBM rate). */
HpMi355xStatus hp_mi355x_v1_mle_fold_profile(HpMi355xContext *ctx,
                                             HpMi355xDeviceAddress polynomial, size_t num_vars,
                                             const uint8_t *r, size_t reps, float *ms_per_fold);
/* Autotune the MSM for layers of 2^log_sizes[i] points (1..24) on this device:
 * candidate window bits around the built-in rule and lane task caps, timed
 * with HIP events over random points and scalars; the fastest go into the
 * context's tuning table (used by layers prepared afterwards) and into
 * out_window / out_cap / out_ms (count entries each). */
HpMi355xStatus hp_mi355x_v1_msm_autotune(HpMi355xContext *ctx, const size_t *log_sizes,
                                         size_t count, uint32_t *out_window, uint32_t *out_cap,
                                         float *out_ms);
/* Install a saved tuning table: window bits (4..20, or 0 for the built-in rule)
 * and lane task cap (0 for the rule) per log2 layer size. */
HpMi355xStatus hp_mi355x_v1_msm_set_tuning(HpMi355xContext *ctx, const size_t *log_sizes,
                                           const uint32_t *window, const uint32_t *cap,
                                           size_t count);
/* One request of multilinear_kzg_batch: a commit (out_points: one packet)
 * or an open at `point` (out_points: num_vars packets, out_evaluation: 32
 * bytes), with the arguments and checks of multilinear_kzg_commit / _open. */
#define HP_MI355X_KZG_COMMIT 0
#define HP_MI355X_KZG_OPEN 1
typedef struct HpMi355xKzgRequest {
  int32_t kind;
  HpMi355xDeviceAddress polynomial;
  size_t num_vars;
  const uint8_t *point;
  uint8_t *out_points;
  uint8_t *out_evaluation;
} HpMi355xKzgRequest;
/* Commits and opens of many polynomials (of any sizes) whose MSMs all run as
 * one batched launch sequence; results identical to the one-at-a-time calls.
 * Every request is checked before any work.  Used to batch the commitments
 * and openings of independent proofs. */
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_batch(HpMi355xContext *ctx,
                                                  const HpMi355xKzgParams *params,
                                                  const HpMi355xKzgRequest *requests,
                                                  size_t count);
/* Asynchronous multilinear_kzg_batch: the same checks (all before any work),
 * then the MSMs are queued on a second stream of the context, behind the work
 * already queued on the context, and the call returns at once with a ticket;
 * other calls on the context (sumcheck rounds of other proofs) then overlap
 * the MSMs on the device.  The requests' output buffers must stay valid until
 * batch_wait, which waits for the work, writes every output (identical to
 * multilinear_kzg_batch) and releases the ticket; call it once per ticket.
 * Input tables and parameters stay alive until then even if released.  Used
 * by the batched prover to overlap commitments with sumchecks. */
typedef struct HpMi355xKzgTicket HpMi355xKzgTicket;
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_batch_submit(HpMi355xContext *ctx,
                                                         const HpMi355xKzgParams *params,
                                                         const HpMi355xKzgRequest *requests,
                                                         size_t count, HpMi355xKzgTicket **out);
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_batch_wait(HpMi355xContext *ctx,
                                                       HpMi355xKzgTicket *ticket);
/* Canonical SRS files (MultilinearUniversalParams serialize_compressed, the
 * format hyperplonk/examples/full_prove_bench.rs:665-760 loads and writes):
 * fill one layer from point_count compressed G1 points (48 bytes each:
 * big-endian x, flag bits 0x80 compressed / 0x40 infinity / 0x20 y > -y on
 * the first byte), decompressed on the device; INVALID_PARAMETERS for a
 * missing compression flag, x >= q or x off the curve (no subgroup check, as
 * deserialize_compressed_unchecked in hyperplonk/benches/bench.rs:53-56). */
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_upload_layer_compressed(
    HpMi355xContext *ctx, HpMi355xKzgParams *params, size_t layer_index,
    const uint8_t *compressed, size_t point_count);
/* Copy one prepared layer back in the compressed encoding (48 bytes a point). */
HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_download_layer_compressed(
    HpMi355xContext *ctx, const HpMi355xKzgParams *params, size_t layer_index,
    uint8_t *compressed, size_t point_count);
/* evaluate_mle that allocates nothing (no counterpart in the reference bridge;
 * for callers that cannot allocate device or pinned memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * Same arguments, checks, statuses and value as hp_mi355x_v1_mle_evaluate;
 * the point, the fold tables and the value live in one static device array
 * per GPU, sized for num_vars <= HP_MI355X_NOALLOC_MAX_NUM_VARS (24 MiB);
 * larger tables return INVALID_PARAMETERS with native_code
 * HP_MI355X_NATIVE_EXCEEDS_SCRATCH.  Every call on a GPU shares that array:
 * calls that overlap on one GPU (contexts of the same device driven from
 * several threads) race. */
#define HP_MI355X_NOALLOC_MAX_NUM_VARS 20
#define HP_MI355X_NATIVE_EXCEEDS_SCRATCH 1001
HpMi355xStatus hp_mi355x_v1_mle_evaluate_noalloc(HpMi355xContext *ctx,
                                                 HpMi355xDeviceAddress polynomial,
                                                 size_t num_vars, const uint8_t *point,
                                                 uint8_t *out_evaluation);
/* Device memory as a prover service admits work by it (no counterpart in the
 * reference): *device_free = free device memory plus the context's cached
 * blocks, *device_total = the device's memory, *context_live = bytes of the
 * context's live buffers (tables, prover states, parameters, work memory of
 * calls in flight), *context_peak = their high-water mark since the last
 * reset (reset_peak != 0 resets it to the current live bytes after
 * reading). */
HpMi355xStatus hp_mi355x_v1_memory_info(HpMi355xContext *ctx, size_t *device_free,
                                        size_t *device_total, size_t *context_live,
                                        size_t *context_peak, int32_t reset_peak);
/* Build provenance (no counterpart in the reference): "sources-sha256=<hex>
 * arch=gfx950" into out (capacity >= 96 bytes), the digest of the sources the
 * library was compiled from (hyperplonk_amd/csrc/build_info.hip). */
HpMi355xStatus hp_mi355x_v1_build_info(char *out, size_t capacity);
/* Roofline denominators measured on this device (no counterpart in the
 * reference): *out_hbm_bytes_per_s from a streaming copy of copy_bytes (read
 * plus write bytes over the kernel time), *out_mad_per_s from a u32 x u32 ->
 * u64 multiply-add chain kernel (v_mad_u64_u32, the field kernels' limb
 * product) on every CU; best of `reps` timed runs each (HIP events on the
 * context stream).  copy_bytes: 1 MiB .. 16 GiB, a multiple of 16. */
HpMi355xStatus hp_mi355x_v1_device_peaks(HpMi355xContext *ctx, size_t copy_bytes, uint32_t reps,
                                         double *out_hbm_bytes_per_s, double *out_mad_per_s);
/* Failure injection for tests of the poisoning contract (no counterpart in the
 * reference bridge, whose tests force device errors through the runtime):
 * the `after`-th HIP call this context checks from now on (1 = the next one)
 * completes normally on the device but is reported as failed with
 * `native_code` (non-zero), exactly as a real device failure would be: the
 * call returns INVALID_STATE with that code, a prover state it was advancing
 * is poisoned, and so is the context.  after = 0 disarms. */
HpMi355xStatus hp_mi355x_v1_debug_inject_failure(HpMi355xContext *ctx, uint32_t after,
                                                 int32_t native_code);
#ifdef __cplusplus
}
#endif
#endif
