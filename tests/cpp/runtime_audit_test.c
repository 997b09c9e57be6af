/* Runtime audit of the C ABI: every HIP runtime call the library makes is
 * counted per entry-point call, the way the reference audits its bridge with
 * linker --wrap shims around the ACL runtime
 * (backend/src/ascend910b3/native/tests/runtime_audit.h:10-38, wrap list at
 * native/CMakeLists.txt:106-117; budgets as in tests/bridge_smoke.cpp:501-512).
 *
 * Linked against the static archive with -Wl,--wrap=<fn> for the allocation,
 * copy, memset, synchronisation and launch entry points of the HIP runtime,
 * so calls made by the archive's objects land in the counters below.  After a
 * warm-up (the context's allocator and pinned staging reach their steady
 * sizes) the online calls must not allocate device or pinned memory, must stay
 * within byte budgets for host <-> device copies, and synchronise the host at
 * most once; buffer_retire must neither copy, synchronise nor launch.  Prints
 * one line of counts per audited call and RUNTIME_AUDIT PASS. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/hp_mi355x.h"

/* HIP runtime types at the C ABI (enums are ints; dim3 is three u32) */
typedef int hipError_t;
typedef struct ihipStream_t *hipStream_t;
typedef struct ihipEvent_t *hipEvent_t;
typedef struct {
  uint32_t x, y, z;
} dim3;
enum { kH2H = 0, kH2D = 1, kD2H = 2, kD2D = 3 };

struct Counts {
  uint64_t malloc_calls, free_calls, host_malloc_calls, host_free_calls;
  uint64_t h2d_calls, h2d_bytes, d2h_calls, d2h_bytes, d2d_calls, other_copy_calls;
  uint64_t memset_calls, stream_syncs, event_syncs, device_syncs, launches;
};
static struct Counts g;

hipError_t __real_hipMalloc(void **p, size_t n);
hipError_t __wrap_hipMalloc(void **p, size_t n) {
  ++g.malloc_calls;
  return __real_hipMalloc(p, n);
}
hipError_t __real_hipFree(void *p);
hipError_t __wrap_hipFree(void *p) {
  ++g.free_calls;
  return __real_hipFree(p);
}
hipError_t __real_hipHostMalloc(void **p, size_t n, unsigned flags);
hipError_t __wrap_hipHostMalloc(void **p, size_t n, unsigned flags) {
  ++g.host_malloc_calls;
  return __real_hipHostMalloc(p, n, flags);
}
hipError_t __real_hipHostFree(void *p);
hipError_t __wrap_hipHostFree(void *p) {
  ++g.host_free_calls;
  return __real_hipHostFree(p);
}
static void count_copy(size_t n, int kind) {
  if (kind == kH2D) {
    ++g.h2d_calls;
    g.h2d_bytes += n;
  } else if (kind == kD2H) {
    ++g.d2h_calls;
    g.d2h_bytes += n;
  } else if (kind == kD2D) {
    ++g.d2d_calls;
  } else {
    ++g.other_copy_calls;
  }
}
hipError_t __real_hipMemcpyAsync(void *dst, const void *src, size_t n, int kind, hipStream_t s);
hipError_t __wrap_hipMemcpyAsync(void *dst, const void *src, size_t n, int kind, hipStream_t s) {
  count_copy(n, kind);
  return __real_hipMemcpyAsync(dst, src, n, kind, s);
}
hipError_t __real_hipMemsetAsync(void *dst, int v, size_t n, hipStream_t s);
hipError_t __wrap_hipMemsetAsync(void *dst, int v, size_t n, hipStream_t s) {
  ++g.memset_calls;
  return __real_hipMemsetAsync(dst, v, n, s);
}
hipError_t __real_hipStreamSynchronize(hipStream_t s);
hipError_t __wrap_hipStreamSynchronize(hipStream_t s) {
  ++g.stream_syncs;
  return __real_hipStreamSynchronize(s);
}
hipError_t __real_hipEventSynchronize(hipEvent_t e);
hipError_t __wrap_hipEventSynchronize(hipEvent_t e) {
  ++g.event_syncs;
  return __real_hipEventSynchronize(e);
}
hipError_t __real_hipLaunchKernel(const void *f, dim3 grid, dim3 block, void **args, size_t shmem,
                                  hipStream_t s);
hipError_t __wrap_hipLaunchKernel(const void *f, dim3 grid, dim3 block, void **args, size_t shmem,
                                  hipStream_t s) {
  ++g.launches;
  return __real_hipLaunchKernel(f, grid, block, args, shmem, s);
}

static int failures = 0;

static void begin(void) { memset(&g, 0, sizeof g); }

static struct Counts end(const char *what) {
  struct Counts c = g;
  printf("%-28s malloc %llu free %llu host_malloc %llu host_free %llu h2d %llu/%lluB d2h %llu/%lluB "
         "d2d %llu other %llu memset %llu syncs %llu+%llu launches %llu\n",
         what, (unsigned long long)c.malloc_calls, (unsigned long long)c.free_calls,
         (unsigned long long)c.host_malloc_calls, (unsigned long long)c.host_free_calls,
         (unsigned long long)c.h2d_calls, (unsigned long long)c.h2d_bytes,
         (unsigned long long)c.d2h_calls, (unsigned long long)c.d2h_bytes,
         (unsigned long long)c.d2d_calls, (unsigned long long)c.other_copy_calls,
         (unsigned long long)c.memset_calls, (unsigned long long)c.stream_syncs,
         (unsigned long long)c.event_syncs, (unsigned long long)c.launches);
  return c;
}

static void require(int ok, const char *what) {
  if (!ok) {
    printf("AUDIT FAIL: %s\n", what);
    ++failures;
  }
}

/* online call: no allocation of any kind, copies within budget, at most one
 * host synchronisation */
static void expect_online(const struct Counts *c, uint64_t max_h2d, uint64_t max_d2h,
                          uint64_t max_launches, const char *what) {
  char msg[160];
  snprintf(msg, sizeof msg, "%s allocated", what);
  require(c->malloc_calls == 0 && c->host_malloc_calls == 0 && c->free_calls == 0 &&
              c->host_free_calls == 0,
          msg);
  snprintf(msg, sizeof msg, "%s H2D budget", what);
  require(c->h2d_bytes <= max_h2d, msg);
  snprintf(msg, sizeof msg, "%s D2H budget", what);
  require(c->d2h_bytes <= max_d2h, msg);
  snprintf(msg, sizeof msg, "%s device-to-device or untyped copy", what);
  require(c->d2d_calls == 0 && c->other_copy_calls == 0, msg);
  snprintf(msg, sizeof msg, "%s synchronised more than once", what);
  require(c->stream_syncs + c->event_syncs + c->device_syncs <= 1, msg);
  snprintf(msg, sizeof msg, "%s launch budget", what);
  require(c->launches <= max_launches, msg);
}

static uint64_t state = 0x0123456789abcdefull;
static uint64_t next64(void) {
  uint64_t z = (state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
/* canonical Fr below 2^253 */
static void rand_fr(uint8_t out[32]) {
  for (int i = 0; i < 4; ++i) {
    uint64_t w = next64();
    if (i == 3) w &= 0x0fffffffffffffffull;
    memcpy(out + 8 * i, &w, 8);
  }
}

#define CHECK(call, code)                                  \
  do {                                                     \
    if ((call).kind != 0) {                                \
      printf("call failed at %s:%d\n", __FILE__, __LINE__); \
      return code;                                         \
    }                                                      \
  } while (0)

enum { NV = 12, N = 1 << NV, DEG = 3 };

static int sumcheck_once(HpMi355xContext *ctx, HpMi355xDeviceAddress *f, int audit) {
  /* c0 f0 f1 f2 + c1 f1 + c2 f2 f0 */
  uint8_t coeffs[3 * 32];
  for (int i = 0; i < 3; ++i) rand_fr(coeffs + 32 * i);
  const size_t offsets[4] = {0, 3, 4, 6};
  const size_t indices[6] = {0, 1, 2, 1, 2, 0};
  HpMi355xSumCheckState *st = NULL;
  if (audit) begin();
  CHECK(hp_mi355x_v1_sumcheck_prover_state_create(ctx, NV, DEG, f, 3, coeffs, offsets, 3, indices, 6,
                                                  &st),
        20);
  if (audit) {
    struct Counts c = end("sumcheck create");
    /* metadata upload only; the state's tables come from the allocator cache */
    expect_online(&c, 4096, 0, 4, "sumcheck create");
  }
  uint8_t msg[(DEG + 1) * 32], ch[32];
  for (int r = 0; r < NV; ++r) {
    rand_fr(ch);
    if (audit) begin();
    CHECK(hp_mi355x_v1_sumcheck_prover_state_round(ctx, st, r ? ch : NULL, msg, DEG + 1), 21);
    if (audit) {
      char what[40];
      snprintf(what, sizeof what, "sumcheck round %d", r);
      struct Counts c = end(what);
      /* launched rounds: challenge by value in the kernel arguments, message
       * written by the kernel into pinned host memory, one synchronisation;
       * rounds in the persistent loops: challenge and message through the
       * state's mailbox in pinned host memory, no synchronisation, the loops
       * launched once (a wide loop, one of fewer blocks, the one-block loop,
       * queued together).  Neither copies nor allocates. */
      expect_online(&c, 0, 0, 3, what);
      require(c.stream_syncs + c.event_syncs <= 1, "a round waits for its message at most once");
    }
  }
  if (audit) begin();
  CHECK(hp_mi355x_v1_sumcheck_prover_state_destroy(ctx, st), 22);
  if (audit) {
    struct Counts c = end("sumcheck destroy");
    require(c.h2d_calls + c.d2h_calls + c.d2d_calls + c.other_copy_calls + c.launches == 0,
            "state destroy copied or launched");
  }
  return 0;
}

int main(void) {
  HpMi355xContext *ctx = NULL;
  CHECK(hp_mi355x_v1_context_create(0, &ctx), 1);
  uint8_t *table = (uint8_t *)malloc((size_t)N * 32);
  HpMi355xDeviceAddress f[3];
  for (int k = 0; k < 3; ++k) {
    for (int i = 0; i < N; ++i) rand_fr(table + 32 * (size_t)i);
    CHECK(hp_mi355x_v1_mle_upload(ctx, NV, table, N, &f[k]), 2);
  }
  HpMi355xKzgParams *params = NULL;
  CHECK(hp_mi355x_v1_multilinear_kzg_params_create(ctx, NV, &params), 3);
  uint8_t trapdoor[NV * 32], point[NV * 32], value[32], commitment[97];
  for (int i = 0; i < NV; ++i) rand_fr(trapdoor + 32 * i);
  CHECK(hp_mi355x_v1_multilinear_kzg_params_generate(ctx, params, trapdoor), 4);
  for (int i = 0; i < NV; ++i) rand_fr(point + 32 * i);
  uint8_t *quotients = (uint8_t *)malloc((size_t)NV * 97);

  /* warm-up: the allocator's blocks and the pinned staging reach their sizes */
  for (int w = 0; w < 2; ++w) {
    CHECK(hp_mi355x_v1_mle_evaluate(ctx, f[0], NV, point, value), 5);
    CHECK(hp_mi355x_v1_multilinear_kzg_commit(ctx, params, f[0], NV, commitment), 6);
    CHECK(hp_mi355x_v1_multilinear_kzg_open(ctx, params, f[0], NV, point, quotients, value), 7);
    if (sumcheck_once(ctx, f, 0)) return 8;
  }

  /* buffer_retire: bookkeeping only */
  HpMi355xDeviceAddress tmp = NULL;
  CHECK(hp_mi355x_v1_mle_upload(ctx, 2, table, 4, &tmp), 9);
  begin();
  CHECK(hp_mi355x_v1_buffer_retire(ctx, tmp), 10);
  {
    struct Counts c = end("buffer_retire");
    require(c.h2d_calls + c.d2h_calls + c.d2d_calls + c.other_copy_calls + c.memset_calls == 0,
            "buffer_retire copied");
    require(c.stream_syncs + c.event_syncs + c.launches + c.free_calls == 0,
            "buffer_retire synchronised, launched or freed");
  }

  begin();
  CHECK(hp_mi355x_v1_mle_evaluate(ctx, f[0], NV, point, value), 11);
  {
    struct Counts c = end("mle_evaluate nv12");
    expect_online(&c, NV * 32, 32, 8, "mle_evaluate");
  }

  begin();
  CHECK(hp_mi355x_v1_multilinear_kzg_commit(ctx, params, f[0], NV, commitment), 12);
  {
    struct Counts c = end("kzg_commit nv12");
    /* the MSM job table up, the level-1 bucket sums down */
    expect_online(&c, 64 * 1024, 512 * 1024, 24, "kzg_commit");
  }

  begin();
  CHECK(hp_mi355x_v1_multilinear_kzg_open(ctx, params, f[0], NV, point, quotients, value), 13);
  {
    struct Counts c = end("kzg_open nv12");
    expect_online(&c, 64 * 1024 + NV * 32, 2 * 1024 * 1024, 24 * NV, "kzg_open");
  }

  if (sumcheck_once(ctx, f, 1)) return 14;

  for (int k = 0; k < 3; ++k) CHECK(hp_mi355x_v1_buffer_retire(ctx, f[k]), 15);
  CHECK(hp_mi355x_v1_multilinear_kzg_params_destroy(ctx, params), 16);
  CHECK(hp_mi355x_v1_context_destroy(ctx), 17);
  free(table);
  free(quotients);
  if (failures) {
    printf("RUNTIME_AUDIT FAIL (%d)\n", failures);
    return 30;
  }
  printf("RUNTIME_AUDIT PASS\n");
  return 0;
}
