// Roofline denominators measured on the device the library runs on:
//   * HBM: a streaming copy (16-byte loads and stores, grid-strided, enough
//     workgroups to fill every CU several times), bytes read + bytes written
//     over the kernel time;
//   * VALU: u32 x u32 -> u64 multiply-adds in the idiom of the field kernels
//     (`acc += (uint64_t)a * b`, one v_mad_u64_u32 each), 16 independent
//     chains per lane so the issue rate, not the chain latency, binds.
// Both are timed with HIP events on the context stream, best of `reps`.
#include <algorithm>

#include "kernels_common.hpp"

namespace hp {
namespace peaks {
namespace {

// Four 16-byte loads in flight per lane before their stores; lanes of a
// wavefront touch consecutive 16-byte words (n a multiple of 4 x the grid's
// thread count is not required: the tail runs one word per lane).
__global__ void __launch_bounds__(256) k_copy(const uint4* src, uint4* dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    dst[i] = a;
    dst[i + stride] = b;
    dst[i + 2 * stride] = c;
    dst[i + 3 * stride] = d;
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

constexpr int kChains = 16;

// Each chain: acc = lo32(acc) * y + acc, a v_mad_u64_u32 whose operands are
// the previous result, so nothing is hoisted out of the loop.
__global__ void __launch_bounds__(256) k_mad_u64(uint32_t iters, uint64_t* sink) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t acc[kChains];
#pragma unroll
  for (int j = 0; j < kChains; ++j) acc[j] = (uint64_t)(tid * 2654435761u + j) << 1;
  const uint32_t y = tid | 1u;
  for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
    for (int j = 0; j < kChains; ++j) acc[j] = (uint64_t)(uint32_t)acc[j] * y + acc[j];
  }
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < kChains; ++j) s ^= acc[j];
  if (s == 0x9e3779b97f4a7c15ull) sink[0] = s;  // keeps the chains live
}

}  // namespace
}  // namespace peaks
}  // namespace hp

using namespace hp;

extern "C" HpMi355xStatus hp_mi355x_v1_device_peaks(HpMi355xContext* ctx, size_t copy_bytes,
                                                    uint32_t reps, double* out_hbm_bytes_per_s,
                                                    double* out_mad_per_s) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out_hbm_bytes_per_s == nullptr || out_mad_per_s == nullptr || reps == 0 || reps > 100)
      return rt::invalid_parameters();
    if (copy_bytes < (size_t(1) << 20) || copy_bytes > (size_t(16) << 30) || copy_bytes % 16)
      return rt::invalid_parameters();
    int cus = 0;
    rt::hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
    rt::BufferRef src = c.allocate(copy_bytes), dst = c.allocate(copy_bytes);
    rt::BufferRef sink = c.allocate(sizeof(uint64_t));
    rt::hip_check(hipMemsetAsync(src->ptr, 0x5a, copy_bytes, c.stream));
    hipEvent_t e0, e1;
    rt::hip_check(hipEventCreate(&e0));
    rt::hip_check(hipEventCreate(&e1));
    auto timed = [&](auto&& launch) {
      float best = 1e30f;
      launch();  // warm-up
      for (uint32_t r = 0; r < reps; ++r) {
        rt::hip_check(hipEventRecord(e0, c.stream));
        launch();
        rt::hip_check(hipEventRecord(e1, c.stream));
        rt::hip_check(hipEventSynchronize(e1));
        float ms = 0.f;
        rt::hip_check(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
      }
      return best;
    };
    const size_t n16 = copy_bytes / 16;
    const unsigned copy_grid = static_cast<unsigned>(std::max(1, cus) * 16);
    const float copy_ms = timed([&] {
      hipLaunchKernelGGL(peaks::k_copy, dim3(copy_grid), dim3(256), 0, c.stream,
                         static_cast<const uint4*>(src->ptr), static_cast<uint4*>(dst->ptr), n16);
      rt::check_launch();
    });
    const uint32_t iters = 4096;
    const unsigned mad_grid = static_cast<unsigned>(std::max(1, cus) * 16);
    const float mad_ms = timed([&] {
      hipLaunchKernelGGL(peaks::k_mad_u64, dim3(mad_grid), dim3(256), 0, c.stream, iters,
                         static_cast<uint64_t*>(sink->ptr));
      rt::check_launch();
    });
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *out_hbm_bytes_per_s = 2.0 * (double)copy_bytes / (copy_ms * 1e-3);
    *out_mad_per_s =
        (double)mad_grid * 256.0 * iters * peaks::kChains / (mad_ms * 1e-3);
    return rt::ok();
  });
}
