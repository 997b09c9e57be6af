// Pippenger G1 MSM on gfx950 against a prepared, window-shifted SRS layer.
//
// Stages (all on the context stream; the host finishes the last level):
//   1. sort      the scalars are made canonical and folded below (r - 1) / 2
//                once; the signed c-bit digits are then sorted by bucket in
//                two counting-sort passes: by partition of 256 buckets
//                (blocks of scalar chunks recompute their digits and write
//                8-byte records in long runs), then inside each partition
//                by one block with LDS counters, leaving base references
//                (w n + i | sign) in bucket-contiguous order.
//   2. tasks     each bucket is cut into tasks of <= task_cap entries.
//   3. accumulate one lane per task (<= cap entries of one bucket): XYZZ +=
//                affine (8M + 2S); 64 independent accumulators per wavefront.
//   4. buckets   per-bucket sum of its task partials.
//   5. reduce    sum_b (b + 1) S_b: blocks of 256 buckets run an LDS suffix
//                scan (U_t = sum_{t' >= t} S_t') and a tree sum of U, giving
//                (A, R) per block; the host adds the blocks with their
//                weights (sum A_k + 256 sum k R_k) on 64-bit limbs.
// The group law is exact in every case, so the result does not depend on the
// order of accumulation.  Work per 2^20 scalars at c = 17: W = 15 windows,
// ~15.7M mixed additions, 2^16 buckets.
//
// The reference computes the same value with ark-ec's msm_unchecked
// (backend/src/cpu/multilinear_kzg.rs:55,131); the reference accelerator
// bridge uses unsigned windows with c = ell - ell/3 + 2
// (backend/src/ascend910b3/native/msm_plan.cpp:119-123).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <vector>

#include "kernels_common.hpp"
#include "mle.hpp"
#include "host_g1.hpp"
#include "msm.hpp"
#include "curve29.hpp"

namespace hp {
namespace msm {
namespace {

constexpr uint32_t kSkip = 0xffffffffu;

struct Task {
  uint32_t start, len, bucket, job;
};

__device__ __forceinline__ uint32_t scalar_bits(const uint32_t* s, uint32_t pos, uint32_t c) {
  const uint32_t k = pos >> 5, sh = pos & 31;
  uint64_t w = s[k];
  if (k + 1 < 8) w |= (uint64_t)s[k + 1] << 32;
  return (uint32_t)(w >> sh) & ((1u << c) - 1u);
}

// Canonical scalar i, folded into the lower half: s > (r - 1) / 2 becomes
// r - s with *neg set (the point is negated instead), so every scalar is
// below 2^254 and W = ceil(255 / c) windows of signed digits end without a
// carry.  Same value: (r - s) (-P) = s P.
__device__ __forceinline__ Fr half_scalar(const Fr* scalars, size_t i, bool& neg) {
  // (r - 1) / 2, little-endian words
  constexpr uint32_t kHalf[8] = {0x80000000u, 0x7fffffffu, 0x7fff2dffu, 0xa9ded201u,
                                 0x04d0ec02u, 0x199cec04u, 0x94cebea4u, 0x39f6d3a9u};
  const Fr s = from_mont(load_fr(scalars, i));
  int cmp = 0;  // sign of s - half, from the top word down
#pragma unroll
  for (int k = 7; k >= 0; --k)
    if (cmp == 0 && s.v[k] != kHalf[k]) cmp = s.v[k] > kHalf[k] ? 1 : -1;
  neg = cmp > 0;
  return neg ? hp::neg(s) : s;
}

// Signed c-bit digit of window w given the running carry; returns the bucket
// key (|d| - 1, or kSkip for a zero digit) and sets *neg for a negative digit.
__device__ __forceinline__ uint32_t window_digit(const uint32_t* s, uint32_t w, uint32_t c,
                                                 uint32_t& carry, bool& neg) {
  const uint32_t half = 1u << (c - 1), full = 1u << c;
  const uint32_t raw = scalar_bits(s, w * c, c) + carry;
  if (raw > half) {
    carry = 1;
    neg = true;
    const uint32_t mag = full - raw;
    return mag ? mag - 1 : kSkip;
  }
  carry = 0;
  neg = false;
  return raw ? raw - 1 : kSkip;
}

// Jobs of one batch share one global bucket space: job j owns global buckets
// [bucket_base, bucket_base + buckets) (bucket_base a multiple of 256, the
// padding buckets stay empty).  Every job's padded bucket range is cut into
// kParts = 256 partitions of 2^part_bits buckets (part_bits = max(0, c - 9)),
// global partition j * 256 + q.  The sort is two counting-sort passes:
//   1. blocks (job, scalar chunk) recompute their digits and place 8-byte
//      records (base reference, bucket within the partition) by partition:
//      (chunk, partition) counters [cc_base, cc_base + chunks * 256),
//      chunk-major, a prefix over each partition's chunks and a scan of the
//      partition totals give every (partition, chunk) its first slot; chunks
//      of ~2048 scalars leave runs of ~100 records per (partition, chunk), so
//      the writes are coalesced;
//   2. one block per partition counts its records per bucket in LDS and writes
//      the base references in bucket order inside the partition's span, with
//      the bucket starts, counts and owners.
// A prepared base: packed Fq29 x, y (96 bytes) padded to 128 so that every
// base is one aligned 128-byte line; a random base read then fetches one line
// instead of straddling two.
struct alignas(128) PreparedBase {
  G1Affine p;
  uint32_t pad[8];
};
static_assert(sizeof(PreparedBase) == 128, "one 128-byte line per base");

__device__ __forceinline__ G1Affine load_base(const PreparedBase* b, size_t i) {
  return load_affine(&b[i].p, 0);
}
__device__ __forceinline__ void store_base(PreparedBase* b, size_t i, const G1Affine& a) {
  store_affine(&b[i].p, 0, a);
}

struct Job {
  const Fr* scalars;
  const PreparedBase* bases;  // W window-shifted copies of the layer, window-major
  uint32_t n, c, windows, buckets;
  uint32_t bucket_base, chunks, cc_base, fold_base;  // fold_base: first folded scalar
  // bucket sums: 2^lanes_log2 lanes per bucket, threads [sum_base, sum_base +
  // buckets << lanes_log2); buckets with more than seq_max partials per lane
  // take the heavy path
  uint32_t lanes_log2, seq_max, sum_base, padded_buckets;  // padded: multiple of 256
  uint32_t part_bits, pad0, pad1, pad2;  // buckets per sort partition: 2^part_bits
};
struct SortItem {
  uint32_t job, chunk;
};

constexpr uint32_t kPartsLog2 = 8, kParts = 1u << kPartsLog2;  // sort partitions per job
constexpr uint32_t kMaxPartBuckets = (1u << 19) >> kPartsLog2;   // c <= 20
constexpr unsigned kSortThreads = 1024;

// Folded canonical scalars of every job of a batch, computed once for both
// sort passes (which read them once per bucket range): eight words, the sign
// of the fold in bit 31 of the top word (folded values are below 2^254).
__global__ void k_fold_scalars(const Job* jobs, uint32_t njobs, uint32_t total, uint4* folded) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = njobs;  // the job with fold_base <= i < fold_base + n
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (jobs[mid].fold_base <= i) lo = mid; else hi = mid;
    }
    bool neg;
    Fr s = half_scalar(jobs[lo].scalars, i - jobs[lo].fold_base, neg);
    if (neg) s.v[7] |= 0x80000000u;
    folded[2 * (size_t)i] = make_uint4(s.v[0], s.v[1], s.v[2], s.v[3]);
    folded[2 * (size_t)i + 1] = make_uint4(s.v[4], s.v[5], s.v[6], s.v[7]);
  }
}

__device__ __forceinline__ bool load_folded(const uint4* folded, size_t i, uint32_t* s) {
  const uint4 a = folded[2 * i], b = folded[2 * i + 1];
  s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w;
  s[4] = b.x; s[5] = b.y; s[6] = b.z; s[7] = b.w & 0x7fffffffu;
  return (b.w >> 31) != 0;
}

// counts[cc_base + e * 256 + q] = digits of chunk e of the job in partition q.
__global__ void __launch_bounds__(kSortThreads)
    k_part_count(const SortItem* items, const Job* jobs, const uint4* folded, uint32_t* counts) {
  __shared__ uint32_t cnt[kParts];
  const SortItem it = items[blockIdx.x];
  const Job j = jobs[it.job];
  for (uint32_t q = threadIdx.x; q < kParts; q += blockDim.x) cnt[q] = 0;
  __syncthreads();
  const size_t i0 = (size_t)j.n * it.chunk / j.chunks, i1 = (size_t)j.n * (it.chunk + 1) / j.chunks;
  for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    uint32_t s[8];
    load_folded(folded, j.fold_base + i, s);
    uint32_t carry = 0;
    for (uint32_t w = 0; w < j.windows; ++w) {
      bool neg;
      const uint32_t key = window_digit(s, w, j.c, carry, neg);
      if (key != kSkip) atomicAdd(&cnt[key >> j.part_bits], 1u);
    }
  }
  __syncthreads();
  uint32_t* row = counts + j.cc_base + (size_t)it.chunk * kParts;
  for (uint32_t q = threadIdx.x; q < kParts; q += blockDim.x) row[q] = cnt[q];
}

// One block per global partition: the first slot of every chunk inside it
// (exclusive prefix over its chunk counts, one thread per chunk, chunks <=
// kMaxChunks) and its record count.
constexpr unsigned kMaxChunks = 512;
__global__ void __launch_bounds__(kMaxChunks)
    k_part_prefix(const Job* jobs, const uint32_t* cc_counts, uint32_t* cc_prefix,
                  uint32_t* part_count) {
  __shared__ uint32_t sc[kMaxChunks];
  const uint32_t gq = blockIdx.x;
  const Job& j = jobs[gq >> kPartsLog2];
  const uint32_t e = threadIdx.x;
  const size_t at = j.cc_base + (gq & (kParts - 1)) + (size_t)e * kParts;
  const uint32_t v = e < j.chunks ? cc_counts[at] : 0u;
  sc[e] = v;
  __syncthreads();
  for (unsigned d = 1; d < kMaxChunks; d <<= 1) {
    const uint32_t add = e >= d ? sc[e - d] : 0u;
    __syncthreads();
    sc[e] += add;
    __syncthreads();
  }
  if (e < j.chunks) cc_prefix[at] = sc[e] - v;
  if (e == kMaxChunks - 1) part_count[gq] = sc[e];
}

// records[slot] = ((w n + i) | sign << 31, bucket within the partition)
__global__ void __launch_bounds__(kSortThreads)
    k_part_scatter(const SortItem* items, const Job* jobs, const uint4* folded,
                   const uint32_t* cc_prefix, const uint32_t* part_start, uint2* records) {
  __shared__ uint32_t cur[kParts];
  const SortItem it = items[blockIdx.x];
  const Job j = jobs[it.job];
  const uint32_t* row = cc_prefix + j.cc_base + (size_t)it.chunk * kParts;
  const uint32_t* first = part_start + (size_t)it.job * kParts;
  for (uint32_t q = threadIdx.x; q < kParts; q += blockDim.x) cur[q] = first[q] + row[q];
  __syncthreads();
  const uint32_t low = (1u << j.part_bits) - 1u;
  const size_t i0 = (size_t)j.n * it.chunk / j.chunks, i1 = (size_t)j.n * (it.chunk + 1) / j.chunks;
  for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    uint32_t s[8];
    const bool sneg = load_folded(folded, j.fold_base + i, s);
    uint32_t carry = 0;
    for (uint32_t w = 0; w < j.windows; ++w) {
      bool neg;
      const uint32_t key = window_digit(s, w, j.c, carry, neg);
      if (key != kSkip) {
        const uint32_t pos = atomicAdd(&cur[key >> j.part_bits], 1u);
        records[pos] = make_uint2((uint32_t)(w * j.n + i) | (neg != sneg ? 0x80000000u : 0u),
                                  key & low);
      }
    }
  }
}

// One block per global partition: counting sort of its records by bucket
// inside the partition's span [part_start, part_start + part_count), which is
// exactly the union of its buckets' spans; writes each bucket's start, count
// and owning job.
__global__ void __launch_bounds__(kSortThreads)
    k_part_sort(const Job* jobs, const uint32_t* part_start, const uint32_t* part_count,
                const uint2* records, uint32_t* sorted, uint32_t* starts, uint32_t* counts,
                uint32_t* owner) {
  __shared__ uint32_t cnt[kMaxPartBuckets], cur[kMaxPartBuckets];
  __shared__ uint32_t share[64];
  const uint32_t gq = blockIdx.x, job = gq >> kPartsLog2;
  const Job& j = jobs[job];
  const uint32_t nb = 1u << j.part_bits;
  const uint32_t s0 = part_start[gq], len = part_count[gq];
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) cnt[b] = 0;
  __syncthreads();
  for (uint32_t x = threadIdx.x; x < len; x += blockDim.x) atomicAdd(&cnt[records[s0 + x].y], 1u);
  __syncthreads();
  // exclusive scan of the nb counters: 64 threads each own a span of
  // ceil(nb / 64) counters, thread 0 scans the 64 span totals
  const uint32_t span = (nb + 63) / 64;
  if (threadIdx.x < 64) {
    uint32_t t = 0;
    for (uint32_t b = threadIdx.x * span; b < min(nb, (threadIdx.x + 1) * span); ++b) t += cnt[b];
    share[threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = s0;
    for (uint32_t k = 0; k < 64; ++k) {
      const uint32_t t = share[k];
      share[k] = run;
      run += t;
    }
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    uint32_t run = share[threadIdx.x];
    for (uint32_t b = threadIdx.x * span; b < min(nb, (threadIdx.x + 1) * span); ++b) {
      cur[b] = run;
      run += cnt[b];
    }
  }
  __syncthreads();
  const uint32_t g0 = j.bucket_base + ((gq & (kParts - 1)) << j.part_bits);
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    starts[g0 + b] = cur[b];
    counts[g0 + b] = cnt[b];
    owner[g0 + b] = job;
  }
  __syncthreads();  // the starts are read before the slots advance
  for (uint32_t x = threadIdx.x; x < len; x += blockDim.x) {
    const uint2 e = records[s0 + x];
    sorted[atomicAdd(&cur[e.y], 1u)] = e.x;
  }
}

// ---- exclusive scan of uint32 (up to 4M elements) -------------------------
constexpr unsigned kScanThreads = 256, kScanPer = 4, kScanTile = kScanThreads * kScanPer;

__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t s[kScanThreads];
  s[threadIdx.x] = v;
  __syncthreads();
  for (unsigned d = 1; d < kScanThreads; d <<= 1) {
    uint32_t add = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  const uint32_t incl = s[threadIdx.x];
  *total = s[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

// out[i] = exclusive prefix within the tile; tile_sums[tile] = tile total.
__global__ void k_scan_tiles(const uint32_t* in, uint32_t* out, size_t n, uint32_t* tile_sums) {
  const size_t base = blockIdx.x * (size_t)kScanTile + threadIdx.x * kScanPer;
  uint32_t v[kScanPer], sum = 0;
#pragma unroll
  for (unsigned j = 0; j < kScanPer; ++j) {
    v[j] = base + j < n ? in[base + j] : 0;
    sum += v[j];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(sum, &total);
#pragma unroll
  for (unsigned j = 0; j < kScanPer; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Exclusive scan of tile sums (<= kScanTile entries) in one block; writes the
// grand total to tile_sums[tiles].
__global__ void k_scan_top(uint32_t* tile_sums, size_t tiles) {
  const size_t base = threadIdx.x * kScanPer;
  uint32_t v[kScanPer], sum = 0;
#pragma unroll
  for (unsigned j = 0; j < kScanPer; ++j) {
    v[j] = base + j < tiles ? tile_sums[base + j] : 0;
    sum += v[j];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(sum, &total);
#pragma unroll
  for (unsigned j = 0; j < kScanPer; ++j) {
    if (base + j < tiles) tile_sums[base + j] = run;
    run += v[j];
  }
  if (threadIdx.x == 0) tile_sums[tiles] = total;
}

__global__ void k_scan_add(uint32_t* out, size_t n, const uint32_t* tile_sums, uint32_t* total_out,
                           size_t tiles) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] += tile_sums[i / kScanTile];
  if (i == 0 && total_out) *total_out = tile_sums[tiles];
}

// Exclusive scan of n <= kSmallScan values by one block (kSmallScan /
// kScanBlock consecutive values per thread), grand total to *total_out.
constexpr unsigned kScanBlock = 1024, kSmallPer = 16, kSmallScan = kScanBlock * kSmallPer;
__global__ void __launch_bounds__(kScanBlock)
    k_scan_small(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total_out) {
  __shared__ uint32_t sh[kScanBlock];
  const size_t base = (size_t)threadIdx.x * kSmallPer;
  uint32_t v[kSmallPer], sum = 0;
#pragma unroll
  for (unsigned j = 0; j < kSmallPer; ++j) {
    v[j] = base + j < n ? in[base + j] : 0u;
    sum += v[j];
  }
  sh[threadIdx.x] = sum;
  __syncthreads();
  for (unsigned d = 1; d < kScanBlock; d <<= 1) {
    const uint32_t add = threadIdx.x >= d ? sh[threadIdx.x - d] : 0u;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t run = sh[threadIdx.x] - sum;
#pragma unroll
  for (unsigned j = 0; j < kSmallPer; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
  if (threadIdx.x == kScanBlock - 1 && total_out) *total_out = sh[threadIdx.x];
}

// ---- tasks ----------------------------------------------------------------
__global__ void k_task_counts(const uint32_t* counts, uint32_t buckets, uint32_t cap,
                              uint32_t* ntask) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < buckets;
       b += gridDim.x * blockDim.x)
    ntask[b] = (counts[b] + cap - 1) / cap;
}

// One thread per task: its bucket is the last b with task_off[b] <= t (a
// binary search), so a hot bucket's thousands of tasks are written in
// parallel rather than by one thread.
__global__ void k_build_tasks(const uint32_t* counts, const uint32_t* starts,
                              const uint32_t* owner, const uint32_t* task_off, uint32_t buckets,
                              uint32_t cap, const uint32_t* ntotal, Task* tasks) {
  const uint32_t total = *ntotal;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = buckets;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (task_off[mid] <= t) lo = mid; else hi = mid;
    }
    const uint32_t done = (t - task_off[lo]) * cap;
    tasks[t] = {starts[lo] + done, min(cap, counts[lo] - done), lo, owner[lo]};
  }
}

using f29::Fq29;
using f29::Xyzz29;

// Xyzz29 = 56 words = 14 x 16 bytes; word w is limb w % 14 of coordinate w / 14.
__device__ __forceinline__ uint32_t& xyzz29_word(Xyzz29& v, int w) {
  f29::Fq29& c = w < 14 ? v.X : w < 28 ? v.Y : w < 42 ? v.ZZ : v.ZZZ;
  return c.l[w % 14];
}
__device__ __forceinline__ uint32_t xyzz29_word(const Xyzz29& v, int w) {
  const f29::Fq29& c = w < 14 ? v.X : w < 28 ? v.Y : w < 42 ? v.ZZ : v.ZZZ;
  return c.l[w % 14];
}

__device__ __forceinline__ void store_xyzz29(Xyzz29* base, size_t i, const Xyzz29& v) {
  uint4* q = reinterpret_cast<uint4*>(base + i);
#pragma unroll
  for (int k = 0; k < 14; ++k)
    q[k] = make_uint4(xyzz29_word(v, 4 * k), xyzz29_word(v, 4 * k + 1),
                      xyzz29_word(v, 4 * k + 2), xyzz29_word(v, 4 * k + 3));
}

__device__ __forceinline__ Xyzz29 load_xyzz29(const Xyzz29* base, size_t i) {
  const uint4* q = reinterpret_cast<const uint4*>(base + i);
  Xyzz29 v;
#pragma unroll
  for (int k = 0; k < 14; ++k) {
    const uint4 a = q[k];
    xyzz29_word(v, 4 * k) = a.x;
    xyzz29_word(v, 4 * k + 1) = a.y;
    xyzz29_word(v, 4 * k + 2) = a.z;
    xyzz29_word(v, 4 * k + 3) = a.w;
  }
  return v;
}

// One lane per task: each task is <= task_cap consecutive entries of one
// bucket, so a wavefront holds 64 independent XYZZ accumulators and no lane
// waits on a cross-lane reduction.  Bases hold packed Fq29 coordinates.
__global__ void __launch_bounds__(256) k_accumulate(const Task* tasks, const uint32_t* ntasks,
                                                    const uint32_t* sorted, const Job* jobs,
                                                    Xyzz29* partials) {
  const uint32_t task = blockIdx.x * blockDim.x + threadIdx.x;
  if (task >= *ntasks) return;
  const Task t = tasks[task];
  const PreparedBase* bases = jobs[t.job].bases;
  Xyzz29 acc = Xyzz29::identity();
  for (uint32_t e = 0; e < t.len; ++e) {
    const uint32_t v = sorted[t.start + e];
    const G1Affine raw = load_base(bases, v & 0x7fffffffu);
    if (raw.is_identity()) continue;
    f29::Aff29 p{f29::unpack(raw.x.v), f29::unpack(raw.y.v)};
    if (v >> 31) p.y = f29::neg_canonical(p.y);
    acc = f29::add_mixed(acc, p);
  }
  store_xyzz29(partials, task, acc);
}

__device__ __forceinline__ Xyzz29 shfl_xor_xyzz(const Xyzz29& v, unsigned mask) {
  Xyzz29 r;
#pragma unroll
  for (int i = 0; i < f29::L; ++i) {
    r.X.l[i] = __shfl_xor(v.X.l[i], mask, 64);
    r.Y.l[i] = __shfl_xor(v.Y.l[i], mask, 64);
    r.ZZ.l[i] = __shfl_xor(v.ZZ.l[i], mask, 64);
    r.ZZZ.l[i] = __shfl_xor(v.ZZZ.l[i], mask, 64);
  }
  return r;
}

// Bucket sum of its task partials by a group of 2^lanes_log2 lanes (per
// job: larger MSMs have more partials per bucket): each lane adds every
// group-th partial, then a butterfly over the group.  Buckets with more than
// seq_max partials per lane (outliers, several times the job's mean) are
// appended to `heavy` instead and summed by the heavy kernels.  W > 1: every
// lane of the group is W threads adding with f29::add_pair (W = 2) or
// f29::add_quad (W = 4), for small batches whose wavefronts leave SIMDs idle
// (a group of 64 / W lanes at most, so it stays inside one wavefront);
// total_threads counts lanes.
constexpr uint32_t kSeqMax = 8;
__device__ __forceinline__ Fq29 shfl_pair(const Fq29& v) {
  Fq29 r;
#pragma unroll
  for (int i = 0; i < f29::L; ++i) r.l[i] = __shfl_xor(v.l[i], 1, 64);
  return r;
}

// Lane K's value to every lane of each quad (DPP quad_perm [K, K, K, K], one
// VALU move per limb): the exchanges of f29::add_quad.
struct QuadBroadcast {
  template <int K>
  __device__ __forceinline__ Fq29 get(const Fq29& v) const {
    Fq29 r;
#pragma unroll
    for (int i = 0; i < f29::L; ++i)
      r.l[i] = (uint32_t)__builtin_amdgcn_mov_dpp((int)v.l[i], K * 0x55, 0xf, 0xf, false);
    return r;
  }
};

// One XYZZ addition by W = 1, 2 or 4 threads (h: the thread's index among
// them; all W hold both operands and get the sum).
template <int W>
__device__ __forceinline__ Xyzz29 lane_add(const Xyzz29& a, const Xyzz29& b, uint32_t h) {
  if constexpr (W == 4) {
    QuadBroadcast bc;
    return f29::add_quad(a, b, h, bc);
  } else if constexpr (W == 2) {
    auto ex = [](const Fq29& v) { return shfl_pair(v); };
    return f29::add_pair(a, b, h, ex);
  } else {
    return f29::add(a, b);
  }
}

template <int W>
__global__ void __launch_bounds__(256) k_bucket_sums(const uint32_t* task_off,
                                                     const uint32_t* ntask,
                                                     const Xyzz29* partials, const Job* jobs,
                                                     uint32_t njobs, uint32_t total_threads,
                                                     Xyzz29* sums, uint32_t* heavy,
                                                     uint32_t* heavy_count) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t gid = tid / W, h = tid % W;
  // job owning this lane (last with sum_base <= gid); wave-uniform since
  // every job's lane range is a multiple of 64 lanes
  uint32_t lo = 0, hi = njobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (jobs[mid].sum_base <= gid) lo = mid; else hi = mid;
  }
  const Job& job = jobs[lo];
  const uint32_t lanes_log2 = job.lanes_log2;
  const uint32_t group = 1u << lanes_log2;
  const uint32_t local = gid - job.sum_base;
  const uint32_t lb = local >> lanes_log2, lane = local & (group - 1);
  const uint32_t b = job.bucket_base + lb;
  bool live = gid < total_threads && lb < job.padded_buckets;  // whole groups share liveness
  Xyzz29 s = Xyzz29::identity();
  if (live) {
    const uint32_t first = task_off[b], cnt = ntask[b];
    if (cnt > job.seq_max * group) {
      if (lane == 0 && h == 0) heavy[atomicAdd(heavy_count, 1u)] = b;
      live = false;
    } else {
      for (uint32_t k = lane; k < cnt; k += group) {
        const Xyzz29 x = load_xyzz29(partials, first + k);
        s = lane_add<W>(s, x, h);
      }
    }
  }
  for (uint32_t m = 1; m < group; m <<= 1) {
    const Xyzz29 o = shfl_xor_xyzz(s, W * m);
    s = lane_add<W>(s, o, h);
  }
  if (live && lane == 0 && h == 0) store_xyzz29(sums, b, s);
}

// Heavy buckets (many partials).  The first kHeavySplitSlots of the heavy
// list are spread over heavy_split(count) <= kHeavySplit workgroups each
// (about 4 partials per thread: strided sums, LDS tree, one block partial per
// workgroup), then k_heavy_final adds the block partials of each with a
// wavefront butterfly; any further heavy bucket is summed by one workgroup
// (k_heavy_sums, grid-strided over the list).
constexpr unsigned kHeavyThreads = 256;
constexpr unsigned kHeavySplit = 16;
constexpr unsigned kHeavySplitSlots = 1024;

__device__ __forceinline__ uint32_t heavy_split(uint32_t cnt) {
  return min(kHeavySplit, max(1u, cnt / (4 * kHeavyThreads)));
}

__device__ Xyzz29 block_tree_sum(Xyzz29 s) {
  __shared__ Xyzz29 lds[kHeavyThreads];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (unsigned d = kHeavyThreads / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) {
      s = f29::add(s, lds[threadIdx.x + d]);
      lds[threadIdx.x] = s;
    }
    __syncthreads();
  }
  return s;
}

__global__ void __launch_bounds__(kHeavyThreads)
    k_heavy_partial(const uint32_t* task_off, const uint32_t* ntask, const Xyzz29* partials,
                    const uint32_t* heavy, const uint32_t* heavy_count, Xyzz29* block_sums) {
  const uint32_t buckets = min(*heavy_count, kHeavySplitSlots);
  const uint32_t total = buckets * kHeavySplit;
  // part-major: the first `buckets` work items are part 0 of every heavy bucket
  for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
    const uint32_t h = w % buckets, part = w / buckets;
    const uint32_t b = heavy[h];
    const uint32_t first = task_off[b], cnt = ntask[b];
    const uint32_t split = heavy_split(cnt);
    if (part >= split) continue;  // uniform per block
    Xyzz29 s = Xyzz29::identity();
    for (uint32_t k = part * kHeavyThreads + threadIdx.x; k < cnt; k += split * kHeavyThreads)
      s = f29::add(s, load_xyzz29(partials, first + k));
    s = block_tree_sum(s);
    if (threadIdx.x == 0) store_xyzz29(block_sums, h * kHeavySplit + part, s);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(64) k_heavy_final(const uint32_t* ntask, const uint32_t* heavy,
                                                    const uint32_t* heavy_count,
                                                    const Xyzz29* block_sums, Xyzz29* sums) {
  const uint32_t total = min(*heavy_count, kHeavySplitSlots);
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t h = blockIdx.x; h < total; h += gridDim.x) {
    const uint32_t split = heavy_split(ntask[heavy[h]]);
    Xyzz29 s = lane < split ? load_xyzz29(block_sums, h * kHeavySplit + lane) : Xyzz29::identity();
    for (unsigned m = 1; m < split; m <<= 1) s = f29::add(s, shfl_xor_xyzz(s, m));
    if (lane == 0) store_xyzz29(sums, heavy[h], s);
  }
}

__global__ void __launch_bounds__(kHeavyThreads)
    k_heavy_sums(const uint32_t* task_off, const uint32_t* ntask, const Xyzz29* partials,
                 const uint32_t* heavy, const uint32_t* heavy_count, Xyzz29* sums) {
  const uint32_t total = *heavy_count;
  for (uint32_t h = kHeavySplitSlots + blockIdx.x; h < total; h += gridDim.x) {
    const uint32_t b = heavy[h];
    const uint32_t first = task_off[b], cnt = ntask[b];
    Xyzz29 s = Xyzz29::identity();
    for (uint32_t k = threadIdx.x; k < cnt; k += kHeavyThreads)
      s = f29::add(s, load_xyzz29(partials, first + k));
    s = block_tree_sum(s);
    if (threadIdx.x == 0) store_xyzz29(sums, b, s);
    __syncthreads();
  }
}

// Level 1: blocks of G buckets.  U_t = sum_{t' >= t} S_t' (suffix scan),
// A = sum_t U_t = sum_t (t + 1) S_t, R = U_0; written in the 32-bit-limb form
// for the host.
// Bucket ranges are padded to kRed1 = 256; small batches reduce in groups of
// 64 (12 dependent additions instead of 16, more groups for the host).
constexpr unsigned kRed1 = 256;
constexpr unsigned kRedSmall = 64;
constexpr uint32_t kPairReduceMax = 65536;  // buckets of a batch reduced on two lanes each
template <unsigned G>
__global__ void __launch_bounds__(G) k_reduce_buckets(const Xyzz29* sums, uint32_t buckets,
                                                      G1Xyzz* pair_a, G1Xyzz* pair_r) {
  __shared__ Xyzz29 u[G];
  const uint32_t t = threadIdx.x;
  const uint32_t b = blockIdx.x * G + t;
  Xyzz29 v = b < buckets ? load_xyzz29(sums, b) : Xyzz29::identity();
  u[t] = v;
  __syncthreads();
  for (unsigned d = 1; d < G; d <<= 1) {
    if (t + d < G) v = f29::add(v, u[t + d]);
    __syncthreads();
    u[t] = v;
    __syncthreads();
  }
  if (t == 0) store_xyzz(pair_r, blockIdx.x, f29::to_xyzz(v));
  // tree sum of U
  for (unsigned d = G / 2; d > 0; d >>= 1) {
    if (t < d) {
      v = f29::add(v, u[t + d]);
      u[t] = v;
    }
    __syncthreads();
  }
  if (t == 0) {
    store_xyzz(pair_a, blockIdx.x, f29::to_xyzz(v));
    __threadfence_system();  // the results may be pinned host memory read after the sync
  }
}

// The same level-1 reduction with W = 2 or 4 threads per bucket
// (f29::add_pair, f29::add_quad): for small batches, whose wavefronts leave
// most SIMDs idle, each of the 2 log2 G dependent additions takes the
// per-thread share of its products.
template <unsigned G, int W>
__global__ void __launch_bounds__(W * G) k_reduce_buckets_pair(const Xyzz29* sums,
                                                               uint32_t buckets, G1Xyzz* pair_a,
                                                               G1Xyzz* pair_r) {
  __shared__ Xyzz29 u[G];
  const uint32_t t = threadIdx.x / W, h = threadIdx.x % W;  // bucket in the group, thread of its lane
  const uint32_t b = blockIdx.x * G + t;
  Xyzz29 v = b < buckets ? load_xyzz29(sums, b) : Xyzz29::identity();
  if (h == 0) u[t] = v;
  __syncthreads();
  for (unsigned d = 1; d < G; d <<= 1) {
    if (t + d < G) v = lane_add<W>(v, u[t + d], h);  // uniform within a lane
    __syncthreads();
    if (h == 0) u[t] = v;
    __syncthreads();
  }
  if (threadIdx.x == 0) store_xyzz(pair_r, blockIdx.x, f29::to_xyzz(v));
  // tree sum of U
  for (unsigned d = G / 2; d > 0; d >>= 1) {
    if (t < d) {
      v = lane_add<W>(v, u[t + d], h);
      if (h == 0) u[t] = v;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    store_xyzz(pair_a, blockIdx.x, f29::to_xyzz(v));
    __threadfence_system();  // the results may be pinned host memory read after the sync
  }
}

// The same level-1 reduction for large batches, two buckets per lane of W
// threads: lane t holds buckets 2t and 2t+1 of the group, R_t = S_2t + S_2t+1
// and A_t = S_2t + 2 S_2t+1.  With U_t the suffix scan of the R_t,
// sum_j (j + 1) S_j = sum_t (A_t + 2 U_t+1), so the scan and the tree run
// over G / 2 values: 18 dependent additions instead of 16, on half the
// wavefronts, where the one-bucket-per-lane launch is bound by the issue
// rate of two wavefronts per SIMD.
template <unsigned G, int W>
__global__ void __launch_bounds__(G / 2 * W) k_reduce_buckets_two(const Xyzz29* sums,
                                                                  uint32_t buckets, G1Xyzz* pair_a,
                                                                  G1Xyzz* pair_r) {
  constexpr unsigned U = G / 2;
  __shared__ Xyzz29 u[U];
  const uint32_t t = threadIdx.x / W, h = threadIdx.x % W;  // bucket pair, thread of its lane
  const uint32_t b = blockIdx.x * G + 2 * t;
  const Xyzz29 s0 = b < buckets ? load_xyzz29(sums, b) : Xyzz29::identity();
  const Xyzz29 s1 = b + 1 < buckets ? load_xyzz29(sums, b + 1) : Xyzz29::identity();
  Xyzz29 v = lane_add<W>(s0, s1, h);  // R_t
  Xyzz29 a = lane_add<W>(v, s1, h);   // A_t
  if (h == 0) u[t] = v;
  __syncthreads();
  for (unsigned d = 1; d < U; d <<= 1) {
    if (t + d < U) v = lane_add<W>(v, u[t + d], h);  // uniform within a lane
    __syncthreads();
    if (h == 0) u[t] = v;
    __syncthreads();
  }
  if (threadIdx.x == 0) store_xyzz(pair_r, blockIdx.x, f29::to_xyzz(v));
  if (t + 1 < U) {
    const Xyzz29 w = u[t + 1];
    a = lane_add<W>(lane_add<W>(a, w, h), w, h);
  }
  __syncthreads();  // every U_t+1 read before the slots take the A_t + 2 U_t+1
  if (h == 0) u[t] = a;
  __syncthreads();
  v = a;
  for (unsigned d = U / 2; d > 0; d >>= 1) {
    if (t < d) {
      v = lane_add<W>(v, u[t + d], h);
      if (h == 0) u[t] = v;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    store_xyzz(pair_a, blockIdx.x, f29::to_xyzz(v));
    __threadfence_system();  // the results may be pinned host memory read after the sync
  }
}

// Prepared bases: 32-bit-limb Montgomery points -> padded packed Fq29 bases,
// and back.
__global__ void k_bases_to29(const G1Affine* pts, PreparedBase* out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    G1Affine p = load_affine(pts, i);
    if (!p.is_identity()) {
      p.x = f29::to29_packed(p.x);
      p.y = f29::to29_packed(p.y);
    }
    store_base(out, i, p);
  }
}

__global__ void k_bases_from29(const PreparedBase* pts, G1Affine* out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    G1Affine p = load_base(pts, i);
    if (!p.is_identity()) {
      p.x = f29::from29_packed(p.x);
      p.y = f29::from29_packed(p.y);
    }
    store_affine(out, i, p);
  }
}

// ---- SRS preparation --------------------------------------------------------
// next[i] = 2^c prev[i] on packed Fq29 coordinates: c XYZZ doublings on
// 14 x 29-bit limbs, normalised with one inversion per thread (Montgomery
// batch inversion over the thread's strided points).
__global__ void k_shift_window29(const PreparedBase* prev, PreparedBase* next, size_t n, uint32_t c,
                                 Xyzz29* scratch, Fq29* prefix) {
  const size_t threads = (size_t)gridDim.x * blockDim.x;
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (tid >= n) return;
  Fq29 acc = f29::one();
  for (size_t i = tid; i < n; i += threads) {
    const G1Affine raw = load_base(prev, i);
    Xyzz29 q = Xyzz29::identity();
    if (!raw.is_identity()) {
      q = f29::dbl_affine(f29::Aff29{f29::unpack(raw.x.v), f29::unpack(raw.y.v)});
      for (uint32_t s = 1; s < c; ++s) q = f29::dbl(q);
    }
    store_xyzz29(scratch, i, q);
    prefix[i] = acc;
    if (!q.is_identity()) acc = f29::mul(acc, f29::mul(q.ZZ, q.ZZZ));
  }
  Fq29 inv = f29::inverse(acc);
  const size_t last = tid + ((n - 1 - tid) / threads) * threads;
  for (size_t i = last;; i -= threads) {
    const Xyzz29 q = load_xyzz29(scratch, i);
    if (q.is_identity()) {
      store_base(next, i, G1Affine::identity());
    } else {
      const Fq29 d_inv = f29::mul(inv, prefix[i]);  // 1 / (ZZ ZZZ)
      inv = f29::mul(inv, f29::mul(q.ZZ, q.ZZZ));
      G1Affine out;
      f29::pack(f29::canonical(f29::mul(q.X, f29::mul(d_inv, q.ZZZ))), out.x.v);
      f29::pack(f29::canonical(f29::mul(q.Y, f29::mul(d_inv, q.ZZ))), out.y.v);
      store_base(next, i, out);
    }
    if (i < threads) break;
  }
}

// Canonical packets (x LE48 | y LE48 as 24 u32, identity = zeros) -> Montgomery
// affine, flagging off-curve finite points.
__global__ void k_points_to_mont(G1Affine* pts, size_t n, unsigned* bad) {
  Fq four = Fq::zero();
  four.v[0] = 4;
  four = to_mont(four);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    G1Affine p = load_affine(pts, i);
    if (p.is_identity()) continue;
    p.x = to_mont(p.x);
    p.y = to_mont(p.y);
    if (sqr(p.y) != sqr(p.x) * p.x + four) atomicOr(bad, 1u);
    store_affine(pts, i, p);
  }
}

__global__ void k_points_from_mont(const G1Affine* pts, G1Affine* out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    G1Affine p = load_affine(pts, i);
    if (!p.is_identity()) {
      p.x = from_mont(p.x);
      p.y = from_mont(p.y);
    }
    store_affine(out, i, p);
  }
}

// ---- compressed encoding (SRS files) ---------------------------------------
// 48 bytes per point = 12 little-endian u32 words w[0..11] of the big-endian
// x, so canonical limb j is bswap(w[11 - j]) and the flag byte is the low
// byte of w[0] (top bits of limb 11).

// canonical y: y > -y  <=>  2 y >= q (y != 0)
__device__ __forceinline__ bool fq_largest(const Fq& canonical) {
  uint32_t d[12];
#pragma unroll
  for (int i = 0; i < 12; ++i)
    d[i] = (canonical.v[i] << 1) | (i ? canonical.v[i - 1] >> 31 : 0u);
  return geq_modulus<FqParams>(d);
}

__global__ void k_decompress(const uint4* in, size_t n, G1Affine* out, unsigned* bad) {
  // (q + 1) / 4, limb-wise from the modulus (q = 3 mod 4, so q + 1 does not carry)
  uint32_t e[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    const uint32_t lo = i == 0 ? FqParams::MOD[0] + 1u : FqParams::MOD[i];
    e[i] = (lo >> 2) | (i < 11 ? FqParams::MOD[i + 1] << 30 : 0u);
  }
  Fq four = Fq::zero();
  four.v[0] = 4;
  four = to_mont(four);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint4 a = in[3 * i], b = in[3 * i + 1], c = in[3 * i + 2];
    const uint32_t w[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
    Fq x;
#pragma unroll
    for (int j = 0; j < 12; ++j) x.v[j] = __builtin_bswap32(w[11 - j]);
    const uint32_t flags = x.v[11] >> 29;  // bit 2 compressed, 1 infinity, 0 largest
    x.v[11] &= 0x1fffffffu;
    if (!(flags & 4u)) {
      atomicOr(bad, 1u);
      continue;
    }
    if (flags & 2u) {  // infinity: every other bit clear
      if ((flags & 1u) || !x.is_zero()) atomicOr(bad, 1u);
      store_affine(out, i, G1Affine::identity());
      continue;
    }
    if (geq_modulus<FqParams>(x.v)) {
      atomicOr(bad, 1u);
      continue;
    }
    x = to_mont(x);
    const Fq rhs = sqr(x) * x + four;
    Fq y = Fq::one();
    for (int l = 11; l >= 0; --l)
      for (int b2 = 31; b2 >= 0; --b2) {
        y = sqr(y);
        if ((e[l] >> b2) & 1u) y = y * rhs;
      }
    if (sqr(y) != rhs) {
      atomicOr(bad, 1u);
      continue;
    }
    if (fq_largest(from_mont(y)) != bool(flags & 1u)) y = neg(y);
    store_affine(out, i, {x, y});
  }
}

__global__ void k_compress(const G1Affine* pts, size_t n, uint4* out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const G1Affine p = load_affine(pts, i);
    uint32_t w[12];
    if (p.is_identity()) {
#pragma unroll
      for (int j = 0; j < 12; ++j) w[j] = 0;
      w[0] = 0xC0u;
    } else {
      const Fq x = from_mont(p.x);
#pragma unroll
      for (int j = 0; j < 12; ++j) w[11 - j] = __builtin_bswap32(x.v[j]);
      w[0] |= fq_largest(from_mont(p.y)) ? 0xA0u : 0x80u;
    }
    out[3 * i] = make_uint4(w[0], w[1], w[2], w[3]);
    out[3 * i + 1] = make_uint4(w[4], w[5], w[6], w[7]);
    out[3 * i + 2] = make_uint4(w[8], w[9], w[10], w[11]);
  }
}

__global__ void k_zero_u32(uint32_t* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// ---- fixed-base generation (SRS) ------------------------------------------
// table[w * 256 + d] = d * 2^(8 w) * g (affine), built by one thread per entry.
__global__ void k_fixed_table(G1Affine g, G1Affine* table) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 32 * 256) return;
  const uint32_t w = idx / 256, d = idx % 256;
  G1Xyzz base = G1Xyzz::from_affine(g);
  for (uint32_t s = 0; s < 8 * w; ++s) base = dbl(base);
  G1Xyzz acc = G1Xyzz::identity();
  for (int b = 7; b >= 0; --b) {
    acc = dbl(acc);
    if ((d >> b) & 1) acc = add(acc, base);
  }
  store_affine(table, idx, to_affine(acc));
}

// out[i] = s_i * g using the byte table; batch-normalised per thread.
__global__ void k_fixed_base(const Fr* scalars, size_t n, const G1Affine* table, G1Affine* out,
                             G1Xyzz* scratch, Fq* prefix) {
  const size_t threads = (size_t)gridDim.x * blockDim.x;
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (tid >= n) return;
  Fq acc = Fq::one();
  for (size_t i = tid; i < n; i += threads) {
    const Fr s = from_mont(load_fr(scalars, i));
    G1Xyzz q = G1Xyzz::identity();
    for (uint32_t w = 0; w < 32; ++w) {
      const uint32_t d = (s.v[w >> 2] >> (8 * (w & 3))) & 0xffu;
      if (d) q = add_mixed(q, load_affine(table, w * 256 + d));
    }
    store_xyzz(scratch, i, q);
    store_fq(&prefix[i], acc);
    if (!q.is_identity()) acc = acc * (q.ZZ * q.ZZZ);
  }
  Fq inv = inverse(acc);
  const size_t last = tid + ((n - 1 - tid) / threads) * threads;
  for (size_t i = last;; i -= threads) {
    const G1Xyzz q = load_xyzz(scratch, i);
    if (q.is_identity()) {
      store_affine(out, i, G1Affine::identity());
    } else {
      const Fq d_inv = inv * load_fq(&prefix[i]);
      inv = inv * (q.ZZ * q.ZZZ);
      store_affine(out, i, {q.X * (d_inv * q.ZZZ), q.Y * (d_inv * q.ZZ)});
    }
    if (i < threads) break;
  }
}

unsigned grid_for(size_t n, unsigned threads = 256, size_t cap = 65535) {
  size_t b = (n + threads - 1) / threads;
  return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(b, cap)));
}

// Exclusive scan of n uint32 values; optional grand total.  Up to
// kSmallScan values (the sort partitions of a batch, the buckets of small
// batches) in one launch of one block; beyond, tile sums are scanned
// recursively when there are more than one tile's worth of tiles.
void exclusive_scan(rt::Context& c, const uint32_t* in, uint32_t* out, size_t n,
                    uint32_t* total_out, std::vector<rt::BufferRef>& keep) {
  if (n <= kSmallScan) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kScanBlock), 0, c.stream, in, out, n, total_out);
    rt::check_launch();
    return;
  }
  const size_t tiles = (n + kScanTile - 1) / kScanTile;
  rt::BufferRef sums_buf = c.allocate((tiles + 1) * sizeof(uint32_t));
  keep.push_back(sums_buf);
  uint32_t* sums = static_cast<uint32_t*>(sums_buf->ptr);
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(kScanThreads), 0, c.stream, in, out,
                     n, sums);
  rt::check_launch();
  if (tiles <= kScanTile) {
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, c.stream, sums, tiles);
    rt::check_launch();
  } else {
    // scan the tile sums in place (recursively), total to sums[tiles]
    rt::BufferRef copy = c.allocate(tiles * sizeof(uint32_t));
    keep.push_back(copy);
    rt::hip_check(hipMemcpyAsync(copy->ptr, sums, tiles * sizeof(uint32_t),
                                 hipMemcpyDeviceToDevice, c.stream));
    exclusive_scan(c, static_cast<const uint32_t*>(copy->ptr), sums, tiles, sums + tiles, keep);
  }
  hipLaunchKernelGGL(k_scan_add, dim3(grid_for(n, 256, 1u << 30)), dim3(256), 0, c.stream, out, n,
                     sums, total_out, tiles);
  rt::check_launch();
}

void zero_u32(rt::Context& c, uint32_t* p, size_t n) {
  hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(n)), dim3(256), 0, c.stream, p, n);
  rt::check_launch();
}

}  // namespace

// Window bits.  All windows share one bucket set, so a window whose digits
// span few values (the top window when c does not divide the scalar width
// well) concentrates n entries in a handful of buckets.  For the folded
// scalars (< 2^254): c = 17 (15 windows, top digits up to 59356 of 65536
// buckets), 15 (17 windows, top window of 14 bits) and 13 (20 windows, 7 bits)
// avoid that.  Measured on MI355X (device time, best task cap): 2^22 c = 17
// 12.2 ms against c = 16 12.8 ms; 2^20 3.42 against 3.60; 2^18 c = 15 1.30
// against 1.37; 2^16 c = 15 0.69 against 0.73 (16384 buckets reduce in
// groups of 64); 2^12 c = 13 0.45 against 0.46 (c = 15).
Plan choose_plan(size_t n, uint32_t override_c) {
  Plan p;
  uint32_t c = override_c;
  if (c == 0) {
    if (n >= (size_t(1) << 20)) {
      c = 17;
    } else if (n >= (size_t(1) << 15)) {
      c = 15;
    } else if (n >= (size_t(1) << 11)) {
      c = 13;
    } else {
      c = 16;
      for (int it = 0; it < 3; ++it) {
        const uint32_t w = (255 + c - 1) / c;
        const double lg = std::log2(std::max<double>(1.0, (double)n * w / 64.0));
        const int cc = (int)std::lround(lg) + 1;
        c = (uint32_t)std::min(20, std::max(4, cc));
      }
    }
  }
  p.c = c;
  p.windows = (255 + c - 1) / c;
  p.buckets = 1u << (c - 1);
  return p;
}

// Bucket sums run on two lanes per addition up to this many lanes (two
// wavefronts per SIMD); beyond it the one-lane launch issues as fast.
constexpr size_t kPairLanes = size_t(1) << 17;

// Four threads per bucket lane (f29::add_quad) in the latency-bound bucket sums
// and level-1 reductions; HP_MI355X_MSM_QUAD=0 keeps two (f29::add_pair).
bool quad_lanes() {
  static const bool on = [] {
    const char* e = std::getenv("HP_MI355X_MSM_QUAD");
    return e == nullptr || std::atoi(e) != 0;
  }();
  return on;
}

// At most 2^sum_lanes_cap() lanes per bucket in the bucket sums of MSMs with
// at least 16384 buckets (HP_MI355X_MSM_SUM_LANES, default 2).  Fewer lanes
// mean longer sequential chains but fewer tree levels over the lanes, and the
// narrower launch then admits two or four threads per addition: measured on
// MI355X, bucket sums at 2^16 (c = 15, ~9.5 partials a bucket) 0.120 ms with 8
// lanes of one thread, 0.094 with 4 lanes of four (whole MSM 0.645 -> 0.615
// ms); 2^18 0.150 against 0.131; 2^20 keeps its 2 lanes.  Applied to the 4096-
// bucket MSMs of an nv-12 proof the cap cost 0.15 ms a proof, and one lane of
// four threads sends too many buckets to the heavy path at 2^12 (0.054 ->
// 0.17 ms), hence the bucket-count condition and the cap of 2.
uint32_t sum_lanes_cap() {
  static const uint32_t v = [] {
    const char* e = std::getenv("HP_MI355X_MSM_SUM_LANES");
    return e ? static_cast<uint32_t>(std::min(6, std::max(0, std::atoi(e)))) : 2u;
  }();
  return v;
}

// Entries per lane task by the largest MSM of a batch (autotuner sweep; for
// batches of 2^12-point MSMs, 4 beats 2 and 8 in whole proofs: the shorter
// bucket-sum chains outweigh the longer lane tasks).
uint32_t default_cap(size_t largest_n) {
  if (largest_n < (size_t(1) << 14)) return 4;
  if (largest_n < (size_t(1) << 18)) return 8;
  if (largest_n < (size_t(1) << 20)) return 16;
  return 32;
}

size_t log2_floor(size_t n) {
  size_t l = 0;
  while ((size_t(2) << l) <= n) ++l;
  return l;
}

Plan plan_for(const rt::Context& c, size_t n, uint32_t override_c) {
  if (override_c == 0 && n > 0) {
    const uint32_t t = c.tuned_window[log2_floor(n)];
    if (t) return choose_plan(n, t);
  }
  return choose_plan(n, override_c);
}

void prepare_layer_from_device(rt::Context& c, Layer& layer, const G1Affine* points, size_t n,
                               uint32_t override_c) {
  layer.n = n;
  layer.plan = plan_for(c, n, override_c);
  const uint32_t W = layer.plan.windows;
  layer.bases = c.allocate(W * n * sizeof(PreparedBase));
  PreparedBase* bases = static_cast<PreparedBase*>(layer.bases->ptr);
  // the bucket stages read packed Fq29 coordinates: window 0 is converted,
  // the shifted copies are built from it in that form
  hipLaunchKernelGGL(k_bases_to29, dim3(grid_for(n)), dim3(256), 0, c.stream, points, bases, n);
  rt::check_launch();
  if (W > 1) {
    const size_t threads = std::max<size_t>(1, std::min<size_t>(n, 65536));
    rt::BufferRef scratch = c.allocate(n * sizeof(Xyzz29));
    rt::BufferRef prefix = c.allocate(n * sizeof(Fq29));
    for (uint32_t w = 1; w < W; ++w) {
      hipLaunchKernelGGL(k_shift_window29, dim3(grid_for(threads, 64)), dim3(64), 0, c.stream,
                         bases + (w - 1) * n, bases + w * n, n, layer.plan.c,
                         static_cast<Xyzz29*>(scratch->ptr), static_cast<Fq29*>(prefix->ptr));
      rt::check_launch();
    }
  }
}

namespace {
// window 0 of a layer back in the 32-bit-limb Montgomery form
rt::BufferRef layer_points(rt::Context& c, const Layer& layer) {
  const size_t n = layer.n;
  rt::BufferRef out = c.allocate(std::max<size_t>(n, 1) * sizeof(G1Affine));
  hipLaunchKernelGGL(k_bases_from29, dim3(grid_for(n)), dim3(256), 0, c.stream,
                     static_cast<const PreparedBase*>(layer.bases->ptr),
                     static_cast<G1Affine*>(out->ptr), n);
  rt::check_launch();
  return out;
}
}  // namespace

bool upload_layer(rt::Context& c, Layer& layer, const uint8_t* packets, size_t n,
                  uint32_t override_c) {
  // host validation: tag, identity coordinates, canonical coordinates
  for (size_t i = 0; i < n; ++i) {
    const uint8_t* p = packets + rt::kG1Bytes * i;
    if (p[0] == 1) {
      for (size_t k = 1; k < rt::kG1Bytes; ++k)
        if (p[k]) return false;
      continue;
    }
    if (p[0] != 0) return false;
    if (!rt::canonical_fq(p + 1) || !rt::canonical_fq(p + 49)) return false;
    bool all_zero = true;
    for (size_t k = 1; k < rt::kG1Bytes && all_zero; ++k) all_zero = p[k] == 0;
    if (all_zero) return false;  // finite-tag (0, 0)
  }
  rt::BufferRef raw = c.allocate(std::max<size_t>(n, 1) * sizeof(G1Affine));
  // pack to 96-byte coordinates (identity -> zeros) through pinned staging
  const size_t chunk = (size_t(32) << 20) / sizeof(G1Affine);
  for (size_t done = 0; done < n; done += chunk) {
    const size_t cnt = std::min(chunk, n - done);
    uint8_t* stage = c.staging(cnt * sizeof(G1Affine));
    for (size_t i = 0; i < cnt; ++i) {
      const uint8_t* p = packets + rt::kG1Bytes * (done + i);
      uint8_t* q = stage + sizeof(G1Affine) * i;
      if (p[0] == 1)
        std::memset(q, 0, sizeof(G1Affine));
      else
        std::memcpy(q, p + 1, 96);
    }
    rt::hip_check(hipMemcpyAsync(static_cast<G1Affine*>(raw->ptr) + done, stage,
                                 cnt * sizeof(G1Affine), hipMemcpyHostToDevice, c.stream));
    c.staging_used();
  }
  rt::BufferRef flag = c.allocate(sizeof(unsigned));
  zero_u32(c, static_cast<uint32_t*>(flag->ptr), 1);
  hipLaunchKernelGGL(k_points_to_mont, dim3(grid_for(n)), dim3(256), 0, c.stream,
                     static_cast<G1Affine*>(raw->ptr), n, static_cast<unsigned*>(flag->ptr));
  rt::check_launch();
  unsigned bad = 0;
  rt::download_bytes(c, flag->ptr, sizeof(unsigned), reinterpret_cast<uint8_t*>(&bad));
  if (bad) return false;
  prepare_layer_from_device(c, layer, static_cast<const G1Affine*>(raw->ptr), n, override_c);
  return true;
}

void Batch::add(const Layer& layer, const Fr* scalars) {
  Entry e;
  e.layer = &layer;
  e.scalars = scalars;
  e.bucket_base = gbuckets_;
  // sort blocks: scalar chunks of >= 1024 scalars, at most 512 per MSM
  const uint32_t B = layer.plan.buckets;
  const uint32_t padded = (B + kRed1 - 1) / kRed1 * kRed1;
  e.chunks = (uint32_t)std::max<size_t>(1, std::min<size_t>(kMaxChunks, layer.n / 1024));
  e.cc_base = cc_total_;
  gbuckets_ += padded;
  cc_total_ += (size_t)kParts * e.chunks;
  m_total_ += layer.n * layer.plan.windows;
  largest_n_ = std::max(largest_n_, layer.n);
  // positions and counter indices are 32-bit on the device
  if (cc_total_ >= (size_t(1) << 32) || m_total_ >= (size_t(1) << 31)) rt::fail_params();
  entries_.push_back(e);
}

// level-1 group size: 64 buckets for batches of up to 16384 buckets
unsigned Batch::red_group() const { return gbuckets_ <= 16384 ? kRedSmall : kRed1; }
size_t Batch::result_slots() const { return 2 * (size_t)(gbuckets_ / red_group()); }

void Batch::enqueue(rt::Context& c, G1Xyzz* result, StageEvents* ev,
                    std::vector<rt::BufferRef>* hold) const {
  auto mark = [&](int i) {
    if (ev) rt::hip_check(hipEventRecord(ev->marks[i], c.stream));
  };
  const uint32_t njobs = (uint32_t)entries_.size();
  if (njobs == 0) return;
  // lane tasks: entries per lane by the largest MSM (4 .. 32), or the
  // context's tuned cap, or an explicit one
  uint32_t cap = default_cap(largest_n_);
  if (cap_override_)
    cap = cap_override_;
  else if (largest_n_ && c.tuned_cap[log2_floor(largest_n_)])
    cap = c.tuned_cap[log2_floor(largest_n_)];
  // job table and sort work items, uploaded together
  std::vector<Job> jobs(njobs);
  std::vector<SortItem> items;
  size_t sum_threads_total = 0, fold_total = 0;
  std::vector<double> per_bucket(njobs);
  for (uint32_t j = 0; j < njobs; ++j) {
    const Entry& e = entries_[j];
    const Plan& p = e.layer->plan;
    jobs[j] = {e.scalars, static_cast<const PreparedBase*>(e.layer->bases->ptr), (uint32_t)e.layer->n,
               p.c, p.windows, p.buckets, e.bucket_base, e.chunks, (uint32_t)e.cc_base,
               (uint32_t)fold_total, 0u, 0u, 0u, 0u};
    fold_total += e.layer->n;
    // bucket-sum shape of this MSM as if alone: lanes per bucket (at most
    // 2^sum_lanes_cap() from 16384 buckets) while each lane keeps >= 2
    // partials and the job's threads stay <= 131072; heavy threshold three
    // times its mean partials per lane
    per_bucket[j] = ((double)e.layer->n * p.windows / cap + 1.0) / p.buckets;
    const uint32_t max_lanes = p.buckets >= 16384 ? sum_lanes_cap() : 6u;
    uint32_t l = 0;
    while (l < max_lanes && ((size_t)p.buckets << (l + 1)) <= 131072 && (double)(2u << l) < per_bucket[j])
      ++l;
    jobs[j].lanes_log2 = l;
    jobs[j].padded_buckets = (p.buckets + kRed1 - 1) / kRed1 * kRed1;
    jobs[j].part_bits = p.c > 9 ? p.c - 9 : 0;  // 256 partitions of the padded range
    for (uint32_t ch = 0; ch < e.chunks; ++ch) items.push_back({j, ch});
  }
  // At most 32 lanes per bucket when that brings the two-lane launch within
  // two wavefronts per SIMD: one more dependent addition per bucket, each at
  // the latency of a two-lane addition instead of a one-lane one.
  auto sum_threads = [&](uint32_t max_l) {
    size_t t = 0;
    for (uint32_t j = 0; j < njobs; ++j)
      t += (((size_t)jobs[j].padded_buckets << std::min(jobs[j].lanes_log2, max_l)) + 63) / 64 * 64;
    return t;
  };
  // At most 16 lanes per bucket when that admits the four-thread lanes within
  // two wavefronts per SIMD (a group of 16 lanes of four threads fills a
  // wavefront): one more dependent addition, each at about 60 % of the
  // two-thread latency.
  const uint32_t max_l = quad_lanes() && 4 * sum_threads(4) <= kPairLanes ? 4
                         : 2 * sum_threads(6) > kPairLanes && 2 * sum_threads(5) <= kPairLanes ? 5
                                                                                                : 6;
  for (uint32_t j = 0; j < njobs; ++j) {
    const uint32_t l = std::min(jobs[j].lanes_log2, max_l);
    jobs[j].lanes_log2 = l;
    jobs[j].seq_max =
        std::max<uint32_t>(kSeqMax, (uint32_t)std::ceil(3.0 * per_bucket[j] / (double)(1u << l)));
    jobs[j].sum_base = sum_threads_total;
    sum_threads_total += (((size_t)jobs[j].padded_buckets << l) + 63) / 64 * 64;
  }
  const size_t job_bytes = njobs * sizeof(Job), item_bytes = items.size() * sizeof(SortItem);
  rt::BufferRef table = c.allocate(job_bytes + item_bytes);
  {
    uint8_t* stage = c.staging(job_bytes + item_bytes);
    std::memcpy(stage, jobs.data(), job_bytes);
    std::memcpy(stage + job_bytes, items.data(), item_bytes);
    rt::hip_check(hipMemcpyAsync(table->ptr, stage, job_bytes + item_bytes, hipMemcpyHostToDevice,
                                 c.stream));
    c.staging_used();
  }
  const Job* djobs = static_cast<const Job*>(table->ptr);
  const SortItem* ditems =
      reinterpret_cast<const SortItem*>(static_cast<const uint8_t*>(table->ptr) + job_bytes);

  const uint32_t G = gbuckets_;
  std::vector<rt::BufferRef> keep;
  rt::BufferRef sorted_buf = c.allocate(std::max<size_t>(m_total_, 1) * sizeof(uint32_t));
  uint32_t* sorted = static_cast<uint32_t*>(sorted_buf->ptr);
  const uint32_t P = njobs * kParts;  // sort partitions
  rt::BufferRef meta =
      c.allocate((2 * cc_total_ + 5 * (size_t)G + 2 * (size_t)P + 4) * sizeof(uint32_t));
  uint32_t* cc_counts = static_cast<uint32_t*>(meta->ptr);
  uint32_t* cc_offsets = cc_counts + cc_total_;
  uint32_t* starts = cc_offsets + cc_total_;
  uint32_t* counts = starts + G;
  uint32_t* owner = counts + G;
  uint32_t* ntask = owner + G;
  uint32_t* task_off = ntask + G;
  uint32_t* part_count = task_off + G;
  uint32_t* part_start = part_count + P;
  uint32_t* totals = part_start + P;  // [0] entries, [1] tasks
  // partition records, 8 bytes per non-zero digit (at most m_total_)
  rt::BufferRef records_buf = c.allocate(std::max<size_t>(m_total_, 1) * sizeof(uint2));
  uint2* records = static_cast<uint2*>(records_buf->ptr);
  keep.push_back(records_buf);

  // folded scalars (fold_total <= m_total_ < 2^31)
  rt::BufferRef folded_buf = c.allocate(std::max<size_t>(fold_total, 1) * 2 * sizeof(uint4));
  const uint4* folded = static_cast<const uint4*>(folded_buf->ptr);
  keep.push_back(folded_buf);

  mark(0);
  hipLaunchKernelGGL(k_fold_scalars, dim3(grid_for(fold_total)), dim3(256), 0, c.stream, djobs,
                     njobs, (uint32_t)fold_total, static_cast<uint4*>(folded_buf->ptr));
  rt::check_launch();
  mark(1);
  hipLaunchKernelGGL(k_part_count, dim3((unsigned)items.size()), dim3(kSortThreads), 0, c.stream,
                     ditems, djobs, folded, cc_counts);
  rt::check_launch();
  hipLaunchKernelGGL(k_part_prefix, dim3(P), dim3(kMaxChunks), 0, c.stream, djobs, cc_counts,
                     cc_offsets, part_count);
  rt::check_launch();
  exclusive_scan(c, part_count, part_start, P, totals, keep);  // totals[0]: entries
  hipLaunchKernelGGL(k_part_scatter, dim3((unsigned)items.size()), dim3(kSortThreads), 0,
                     c.stream, ditems, djobs, folded, cc_offsets, part_start, records);
  rt::check_launch();
  hipLaunchKernelGGL(k_part_sort, dim3(P), dim3(kSortThreads), 0, c.stream, djobs, part_start,
                     part_count, records, sorted, starts, counts, owner);
  rt::check_launch();
  mark(2);
  hipLaunchKernelGGL(k_task_counts, dim3(grid_for(G)), dim3(256), 0, c.stream, counts, G, cap,
                     ntask);
  rt::check_launch();
  exclusive_scan(c, ntask, task_off, G, totals + 1, keep);
  const size_t max_tasks = m_total_ / cap + G + 1;
  rt::BufferRef tasks_buf = c.allocate(max_tasks * sizeof(Task));
  Task* tasks = static_cast<Task*>(tasks_buf->ptr);
  hipLaunchKernelGGL(k_build_tasks, dim3(grid_for(max_tasks)), dim3(256), 0, c.stream, counts,
                     starts, owner, task_off, G, cap, totals + 1, tasks);
  rt::check_launch();
  rt::BufferRef partials_buf = c.allocate(max_tasks * sizeof(Xyzz29));
  Xyzz29* partials = static_cast<Xyzz29*>(partials_buf->ptr);
  mark(3);
  hipLaunchKernelGGL(k_accumulate, dim3((unsigned)((max_tasks + 255) / 256)), dim3(256), 0,
                     c.stream, tasks, totals + 1, sorted, djobs, partials);
  rt::check_launch();
  rt::BufferRef sums_buf = c.allocate((size_t)G * sizeof(Xyzz29));
  Xyzz29* sums = static_cast<Xyzz29*>(sums_buf->ptr);
  mark(4);
  // lanes per bucket: enough groups for ~2 wavefronts per SIMD, and no more
  // lanes than about half the expected partials per bucket
  if (sum_threads_total >= (size_t(1) << 32)) rt::fail_params();
  rt::BufferRef heavy_buf = c.allocate(((size_t)G + 1) * sizeof(uint32_t));
  uint32_t* heavy_count = static_cast<uint32_t*>(heavy_buf->ptr);
  uint32_t* heavy = heavy_count + 1;
  zero_u32(c, heavy_count, 1);
  // pairs of threads per lane where the launch would leave SIMDs idle (at most
  // two wavefronts per SIMD when doubled) and every lane group fits a
  // wavefront as pairs
  bool pair = 2 * sum_threads_total <= kPairLanes;
  for (uint32_t j = 0; j < njobs; ++j) pair = pair && jobs[j].lanes_log2 <= 5;
  bool quad = pair && quad_lanes() && 4 * sum_threads_total <= kPairLanes;
  for (uint32_t j = 0; j < njobs; ++j) quad = quad && jobs[j].lanes_log2 <= 4;
  auto launch_sums = [&](auto kernel, size_t threads_per_lane) {
    hipLaunchKernelGGL(kernel, dim3((unsigned)((threads_per_lane * sum_threads_total + 255) / 256)),
                       dim3(256), 0, c.stream, task_off, ntask, partials, djobs, njobs,
                       (uint32_t)sum_threads_total, sums, heavy, heavy_count);
  };
  if (quad)
    launch_sums(k_bucket_sums<4>, 4);
  else if (pair)
    launch_sums(k_bucket_sums<2>, 2);
  else
    launch_sums(k_bucket_sums<1>, 1);
  rt::check_launch();
  rt::BufferRef heavy_parts = c.allocate((size_t)kHeavySplitSlots * kHeavySplit * sizeof(Xyzz29));
  Xyzz29* block_sums = static_cast<Xyzz29*>(heavy_parts->ptr);
  hipLaunchKernelGGL(k_heavy_partial, dim3(1024), dim3(kHeavyThreads), 0, c.stream, task_off, ntask,
                     partials, heavy, heavy_count, block_sums);
  rt::check_launch();
  hipLaunchKernelGGL(k_heavy_final, dim3(256), dim3(64), 0, c.stream, ntask, heavy, heavy_count,
                     block_sums, sums);
  rt::check_launch();
  hipLaunchKernelGGL(k_heavy_sums, dim3(256), dim3(kHeavyThreads), 0, c.stream, task_off, ntask,
                     partials, heavy, heavy_count, sums);
  rt::check_launch();
  keep.push_back(std::move(heavy_parts));
  // level 1 of the weighted bucket reduction: (A_k, R_k) per red_group() buckets
  const unsigned rg = red_group();
  const size_t groups = G / rg;
  mark(5);
  // two lanes per bucket while the doubled launch stays within two wavefronts
  // per SIMD (the one-lane launch leaves SIMDs idle or is latency-bound)
  if (rg == kRedSmall && quad_lanes())
    hipLaunchKernelGGL((k_reduce_buckets_pair<kRedSmall, 4>), dim3((unsigned)groups),
                       dim3(4 * kRedSmall), 0, c.stream, sums, G, result, result + groups);
  else if (rg == kRedSmall)
    hipLaunchKernelGGL((k_reduce_buckets_pair<kRedSmall, 2>), dim3((unsigned)groups),
                       dim3(2 * kRedSmall), 0, c.stream, sums, G, result, result + groups);
  else if (G <= kPairReduceMax && quad_lanes())
    hipLaunchKernelGGL((k_reduce_buckets_two<kRed1, 4>), dim3((unsigned)groups), dim3(2 * kRed1), 0,
                       c.stream, sums, G, result, result + groups);
  else if (G <= kPairReduceMax)
    hipLaunchKernelGGL((k_reduce_buckets_two<kRed1, 2>), dim3((unsigned)groups), dim3(kRed1), 0,
                       c.stream, sums, G, result, result + groups);
  else
    hipLaunchKernelGGL(k_reduce_buckets<kRed1>, dim3((unsigned)groups), dim3(kRed1), 0, c.stream,
                       sums, G, result, result + groups);
  rt::check_launch();
  if (ev) {
    mark(6);
    if (ev->entries_out)
      rt::hip_check(hipMemcpyAsync(ev->entries_out, totals, sizeof(uint32_t),
                                   hipMemcpyDeviceToDevice, c.stream));
  }
  if (hold) {
    for (rt::BufferRef* b : {&table, &sorted_buf, &meta, &tasks_buf, &partials_buf, &sums_buf, &heavy_buf})
      hold->push_back(*b);
    hold->insert(hold->end(), keep.begin(), keep.end());
  }
}

// Host end of every MSM of the batch: MSM j is sum_k A_k + g sum_k k R_k (g the group size)
// over its level-1 groups k.  The groups are cut into chunks of kFinishChunk
// (c covers [c L, c L + L)) and combined per MSM:
//   sum_k k R_k = sum_c (W_c - S_c) + L sum_c c S_c,
// with S_c = sum R_k and W_c = sum (k - c L + 1) R_k over chunk c.  Every
// step is a chain of dependent additions (~1 us each on one core), so the
// chains run side by side on the worker pool: per chunk, the A sum and the
// (S, W) running sums; per MSM, sum_c A_c, sum_c (W_c - S_c) and
// sum_c c S_c.
namespace {
constexpr size_t kFinishChunk = 16;

struct ChunkSums {
  host::Xyzz64 a{}, s{}, w{};  // zero ZZ: the identity
};

void chunk_a(const G1Xyzz* A, size_t lo, size_t hi, ChunkSums& c) {
  for (size_t k = lo; k < hi; ++k) c.a = host::add(c.a, host::from_device(A[k]));
}

void chunk_sw(const G1Xyzz* R, size_t lo, size_t hi, ChunkSums& c) {
  for (size_t k = hi; k-- > lo;) {
    c.s = host::add(c.s, host::from_device(R[k]));
    c.w = host::add(c.w, c.s);
  }
}

// The three chains of one MSM's combine: 0 sum A_c, 1 sum (W_c - S_c),
// 2 L sum_c c S_c.
host::Xyzz64 combine_part(const ChunkSums* cs, size_t n, int part) {
  host::Xyzz64 acc{}, run{};
  if (part == 0) {
    for (size_t c = 0; c < n; ++c) acc = host::add(acc, cs[c].a);
  } else if (part == 1) {
    for (size_t c = 0; c < n; ++c) acc = host::add(acc, host::add(cs[c].w, host::neg(cs[c].s)));
  } else {
    for (size_t c = n; c-- > 1;) {
      run = host::add(run, cs[c].s);
      acc = host::add(acc, run);
    }
    for (size_t m = 1; m < kFinishChunk; m <<= 1) acc = host::dbl(acc);
  }
  return acc;
}
}  // namespace

void Batch::finish_packets(const G1Xyzz* slots, uint8_t* out97) const {
  const unsigned rg = red_group();
  const size_t all = gbuckets_ / rg, njobs = entries_.size();
  struct Item {
    size_t job, lo, hi;
  };
  std::vector<Item> items;
  std::vector<size_t> first(njobs + 1);
  for (size_t j = 0; j < njobs; ++j) {
    first[j] = items.size();
    const size_t groups = (entries_[j].layer->plan.buckets + kRed1 - 1) / kRed1 * (kRed1 / rg);
    for (size_t lo = 0; lo < groups; lo += kFinishChunk)
      items.push_back({j, lo, std::min(groups, lo + kFinishChunk)});
  }
  first[njobs] = items.size();
  std::vector<ChunkSums> sums(items.size());
  // (S, W) chains first: they are twice as long as the A chains
  rt::parallel_for(2 * items.size(), [&](size_t t) {
    const Item& it = items[t % items.size()];
    const size_t g0 = entries_[it.job].bucket_base / rg;
    if (t < items.size())
      chunk_sw(slots + all + g0, it.lo, it.hi, sums[t]);
    else
      chunk_a(slots + g0, it.lo, it.hi, sums[t - items.size()]);
  });
  std::vector<host::Xyzz64> parts(3 * njobs);
  rt::parallel_for(3 * njobs, [&](size_t t) {
    const size_t j = t / 3;
    parts[t] = combine_part(sums.data() + first[j], first[j + 1] - first[j], (int)(t % 3));
  });
  rt::parallel_for(njobs, [&](size_t j) {
    host::Xyzz64 acc = host::add(parts[3 * j + 1], parts[3 * j + 2]);
    for (unsigned m = 1; m < rg; m <<= 1) acc = host::dbl(acc);
    host::to_packet(host::add(parts[3 * j], acc), out97 + rt::kG1Bytes * j);
  });
}

void affine_to_packet(const G1Affine& a, uint8_t* out) {
  std::memset(out, 0, rt::kG1Bytes);
  if (a.is_identity()) {
    out[0] = 1;
    return;
  }
  const Fq x = from_mont(a.x), y = from_mont(a.y);
  std::memcpy(out + 1, x.v, 48);
  std::memcpy(out + 49, y.v, 48);
}

void to_packet(const G1Xyzz& p, uint8_t* out) { host::to_packet(host::from_device(p), out); }

void download_layer(rt::Context& c, const Layer& layer, uint8_t* packets) {
  const size_t n = layer.n;
  rt::BufferRef tmp = layer_points(c, layer);
  hipLaunchKernelGGL(k_points_from_mont, dim3(grid_for(n)), dim3(256), 0, c.stream,
                     static_cast<const G1Affine*>(tmp->ptr), static_cast<G1Affine*>(tmp->ptr), n);
  rt::check_launch();
  std::vector<uint8_t> raw(n * sizeof(G1Affine));
  rt::download_bytes(c, tmp->ptr, raw.size(), raw.data());
  for (size_t i = 0; i < n; ++i) {
    uint8_t* out = packets + rt::kG1Bytes * i;
    const uint8_t* q = raw.data() + sizeof(G1Affine) * i;
    bool zero = std::all_of(q, q + 96, [](uint8_t v) { return v == 0; });
    out[0] = zero ? 1 : 0;
    if (zero)
      std::memset(out + 1, 0, 96);
    else
      std::memcpy(out + 1, q, 96);
  }
}

bool upload_layer_compressed(rt::Context& c, Layer& layer, const uint8_t* compressed, size_t n,
                             uint32_t override_c) {
  constexpr size_t kBytes = 48;
  rt::BufferRef packed = c.allocate(std::max<size_t>(n, 1) * kBytes);
  const size_t chunk = (size_t(32) << 20) / kBytes;
  for (size_t done = 0; done < n; done += chunk) {
    const size_t cnt = std::min(chunk, n - done);
    uint8_t* stage = c.staging(cnt * kBytes);
    std::memcpy(stage, compressed + kBytes * done, cnt * kBytes);
    rt::hip_check(hipMemcpyAsync(static_cast<uint8_t*>(packed->ptr) + kBytes * done, stage,
                                 cnt * kBytes, hipMemcpyHostToDevice, c.stream));
    c.staging_used();
  }
  rt::BufferRef raw = c.allocate(std::max<size_t>(n, 1) * sizeof(G1Affine));
  rt::BufferRef flag = c.allocate(sizeof(unsigned));
  zero_u32(c, static_cast<uint32_t*>(flag->ptr), 1);
  hipLaunchKernelGGL(k_decompress, dim3(grid_for(n)), dim3(256), 0, c.stream,
                     static_cast<const uint4*>(packed->ptr), n, static_cast<G1Affine*>(raw->ptr),
                     static_cast<unsigned*>(flag->ptr));
  rt::check_launch();
  unsigned bad = 0;
  rt::download_bytes(c, flag->ptr, sizeof(unsigned), reinterpret_cast<uint8_t*>(&bad));
  if (bad) return false;
  prepare_layer_from_device(c, layer, static_cast<const G1Affine*>(raw->ptr), n, override_c);
  return true;
}

void download_layer_compressed(rt::Context& c, const Layer& layer, uint8_t* compressed) {
  const size_t n = layer.n;
  rt::BufferRef pts = layer_points(c, layer);
  rt::BufferRef tmp = c.allocate(std::max<size_t>(n, 1) * 48);
  hipLaunchKernelGGL(k_compress, dim3(grid_for(n)), dim3(256), 0, c.stream,
                     static_cast<const G1Affine*>(pts->ptr), n, static_cast<uint4*>(tmp->ptr));
  rt::check_launch();
  rt::download_bytes(c, tmp->ptr, n * 48, compressed);
}

rt::BufferRef make_fixed_table(rt::Context& c) {
  // standard generator (canonical) -> Montgomery on the host
  static const uint32_t gx[12] = {0xdb22c6bbu, 0xfb3af00au, 0xf97a1aefu, 0x6c55e83fu,
                                  0x171bac58u, 0xa14e3a3fu, 0x9774b905u, 0xc3688c4fu,
                                  0x4fa9ac0fu, 0x2695638cu, 0x3197d794u, 0x17f1d3a7u};
  static const uint32_t gy[12] = {0x46c5e7e1u, 0x0caa2329u, 0xa2888ae4u, 0xd03cc744u,
                                  0x2c04b3edu, 0x00db18cbu, 0xd5d00af6u, 0xfcf5e095u,
                                  0x741d8ae4u, 0xa09e30edu, 0xe3aaa0f1u, 0x08b3f481u};
  G1Affine g;
  std::memcpy(g.x.v, gx, 48);
  std::memcpy(g.y.v, gy, 48);
  g.x = to_mont(g.x);
  g.y = to_mont(g.y);
  rt::BufferRef table = c.allocate(32 * 256 * sizeof(G1Affine));
  hipLaunchKernelGGL(k_fixed_table, dim3(32), dim3(256), 0, c.stream, g,
                     static_cast<G1Affine*>(table->ptr));
  rt::check_launch();
  return table;
}

void generate_points(rt::Context& c, const Fr* scalars, size_t n, G1Affine* out,
                     const G1Affine* table) {
  const size_t threads = std::max<size_t>(1, std::min<size_t>(n, 65536));
  rt::BufferRef scratch = c.allocate(std::max<size_t>(n, 1) * sizeof(G1Xyzz));
  rt::BufferRef prefix = c.allocate(std::max<size_t>(n, 1) * sizeof(Fq));
  hipLaunchKernelGGL(k_fixed_base, dim3(grid_for(threads, 64)), dim3(64), 0, c.stream, scalars, n,
                     table, out, static_cast<G1Xyzz*>(scratch->ptr),
                     static_cast<Fq*>(prefix->ptr));
  rt::check_launch();
}

namespace {

// Median device time of `reps` single-MSM batches (after one warm-up).
float time_msm(rt::Context& c, const Layer& layer, const Fr* scalars, uint32_t cap, int reps) {
  Batch b;
  b.add(layer, scalars);
  b.set_cap(cap);
  rt::BufferRef res = c.allocate(std::max<size_t>(b.result_slots(), 1) * sizeof(G1Xyzz));
  // two untimed runs, then the timed ones back to back (no host wait between
  // them, so the clock stays at its loaded level); median of the reps
  std::vector<hipEvent_t> ev(2 * (size_t)reps);
  for (auto& e : ev) rt::hip_check(hipEventCreate(&e));
  b.enqueue(c, static_cast<G1Xyzz*>(res->ptr));
  b.enqueue(c, static_cast<G1Xyzz*>(res->ptr));
  for (int r = 0; r < reps; ++r) {
    rt::hip_check(hipEventRecord(ev[2 * r], c.stream));
    b.enqueue(c, static_cast<G1Xyzz*>(res->ptr));
    rt::hip_check(hipEventRecord(ev[2 * r + 1], c.stream));
  }
  rt::hip_check(hipEventSynchronize(ev.back()));
  std::vector<float> t;
  for (int r = 0; r < reps; ++r) {
    float ms = 0;
    rt::hip_check(hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]));
    t.push_back(ms);
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}

// canonical uniform Fr values from a SplitMix64 stream (rejection of 255-bit draws)
std::vector<uint8_t> random_fr(size_t n, uint64_t seed) {
  static const uint64_t r[4] = {0xffffffff00000001ULL, 0x53bda402fffe5bfeULL, 0x3339d80809a1d805ULL,
                                0x73eda753299d7d48ULL};
  std::vector<uint8_t> out(n * rt::kFrBytes);
  uint64_t s = seed;
  auto next = [&] {
    s += 0x9E3779B97F4A7C15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  };
  for (size_t i = 0; i < n;) {
    uint64_t w[4];
    for (int k = 0; k < 4; ++k) w[k] = next();
    w[3] &= 0x7FFFFFFFFFFFFFFFULL;
    bool less = false;
    for (int k = 3; k >= 0; --k)
      if (w[k] != r[k]) {
        less = w[k] < r[k];
        break;
      }
    if (!less) continue;
    std::memcpy(out.data() + i * rt::kFrBytes, w, rt::kFrBytes);
    ++i;
  }
  return out;
}

}  // namespace

void autotune(rt::Context& c, size_t log_n, uint32_t* best_c, uint32_t* best_cap, float* best_ms) {
  const size_t n = size_t(1) << log_n;
  rt::BufferRef table = make_fixed_table(c);
  const auto ks = random_fr(n, 0x7475e3ULL + log_n);
  rt::BufferRef kdev = rt::upload_fr_table(c, ks.data(), n, log_n);
  rt::BufferRef pts = c.allocate(n * sizeof(G1Affine));
  generate_points(c, static_cast<const Fr*>(kdev->ptr), n, static_cast<G1Affine*>(pts->ptr),
                  static_cast<const G1Affine*>(table->ptr));
  const auto ss = random_fr(n, 0x5ca1a75ULL + log_n);
  rt::BufferRef sdev = rt::upload_fr_table(c, ss.data(), n, log_n);
  const Fr* scalars = static_cast<const Fr*>(sdev->ptr);
  const uint32_t rule = choose_plan(n, 0).c;
  std::vector<uint32_t> cands;
  for (int d = -2; d <= 2; ++d) {
    const int cc = (int)rule + d;
    if (cc >= 4 && cc <= 20) cands.push_back((uint32_t)cc);
  }
  for (uint32_t extra : {8u, 16u})
    if (extra + 3 >= rule && extra <= rule + 3 &&
        std::find(cands.begin(), cands.end(), extra) == cands.end())
      cands.push_back(extra);
  // A candidate replaces the built-in choice only when it is faster by more
  // than kMargin: near ties (within the run-to-run spread) keep the rule, so
  // the plan does not flip between runs on timing noise.
  constexpr float kMargin = 0.985f;
  float best = 1e30f;
  uint32_t bc = rule;
  Layer best_layer;
  float rule_ms = 1e30f;
  for (uint32_t cc : cands) {
    Layer layer;
    prepare_layer_from_device(c, layer, static_cast<const G1Affine*>(pts->ptr), n, cc);
    const float ms = time_msm(c, layer, scalars, default_cap(n), 3);
    if (cc == rule) rule_ms = ms;
    if (ms < best) {
      best = ms;
      bc = cc;
      best_layer = std::move(layer);
    }
  }
  if (bc != rule && best > rule_ms * kMargin) {
    bc = rule;
    best = rule_ms;
    best_layer = Layer();
    prepare_layer_from_device(c, best_layer, static_cast<const G1Affine*>(pts->ptr), n, rule);
  }
  // caps: two passes (the second in reverse order), each cap's better median
  const uint32_t caps[] = {2u, 4u, 8u, 16u, 32u, 64u};
  constexpr int kCaps = 6;
  float cap_ms[kCaps];
  for (int i = 0; i < kCaps; ++i) cap_ms[i] = time_msm(c, best_layer, scalars, caps[i], 5);
  for (int i = kCaps - 1; i >= 0; --i)
    cap_ms[i] = std::min(cap_ms[i], time_msm(c, best_layer, scalars, caps[i], 5));
  uint32_t bcap = 0;
  float best2 = 1e30f, def_ms = 1e30f;
  const uint32_t def_cap = default_cap(n);
  for (int i = 0; i < kCaps; ++i) {
    if (caps[i] == def_cap) def_ms = cap_ms[i];
    if (cap_ms[i] < best2) {
      best2 = cap_ms[i];
      bcap = caps[i];
    }
  }
  if (bcap != def_cap && def_ms < 1e29f && best2 > def_ms * kMargin) {
    bcap = def_cap;
    best2 = def_ms;
  }
  c.tuned_window[log_n] = static_cast<uint8_t>(bc);
  c.tuned_cap[log_n] = static_cast<uint8_t>(bcap);
  *best_c = bc;
  *best_cap = bcap;
  *best_ms = std::min(best, best2);
}

}  // namespace msm
}  // namespace hp
