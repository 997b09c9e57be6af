// SumCheck prover state (SumCheckProver, backend/src/sum_check.rs:17-38).
//
// Round k of a state over n variables and maximum degree d:
//   k > 0: fold every distinct factor with the previous challenge c,
//          new[b] = old[2b] + c (old[2b+1] - old[2b])  (cpu/sum_check.rs:103-118);
//   then   g(t) = sum_p coeff_p sum_b prod_{f in p} (f[2b] + t (f[2b+1] - f[2b]))
//          for t = 0..d, evaluated directly at every t (the CPU backend
//          evaluates t <= arity and extrapolates, cpu/sum_check.rs:125-177:
//          the same polynomial values).
// Tables are HBM-resident Montgomery arrays.  A round is one launch
// (k_round_jobs) that folds with the previous challenge, evaluates the message
// with per-block partial sums and lets the last block write the canonical
// message; rounds of many states share launches (round_batch).  Products are
// (coefficient, factor slots) with repeats kept; the coefficient is applied to
// the first factor's line once per pair, not once per t.  Degrees above 8 take
// a generic path (separate fold, message in chunks of 8 values, reduction).
//
// Validation and statuses follow the reference bridge
// (backend/src/ascend910b3/native/bridge_sumcheck.cpp:210-331 create,
// :519-550 round).
#include <algorithm>
#include <chrono>
#include <memory>
#include <map>
#include <thread>
#include <deque>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "field29.hpp"
#include "kernels_common.hpp"

namespace hp {
namespace sc {
namespace {

constexpr unsigned kThreads = 256;
constexpr unsigned kMaxBlocks = 1024;

// Warp-level sum of an Fr across the 64 lanes of a wavefront.
__device__ __forceinline__ Fr wave_sum(Fr v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) {
    Fr o;
#pragma unroll
    for (int i = 0; i < 8; ++i) o.v[i] = __shfl_xor(v.v[i], offset, 64);
    v = v + o;
  }
  return v;
}

// Sum acc[t] over the block; thread 0 writes partial[block * stride + t].
template <int T>
__device__ void block_reduce_store(Fr (&acc)[T], int count, Fr* partial, size_t stride) {
  __shared__ Fr lds[kThreads / 64][T];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < T; ++t) {
    if (t < count) {
      Fr s = wave_sum(acc[t]);
      if (lane == 0) lds[wave][t] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)count) {
    const int t = threadIdx.x;
    Fr s = Fr::zero();
#pragma unroll
    for (int w = 0; w < (int)(kThreads / 64); ++w) s = s + lds[w][t];
    store_fr(partial, blockIdx.x * stride + t, s);
  }
}

// new[f][b] = old[f][2b] + c (old[f][2b+1] - old[f][2b]) for every factor f.
__global__ void k_fold_factors(const Fr* const* src, Fr* const* dst, size_t factors, Fr c,
                               size_t half) {
  const size_t total = factors * half;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t f = i / half, b = i % half;
    const Fr* s = src[f];
    const Fr lo = load_fr(s, 2 * b), hi = load_fr(s, 2 * b + 1);
    store_fr(dst[f], b, lo + c * (hi - lo));
  }
}

struct Products {
  const Fr* coeffs;         // product_count, Montgomery
  const uint32_t* offsets;  // product_count + 1
  const uint32_t* slots;    // flattened distinct-factor slots
  uint32_t count;
};

// ---- one round for a list of states (jobs) in one launch -----------------
//
// A job is one state's round.  Its blocks are a contiguous range of the grid
// (block_begin, blocks); the last of them to finish (per-job ticket) adds the
// job's block partials and writes the canonical message to `out`.
// Two work decompositions:
//   pairs    thread per pair b: fold every distinct factor on the pair (when
//            folding), then evaluate every product on it; coalesced streaming,
//            used while the tables are large;
//   items    thread per (product p, pair b), product-major so a wavefront
//            shares p: the thread folds only the factors of p on the fly (the
//            first occurrence of each factor writes the folded pair) and
//            evaluates p; many short chains instead of one long one, used once
//            the tables are small and latency dominates.
struct RoundJob {
  const Fr* const* src;    // tables read (inputs or the previous fold slot)
  Fr* const* dst;          // fold destination tables (folding rounds)
  const Fr* coeffs;        // product coefficients, Montgomery
  const Fr* coeffs29;      // the same times 2^(5 m_p) for a product of m_p factors
  const uint32_t* offsets; // product_count + 1
  const uint32_t* slots;   // flattened factor slots
  const uint8_t* owner;    // per flattened entry: 1 = writes its factor's fold
  Fr* partial;             // blocks x (degree + 1)
  unsigned* ticket;
  Fr* out;                 // degree + 1 canonical values
  Fr challenge;            // Montgomery
  uint32_t pairs, products, factors, degree, fold, items_mode, block_begin, blocks;
  // product classes (mode 3): per class {m, output offset, member begin, member
  // end}, then the member product indices; class coefficients (Montgomery)
  // times 2^(5 (m - 1))
  const uint32_t* classes;
  const Fr* class_k;
  uint32_t class_count, class_width;
  uint32_t entries;  // flattened factor entries (offsets[products])
  // class finish: the (class, t) sums as stamped words (kMaxClassWidth x 8,
  // zeroed when the state is created) and this round's stamp (round + 1)
  uint64_t* class_sums;
  uint32_t class_stamp;
};

constexpr int kJobsPerLaunch = 16;
struct JobPack {
  RoundJob job[kJobsPerLaunch];
  uint32_t count;
  // 1: multi-block jobs only store their block partials and k_finish_split,
  // launched next, adds them; 0: the last block of each job (a ticket) adds them
  uint32_t split;
};

// term[t] = coeff * prod_k (lo_k + t d_k) for t = 0..deg, streaming over the
// factors: the first factor's line is scaled by the coefficient (2 products),
// every further factor costs deg + 1 products and its line values are reached
// by repeated addition of d_k.
//
// Fr29 arithmetic on the stored form (A = a 2^256 read as an integer): each
// Fr29 product divides by 2^261, so a product of m factors starts from the
// coefficient times 2^(256 + 5 m) (coeffs29) and ends in the stored form.
// Bounds (r = the modulus; lo, hi canonical, t <= 8): d < 2r, lines < 17r, the
// first factor's terms < 9.3r, every later product < 3.3r (R / r = 70.7).
using f29::Fr29;

// Every job of a launch has degree D exactly (launches group jobs by degree),
// so the point count is a compile-time constant.
template <int D>
struct ProductTerms {
  Fr29 term[D + 1];
  __device__ __forceinline__ void first(const Fr29& coeff, const Fr29& lo, const Fr29& hi) {
    Fr29 line = f29::mul(coeff, lo);
    const Fr29 d = f29::mul(coeff, f29::sub1p(hi, lo));
#pragma unroll
    for (int t = 0; t <= D; ++t) {
      term[t] = line;
      if (t < D) line = f29::add(line, d);
    }
  }
  __device__ __forceinline__ void times(const Fr29& lo, const Fr29& hi) {
    Fr29 line = lo;
    const Fr29 d = f29::sub1p(hi, lo);
#pragma unroll
    for (int t = 0; t <= D; ++t) {
      term[t] = f29::mul(term[t], line);
      if (t < D) line = f29::add(line, d);
    }
  }
};

// lo + c (hi - lo) for canonical lo, hi and c29 = c 2^261: canonical result
__device__ __forceinline__ Fr29 fold29(const Fr29& c29, const Fr29& lo, const Fr29& hi) {
  const Fr29 f = f29::add(lo, f29::mul(c29, f29::sub1p(hi, lo)));  // < 2.03 r
  return f29::cond_sub(f29::cond_sub(f, f29::FrP::P), f29::FrP::P);
}

__device__ __forceinline__ Fr29 load29(const Fr* base, size_t i) {
  return f29::unpack_fr(load_fr(base, i));
}

// Between reductions an accumulator holds a reduced value (< 3 r) plus up to
// four product terms; a product of one factor ends at its first factor's
// terms (< 9.3 r), so the sum stays below 3 r + 4 x 9.3 r < 41 r < 2^261, the
// input bound of reduce_est.
template <int D>
__device__ __forceinline__ void reduce_acc(Fr29 (&acc)[D + 1]) {
#pragma unroll
  for (int t = 0; t <= D; ++t) acc[t] = f29::reduce_est(acc[t]);
}

template <int D>
__device__ __forceinline__ void to_stored(const Fr29 (&acc)[D + 1], Fr (&out)[D + 1]) {
#pragma unroll
  for (int t = 0; t <= D; ++t) out[t] = f29::pack_fr(acc[t]);
}

__device__ __forceinline__ uint32_t job_of_block(const JobPack& pack) {
  uint32_t j = 0;
  while (j + 1 < pack.count && blockIdx.x >= pack.job[j + 1].block_begin) ++j;
  return j;
}

// Block partials of a job, then the last block of the job adds them and
// writes the canonical message.
template <int D>
__device__ void finish_job(const RoundJob& job, uint32_t lblock, Fr (&acc)[D + 1], bool split) {
  constexpr uint32_t deg = D;
  if (job.blocks == 1) {
    // One block holds the whole message: no block partials, no ticket and no
    // device-scope fences (on this multi-XCD part each one writes back the
    // L2, several microseconds apiece, which dominated the small rounds).
    // Thread t reads back the value it wrote itself.
    __shared__ __align__(16) Fr msg[D + 1];
    block_reduce_store<D + 1>(acc, deg + 1, msg, 0);
    if (threadIdx.x <= deg) {
      store_fr(job.out, threadIdx.x, from_mont(msg[threadIdx.x]));
      __threadfence_system();  // `out` may be pinned host memory read after the sync
    }
    return;
  }
  block_reduce_store<D + 1>(acc, deg + 1, job.partial + (size_t)lblock * (deg + 1), 0);
  // the kernel boundary orders the partials before k_finish_split reads them,
  // with no device-scope fence or atomic here
  if (split) return;
  __shared__ unsigned last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(job.ticket, 1u) == job.blocks - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  // every thread of the last block sums a strided share of the block partials
  // of one point t (groups of deg + 1 threads), then one thread per t adds the
  // groups: independent loads instead of one chain over all blocks
  const uint32_t width = deg + 1, groups = kThreads / width;
  const uint32_t t = threadIdx.x % width, g = threadIdx.x / width;
  Fr s = Fr::zero();
  if (g < groups)
    for (uint32_t blk = g; blk < job.blocks; blk += groups)
      s = s + load_fr(job.partial, (size_t)blk * width + t);
  __shared__ Fr red[kThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t hop = 1; hop < groups; hop <<= 1) {
    if (g < groups && (g & (2 * hop - 1)) == 0 && g + hop < groups)
      red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + hop * width];
    __syncthreads();
  }
  if (threadIdx.x < width) {
    store_fr(job.out, threadIdx.x, from_mont(red[threadIdx.x]));
    __threadfence_system();  // `out` may be pinned host memory read after the sync
  }
  if (threadIdx.x == 0) *job.ticket = 0;
}

// Second phase of a split launch: block j adds the block partials of job j of
// the pack (strided shares of each point t, then one thread per t) and writes
// its canonical message.  Single-block jobs finished in the round kernel.
__device__ __forceinline__ void finish_split_body(const JobPack& pack) {
  const RoundJob& job = pack.job[blockIdx.x];
  if (job.blocks <= 1) return;
  const uint32_t width = job.degree + 1, groups = kThreads / width;
  const uint32_t t = threadIdx.x % width, g = threadIdx.x / width;
  Fr s = Fr::zero();
  if (g < groups)
    for (uint32_t blk = g; blk < job.blocks; blk += groups)
      s = s + load_fr(job.partial, (size_t)blk * width + t);
  __shared__ Fr red[kThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  // tree over the groups (log2 of ~40-130 dependent additions instead of a chain)
  for (uint32_t hop = 1; hop < groups; hop <<= 1) {
    if (g < groups && (g & (2 * hop - 1)) == 0 && g + hop < groups)
      red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + hop * width];
    __syncthreads();
  }
  if (threadIdx.x < width) {
    store_fr(job.out, threadIdx.x, from_mont(red[threadIdx.x]));
    __threadfence_system();  // `out` may be pinned host memory read after the sync
  }
}
__global__ void __launch_bounds__(kThreads) k_finish_split(const JobPack pack) { finish_split_body(pack); }
// the same kernel over a pack in device memory (replayed round graphs)
__global__ void __launch_bounds__(kThreads) k_finish_split_dev(const JobPack* pack) { finish_split_body(*pack); }

// Thread per pair (grid-strided): fold every distinct factor on the pair, then
// every product.
template <int D>
__device__ __forceinline__ void round_pairs_body(const JobPack& pack) {
  const RoundJob& job = pack.job[job_of_block(pack)];
  const uint32_t lblock = blockIdx.x - job.block_begin;
  const size_t stride = (size_t)job.blocks * kThreads;
  Fr29 acc[D + 1];
#pragma unroll
  for (int t = 0; t <= D; ++t) acc[t] = f29::zero_t<f29::FrP>();
  const Fr29 c29 = f29::prescale_fr(job.challenge);
  const Fr* const* tbl = job.fold ? job.dst : job.src;
  for (size_t b = (size_t)lblock * kThreads + threadIdx.x; b < job.pairs; b += stride) {
    if (job.fold) {
      for (uint32_t f = 0; f < job.factors; ++f) {
        const Fr* o = job.src[f];
        const Fr29 a0 = load29(o, 4 * b), a1 = load29(o, 4 * b + 1);
        const Fr29 a2 = load29(o, 4 * b + 2), a3 = load29(o, 4 * b + 3);
        store_fr(job.dst[f], 2 * b, f29::pack_fr_canonical(fold29(c29, a0, a1)));
        store_fr(job.dst[f], 2 * b + 1, f29::pack_fr_canonical(fold29(c29, a2, a3)));
      }
    }
    for (uint32_t p = 0; p < job.products; ++p) {
      const uint32_t begin = job.offsets[p], end = job.offsets[p + 1];
      ProductTerms<D> pt;
      const Fr* t0 = tbl[job.slots[begin]];
      pt.first(load29(job.coeffs29, p), load29(t0, 2 * b), load29(t0, 2 * b + 1));
      for (uint32_t e = begin + 1; e < end; ++e) {
        const Fr* t = tbl[job.slots[e]];
        pt.times(load29(t, 2 * b), load29(t, 2 * b + 1));
      }
#pragma unroll
      for (int t = 0; t <= D; ++t) acc[t] = f29::add(acc[t], pt.term[t]);
      if ((p & 3) == 3) reduce_acc<D>(acc);
    }
    reduce_acc<D>(acc);
  }
  Fr out[D + 1];
  to_stored<D>(acc, out);
  finish_job<D>(job, lblock, out, pack.split != 0);
}
template <int D>
__global__ void __launch_bounds__(kThreads) k_round_pairs(const JobPack pack) { round_pairs_body<D>(pack); }
// the same kernel over a pack in device memory (replayed round graphs)
template <int D>
__global__ void __launch_bounds__(kThreads) k_round_pairs_dev(const JobPack* pack) { round_pairs_body<D>(*pack); }

// ---- class-major pairs decomposition (large tables) -------------------------
//
// Products of equal arity m and equal coefficient form a class.  A product of
// m lines is a polynomial of degree m in t, so a class's products are
// evaluated without the coefficient at t = 0..m only: every factor costs m + 1
// products instead of d + 1, and the two coefficient products per product and
// pair disappear.  Each thread sums its pairs per (class, t), the block adds
// the threads' sums, and k_finish_classes adds the blocks, multiplies each
// class's m + 1 sums by its coefficient, extends them to t = 0..d by finite
// differences (exact for a polynomial of degree m) and adds the classes.
constexpr uint32_t kMaxClasses = 16, kMaxClassWidth = 64;

// Sum over the thread's pairs of the class's products at t = 0..M, below 3r.
// Fr29 on the stored form: a product of m stored values (a 2^256) through
// m - 1 Fr29 products is prod(a) 2^(256 - 5 (m - 1)); the class coefficient
// carries the missing 2^(5 (m - 1)).
// Bounds (r the modulus): lo, hi canonical, d < 2r, lines < (2M + 1) r <= 17r,
// a one-factor term < 3r, a product term < 5.1r (17^2 / 70.7 + 1); eight
// terms on a reduced accumulator (< 3r) stay below 44r < 2^261.
template <int M>
__device__ __forceinline__ void class_pass_plain(const RoundJob& job, const Fr* const* tbl, uint32_t mb,
                                           uint32_t me, size_t b0, size_t stride,
                                           Fr29 (&out)[M + 1]) {
  Fr29 acc[M + 1];
#pragma unroll
  for (int t = 0; t <= M; ++t) acc[t] = f29::zero_t<f29::FrP>();
  const uint32_t* members = job.classes + 4 * job.class_count;
  uint32_t pending = 0;
  for (size_t b = b0; b < job.pairs; b += stride) {
    for (uint32_t q = mb; q < me; ++q) {
      const uint32_t begin = job.offsets[members[q]];
      Fr29 term[M + 1];
      {
        const Fr* t0 = tbl[job.slots[begin]];
        const Fr29 lo = load29(t0, 2 * b), hi = load29(t0, 2 * b + 1);
        const Fr29 d = f29::sub1p(hi, lo);
        Fr29 line = lo;
#pragma unroll
        for (int t = 0; t <= M; ++t) {
          term[t] = line;
          if (t < M) line = f29::add(line, d);
        }
      }
#pragma unroll 1
      for (int e = 1; e < M; ++e) {
        const Fr* tt = tbl[job.slots[begin + e]];
        const Fr29 lo = load29(tt, 2 * b), hi = load29(tt, 2 * b + 1);
        const Fr29 d = f29::sub1p(hi, lo);
        Fr29 line = lo;
#pragma unroll
        for (int t = 0; t <= M; ++t) {
          term[t] = f29::mul(term[t], line);
          if (t < M) line = f29::add(line, d);
        }
      }
#pragma unroll
      for (int t = 0; t <= M; ++t) acc[t] = f29::add(acc[t], term[t]);
      if (++pending == 8) {
#pragma unroll
        for (int t = 0; t <= M; ++t) acc[t] = f29::reduce_est(acc[t]);
        pending = 0;
      }
    }
  }
#pragma unroll
  for (int t = 0; t <= M; ++t) out[t] = f29::reduce_est(acc[t]);
}

template <int M>
__device__ __forceinline__ void class_pass_ahead(const RoundJob& job, const Fr* const* tbl, uint32_t mb,
                                           uint32_t me, size_t b0, size_t stride,
                                           Fr29 (&out)[M + 1]) {
  Fr29 acc[M + 1];
#pragma unroll
  for (int t = 0; t <= M; ++t) acc[t] = f29::zero_t<f29::FrP>();
  const uint32_t* members = job.classes + 4 * job.class_count;
  uint32_t pending = 0;
  // The thread's factor visits in order (pair b, member q, factor e), each
  // visit's two table values loaded one visit ahead: the loads of the next
  // factor are in flight while this one is multiplied.
  auto at = [&](size_t b, uint32_t q, uint32_t e) {
    return tbl[job.slots[job.offsets[members[q]] + e]] + 2 * b;
  };
  if (b0 < job.pairs && mb < me) {
    size_t b = b0;
    uint32_t q = mb, e = 0;
    Fr cur_lo = load_fr(at(b, q, 0), 0), cur_hi = load_fr(at(b, q, 0), 1);
    Fr29 term[M + 1];
    for (;;) {
      size_t nb = b;
      uint32_t nq = q, ne = e + 1;
      if (ne == M) {
        ne = 0;
        if (++nq == me) {
          nq = mb;
          nb = b + stride;
        }
      }
      const bool more = nb < job.pairs;
      Fr nxt_lo, nxt_hi;
      if (more) {
        const Fr* a = at(nb, nq, ne);
        nxt_lo = load_fr(a, 0);
        nxt_hi = load_fr(a, 1);
      }
      {
        const Fr29 lo = f29::unpack_fr(cur_lo), hi = f29::unpack_fr(cur_hi);
        const Fr29 d = f29::sub1p(hi, lo);
        Fr29 line = lo;
        if (e == 0) {
#pragma unroll
          for (int t = 0; t <= M; ++t) {
            term[t] = line;
            if (t < M) line = f29::add(line, d);
          }
        } else {
#pragma unroll
          for (int t = 0; t <= M; ++t) {
            term[t] = f29::mul(term[t], line);
            if (t < M) line = f29::add(line, d);
          }
        }
      }
      if (e + 1 == M) {
#pragma unroll
        for (int t = 0; t <= M; ++t) acc[t] = f29::add(acc[t], term[t]);
        if (++pending == 8) {
#pragma unroll
          for (int t = 0; t <= M; ++t) acc[t] = f29::reduce_est(acc[t]);
          pending = 0;
        }
      }
      if (!more) break;
      cur_lo = nxt_lo;
      cur_hi = nxt_hi;
      b = nb;
      q = nq;
      e = ne;
    }
  }
#pragma unroll
  for (int t = 0; t <= M; ++t) out[t] = f29::reduce_est(acc[t]);
}

// A class's sums over the thread's pairs.  Degrees >= 5 load each factor one
// visit ahead: that kernel then fits 228 VGPRs without spilling (it spilled
// at 256) and its rounds run 1-2 % faster; at degree 2 the look-ahead's
// registers cost a wavefront per SIMD and the rounds ran 3-12 % slower.
template <int M, bool kAhead>
__device__ __forceinline__ void class_pass(const RoundJob& job, const Fr* const* tbl, uint32_t mb,
                                           uint32_t me, size_t b0, size_t stride,
                                           Fr29 (&out)[M + 1]) {
  if constexpr (kAhead)
    class_pass_ahead<M>(job, tbl, mb, me, b0, stride, out);
  else
    class_pass_plain<M>(job, tbl, mb, me, b0, stride, out);
}

__device__ __forceinline__ Fr29 shfl_xor29(const Fr29& v, int offset) {
  Fr29 r;
#pragma unroll
  for (int i = 0; i < f29::FrP::L; ++i) r.l[i] = __shfl_xor(v.l[i], offset, 64);
  return r;
}

// Sums of W values over the 64 lanes of a wavefront, stored canonical at
// row[idx].  Reduce-scatter first: at each of the first log2(Wp) levels (Wp
// the power of two >= W) a lane keeps half of its values and adds its
// partner's copy of that half, Wp - 1 additions in all instead of W per level;
// then the lanes that hold the same value add over the remaining levels.
// Lane l ends with value idx(l) = sum_k bit_(5-k)(l) Wp / 2^(k+1); the lanes
// whose low 6 - log2(Wp) bits are zero store it.  Inputs below 3r; a
// reduction after the fourth level keeps every partial sum below 48r < 2^261.
template <int W>
__device__ __forceinline__ void wave_reduce_store(const Fr29 (&v)[W], Fr* row) {
  constexpr int Wp = W <= 1 ? 1 : W <= 2 ? 2 : W <= 4 ? 4 : W <= 8 ? 8 : 16;
  constexpr int LW = Wp == 1 ? 0 : Wp == 2 ? 1 : Wp == 4 ? 2 : Wp == 8 ? 3 : 4;
  const int lane = threadIdx.x & 63;
  Fr29 x[Wp];
#pragma unroll
  for (int i = 0; i < Wp; ++i) x[i] = i < W ? v[i] : f29::zero_t<f29::FrP>();
#pragma unroll
  for (int k = 0; k < LW; ++k) {
    const int o = 32 >> k, half = Wp >> (k + 1);
    const bool upper = (lane & o) != 0;
#pragma unroll
    for (int i = 0; i < half; ++i) {
      Fr29 send, keep;
#pragma unroll
      for (int q = 0; q < f29::FrP::L; ++q) {
        send.l[q] = upper ? x[i].l[q] : x[i + half].l[q];
        keep.l[q] = upper ? x[i + half].l[q] : x[i].l[q];
      }
      x[i] = f29::add(keep, shfl_xor29(send, o));
      if (k == 3) x[i] = f29::reduce_est(x[i]);
    }
  }
#pragma unroll
  for (int k = LW; k < 6; ++k) {
    x[0] = f29::add(x[0], shfl_xor29(x[0], 32 >> k));
    if (k == 3) x[0] = f29::reduce_est(x[0]);
  }
  int idx = 0;
#pragma unroll
  for (int k = 0; k < LW; ++k)
    if (lane & (32 >> k)) idx += Wp >> (k + 1);
  if ((lane & ((64 >> LW) - 1)) == 0 && idx < W) row[idx] = f29::pack_fr_reduced(f29::reduce_est(x[0]));
}

template <int M, bool kAhead>
__device__ __forceinline__ void class_run(const RoundJob& job, const Fr* const* tbl, uint32_t c,
                                          size_t b0, size_t stride, Fr (*lds)[kMaxClassWidth]) {
  Fr29 v[M + 1];
  class_pass<M, kAhead>(job, tbl, job.classes[4 * c + 2], job.classes[4 * c + 3], b0, stride, v);
  wave_reduce_store<M + 1>(v, &lds[threadIdx.x >> 6][job.classes[4 * c + 1]]);
}

// Thread per pair (grid-strided): fold every distinct factor on the pair
// (folding rounds), then every class over the thread's pairs; the block's sums
// per (class, t) go to job.partial[block][class offset + t].
// At most 256 registers per lane (VGPRs and AGPRs together): two waves per
// SIMD, which the degree-5 passes would otherwise lose by a few registers.
template <int D>
__device__ __forceinline__ void round_classes_body(const JobPack& pack) {
  const RoundJob& job = pack.job[job_of_block(pack)];
  const uint32_t lblock = blockIdx.x - job.block_begin;
  const size_t stride = (size_t)job.blocks * kThreads;
  const size_t b0 = (size_t)lblock * kThreads + threadIdx.x;
  const Fr* const* tbl = job.fold ? job.dst : job.src;
  constexpr bool kAhead = D >= 5;
  if (!kAhead && job.fold) {
    const Fr29 c29 = f29::prescale_fr(job.challenge);
    for (size_t b = b0; b < job.pairs; b += stride) {
      for (uint32_t f = 0; f < job.factors; ++f) {
        const Fr* o = job.src[f];
        const Fr29 a0 = load29(o, 4 * b), a1 = load29(o, 4 * b + 1);
        const Fr29 a2 = load29(o, 4 * b + 2), a3 = load29(o, 4 * b + 3);
        store_fr(job.dst[f], 2 * b, f29::pack_fr_canonical(fold29(c29, a0, a1)));
        store_fr(job.dst[f], 2 * b + 1, f29::pack_fr_canonical(fold29(c29, a2, a3)));
      }
    }
    // each thread reads back only the folded pairs it wrote itself
  }
  if (kAhead && job.fold && b0 < job.pairs && job.factors > 0) {
    const Fr29 c29 = f29::prescale_fr(job.challenge);
    // visits (pair b, factor f) in order, each visit's four values loaded one
    // visit ahead (as in class_pass)
    size_t b = b0;
    uint32_t f = 0;
    Fr cur[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) cur[i] = load_fr(job.src[0], 4 * b + i);
    for (;;) {
      size_t nb = b;
      uint32_t nf = f + 1;
      if (nf == job.factors) {
        nf = 0;
        nb = b + stride;
      }
      const bool more = nb < job.pairs;
      Fr nxt[4];
      if (more) {
        const Fr* o = job.src[nf];
#pragma unroll
        for (int i = 0; i < 4; ++i) nxt[i] = load_fr(o, 4 * nb + i);
      }
      const Fr29 a0 = f29::unpack_fr(cur[0]), a1 = f29::unpack_fr(cur[1]);
      const Fr29 a2 = f29::unpack_fr(cur[2]), a3 = f29::unpack_fr(cur[3]);
      store_fr(job.dst[f], 2 * b, f29::pack_fr_canonical(fold29(c29, a0, a1)));
      store_fr(job.dst[f], 2 * b + 1, f29::pack_fr_canonical(fold29(c29, a2, a3)));
      if (!more) break;
#pragma unroll
      for (int i = 0; i < 4; ++i) cur[i] = nxt[i];
      b = nb;
      f = nf;
    }
    // each thread reads back only the folded pairs it wrote itself
  }
  __shared__ Fr lds[kThreads / 64][kMaxClassWidth];
  for (uint32_t c = 0; c < job.class_count; ++c) {
    switch (job.classes[4 * c]) {
      case 1: class_run<1, kAhead>(job, tbl, c, b0, stride, lds); break;
      case 2: if constexpr (D >= 2) class_run<2, kAhead>(job, tbl, c, b0, stride, lds); break;
      case 3: if constexpr (D >= 3) class_run<3, kAhead>(job, tbl, c, b0, stride, lds); break;
      case 4: if constexpr (D >= 4) class_run<4, kAhead>(job, tbl, c, b0, stride, lds); break;
      case 5: if constexpr (D >= 5) class_run<5, kAhead>(job, tbl, c, b0, stride, lds); break;
      case 6: if constexpr (D >= 6) class_run<6, kAhead>(job, tbl, c, b0, stride, lds); break;
      case 7: if constexpr (D >= 7) class_run<7, kAhead>(job, tbl, c, b0, stride, lds); break;
      case 8: if constexpr (D >= 8) class_run<8, kAhead>(job, tbl, c, b0, stride, lds); break;
      default: break;
    }
  }
  __syncthreads();
  if (threadIdx.x < job.class_width) {
    Fr s = Fr::zero();
#pragma unroll
    for (int w = 0; w < (int)(kThreads / 64); ++w) s = s + lds[w][threadIdx.x];
    store_fr(job.partial, (size_t)lblock * job.class_width + threadIdx.x, s);
  }
}
template <int D>
__global__ void __launch_bounds__(kThreads, 2) k_round_classes(const JobPack pack) { round_classes_body<D>(pack); }
// the same kernel over a pack in device memory (replayed round graphs)
template <int D>
__global__ void __launch_bounds__(kThreads, 2) k_round_classes_dev(const JobPack* pack) { round_classes_body<D>(*pack); }

// Second phase of a class launch, block (i, j) per job j and (class, t) sum
// i: the block adds value i of every block partial of job j (four rows a
// thread, loaded at once), stores the sum as stamped 64-bit words and takes the
// job's ticket; the job's last block waits until every sum carries the stamp
// (agent-scope loads; no fence, whose L2 write-back costs more than the whole
// finish), multiplies each class's m + 1 sums by its coefficient, extends the
// class from t = 0..m to t = 0..d and writes the canonical message.  Extension
// by the trailing difference diagonal e_j = nabla^j f(m): nabla^m f is
// constant, and f(t + 1) follows from e_j += e_(j+1) for j = m - 1 .. 0.  A
// single block reading all partials took 13-40 us a round (up to ~100 partial
// loads in a row per thread); this takes a few.
__device__ __forceinline__ void finish_classes_body(const JobPack& pack) {
  const RoundJob& job = pack.job[blockIdx.y];
  const uint32_t width = job.class_width, i = blockIdx.x;
  if (i >= width) return;
  Fr s = Fr::zero();
  {
    constexpr uint32_t kU = 4;  // kMaxBlocks / kThreads
    Fr v[kU];
#pragma unroll
    for (uint32_t u = 0; u < kU; ++u) {
      const uint32_t b = threadIdx.x + u * kThreads;
      v[u] = b < job.blocks ? load_fr(job.partial, (size_t)b * width + i) : Fr::zero();
    }
    s = (v[0] + v[1]) + (v[2] + v[3]);
    for (uint32_t b = threadIdx.x + kU * kThreads; b < job.blocks; b += kThreads)
      s = s + load_fr(job.partial, (size_t)b * width + i);
  }
#pragma unroll
  for (uint32_t off = 32; off > 0; off >>= 1) {
    Fr o;
#pragma unroll
    for (int q = 0; q < 8; ++q) o.v[q] = __shfl_xor(s.v[q], off, 64);
    s = s + o;
  }
  __shared__ Fr wsum[kThreads / 64];
  __shared__ uint32_t last;
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  const uint64_t stamp = static_cast<uint64_t>(job.class_stamp) << 32;
  if (threadIdx.x < 64) {
    Fr t = wsum[0];
    for (uint32_t w = 1; w < kThreads / 64; ++w) t = t + wsum[w];
    if (threadIdx.x < 8)
      __hip_atomic_store(&job.class_sums[i * 8 + threadIdx.x], stamp | t.v[threadIdx.x], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0)
      last = __hip_atomic_fetch_add(job.ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == width - 1;
  }
  __syncthreads();
  if (!last) return;
  // the job's last block: every (class, t) sum, once stamped (each was
  // stored before its block's ticket, so it arrives; the wait is bounded all
  // the same, and a sum that never came leaves an all-ones word, which no
  // canonical value has and the host rejects)
  __shared__ Fr red[kMaxClassWidth];
  __shared__ uint32_t missing;
  if (threadIdx.x == 0) missing = 0;
  __syncthreads();
  for (uint32_t idx = threadIdx.x; idx < width * 8; idx += kThreads) {
    uint64_t* w = &job.class_sums[idx];
    uint64_t x = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (uint32_t spin = 0; (x & ~0xffffffffull) != stamp; ++spin) {
      if (spin == (1u << 22)) {
        missing = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      x = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    red[idx / 8].v[idx % 8] = static_cast<uint32_t>(x);
  }
  if (threadIdx.x == 0) __hip_atomic_store(job.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (missing) {
    if (threadIdx.x <= job.degree) {
      Fr bad;
#pragma unroll
      for (int q = 0; q < 8; ++q) bad.v[q] = 0xffffffffu;
      store_fr(job.out, threadIdx.x, bad);
      __threadfence_system();
    }
    return;
  }
  constexpr uint32_t kW = 9;  // degree <= 8
  __shared__ Fr val[kMaxClasses][kW], dif[kMaxClasses][kW], diag[kMaxClasses][kW];
  const uint32_t deg = job.degree;
  // the class coefficient times each sum: one thread per (class, t) value
  // (one product of latency instead of m + 1 in a row)
  static_assert(kMaxClasses * kW <= kThreads, "one value a thread");
  if (threadIdx.x < job.class_count * kW) {
    const uint32_t c = threadIdx.x / kW, j = threadIdx.x % kW;
    if (j <= job.classes[4 * c]) val[c][j] = red[job.classes[4 * c + 1] + j] * load_fr(job.class_k, c);
  }
  __syncthreads();
  if (threadIdx.x < job.class_count) {
    const uint32_t c = threadIdx.x;
    const uint32_t m = job.classes[4 * c];
    for (uint32_t j = 0; j <= m; ++j) dif[c][j] = val[c][j];
    diag[c][0] = val[c][m];
    for (uint32_t j = 1; j <= m; ++j) {
      for (uint32_t q = 0; q + j <= m; ++q) dif[c][q] = dif[c][q + 1] - dif[c][q];
      diag[c][j] = dif[c][m - j];
    }
    for (uint32_t t = m + 1; t <= deg; ++t) {
      for (uint32_t j = m; j-- > 0;) diag[c][j] = diag[c][j] + diag[c][j + 1];
      val[c][t] = diag[c][0];
    }
  }
  __syncthreads();
  if (threadIdx.x <= deg) {
    Fr sum = Fr::zero();
    for (uint32_t c = 0; c < job.class_count; ++c) sum = sum + val[c][threadIdx.x];
    store_fr(job.out, threadIdx.x, from_mont(sum));
    __threadfence_system();  // `out` may be pinned host memory read after the sync
  }
}
__global__ void __launch_bounds__(kThreads) k_finish_classes(const JobPack pack) { finish_classes_body(pack); }
// the same kernel over a pack in device memory (replayed round graphs)
__global__ void __launch_bounds__(kThreads) k_finish_classes_dev(const JobPack* pack) { finish_classes_body(*pack); }

// ---- latency-bound decompositions (small tables) ----------------------------
//
// A small round is a chain of dependent steps per thread; the metadata
// lookups (offsets -> slots -> table pointer -> table) used to add three
// dependent global loads per factor.  The block first copies its job's
// offsets, coefficients and per-entry table pointers into LDS (one round of
// global loads for all of them), and each thread loads the next factor's table
// values while it computes on the current one.
constexpr uint32_t kMetaProducts = 64, kMetaEntries = 256;

struct SmallMeta {
  uint32_t off[kMetaProducts + 1];
  const Fr* src[kMetaEntries];
  Fr* dst[kMetaEntries];
  uint8_t own[kMetaEntries];
};

// Copies the job's metadata into `m` when it fits (uniform per block; every
// thread reaches the barrier).  Returns whether it did.
__device__ __forceinline__ bool load_meta(const RoundJob& job, SmallMeta& m) {
  const bool fits = job.products <= kMetaProducts && job.entries <= kMetaEntries;
  if (fits) {
    for (uint32_t i = threadIdx.x; i <= job.products; i += kThreads) m.off[i] = job.offsets[i];
    for (uint32_t e = threadIdx.x; e < job.entries; e += kThreads) {
      const uint32_t f = job.slots[e];
      m.src[e] = job.src[f];
      if (job.fold) {
        m.dst[e] = job.dst[f];
        m.own[e] = job.owner[e];
      }
    }
  }
  __syncthreads();
  return fits;
}

// The table values one entry contributes on pair b: the four old values of the
// folding rounds, the two current ones otherwise.
template <bool FOLD>
struct EntryVals {
  Fr v[FOLD ? 4 : 2];
};

template <bool FOLD>
__device__ __forceinline__ void load_entry(const Fr* t, size_t b, EntryVals<FOLD>& x) {
#pragma unroll
  for (int i = 0; i < (FOLD ? 4 : 2); ++i) x.v[i] = load_fr(t, (FOLD ? 4 : 2) * b + i);
}

// lo, hi of the entry on pair b (folded with c29 in folding rounds, the fold
// stored to dst when dst is not null).
template <bool FOLD>
__device__ __forceinline__ void entry_line(const EntryVals<FOLD>& x, const Fr29& c29, Fr* dst,
                                           size_t b, Fr29& lo, Fr29& hi) {
  if constexpr (FOLD) {
    lo = fold29(c29, f29::unpack_fr(x.v[0]), f29::unpack_fr(x.v[1]));
    hi = fold29(c29, f29::unpack_fr(x.v[2]), f29::unpack_fr(x.v[3]));
    if (dst != nullptr) {
      store_fr(dst, 2 * b, f29::pack_fr_canonical(lo));
      store_fr(dst, 2 * b + 1, f29::pack_fr_canonical(hi));
    }
  } else {
    lo = f29::unpack_fr(x.v[0]);
    hi = f29::unpack_fr(x.v[1]);
  }
}

// Metadata accessors: the LDS copy when the job's metadata fits, else global.
struct MetaView {
  const RoundJob& job;
  const SmallMeta& meta;
  bool cached;
  __device__ __forceinline__ uint32_t off(uint32_t p) const {
    return cached ? meta.off[p] : job.offsets[p];
  }
  __device__ __forceinline__ const Fr* src(uint32_t e) const {
    return cached ? meta.src[e] : job.src[job.slots[e]];
  }
  __device__ __forceinline__ Fr* dst(uint32_t e) const {
    return cached ? meta.dst[e] : job.dst[job.slots[e]];
  }
  __device__ __forceinline__ bool own(uint32_t e) const {
    return cached ? meta.own[e] != 0 : job.owner[e] != 0;
  }
};

// Product p on pair b at every t (items) or at t only (points, ALL = false),
// the factors' table values loaded one factor ahead of the arithmetic.
template <int D, bool FOLD, bool ALL>
__device__ __forceinline__ void eval_product(const MetaView& mv, uint32_t p, size_t b, uint32_t t,
                                             bool store, ProductTerms<D>& pt, Fr29& term) {
  const RoundJob& job = mv.job;
  const Fr29 c29 = FOLD ? f29::prescale_fr(job.challenge) : f29::zero_t<f29::FrP>();
  const uint32_t begin = mv.off(p);
  const uint32_t m = mv.off(p + 1) - begin;
  EntryVals<FOLD> buf[2];
  load_entry<FOLD>(mv.src(begin), b, buf[0]);
  const Fr29 coeff = load29(job.coeffs29, p);
  if constexpr (!ALL) term = coeff;
  // fully unrolled: buf is indexed by compile-time constants only
#pragma clang loop unroll(full)
  for (int i = 0; i < D; ++i) {
    if ((uint32_t)i < m) {
      const uint32_t e = begin + i;
      if ((uint32_t)i + 1 < m) load_entry<FOLD>(mv.src(e + 1), b, buf[(i + 1) & 1]);
      Fr* dst = FOLD && store && mv.own(e) ? mv.dst(e) : nullptr;
      Fr29 lo, hi;
      entry_line<FOLD>(buf[i & 1], c29, dst, b, lo, hi);
      if constexpr (ALL) {
        if (i == 0)
          pt.first(coeff, lo, hi);
        else
          pt.times(lo, hi);
      } else {
        // lo + t (hi - lo) < 17 r
        const Fr29 d = f29::sub1p(hi, lo);
        Fr29 line = lo;
        for (uint32_t k = 0; k < t; ++k) line = f29::add(line, d);
        term = f29::mul(term, line);
      }
    }
  }
}

// Thread per (product p, pair b), product-major, one item per thread: the
// factors of p are folded on the fly (the first occurrence of a factor writes
// its folded pair) and p is evaluated at every t; short independent chains for
// the mid-size rounds where latency dominates.
template <int D, bool FOLD>
__device__ __forceinline__ void items_body(const JobPack& pack, const RoundJob& job,
                                           const MetaView& mv) {
  const uint32_t lblock = blockIdx.x - job.block_begin;
  const size_t items = (size_t)job.pairs * job.products;
  const size_t it = (size_t)lblock * kThreads + threadIdx.x;
  ProductTerms<D> pt;
#pragma unroll
  for (int t = 0; t <= D; ++t) pt.term[t] = f29::zero_t<f29::FrP>();
  if (it < items) {
    const uint32_t p = static_cast<uint32_t>(it / job.pairs);
    const size_t b = it - (size_t)p * job.pairs;
    Fr29 unused;
    eval_product<D, FOLD, true>(mv, p, b, 0, FOLD, pt, unused);
  }
  Fr out[D + 1];
  to_stored<D>(pt.term, out);
  finish_job<D>(job, lblock, out, pack.split != 0);
}

template <int D>
__device__ __forceinline__ void round_items_body(const JobPack& pack) {
  const RoundJob& job = pack.job[job_of_block(pack)];
  __shared__ SmallMeta meta;
  const MetaView mv{job, meta, load_meta(job, meta)};
  if (job.fold)
    items_body<D, true>(pack, job, mv);
  else
    items_body<D, false>(pack, job, mv);
}
template <int D>
__global__ void __launch_bounds__(kThreads) k_round_items(const JobPack pack) { round_items_body<D>(pack); }
// the same kernel over a pack in device memory (replayed round graphs)
template <int D>
__global__ void __launch_bounds__(kThreads) k_round_items_dev(const JobPack* pack) { round_items_body<D>(*pack); }

// ---- the smallest rounds: (pair, product, point) items in three phases ------
//
// A block takes whole pairs (or, when a pair has more than a block's items,
// a range of its products) and runs
//   1. fold: thread per (distinct factor, pair) folds the factor on the pair
//      (the first block of the pair stores the fold for the next round) and
//      leaves lo, hi in LDS;
//   2. evaluate: thread per (pair, product, t) multiplies the coefficient and
//      the product's lines at t as a tree (depth ceil(log2(m + 1)) for m
//      factors instead of a chain of 3 m products);
//   3. reduce: per t, a tree over the block's values in LDS.
// Threads are t-major within the block (thread = t G + g for G = threads /
// (d + 1) slots g), so every t's values are contiguous.  Needs the job's
// metadata in LDS (at most kMetaProducts products and kMetaEntries entries,
// hence at most kMetaEntries distinct factors).
struct PointsMeta {
  uint32_t off[kMetaProducts + 1];
  uint8_t slot[kMetaEntries];
  const Fr* src[kMetaEntries];
  Fr* dst[kMetaEntries];
};

// Block shape of a points job: G slots per t, pairs per block (0 when one
// pair's products span blocks) and blocks per pair.
struct PointsShape {
  uint32_t slots, pairs_per_block, blocks_per_pair;
};

__host__ __device__ inline PointsShape points_shape(uint32_t products, uint32_t degree) {
  PointsShape sh;
  sh.slots = kThreads / (degree + 1);
  if (products <= sh.slots) {
    sh.pairs_per_block = sh.slots / products;
    sh.blocks_per_pair = 1;
  } else {
    sh.pairs_per_block = 0;
    sh.blocks_per_pair = (products + sh.slots - 1) / sh.slots;
  }
  return sh;
}

template <int D>
__device__ __forceinline__ void round_points_body(const JobPack& pack) {
  const RoundJob& job = pack.job[job_of_block(pack)];
  const uint32_t lblock = blockIdx.x - job.block_begin;
  constexpr uint32_t width = D + 1;
  const PointsShape sh = points_shape(job.products, D);
  const uint32_t G = sh.slots;
  const uint32_t bpp = sh.blocks_per_pair, sub = lblock % bpp;
  const size_t b0 = sh.pairs_per_block ? (size_t)lblock * sh.pairs_per_block : (size_t)(lblock / bpp);
  const uint32_t npairs = sh.pairs_per_block
                              ? static_cast<uint32_t>(min((size_t)sh.pairs_per_block, job.pairs - b0))
                              : 1u;
  const uint32_t F = job.factors;

  __shared__ PointsMeta meta;
  __shared__ Fr29 lo_s[kMetaEntries], hi_s[kMetaEntries];
  __shared__ Fr29 red[kThreads];
  for (uint32_t i = threadIdx.x; i <= job.products; i += kThreads) meta.off[i] = job.offsets[i];
  for (uint32_t e = threadIdx.x; e < job.entries; e += kThreads) meta.slot[e] = (uint8_t)job.slots[e];
  for (uint32_t f = threadIdx.x; f < F; f += kThreads) {
    meta.src[f] = job.src[f];
    if (job.fold) meta.dst[f] = job.dst[f];
  }
  __syncthreads();

  // 1. fold (or load) every distinct factor on the block's pairs
  for (uint32_t u = threadIdx.x; u < F * npairs; u += kThreads) {
    const uint32_t f = u / npairs, q = u - f * npairs;
    const size_t b = b0 + q;
    const Fr* o = meta.src[f];
    Fr29 lo, hi;
    if (job.fold) {
      const Fr a0 = load_fr(o, 4 * b), a1 = load_fr(o, 4 * b + 1);
      const Fr a2 = load_fr(o, 4 * b + 2), a3 = load_fr(o, 4 * b + 3);
      const Fr29 c29 = f29::prescale_fr(job.challenge);
      lo = fold29(c29, f29::unpack_fr(a0), f29::unpack_fr(a1));
      hi = fold29(c29, f29::unpack_fr(a2), f29::unpack_fr(a3));
      if (sub == 0) {
        store_fr(meta.dst[f], 2 * b, f29::pack_fr_canonical(lo));
        store_fr(meta.dst[f], 2 * b + 1, f29::pack_fr_canonical(hi));
      }
    } else {
      lo = load29(o, 2 * b);
      hi = load29(o, 2 * b + 1);
    }
    lo_s[u] = lo;
    hi_s[u] = hi;
  }
  __syncthreads();

  // 2. evaluate: thread = t G + g
  const uint32_t t = threadIdx.x / G, g = threadIdx.x - t * G;
  Fr29 val = f29::zero_t<f29::FrP>();
  uint32_t q = 0, p = 0;
  bool active = t < width;
  if (sh.pairs_per_block) {
    q = g / job.products;
    p = g - q * job.products;
    active = active && q < npairs;
  } else {
    p = sub * G + g;
    active = active && p < job.products;
  }
  if (active) {
    const uint32_t begin = meta.off[p], m = meta.off[p + 1] - begin;
    Fr29 v[D + 1];
    v[0] = load29(job.coeffs29, p);
#pragma unroll
    for (int i = 0; i < D; ++i) {
      if ((uint32_t)i < m) {
        const uint32_t u = (uint32_t)meta.slot[begin + i] * npairs + q;
        const Fr29 lo = lo_s[u];
        // lo + t (hi - lo) < 17 r
        const Fr29 d = f29::sub1p(hi_s[u], lo);
        Fr29 line = lo;
        for (uint32_t k = 0; k < t; ++k) line = f29::add(line, d);
        v[i + 1] = line;
      }
    }
    // tree over the m + 1 leaves (products stay below 5.1 r)
#pragma unroll
    for (int h = 1; h <= D; h <<= 1)
#pragma unroll
      for (int i = 0; i + h <= D; i += 2 * h)
        if ((uint32_t)(i + h) <= m) v[i] = f29::mul(v[i], v[i + h]);
    val = f29::reduce_est(v[0]);  // < 3 r
  }
  red[threadIdx.x] = val;
  __syncthreads();
  // 3. per t, a tree over the G slots: each level adds two values below 3 r
  // and reduces again
  for (uint32_t h = 1; h < G; h <<= 1) {
    if (t < width && (g & (2 * h - 1)) == 0 && g + h < G)
      red[threadIdx.x] = f29::reduce_est(f29::add(red[threadIdx.x], red[threadIdx.x + h]));
    __syncthreads();
  }
  // red[k G] is the block's value at t = k
  if (job.blocks == 1 || pack.split) {
    if (threadIdx.x < width) {
      const Fr v = f29::pack_fr_reduced(red[threadIdx.x * G]);
      if (job.blocks == 1) {
        store_fr(job.out, threadIdx.x, from_mont(v));
        __threadfence_system();  // `out` may be pinned host memory read after the sync
      } else {
        store_fr(job.partial, (size_t)lblock * width + threadIdx.x, v);
      }
    }
    return;
  }
  // ticket finish (replayed round graphs): thread k carries the value at t = k
  Fr out[D + 1];
#pragma unroll
  for (int k = 0; k <= D; ++k) out[k] = Fr::zero();
  if (threadIdx.x < width) {
    const Fr v = f29::pack_fr_reduced(red[threadIdx.x * G]);
#pragma unroll
    for (int k = 0; k <= D; ++k)
      if ((uint32_t)k == threadIdx.x) out[k] = v;
  }
  finish_job<D>(job, lblock, out, false);
}
template <int D>
__global__ void __launch_bounds__(kThreads) k_round_points(const JobPack pack) { round_points_body<D>(pack); }
// the same kernel over a pack in device memory (replayed round graphs)
template <int D>
__global__ void __launch_bounds__(kThreads) k_round_points_dev(const JobPack* pack) { round_points_body<D>(*pack); }

// ---- persistent round loops for the last (small) rounds -------------------
//
// Once a state's table is small, a round's arithmetic takes a few
// microseconds and a launch per round (launch, end-of-kernel write-back, the
// host waking on the stream) costs several times that.  Blocks then run the
// remaining rounds in a loop and talk to the host through a mailbox in pinned
// host memory: each message is published there and the blocks wait for the
// host's next challenge, with a bounded wait (a wall-clock limit and an
// iteration cap) and an abort flag, so every wave always reaches the exit:
// all rounds done, abort from the host, or the wait limit (the host then
// relaunches from the round it is at).
//
// One state may run on G = 2^b blocks (a wide loop, one state only): block g
// owns pairs [g P, (g + 1) P) of every round (P = pairs / G).  Pair i of a
// round folds entries 4i .. 4i + 3 of the previous round's table into entries
// 2i, 2i + 1, so a block only ever reads what it wrote itself and the blocks
// exchange nothing but their partial messages.  A wide loop runs in up to
// three stages of G, then fewer, then one block (every block keeps at least
// one pair); at a stage change the blocks beyond the new count exit and the
// others take over their pairs, reading what those blocks wrote with
// agent-scope loads.  A wide loop stores its folds through to memory
// (agent-scope stores): its later rounds rewrite entries that other blocks,
// on other XCDs, wrote in earlier rounds, and with plain stores two L2s could
// hold such a line dirty, the stale copy written back last.
//
// Nothing here relies on the order in which two stores become visible: every
// word that crosses from one block to another, or between the device and the
// host, carries the round it belongs to.  The host's challenge words carry a
// round tag in their spare top bits; the message words and the wide loop's
// partial words are 64-bit atomic stores of (stamp << 32 | word), and the
// reader waits until every word it needs has its round's stamp.  So no fence
// is needed (a device-scope fence writes back the L2, several microseconds on
// this part), and an earlier value of a word can never pass for a later one.
constexpr unsigned kTailThreads = 1024;
constexpr uint32_t kTailRunning = 0, kTailDone = 1, kTailTimedOut = 2, kTailAborted = 3;
constexpr uint32_t kTailMaxWidth = 9;  // degree <= 8
// metadata limits of a state that runs its rounds in the loop (kept in LDS)
constexpr uint32_t kTailMaxProducts = 32, kTailMaxEntries = 128, kTailMaxFactors = 64;
// (distinct factor, pair) lines kept in LDS at once (2 x 32 bytes each); a
// block takes its pairs in chunks of at most this many lines
constexpr uint32_t kTailMaxLines = 768;
// blocks of one wide loop (at most one per CU, so all are resident)
constexpr uint32_t kWideMaxBlocks = 256;

struct TailJob {
  const Fr* const* tables[3];  // inputs, fold slot A, fold slot B (device pointer arrays)
  const Fr* coeffs29;
  const uint32_t* offsets;
  const uint32_t* slots;
  const uint8_t* owner;
  uint32_t products, degree, num_vars, first_round;
  uint32_t entries, factors, pad0, pad1;
  uint32_t challenge29[f29::FrP::L];  // the first round's challenge c as c 2^261 in Fr29 limbs
};

// The stages of a wide launch (one state): stage i runs rounds
// [end[i - 1], end[i]) on 2^blocks_log2[i] blocks (end[2] is the state's
// round count).  Words exchanged between blocks in round k carry the stamp
// stamp_base + k + 1.  All zero: a one-block launch of several states.
struct TailStage {
  uint32_t end[3];
  uint32_t blocks_log2[3];
  uint32_t stamp_base, wide;
};

// The host's challenge for round r is the prescaled challenge's nine limb
// words, each carrying tail_tag(r) in its top three bits (limbs are below
// 2^29), so one load instruction per poll reads a whole challenge and knows
// it is whole: a word is either the previous post's or the new one's, and
// consecutive rounds have different tags.  Tag 0 marks cleared words.
__host__ __device__ constexpr uint32_t tail_tag(uint32_t round) { return round % 7 + 1; }
constexpr uint32_t kTailAbortWord = f29::FrP::L;

struct TailMailbox {
  alignas(64) TailJob job;  // written by the host before the launch
  // host -> device, one 64-byte line: [0, 9) challenge limbs | tag << 29,
  // [9] abort (1: stop at the next wait)
  alignas(64) uint32_t post[16];
  // device -> host
  alignas(64) uint32_t status;  // kTailRunning / Done / TimedOut / Aborted
  // round r's message, stored (Montgomery) form: word i as (r + 1) << 32 | w
  alignas(64) uint64_t message[kTailMaxWidth * 8];
};

// Device memory of a wide loop: the blocks' partial messages, word i of
// block g's point t at (g kTailMaxWidth + t) 8 + i, and the challenge block 0
// passes on to the other blocks (limb words 0..8, the stop code in word 9),
// each word as stamp << 32 | w.
struct WideArea {
  uint64_t* part;   // kWideMaxBlocks x kTailMaxWidth x 8 words
  uint64_t* bcast;  // 16 words
};

// Host-written words are read with relaxed system-scope loads (they bypass the
// caches): an acquire would also invalidate the L2, and every table load of
// the next round would then go to HBM.
__device__ __forceinline__ uint32_t sys_load_relaxed(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void sys_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void sys_store64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// Agent-scope relaxed atomics are coherent across the XCDs' L2 caches
// without the L2 write-back or invalidation of a device-scope fence.
__device__ __forceinline__ void agent_store64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t agent_load64(uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Entry i of a table that other blocks stored through to memory, read with
// agent-scope loads (this L2 may hold an older copy of its line).
__device__ __forceinline__ Fr29 load29_agent(const Fr* base, size_t i) {
  uint64_t* p = reinterpret_cast<uint64_t*>(const_cast<Fr*>(base) + i);
  Fr v;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint64_t x = agent_load64(p + q);
    v.v[2 * q] = static_cast<uint32_t>(x);
    v.v[2 * q + 1] = static_cast<uint32_t>(x >> 32);
  }
  return f29::unpack_fr(v);
}

// A fold output of a wide loop, stored through to memory (agent-scope
// stores leave no dirty line in this XCD's L2).  A wide loop's later rounds
// rewrite entries that other blocks, on other XCDs, wrote in earlier rounds;
// with plain stores both L2s could hold the line dirty, and the stale copy,
// written back last, would overwrite the new one.
__device__ __forceinline__ void store_fr_through(Fr* base, size_t i, const Fr& v) {
  uint64_t* p = reinterpret_cast<uint64_t*>(base + i);
#pragma unroll
  for (int q = 0; q < 4; ++q)
    agent_store64(p + q, static_cast<uint64_t>(v.v[2 * q]) | static_cast<uint64_t>(v.v[2 * q + 1]) << 32);
}

// Wavefront 0 of a block waits for round `round`'s challenge (bounded;
// abort from the host); returns the stop code (kTailRunning: `c29` holds the
// challenge's limbs).  Lanes 0..8 load the limb words and lane 9 the abort
// word, one load instruction per poll.
__device__ uint32_t tail_wait_challenge(TailMailbox* box, uint32_t round, uint64_t wait_ticks,
                                        uint32_t spin_cap, uint32_t* c29) {
  const uint32_t lane = threadIdx.x & 63, want = tail_tag(round);
  const uint64_t start = static_cast<uint64_t>(wall_clock64());
  for (uint32_t spin = 0;; ++spin) {
    const uint32_t w = lane <= kTailAbortWord ? sys_load_relaxed(&box->post[lane]) : 0;
    const bool ready = lane >= kTailAbortWord || (w >> 29) == want;
    const uint32_t abort_word = __shfl(w, kTailAbortWord, 64);
    if (abort_word != 0) return kTailAborted;
    if (__ballot(ready) == ~0ull) {
      if (lane < kTailAbortWord) c29[lane] = w & 0x1fffffffu;
      return kTailRunning;
    }
    if (spin >= spin_cap || static_cast<uint64_t>(wall_clock64()) - start > wait_ticks)
      return kTailTimedOut;
    __builtin_amdgcn_s_sleep(2);
  }
}

// Wavefront 0 of every block of a loop: round r's challenge into c29, or the
// stop code.  Only block 0 polls the host's mailbox (host memory read by
// hundreds of blocks at once slowed every round to hundreds of
// microseconds); in a wide loop it passes the challenge, or its stop code,
// on through device memory, stamped with `stamp`, and the other blocks poll
// that (bounded).
__device__ uint32_t loop_wait(TailMailbox* box, uint32_t r, uint32_t g, uint32_t G, WideArea wide,
                              uint32_t stamp, uint64_t wait_ticks, uint32_t spin_cap, uint32_t* c29) {
  const uint32_t lane = threadIdx.x & 63;
  if (g == 0) {
    const uint32_t stop = tail_wait_challenge(box, r, wait_ticks, spin_cap, c29);
    if (G > 1 && lane <= kTailAbortWord)
      agent_store64(&wide.bcast[lane], static_cast<uint64_t>(stamp) << 32 |
                                           (lane < kTailAbortWord ? c29[lane] : stop));
    return stop;
  }
  const uint64_t start = static_cast<uint64_t>(wall_clock64());
  for (uint32_t spin = 0;; ++spin) {
    const uint64_t x = lane <= kTailAbortWord ? agent_load64(&wide.bcast[lane]) : 0;
    const bool ready = lane > kTailAbortWord || (x >> 32) == stamp;
    if (__ballot(ready) == ~0ull) {
      if (lane < kTailAbortWord) c29[lane] = static_cast<uint32_t>(x);
      return __shfl(static_cast<uint32_t>(x), kTailAbortWord, 64);
    }
    if (spin >= spin_cap || static_cast<uint64_t>(wall_clock64()) - start > wait_ticks) return kTailTimedOut;
    __builtin_amdgcn_s_sleep(1);
  }
}

// Block k of the launch (one state per block), or every block of a wide
// launch (blocks_log2 = b > 0, one state), runs its rounds.  Each round runs,
// chunk by chunk over the block's pairs, in three phases:
//   1. thread per (distinct factor f, pair b) folds f on b with the previous
//      challenge (or loads it in round 0), stores the fold for the next round
//      and leaves the line's ends lo, hi in LDS: every factor is folded once
//      per pair, however many products and points use it;
//   2. the block's 16 wavefronts split over the points, 16 / (d + 1) per
//      point (at least one; wavefront w takes point w / ws, w >= ws (d + 1)
//      idle): the lanes of the ws wavefronts of t take its (product, pair)
//      items, multiply the coefficient by the product's lines at t from LDS,
//      and add to their running sums, kept across chunks; pairs is a power
//      of two, so an item splits by shifts;
//   3. after the last chunk, wavefront sums, then one thread per point adds
//      its ws of them: the block's message (one block) or partial (wide).
// A wide loop's blocks store their stamped partials; block 0 waits for all of
// them (bounded), adds them and publishes the message.
// The fold writes of round k are read by other threads of the same block in
// round k + 1 after the block barriers.
__global__ void __launch_bounds__(kTailThreads) k_round_tail(TailMailbox* const* boxes,
                                                             uint64_t wait_ticks,
                                                             uint32_t spin_cap,
                                                             TailStage stage, WideArea wide) {
  const bool staged = stage.wide != 0;
  const uint32_t g = staged ? blockIdx.x : 0;
  TailMailbox* box = boxes[staged ? 0 : blockIdx.x];
  __shared__ TailJob job;
  __shared__ uint32_t s_c29[f29::FrP::L];
  __shared__ uint32_t s_stop;
  __shared__ Fr msg[kTailMaxWidth];
  __shared__ Fr part[kTailThreads / 64];  // wavefront sums
  // the state's metadata, kept in LDS for all its rounds
  __shared__ uint32_t s_off[kTailMaxProducts + 1];
  __shared__ uint32_t s_slot[kTailMaxEntries];
  __shared__ Fr s_coeff[kTailMaxProducts];
  __shared__ const Fr* s_tab[3][kTailMaxFactors];
  // line ends of every (distinct factor, pair) of a chunk, stored form
  __shared__ Fr s_lo[kTailMaxLines], s_hi[kTailMaxLines];
  if (threadIdx.x < sizeof(TailJob) / 4)
    reinterpret_cast<uint32_t*>(&job)[threadIdx.x] =
        reinterpret_cast<const uint32_t*>(&box->job)[threadIdx.x];
  __syncthreads();
  if (threadIdx.x < f29::FrP::L) s_c29[threadIdx.x] = job.challenge29[threadIdx.x];
  if (threadIdx.x == 0) s_stop = kTailRunning;
  if (threadIdx.x <= job.products) s_off[threadIdx.x] = job.offsets[threadIdx.x];
  if (threadIdx.x < job.products) s_coeff[threadIdx.x] = load_fr(job.coeffs29, threadIdx.x);
  if (threadIdx.x < job.entries) s_slot[threadIdx.x] = job.slots[threadIdx.x];
  if (threadIdx.x < job.factors)
    for (int t = 0; t < 3; ++t) s_tab[t][threadIdx.x] = job.tables[t][threadIdx.x];
  __syncthreads();
  const uint32_t n = job.num_vars, width = job.degree + 1, F = job.factors;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // pairs per chunk: the largest power of two whose F lines fit the LDS
  const uint32_t chunk_cap = 31u - __clz(kTailMaxLines / F);
  constexpr uint32_t kWaves = kTailThreads / 64;
  for (uint32_t k = job.first_round;; ++k) {
    Fr29 c29;
#pragma unroll
    for (int i = 0; i < f29::FrP::L; ++i) c29.l[i] = s_c29[i];
    // this round's stage: blocks beyond its count are done
    uint32_t si = 0;
    if (staged)
      while (k >= stage.end[si]) ++si;
    const uint32_t blocks_log2 = stage.blocks_log2[si], G = 1u << blocks_log2;
    if (g >= G) return;
    // the first round of a later stage reads folds other blocks stored
    const bool handed = si > 0 && k == stage.end[si - 1];
    const uint32_t lb = n - k - 1 - blocks_log2;  // log2 of the block's pairs (the host keeps it >= 0)
    const uint32_t lc = lb < chunk_cap ? lb : chunk_cap;  // log2 of a chunk's pairs
    const uint32_t cpairs = 1u << lc, base = g << lb;
    const bool fold = k > 0;
    const Fr* const* src = k <= 1 ? s_tab[0] : s_tab[1 + ((k - 2) & 1)];
    const Fr* const* dst = k == 0 ? nullptr : s_tab[1 + ((k - 1) & 1)];
    // phase 2's split: ws wavefronts per point over a chunk's items
    const uint32_t items = job.products << lc;
    const uint32_t ws_fit = width >= kWaves ? 1u : kWaves / width, ws_need = (items + 63) / 64;
    const uint32_t ws = ws_fit < ws_need ? ws_fit : ws_need;  // >= 1: items >= 1
    const uint32_t held = items < 64 ? items : 64;  // lanes with items in a wavefront
    const bool active = wave < ws * width;           // at most one (point, share) per wavefront
    const uint32_t t = active ? wave / ws : 0, j = active ? wave - t * ws : 0;
    // lazy Fr29 sums (each term < 1.3 r; reduced every 16 terms, below 24 r)
    Fr29 acc = f29::zero_t<f29::FrP>();
    uint32_t terms = 0;
    for (uint32_t cb = base; cb < base + (1u << lb); cb += cpairs) {
      // 1. lines of every distinct factor on every pair of the chunk
      for (uint32_t u = threadIdx.x; u < (F << lc); u += kTailThreads) {
        const uint32_t f = u >> lc, b = cb + (u & (cpairs - 1));
        const Fr* o = src[f];
        if (fold) {
          Fr29 a0, a1, a2, a3;
          if (handed) {
            a0 = load29_agent(o, 4 * b), a1 = load29_agent(o, 4 * b + 1);
            a2 = load29_agent(o, 4 * b + 2), a3 = load29_agent(o, 4 * b + 3);
          } else {
            a0 = load29(o, 4 * b), a1 = load29(o, 4 * b + 1);
            a2 = load29(o, 4 * b + 2), a3 = load29(o, 4 * b + 3);
          }
          const Fr lo = f29::pack_fr_canonical(fold29(c29, a0, a1));
          const Fr hi = f29::pack_fr_canonical(fold29(c29, a2, a3));
          if (G > 1) {
            store_fr_through(const_cast<Fr*>(dst[f]), 2 * b, lo);
            store_fr_through(const_cast<Fr*>(dst[f]), 2 * b + 1, hi);
          } else {
            store_fr(const_cast<Fr*>(dst[f]), 2 * b, lo);
            store_fr(const_cast<Fr*>(dst[f]), 2 * b + 1, hi);
          }
          s_lo[u] = lo;
          s_hi[u] = hi;
        } else {
          s_lo[u] = load_fr(o, 2 * b);
          s_hi[u] = load_fr(o, 2 * b + 1);
        }
      }
      __syncthreads();  // also orders this round's fold writes before the next round's reads
      // 2. the chunk's items of this wavefront's point
      if (active)
        for (uint32_t it = j * 64 + lane; it < items; it += ws * 64) {
          const uint32_t p = it >> lc, b = it & (cpairs - 1);
          Fr29 term = f29::unpack_fr(s_coeff[p]);
          for (uint32_t e = s_off[p]; e < s_off[p + 1]; ++e) {
            const uint32_t u = (s_slot[e] << lc) | b;
            const Fr29 lo = f29::unpack_fr(s_lo[u]);
            const Fr29 d = f29::sub1p(f29::unpack_fr(s_hi[u]), lo);
            Fr29 line = lo;  // lo + t (hi - lo) < 17 r
            for (uint32_t q = 0; q < t; ++q) line = f29::add(line, d);
            term = f29::mul(term, line);
          }
          acc = f29::add(acc, term);
          if ((++terms & 15u) == 0) acc = f29::reduce_est(acc);
        }
      __syncthreads();  // the next chunk overwrites the lines
    }
    // 3. a butterfly over the lanes that hold items (Fr additions on the
    // packed values), then one thread per point
    if (active) {
      Fr v = f29::pack_fr_reduced(f29::reduce_est(acc));
#pragma unroll
      for (uint32_t off = 32; off > 0; off >>= 1) {
        if (off < held) {  // uniform: lanes >= held hold zero
          Fr o;
#pragma unroll
          for (int i = 0; i < 8; ++i) o.v[i] = __shfl_xor(v.v[i], off, 64);
          v = v + o;
        }
      }
      if (lane == 0) part[wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < width) {
      Fr sum = part[threadIdx.x * ws];
      for (uint32_t q = 1; q < ws; ++q) sum = sum + part[threadIdx.x * ws + q];
      msg[threadIdx.x] = sum;  // stored form: the host converts to canonical
    }
    __syncthreads();
    if (G > 1) {
      // the block's partial, stamped; block 0 adds the G partials of each
      // point: thread q waits for block q's words (bounded), a butterfly per
      // wavefront, then one thread per point adds the wavefront sums (kept in
      // s_lo, free after the chunks)
      const uint64_t stamp = static_cast<uint64_t>(stage.stamp_base + k + 1) << 32;
      if (threadIdx.x < 64)
        for (uint32_t i = threadIdx.x; i < width * 8; i += 64)
          agent_store64(&wide.part[(g * kTailMaxWidth + i / 8) * 8 + i % 8], stamp | msg[i / 8].v[i % 8]);
      if (g == 0) {
        // item j = t G + q (block q's partial at point t), 1024 items a pass:
        // each thread loads its item's eight words at once (waiting only for
        // words not yet stamped), then a butterfly over segments of
        // min(G, 64) lanes (one point each) and one LDS slot per segment
        bool late = false;
        const uint64_t start = static_cast<uint64_t>(wall_clock64());
        const uint32_t seg = G < 64 ? G : 64, nseg = G < 64 ? 1u : G / 64;
        for (uint32_t j0 = 0; j0 < width * G; j0 += kTailThreads) {
          const uint32_t j = j0 + threadIdx.x;
          const uint32_t tt = j >> blocks_log2, q = j & (G - 1);
          Fr v = Fr::zero();
          if (j < width * G) {
            // all eight words again while any is unstamped: one memory round
            // trip a poll (a word at a time cost up to eight in a row)
            uint64_t* w = &wide.part[(q * kTailMaxWidth + tt) * 8];
            uint64_t x[8];
            for (uint32_t spin = 0;; ++spin) {
#pragma unroll
              for (int i = 0; i < 8; ++i) x[i] = agent_load64(w + i);
              bool ok = true;
#pragma unroll
              for (int i = 0; i < 8; ++i) ok = ok && (x[i] & ~0xffffffffull) == stamp;
              if (ok) break;
              if (spin >= spin_cap || static_cast<uint64_t>(wall_clock64()) - start > wait_ticks) {
                late = true;
                break;
              }
              __builtin_amdgcn_s_sleep(1);
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) v.v[i] = static_cast<uint32_t>(x[i]);
          }
#pragma unroll
          for (uint32_t off = 32; off > 0; off >>= 1) {
            if (off < seg) {  // uniform
              Fr o;
#pragma unroll
              for (int i = 0; i < 8; ++i) o.v[i] = __shfl_xor(v.v[i], off, 64);
              v = v + o;
            }
          }
          if (j < width * G && (q & (seg - 1)) == 0) s_lo[tt * nseg + q / 64] = v;
        }
        if (late) s_stop = kTailTimedOut;  // a block never delivered
        __syncthreads();
        if (threadIdx.x < width) {
          Fr sum = s_lo[threadIdx.x * nseg];
          for (uint32_t q = 1; q < nseg; ++q) sum = sum + s_lo[threadIdx.x * nseg + q];
          msg[threadIdx.x] = sum;
        }
        __syncthreads();
        if (s_stop != kTailRunning) {
          if (threadIdx.x == 0) sys_store(&box->status, kTailTimedOut);
          return;
        }
      }
    }
    const bool last_round = k + 1 == n;
    // the block holding the message: wavefront 0 writes its stamped words
    // (64-bit atomic stores); the host takes them once all carry round k's
    // stamp
    if (g == 0 && threadIdx.x < 64) {
      const uint64_t stamp = static_cast<uint64_t>(k + 1) << 32;
      if (last_round && threadIdx.x == 0) sys_store(&box->status, kTailDone);
      for (uint32_t i = threadIdx.x; i < width * 8; i += 64)
        sys_store64(&box->message[i], stamp | msg[i / 8].v[i % 8]);
    }
    if (last_round) return;
    // blocks without pairs in round k + 1 are done
    uint32_t next_si = si;
    if (staged)
      while (k + 1 >= stage.end[next_si]) ++next_si;
    const uint32_t next_G = 1u << stage.blocks_log2[next_si];
    if (g >= next_G) return;
    // the challenge of round k + 1
    if (threadIdx.x < 64) {
      const uint32_t stop = loop_wait(box, k + 1, g, next_G, wide, stage.stamp_base + k + 2, wait_ticks,
                                      spin_cap, s_c29);
      if (threadIdx.x == 0) {
        if (stop != kTailRunning) sys_store(&box->status, stop);
        s_stop = stop;
      }
    }
    __syncthreads();
    if (s_stop != kTailRunning) return;
  }
}

// Any degree: message values for t in [t0, t0 + 8) only (host loops over t0).
__global__ void __launch_bounds__(kThreads) k_message_chunk(const Fr* const* tbl, Products prods,
                                                            size_t pairs, uint32_t t0,
                                                            uint32_t tcount, Fr* partial,
                                                            size_t stride) {
  constexpr int T = 8;
  Fr acc[T];
#pragma unroll
  for (int t = 0; t < T; ++t) acc[t] = Fr::zero();
  for (size_t b = blockIdx.x * (size_t)blockDim.x + threadIdx.x; b < pairs;
       b += (size_t)gridDim.x * blockDim.x) {
    for (uint32_t p = 0; p < prods.count; ++p) {
      const uint32_t begin = prods.offsets[p];
      const uint32_t arity = prods.offsets[p + 1] - begin;
      const Fr coeff = load_fr(prods.coeffs, p);
#pragma unroll
      for (int t = 0; t < T; ++t) {
        if (t < (int)tcount) {
          const Fr tt = from_u64<FrParams>(t0 + t);
          Fr term = coeff;
          for (uint32_t k = 0; k < arity; ++k) {
            const Fr* tab = tbl[prods.slots[begin + k]];
            const Fr lo = load_fr(tab, 2 * b), hi = load_fr(tab, 2 * b + 1);
            term = term * (lo + tt * (hi - lo));
          }
          acc[t] = acc[t] + term;
        }
      }
    }
  }
  // partial layout: [block][t] with row stride `stride` (= d + 1), offset t0
  __shared__ Fr lds[kThreads / 64][T];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < T; ++t) {
    if (t < (int)tcount) {
      Fr s = wave_sum(acc[t]);
      if (lane == 0) lds[wave][t] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x < tcount) {
    Fr s = Fr::zero();
    for (int w = 0; w < (int)(kThreads / 64); ++w) s = s + lds[w][threadIdx.x];
    store_fr(partial, blockIdx.x * stride + t0 + threadIdx.x, s);
  }
}

// out[t] = canonical(sum_blocks partial[block][t])
__global__ void k_reduce_partials(const Fr* partial, size_t blocks, size_t width, Fr* out) {
  for (size_t t = threadIdx.x; t < width; t += blockDim.x) {
    Fr s = Fr::zero();
    for (size_t b = 0; b < blocks; ++b) s = s + load_fr(partial, b * width + t);
    store_fr(out, t, from_mont(s));
  }
}

__global__ void k_zero_fr(Fr* out, size_t n) {
  for (size_t t = threadIdx.x; t < n; t += blockDim.x) store_fr(out, t, Fr::zero());
}

}  // namespace

struct State {
  size_t num_vars = 0, max_degree = 0, round = 0;
  bool failed = false;
  int32_t failure_code = 0;
  std::vector<rt::BufferRef> inputs;  // distinct factors (shared, never written)
  rt::BufferRef slots;                // ping-pong fold slots: 2 per distinct factor
  rt::BufferRef meta;                 // coeffs | table pointers | offsets | slots | owner
  Products prods{};
  const Fr* coeffs29 = nullptr;  // coefficients times 2^(5 m_p) (Fr29 product chains)
  const uint8_t* owner = nullptr;
  const Fr** table_ptrs[3] = {nullptr, nullptr, nullptr};  // inputs, slot A, slot B
  size_t distinct = 0, product_count = 0, index_count = 0;
  // product classes of the class-major decomposition (class_count 0: not used)
  const uint32_t* classes = nullptr;
  const Fr* class_k = nullptr;
  uint32_t class_count = 0, class_width = 0;
  rt::BufferRef partial, result, ticket;
  uint64_t* class_sums = nullptr;  // class finish sums (stamped words) in `ticket`'s block
  Fr* slot_base[2] = {nullptr, nullptr};
  // persistent round loop: this state's mailbox (pinned host memory, from the
  // context's pool) and whether its block may still wait on the host
  TailMailbox* box = nullptr;
  TailMailbox* dbox = nullptr;
  bool tail_active = false;
};

}  // namespace sc
}  // namespace hp

struct HpMi355xSumCheckState {
  HpMi355xContext* ctx;
  std::shared_ptr<hp::sc::State> impl;
};

using namespace hp;

namespace {

constexpr unsigned kMaxJobDegree = 8;
// class-major launches up to this degree (above it the class passes' terms
// and sums exceed the register file)
constexpr unsigned kMaxClassDegree = 6;

unsigned message_blocks(size_t pairs) {
  size_t b = (pairs + sc::kThreads - 1) / sc::kThreads;
  return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(b, sc::kMaxBlocks)));
}

// tables at or below this many pairs per round use the per-(product, pair)
// decomposition; at or below points_below() (product, pair) items, the
// per-(product, pair, point) one
size_t points_below() {
  static const size_t v = [] {
    const char* e = std::getenv("HP_MI355X_SC_POINTS_BELOW");
    return e ? static_cast<size_t>(std::strtoull(e, nullptr, 0)) : size_t(1) << 14;
  }();
  return v;
}

size_t items_below() {
  static const size_t v = [] {
    const char* e = std::getenv("HP_MI355X_SC_ITEMS_BELOW");
    return e ? static_cast<size_t>(std::strtoull(e, nullptr, 0)) : size_t(1) << 15;
  }();
  return v;
}

// grid blocks and decomposition (0 pairs, 1 items, 2 points) of round k of a
// state (same rule at create time and at run time)
// The class-major decomposition of large tables (mode 3) for states whose
// products form few enough classes; HP_MI355X_SC_CLASSES=0 keeps the
// per-product one (mode 0).
bool classes_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("HP_MI355X_SC_CLASSES");
    return e == nullptr || std::atoi(e) != 0;
  }();
  return on;
}

// Pairs per thread of a class launch, HP_MI355X_SC_CLASS_PAIRS (default 1:
// at n = 20, 2 and 4 pairs per thread measured 4 % and 20 % slower, the
// smaller grids leaving SIMDs idle, against the amortised class reductions).
size_t class_pairs_per_thread() {
  static const size_t v = [] {
    const char* e = std::getenv("HP_MI355X_SC_CLASS_PAIRS");
    const size_t x = e ? static_cast<size_t>(std::strtoull(e, nullptr, 0)) : size_t(1);
    return std::max<size_t>(1, x);
  }();
  return v;
}

unsigned job_blocks(const sc::State& s, size_t k, uint32_t* mode) {
  const size_t pairs = size_t(1) << (s.num_vars - k - 1);
  if (s.product_count != 0 && pairs >= items_below()) {
    *mode = s.class_count != 0 && classes_enabled() ? 3 : 0;
    if (*mode == 3)
      return message_blocks((pairs + class_pairs_per_thread() - 1) / class_pairs_per_thread());
    return message_blocks(pairs);
  }
  const size_t items = pairs * s.product_count;
  *mode = s.product_count != 0 && items <= points_below() && s.max_degree <= kMaxJobDegree &&
                  s.product_count <= sc::kMetaProducts && s.index_count <= sc::kMetaEntries
              ? 2
              : 1;
  if (*mode == 2) {
    const sc::PointsShape sh = sc::points_shape(static_cast<uint32_t>(s.product_count),
                                                static_cast<uint32_t>(s.max_degree));
    // the fold phase keeps (distinct factor, pair) lines of the block in LDS
    if (s.distinct * std::max<uint32_t>(1, sh.pairs_per_block) > sc::kMetaEntries) {
      *mode = 1;
      return static_cast<unsigned>(std::max<size_t>(1, (items + sc::kThreads - 1) / sc::kThreads));
    }
    const size_t blocks = sh.pairs_per_block ? (pairs + sh.pairs_per_block - 1) / sh.pairs_per_block
                                             : pairs * sh.blocks_per_pair;
    return static_cast<unsigned>(blocks);
  }
  return static_cast<unsigned>(std::max<size_t>(1, (items + sc::kThreads - 1) / sc::kThreads));
}

// Generic path (degree > kMaxJobDegree): message for t in chunks of 8 values.
void launch_message(rt::Context& c, sc::State& s, const Fr* const* tables, size_t pairs) {
  const unsigned blocks = message_blocks(pairs);
  const size_t width = s.max_degree + 1;
  Fr* partial = static_cast<Fr*>(s.partial->ptr);
  for (size_t t0 = 0; t0 < width; t0 += 8) {
    const uint32_t tc = static_cast<uint32_t>(std::min<size_t>(8, width - t0));
    hipLaunchKernelGGL(sc::k_message_chunk, dim3(blocks), dim3(sc::kThreads), 0, c.stream,
                       tables, s.prods, pairs, static_cast<uint32_t>(t0), tc, partial, width);
    rt::check_launch();
  }
  hipLaunchKernelGGL(sc::k_reduce_partials, dim3(1), dim3(64), 0, c.stream, partial,
                     (size_t)blocks, width, static_cast<Fr*>(s.result->ptr));
  rt::check_launch();
}

// Tables read and written by round k: round k folds into slot (k-1)%2 (slot A
// holds 2^(n-1) entries per factor, slot B 2^(n-2)) from the inputs (k = 1) or
// slot (k-2)%2.
void round_tables(const sc::State& s, size_t k, const Fr* const** src, Fr* const** dst) {
  *src = k <= 1 ? s.table_ptrs[0] : s.table_ptrs[1 + ((k - 2) & 1)];
  *dst = k == 0 ? nullptr : const_cast<Fr* const*>(s.table_ptrs[1 + ((k - 1) & 1)]);
}

sc::RoundJob make_job(const sc::State& s, const Fr& ch, Fr* out) {
  const size_t k = s.round;
  sc::RoundJob j{};
  round_tables(s, k, &j.src, &j.dst);
  j.coeffs = s.prods.coeffs;
  j.coeffs29 = s.coeffs29;
  j.offsets = s.prods.offsets;
  j.slots = s.prods.slots;
  j.owner = s.owner;
  j.partial = static_cast<Fr*>(s.partial->ptr);
  j.ticket = static_cast<unsigned*>(s.ticket->ptr);
  j.out = out;
  j.challenge = ch;
  j.pairs = static_cast<uint32_t>(size_t(1) << (s.num_vars - k - 1));
  j.products = static_cast<uint32_t>(s.product_count);
  j.factors = static_cast<uint32_t>(s.distinct);
  j.degree = static_cast<uint32_t>(s.max_degree);
  j.fold = (k > 0 && s.product_count > 0) ? 1 : 0;
  uint32_t mode = 0;
  j.blocks = job_blocks(s, k, &mode);
  j.items_mode = mode;
  j.classes = s.classes;
  j.class_k = s.class_k;
  j.class_count = s.class_count;
  j.class_width = s.class_width;
  j.class_sums = s.class_sums;
  j.class_stamp = static_cast<uint32_t>(k) + 1;
  j.entries = static_cast<uint32_t>(s.index_count);
  return j;
}

struct PackLaunch {
  uint32_t degree;
  uint32_t mode;  // 0 pairs, 1 items, 2 points
  unsigned grid;
  sc::JobPack pack;
};

// Group jobs (degree <= kMaxJobDegree) into launches: one per (degree,
// decomposition) and per kJobsPerLaunch jobs; sets each job's block_begin.
std::vector<PackLaunch> plan_packs(const std::vector<sc::RoundJob>& jobs) {
  auto key = [&](size_t i) {
    return std::make_pair(std::max<uint32_t>(jobs[i].degree, 1), jobs[i].items_mode);
  };
  std::vector<size_t> order(jobs.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return key(a) < key(b); });
  std::vector<PackLaunch> out;
  size_t i = 0;
  while (i < order.size()) {
    const auto k = key(order[i]);
    PackLaunch l{};
    l.degree = k.first;
    l.mode = k.second;
    while (i < order.size() && key(order[i]) == k && l.pack.count < (uint32_t)sc::kJobsPerLaunch) {
      sc::RoundJob j = jobs[order[i]];
      j.block_begin = l.grid;
      l.grid += j.blocks;
      l.pack.job[l.pack.count++] = j;
      ++i;
    }
    out.push_back(l);
  }
  return out;
}

// dev: the variant that reads its pack from device memory (round graphs)
template <int D>
const void* kernel_for(uint32_t mode, bool dev) {
  auto f = [](auto by_value, auto by_pointer, bool d) {
    return d ? reinterpret_cast<const void*>(by_pointer) : reinterpret_cast<const void*>(by_value);
  };
  if constexpr (D <= kMaxClassDegree) {
    if (mode == 3) return f(&sc::k_round_classes<D>, &sc::k_round_classes_dev<D>, dev);
  }
  return mode == 2   ? f(&sc::k_round_points<D>, &sc::k_round_points_dev<D>, dev)
         : mode == 1 ? f(&sc::k_round_items<D>, &sc::k_round_items_dev<D>, dev)
                     : f(&sc::k_round_pairs<D>, &sc::k_round_pairs_dev<D>, dev);
}

const void* kernel_of(uint32_t d, uint32_t mode, bool dev) {
  switch (d) {
    case 1: return kernel_for<1>(mode, dev);
    case 2: return kernel_for<2>(mode, dev);
    case 3: return kernel_for<3>(mode, dev);
    case 4: return kernel_for<4>(mode, dev);
    case 5: return kernel_for<5>(mode, dev);
    case 6: return kernel_for<6>(mode, dev);
    case 7: return kernel_for<7>(mode, dev);
    case 8: return kernel_for<8>(mode, dev);
    default: rt::fail_state(-3);
  }
}

// The second kernel of a launch (0 if none): the class finish, or the split
// finish when a job has several blocks; its grid in *grid.
const void* finish_of(const PackLaunch& l, bool dev, dim3* grid) {
  if (l.mode == 3) {
    uint32_t width = 0;
    for (uint32_t j = 0; j < l.pack.count; ++j) width = std::max(width, l.pack.job[j].class_width);
    *grid = dim3(width, l.pack.count);
    return dev ? reinterpret_cast<const void*>(&sc::k_finish_classes_dev)
               : reinterpret_cast<const void*>(&sc::k_finish_classes);
  }
  bool multi = false;
  for (uint32_t j = 0; j < l.pack.count; ++j) multi = multi || l.pack.job[j].blocks > 1;
  if (!multi) return nullptr;
  *grid = dim3(l.pack.count);
  return dev ? reinterpret_cast<const void*>(&sc::k_finish_split_dev)
             : reinterpret_cast<const void*>(&sc::k_finish_split);
}

// Split launches: a round kernel whose multi-block jobs only store their
// block partials, then k_finish_split (one block per job) adds them.  The
// ticket path's two device-scope fences (each an L2 write-back on this
// multi-XCD part) cost more than the second launch.
void launch_jobs(rt::Context& c, std::vector<sc::RoundJob>& jobs) {
  for (PackLaunch& l : plan_packs(jobs)) {
    l.pack.split = 1;
    void* args[] = {&l.pack};
    rt::hip_check(hipLaunchKernel(kernel_of(l.degree, l.mode, false), dim3(l.grid), dim3(sc::kThreads),
                                  args, 0, c.stream));
    dim3 fgrid;
    if (const void* fin = finish_of(l, false, &fgrid))
      rt::hip_check(hipLaunchKernel(fin, fgrid, dim3(sc::kThreads), args, 0, c.stream));
  }
}

// ---- persistent round loops: host side -------------------------------------
//
// A context keeps a pool of mailboxes and the list of states whose blocks may
// still wait for a challenge.  At most one round-loop launch is active per
// context: a new one (or any other work on the context stream) first stops
// the running one (abort flags, then a stream drain).
struct TailPool {
  std::vector<sc::TailMailbox*> chunks;  // pinned blocks of kChunk mailboxes
  std::vector<std::pair<sc::TailMailbox*, sc::TailMailbox*>> free;  // host, device
  std::vector<sc::State*> active;
  sc::TailMailbox** ptrs = nullptr;  // pinned: device mailbox pointers of a launch
  sc::TailMailbox** dptrs = nullptr;
  size_t ptrs_cap = 0;
  uint64_t wait_ticks = 0;
  // wide loops: the partial words (device memory, allocated and zeroed at the
  // first wide launch) and the stamp base of the next launch (+64 a launch:
  // a launch covers at most 32 rounds, so no stamp repeats until the counter
  // wraps, when the words are zeroed again)
  void* wide_mem = nullptr;
  uint32_t wide_stamp = 64;
  uint32_t cus_log2 = 0;  // a wide loop's blocks must all be resident: at most one per CU
  static constexpr size_t kChunk = 64;
};

// The device waits this long for the host's next challenge before its block
// stops (the host then relaunches from its round); HP_MI355X_SC_TAIL_WAIT_US.
uint64_t tail_wait_us() {
  static const uint64_t v = [] {
    const char* e = std::getenv("HP_MI355X_SC_TAIL_WAIT_US");
    return e ? static_cast<uint64_t>(std::strtoull(e, nullptr, 0)) : uint64_t(100000);
  }();
  return v;
}

// A state runs its rounds in the loop once a round has at most this many
// (product, pair, point) items (and its (distinct factor, pair) lines fit the
// block's LDS): HP_MI355X_SC_TAIL_ITEMS for rounds of several states at once
// (round_batch; default 4096: measured ~5 % faster on the 64-sumcheck batch),
// HP_MI355X_SC_TAIL_SINGLE_ITEMS for one state's rounds (default 16384: at
// n = 20, degree 2, rounds 12-19 run in the loop at ~9 us each against
// 21-29 us per launched round, the whole sumcheck 0.97 -> 0.88 ms); 0 turns
// the loop off.
// A single state whose rounds do not fit one block yet runs them in a wide
// loop (one block per CU, then fewer) once a round has at most this many
// (product, pair, point) items: HP_MI355X_SC_WIDE_ITEMS (0 turns it off), with
// at most HP_MI355X_SC_WIDE_BLOCKS blocks (a power of two, <= 256).  Measured
// at n = 20: at 2^19 the degree-2 loop starts at round 4 (round 3 in the loop
// 71 us against 61 us launched), the degree-5 one at round 6.
size_t wide_items() {
  static const size_t v = [] {
    const char* e = std::getenv("HP_MI355X_SC_WIDE_ITEMS");
    return e ? static_cast<size_t>(std::strtoull(e, nullptr, 0)) : size_t(1) << 19;
  }();
  return v;
}

// A wide loop hands its rounds to one block once a round's factor products
// (pairs x points x the products' total arity) are at most this many,
// HP_MI355X_SC_WIDE_TAIL_WORK: one CU takes about 4 us per 1000 of them, a
// wide round costs ~9 us more than a one-block round in exchanges.
size_t wide_tail_work() {
  static const size_t v = [] {
    const char* e = std::getenv("HP_MI355X_SC_WIDE_TAIL_WORK");
    return e ? static_cast<size_t>(std::strtoull(e, nullptr, 0)) : size_t(2048);
  }();
  return v;
}

uint32_t wide_blocks_log2() {
  static const uint32_t v = [] {
    const char* e = std::getenv("HP_MI355X_SC_WIDE_BLOCKS");
    size_t b = e ? static_cast<size_t>(std::strtoull(e, nullptr, 0)) : sc::kWideMaxBlocks;
    b = std::min<size_t>(std::max<size_t>(b, 2), sc::kWideMaxBlocks);
    uint32_t l = 0;
    while ((size_t(2) << l) <= b) ++l;
    return l;
  }();
  return v;
}

size_t tail_items(bool batch) {
  static const size_t batch_v = [] {
    const char* e = std::getenv("HP_MI355X_SC_TAIL_ITEMS");
    return e ? static_cast<size_t>(std::strtoull(e, nullptr, 0)) : size_t(4096);
  }();
  static const size_t single_v = [] {
    const char* e = std::getenv("HP_MI355X_SC_TAIL_SINGLE_ITEMS");
    return e ? static_cast<size_t>(std::strtoull(e, nullptr, 0)) : size_t(16384);
  }();
  return batch ? batch_v : single_v;
}

constexpr size_t kTailMaxStates = 128;

// the state's metadata fits a loop block's LDS
bool loop_meta_ok(const sc::State& s) {
  return s.max_degree <= kMaxJobDegree && s.product_count > 0 && s.product_count <= sc::kTailMaxProducts &&
         s.index_count <= sc::kTailMaxEntries && s.distinct <= sc::kTailMaxFactors;
}

// round k of the state fits the one-block loop
bool tail_ok_at(const sc::State& s, size_t k, bool batch) {
  if (!loop_meta_ok(s) || k >= s.num_vars) return false;
  const size_t pairs = size_t(1) << (s.num_vars - k - 1);
  if (s.distinct * pairs > sc::kTailMaxLines) return false;
  return pairs * s.product_count * (s.max_degree + 1) <= tail_items(batch);
}

bool tail_ok(const sc::State& s, bool batch) { return tail_ok_at(s, s.round, batch); }

// the state's current round starts (or continues) a wide loop
// States of degree 5 and more enter the wide loop at four times the item
// count: their rounds carry more work per pair, so the loop's exchanges weigh
// less (n = 20 degree 5 from round 4 instead of 6: 2.02-2.04 -> 1.98-2.01 ms;
// degree 2 from round 3 instead of 4 was 8 % slower, so it keeps the base).
bool wide_ok(const sc::State& s) {
  if (!loop_meta_ok(s) || s.round >= s.num_vars || wide_items() == 0) return false;
  const size_t pairs = size_t(1) << (s.num_vars - s.round - 1);
  const size_t limit = wide_items() * (s.max_degree >= 5 ? 4 : 1);
  return pairs * s.product_count * (s.max_degree + 1) <= limit;
}

void quiesce_tails(rt::Context& c) {
  auto* pool = static_cast<TailPool*>(c.tails);
  if (pool == nullptr || pool->active.empty()) return;
  for (sc::State* s : pool->active) __atomic_store_n(&s->box->post[sc::kTailAbortWord], 1u, __ATOMIC_RELEASE);
  for (sc::State* s : pool->active) s->tail_active = false;
  pool->active.clear();
  rt::hip_check(hipStreamSynchronize(c.stream));  // every block has seen its abort and exited
}

void release_tail_pool(rt::Context& c) {
  auto* pool = static_cast<TailPool*>(c.tails);
  if (pool == nullptr) return;
  for (sc::TailMailbox* m : pool->chunks) (void)hipHostFree(m);
  if (pool->ptrs) (void)hipHostFree(pool->ptrs);
  if (pool->wide_mem) (void)hipFree(pool->wide_mem);
  delete pool;
  c.tails = nullptr;
  c.quiesce = nullptr;
  c.release_tails = nullptr;
}

TailPool& tail_pool(rt::Context& c) {
  if (c.tails == nullptr) {
    auto* pool = new TailPool();
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c.device) != hipSuccess || khz <= 0) {
      (void)hipGetLastError();
      khz = 100000;  // 100 MHz
    }
    pool->wait_ticks = tail_wait_us() * static_cast<uint64_t>(khz) / 1000;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess || cus <= 0) {
      (void)hipGetLastError();
      cus = 2;
    }
    while ((2 << pool->cus_log2) <= cus) ++pool->cus_log2;
    c.tails = pool;
    c.quiesce = &quiesce_tails;
    c.release_tails = &release_tail_pool;
  }
  return *static_cast<TailPool*>(c.tails);
}

void tail_mailbox(TailPool& pool, sc::State& s) {
  if (s.box) return;
  if (pool.free.empty()) {
    void* raw = nullptr;
    rt::hip_check(hipHostMalloc(&raw, TailPool::kChunk * sizeof(sc::TailMailbox),
                                hipHostMallocCoherent | hipHostMallocMapped));
    pool.chunks.push_back(static_cast<sc::TailMailbox*>(raw));
    void* draw = nullptr;
    rt::hip_check(hipHostGetDevicePointer(&draw, raw, 0));
    auto* h = static_cast<sc::TailMailbox*>(raw);
    auto* d = static_cast<sc::TailMailbox*>(draw);
    for (size_t i = 0; i < TailPool::kChunk; ++i) pool.free.emplace_back(h + i, d + i);
  }
  s.box = pool.free.back().first;
  s.dbox = pool.free.back().second;
  pool.free.pop_back();
}

void tail_return_mailbox(rt::Context& c, sc::State& s) {
  if (s.box == nullptr) return;
  auto* pool = static_cast<TailPool*>(c.tails);
  if (pool) pool->free.emplace_back(s.box, s.dbox);
  s.box = s.dbox = nullptr;
}

void tail_ptrs(TailPool& pool, size_t count) {
  if (pool.ptrs_cap < count) {
    if (pool.ptrs) rt::hip_check(hipHostFree(pool.ptrs));
    pool.ptrs = nullptr;
    pool.ptrs_cap = 0;
    const size_t cap = std::max<size_t>(count, 64);
    void* raw = nullptr;
    rt::hip_check(hipHostMalloc(&raw, cap * sizeof(void*), hipHostMallocCoherent | hipHostMallocMapped));
    pool.ptrs = static_cast<sc::TailMailbox**>(raw);
    void* draw = nullptr;
    rt::hip_check(hipHostGetDevicePointer(&draw, raw, 0));
    pool.dptrs = static_cast<sc::TailMailbox**>(draw);
    pool.ptrs_cap = cap;
  }
}

// The state's loop job from its current round with the challenge (stored
// form; ignored at round 0); the mailbox cleared for a new launch.
void tail_job(TailPool& pool, sc::State& s, const Fr& challenge) {
  tail_mailbox(pool, s);
  sc::TailJob& j = s.box->job;
  for (int t = 0; t < 3; ++t) j.tables[t] = s.table_ptrs[t];
  j.coeffs29 = s.coeffs29;
  j.offsets = s.prods.offsets;
  j.slots = s.prods.slots;
  j.owner = s.owner;
  j.products = static_cast<uint32_t>(s.product_count);
  j.degree = static_cast<uint32_t>(s.max_degree);
  j.num_vars = static_cast<uint32_t>(s.num_vars);
  j.first_round = static_cast<uint32_t>(s.round);
  j.entries = static_cast<uint32_t>(s.index_count);
  j.factors = static_cast<uint32_t>(s.distinct);
  const f29::Fr29 c29 = f29::prescale_fr(challenge);
  std::memcpy(j.challenge29, c29.l, sizeof j.challenge29);
  for (uint32_t& w : s.box->post) w = 0;  // no challenge (tag 0), no abort
  for (uint64_t& w : s.box->message) w = 0;  // no round's stamp
  s.box->status = sc::kTailRunning;
}

constexpr uint32_t kTailSpinCap = 50000000u;

// One launch, one block per state, from each state's current round with its
// challenge.  Precondition: no active loop.
void launch_tails(rt::Context& c, TailPool& pool, const std::vector<sc::State*>& states,
                  const std::vector<Fr>& challenges) {
  tail_ptrs(pool, states.size());
  for (size_t i = 0; i < states.size(); ++i) {
    tail_job(pool, *states[i], challenges[i]);
    pool.ptrs[i] = states[i]->dbox;
  }
  __atomic_thread_fence(__ATOMIC_SEQ_CST);
  hipLaunchKernelGGL(sc::k_round_tail, dim3(static_cast<unsigned>(states.size())), dim3(sc::kTailThreads),
                     0, c.stream, pool.dptrs, pool.wait_ticks, kTailSpinCap, sc::TailStage{},
                     sc::WideArea{nullptr, nullptr});
  rt::check_launch();
  for (sc::State* s : states) {
    s->tail_active = true;
    pool.active.push_back(s);
  }
}

// A wide launch of one state from its current round, in up to three stages:
// G blocks (the most the CUs, the maximum and 32 pairs a block in the first
// round allow) while every block keeps a pair, then, if that stops short of the one-block loop,
// as many blocks as the round before the one-block loop has pairs, then one
// block.  Precondition: no active loop, wide_ok(s), !tail_ok(s).
void launch_wide(rt::Context& c, TailPool& pool, sc::State& s, const Fr& challenge) {
  const size_t n = s.num_vars;
  size_t tail_start = s.round;  // n: the one-block loop is switched off
  auto work = [&](size_t k) { return (size_t(1) << (n - k - 1)) * s.index_count * (s.max_degree + 1); };
  while (tail_start < n && !(tail_ok_at(s, tail_start, false) && work(tail_start) <= wide_tail_work()))
    ++tail_start;
  struct Stage {
    uint32_t first, end, blocks_log2;
  };
  Stage st[3];
  int count = 0;
  // at least 2^kWideMinPairsLog2 pairs a block in the first round: with fewer,
  // the blocks' exchanges outweigh their work (nv-12 proofs: 64 blocks from
  // round 0, 2.92-2.96 ms, against 256 blocks of 8 pairs, 2.98-3.01 ms; the
  // n = 20 loops start with 2^15 pairs or more and keep 256 blocks)
  constexpr uint32_t kWideMinPairsLog2 = 5;
  const uint32_t pairs_log2 = static_cast<uint32_t>(n - s.round - 1);
  const uint32_t b1 = std::min<uint32_t>(
      std::min(wide_blocks_log2(), pool.cus_log2),
      pairs_log2 > kWideMinPairsLog2 ? pairs_log2 - kWideMinPairsLog2 : std::min<uint32_t>(pairs_log2, 1));
  const size_t end1 = std::min<size_t>(tail_start, n - b1);  // round k has 2^(n-k-1) pairs
  st[count++] = {static_cast<uint32_t>(s.round), static_cast<uint32_t>(end1), b1};
  if (end1 < tail_start)
    st[count++] = {static_cast<uint32_t>(end1), static_cast<uint32_t>(tail_start),
                   static_cast<uint32_t>(n - tail_start)};
  if (tail_start < n) st[count++] = {static_cast<uint32_t>(tail_start), static_cast<uint32_t>(n), 0};
  const size_t part_bytes = (size_t(sc::kWideMaxBlocks) * sc::kTailMaxWidth * 8 + 16) * sizeof(uint64_t);
  bool zero = false;
  if (pool.wide_mem == nullptr) {
    rt::hip_check(hipMalloc(&pool.wide_mem, part_bytes));
    zero = true;
  }
  if (pool.wide_stamp > UINT32_MAX - 64) {  // the stamps would wrap: start again from zeroed words
    pool.wide_stamp = 64;
    zero = true;
  }
  if (zero) rt::hip_check(hipMemsetAsync(pool.wide_mem, 0, part_bytes, c.stream));
  const uint32_t stamp_base = pool.wide_stamp;
  pool.wide_stamp += 64;
  const sc::WideArea area{static_cast<uint64_t*>(pool.wide_mem),
                          static_cast<uint64_t*>(pool.wide_mem) + size_t(sc::kWideMaxBlocks) * sc::kTailMaxWidth * 8};
  tail_ptrs(pool, 1);
  tail_job(pool, s, challenge);
  pool.ptrs[0] = s.dbox;
  sc::TailStage stage{};
  for (int i = 0; i < 3; ++i) {  // unused stages are empty: [n, n)
    stage.end[i] = i < count ? st[i].end : static_cast<uint32_t>(n);
    stage.blocks_log2[i] = i < count ? st[i].blocks_log2 : 0;
  }
  stage.stamp_base = stamp_base;
  stage.wide = 1;
  __atomic_thread_fence(__ATOMIC_SEQ_CST);
  hipLaunchKernelGGL(sc::k_round_tail, dim3(1u << st[0].blocks_log2), dim3(sc::kTailThreads), 0, c.stream,
                     pool.dptrs, pool.wait_ticks, kTailSpinCap, stage, area);
  rt::check_launch();
  s.tail_active = true;
  pool.active.push_back(&s);
}

void tail_post(sc::State& s, const Fr& challenge) {
  const f29::Fr29 c29 = f29::prescale_fr(challenge);
  const uint32_t tag = sc::tail_tag(static_cast<uint32_t>(s.round)) << 29;
  for (int i = 0; i < f29::FrP::L; ++i) __atomic_store_n(&s.box->post[i], c29.l[i] | tag, __ATOMIC_RELEASE);
}

// Wait for the message of the state's current round: true with it in `out`
// (width canonical Fr), false if the block stopped first (wait limit or abort).
// A word is taken once it carries the round's stamp; it is not written again
// before the next challenge is posted.
bool tail_wait(rt::Context& c, sc::State& s, uint8_t* out) {
  const uint64_t want = static_cast<uint64_t>(s.round + 1) << 32;
  const size_t words = (s.max_degree + 1) * 8;
  uint32_t w[sc::kTailMaxWidth * 8];
  size_t got = 0;
  auto take = [&] {
    while (got < words) {
      const uint64_t x = __atomic_load_n(&s.box->message[got], __ATOMIC_ACQUIRE);
      if ((x & ~uint64_t(0xffffffff)) != want) return false;
      w[got++] = static_cast<uint32_t>(x);
    }
    return true;
  };
  const auto start = std::chrono::steady_clock::now();
  for (uint32_t spin = 0;; ++spin) {
    if (take()) break;
    const uint32_t st = __atomic_load_n(&s.box->status, __ATOMIC_ACQUIRE);
    if (st == sc::kTailTimedOut || st == sc::kTailAborted) {
      // the block may have published this round just before stopping
      if (take()) break;
      return false;
    }
    if ((spin & 1023u) == 1023u) {
      if (std::chrono::steady_clock::now() - start > std::chrono::seconds(60)) {
        // the device never answered: stop it and report a device failure
        quiesce_tails(c);
        rt::fail_state(-4);
      }
      std::this_thread::yield();
    } else {
      __builtin_ia32_pause();
    }
  }
  // stored (Montgomery) form -> canonical
  for (size_t t = 0; t <= s.max_degree; ++t) {
    Fr v;
    std::memcpy(v.v, w + 8 * t, rt::kFrBytes);
    v = from_mont(v);
    std::memcpy(out + rt::kFrBytes * t, v.v, rt::kFrBytes);
  }
  return true;
}

// After a state's round came back through its loop.
void tail_advance(TailPool& pool, sc::State& s) {
  s.round += 1;
  if (s.round == s.num_vars) {  // its block returned after the last message
    s.tail_active = false;
    pool.active.erase(std::remove(pool.active.begin(), pool.active.end(), &s), pool.active.end());
  }
}

// Rounds of `states` (all tail_ok, or one wide_ok state) through the loop:
// post or launch, wait.
// Returns false (nothing advanced, no loop active) when a block stopped
// before answering twice in a row; the caller then uses the launch path.
bool tail_rounds(rt::Context& c, const std::vector<sc::State*>& states, const std::vector<Fr>& chs,
                 const std::vector<uint8_t*>& outs) {
  TailPool& pool = tail_pool(c);
  bool launch = false;
  for (sc::State* s : states) launch = launch || !s->tail_active;
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (launch) {
      quiesce_tails(c);
      if (states.size() == 1 && !tail_ok(*states[0], false))
        launch_wide(c, pool, *states[0], chs[0]);
      else
        launch_tails(c, pool, states, chs);
    } else {
      for (size_t i = 0; i < states.size(); ++i) tail_post(*states[i], chs[i]);
      // the runtime's error state, checked where a launched round checks its
      // launch (a loop round makes no other HIP call)
      rt::check_launch();
    }
    bool all = true;
    for (size_t i = 0; i < states.size() && all; ++i) all = tail_wait(c, *states[i], outs[i]);
    // and where a launched round checks its synchronisation
    if (all) rt::check_launch();
    if (all) {
      for (sc::State* s : states) tail_advance(pool, *s);
      return true;
    }
    launch = true;  // a block stopped (wait limit): relaunch every state once
  }
  quiesce_tails(c);
  return false;
}

// ---- round launches as hipGraphs, replayed with new challenges ----------
//
// A batch round whose launch shapes (per launch: degree class, decomposition,
// grid, jobs and their block counts and class widths) were seen before
// replays the graph built for them, with no node updates: the graph's first
// node copies this round's job packs (tables, challenges, stamps) from its
// pinned staging buffer into device memory, its kernels (the *_dev variants)
// read their pack there, the same launches as launch_jobs (round kernel, then
// the class or split finish), and the finishes write the messages straight
// into the graph's pinned buffer.  Graphs hold no memset nodes; the per-job
// tickets are reset by the kernels.
struct RoundGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint8_t* host = nullptr;  // pinned: the messages
  size_t bytes = 0;
  rt::BufferRef packs;               // the launches' packs, device
  sc::JobPack* packs_host = nullptr;  // pinned staging, copied by the first node
  std::vector<const sc::JobPack*> pack_ptrs;  // kernel arguments (device addresses)
  ~RoundGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (host) (void)hipHostFree(host);
    if (packs_host) (void)hipHostFree(packs_host);
  }
};

struct RoundGraphCache {
  std::map<std::vector<uint32_t>, std::unique_ptr<RoundGraph>> graphs;
  std::deque<std::vector<uint32_t>> order;  // oldest first
  static constexpr size_t kMax = 512;
};

int g_round_graph_key;

// Off by default (HP_MI355X_ROUND_GRAPHS=1 turns it on): the launched rounds
// are the large ones, whose kernels outlast their launches, and the small
// rounds run in the persistent round loops.
bool round_graphs_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("HP_MI355X_ROUND_GRAPHS");
    return e != nullptr && std::atoi(e) != 0;
  }();
  return on;
}

RoundGraphCache& graph_cache(rt::Context& c) {
  auto it = c.resources.find(&g_round_graph_key);
  if (it == c.resources.end())
    it = c.resources
             .emplace(&g_round_graph_key,
                      rt::Context::Resource{rt::Context::Kind::internal,
                                            std::make_shared<RoundGraphCache>()})
             .first;
  return *static_cast<RoundGraphCache*>(it->second.object.get());
}

// Run the jobs (outputs at offsets out_offset[i] of a total_bytes buffer)
// through a cached or new graph and copy the messages to `host_out`.
void run_jobs_graph(rt::Context& c, std::vector<sc::RoundJob>& jobs,
                    const std::vector<size_t>& out_offset, size_t total_bytes, uint8_t* host_out) {
  // launch shapes, independent of addresses
  std::vector<uint32_t> key;
  {
    std::vector<PackLaunch> shape = plan_packs(jobs);
    for (const PackLaunch& l : shape) {
      key.push_back(l.degree);
      key.push_back(l.mode);
      key.push_back(l.grid);
      key.push_back(l.pack.count);
      for (uint32_t j = 0; j < l.pack.count; ++j) {
        key.push_back(l.pack.job[j].degree);
        key.push_back(l.pack.job[j].blocks);
        key.push_back(l.pack.job[j].class_width);
      }
    }
    key.push_back(static_cast<uint32_t>(total_bytes));
  }
  RoundGraphCache& cache = graph_cache(c);
  auto it = cache.graphs.find(key);
  RoundGraph* g = it == cache.graphs.end() ? nullptr : it->second.get();
  bool fresh = false;
  if (g == nullptr) {
    if (cache.graphs.size() >= RoundGraphCache::kMax) {
      cache.graphs.erase(cache.order.front());
      cache.order.pop_front();
    }
    auto ng = std::make_unique<RoundGraph>();
    ng->bytes = total_bytes;
    rt::hip_check(hipHostMalloc(reinterpret_cast<void**>(&ng->host), total_bytes, hipHostMallocDefault));
    g = ng.get();
    cache.graphs.emplace(key, std::move(ng));
    cache.order.push_back(key);
    fresh = true;
  }
  Fr* out_base = reinterpret_cast<Fr*>(g->host);
  for (size_t i = 0; i < jobs.size(); ++i) jobs[i].out = out_base + out_offset[i];
  std::vector<PackLaunch> packs = plan_packs(jobs);
  for (PackLaunch& l : packs) l.pack.split = 1;
  const size_t pack_bytes = packs.size() * sizeof(sc::JobPack);
  if (fresh) {
    g->packs = c.allocate(pack_bytes);
    rt::hip_check(hipHostMalloc(reinterpret_cast<void**>(&g->packs_host), pack_bytes, hipHostMallocDefault));
    const sc::JobPack* dev = static_cast<const sc::JobPack*>(g->packs->ptr);
    for (size_t k = 0; k < packs.size(); ++k) g->pack_ptrs.push_back(dev + k);
    rt::hip_check(hipGraphCreate(&g->graph, 0));
    hipGraphNode_t prev;
    rt::hip_check(hipGraphAddMemcpyNode1D(&prev, g->graph, nullptr, 0, g->packs->ptr, g->packs_host,
                                          pack_bytes, hipMemcpyHostToDevice));
    auto add_kernel = [&](const void* f, dim3 grid, const sc::JobPack** arg) {
      void* args[] = {arg};
      hipKernelNodeParams p{};
      p.func = const_cast<void*>(f);
      p.gridDim = grid;
      p.blockDim = dim3(sc::kThreads);
      p.sharedMemBytes = 0;
      p.kernelParams = args;
      p.extra = nullptr;
      hipGraphNode_t node;
      rt::hip_check(hipGraphAddKernelNode(&node, g->graph, &prev, 1, &p));
      prev = node;
    };
    for (size_t k = 0; k < packs.size(); ++k) {
      const PackLaunch& l = packs[k];
      add_kernel(kernel_of(l.degree, l.mode, true), dim3(l.grid), &g->pack_ptrs[k]);
      dim3 fgrid;
      if (const void* fin = finish_of(l, true, &fgrid)) add_kernel(fin, fgrid, &g->pack_ptrs[k]);
    }
    rt::hip_check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  }
  // the previous replay of this graph has ended (each call synchronises), so
  // its staging buffer is free
  for (size_t k = 0; k < packs.size(); ++k) g->packs_host[k] = packs[k].pack;
  rt::hip_check(hipGraphLaunch(g->exec, c.stream));
  rt::hip_check(hipStreamSynchronize(c.stream));
  std::memcpy(host_out, g->host, total_bytes);
}

// Checks of one round call before any side effect (bridge_sumcheck.cpp:519-550).
HpMi355xStatus check_round(HpMi355xContext* ctx, rt::Context& c, HpMi355xSumCheckState* state,
                           const uint8_t* challenge, const uint8_t* out, size_t count) {
  // membership first: nothing is read through a handle this context does not hold
  if (state == nullptr || !c.holds(state, rt::Context::Kind::sumcheck_state))
    return rt::invalid_parameters();
  if (state->ctx != ctx) return rt::invalid_parameters();
  const sc::State& s = *state->impl;
  if (s.failed) return rt::invalid_state(s.failure_code);
  if (out == nullptr) return rt::invalid_parameters();
  if (count != s.max_degree + 1) return rt::invalid_parameters();
  if (s.round >= s.num_vars) return rt::invalid_state();
  if (s.round == 0) {
    if (challenge != nullptr) return rt::invalid_state();
  } else {
    if (challenge == nullptr) return rt::invalid_state();
    if (!rt::canonical_fr(challenge)) return rt::invalid_parameters();
  }
  return rt::ok();
}

Fr challenge_of(const sc::State& s, const uint8_t* bytes) {
  Fr ch = Fr::zero();
  if (s.round > 0) {
    std::memcpy(ch.v, bytes, rt::kFrBytes);
    ch = to_mont(ch);
  }
  return ch;
}

// Generic round of one state of degree > kMaxJobDegree into s.result.
void generic_round(rt::Context& c, sc::State& s, const Fr& ch) {
  const size_t k = s.round;
  const size_t pairs = size_t(1) << (s.num_vars - k - 1);
  if (s.product_count == 0) {
    hipLaunchKernelGGL(sc::k_zero_fr, dim3(1), dim3(64), 0, c.stream,
                       static_cast<Fr*>(s.result->ptr), s.max_degree + 1);
    rt::check_launch();
    return;
  }
  const Fr* const* src;
  Fr* const* dst;
  round_tables(s, k, &src, &dst);
  const Fr* const* tables = src;
  if (k > 0) {
    const size_t half = 2 * pairs;
    const size_t total = s.distinct * half;
    unsigned blocks =
        static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 65535)));
    hipLaunchKernelGGL(sc::k_fold_factors, dim3(blocks), dim3(256), 0, c.stream, src, dst,
                       s.distinct, ch, half);
    rt::check_launch();
    tables = dst;
  }
  launch_message(c, s, tables, pairs);
}

void mark_failed(sc::State& s, const rt::Failure& f) {
  if (f.poison) {
    s.failed = true;
    s.failure_code = f.status.native_code ? f.status.native_code : -1;
  }
}

}  // namespace

extern "C" HpMi355xStatus hp_mi355x_v1_sumcheck_prover_state_create(
    HpMi355xContext* ctx, size_t num_vars, size_t max_degree,
    const HpMi355xDeviceAddress* factors, size_t factor_count, const uint8_t* coefficients,
    const size_t* product_offsets, size_t product_count, const size_t* factor_indices,
    size_t index_count, HpMi355xSumCheckState** out) {
  if (out) *out = nullptr;
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out == nullptr) return rt::invalid_parameters();
    // ---- pure validation (no device side effects) ----
    if (num_vars == 0 || max_degree == 0) return rt::invalid_parameters();
    if (num_vars >= rt::kMaxNumVars) return rt::invalid_parameters();
    if (max_degree > 4096) return rt::invalid_parameters();
    if (product_count > (size_t(1) << 24) || index_count > (size_t(1) << 26))
      return rt::invalid_parameters();
    if (product_offsets == nullptr || product_offsets[0] != 0) return rt::invalid_parameters();
    for (size_t p = 0; p < product_count; ++p) {
      const size_t b = product_offsets[p], e = product_offsets[p + 1];
      if (e < b) return rt::invalid_parameters();
      if (e - b == 0 || e - b > max_degree) return rt::invalid_parameters();
    }
    if (product_offsets[product_count] != index_count) return rt::invalid_parameters();
    if (factor_count > 0 && factors == nullptr) return rt::invalid_parameters();
    if (index_count > 0 && factor_indices == nullptr) return rt::invalid_parameters();
    if (product_count > 0) {
      if (coefficients == nullptr) return rt::invalid_parameters();
      if (!rt::canonical_fr_array(coefficients, product_count)) return rt::invalid_parameters();
    }
    for (size_t i = 0; i < index_count; ++i)
      if (factor_indices[i] >= factor_count) return rt::invalid_parameters();
    // stable dedup of factor handles by address; repeats keep multiplicity
    std::vector<rt::BufferRef> distinct;
    std::vector<uint32_t> slot_of(factor_count);
    std::unordered_map<void*, uint32_t> seen;
    for (size_t i = 0; i < factor_count; ++i) {
      void* a = factors[i];
      if (a == nullptr) return rt::invalid_parameters();
      auto it = seen.find(a);
      if (it != seen.end()) {
        slot_of[i] = it->second;
        continue;
      }
      rt::BufferRef ref = c.lookup(a, num_vars);
      if (!ref) return rt::invalid_parameters();
      slot_of[i] = static_cast<uint32_t>(distinct.size());
      seen.emplace(a, slot_of[i]);
      distinct.push_back(std::move(ref));
    }

    // ---- resources ----
    auto s = std::make_shared<sc::State>();
    s->num_vars = num_vars;
    s->max_degree = max_degree;
    s->inputs = distinct;
    s->distinct = distinct.size();
    s->product_count = product_count;
    s->index_count = index_count;
    const size_t F = distinct.size();
    if (F > 0 && num_vars >= 2) {
      const size_t a = size_t(1) << (num_vars - 1), b = size_t(1) << (num_vars - 2);
      s->slots = c.allocate(F * (a + b) * rt::kFrBytes);
      s->slot_base[0] = static_cast<Fr*>(s->slots->ptr);
      s->slot_base[1] = s->slot_base[0] + F * a;
    }
    // metadata block: coeffs (Fr) | coeffs29 (Fr) | table ptrs x3 (F each) | offsets | slots |
    // owner
    const size_t coeff_bytes = std::max<size_t>(product_count, 1) * rt::kFrBytes;
    const size_t ptr_bytes = 3 * std::max<size_t>(F, 1) * sizeof(void*);
    const size_t off_bytes = (product_count + 1) * sizeof(uint32_t);
    const size_t idx_bytes = std::max<size_t>(index_count, 1) * sizeof(uint32_t);
    const size_t own_bytes = std::max<size_t>(index_count, 1);
    const size_t cb = 2 * coeff_bytes;
    // product classes (equal arity, equal coefficient) in order of first
    // appearance, for the class-major decomposition of large tables; none
    // past kMaxClasses classes or kMaxClassWidth values (m + 1 per class)
    std::vector<uint32_t> cls_m, cls_first;
    std::vector<std::vector<uint32_t>> cls_members;
    bool use_classes = product_count > 0 && max_degree <= kMaxClassDegree;
    uint32_t cls_width = 0;
    for (size_t p = 0; p < product_count && use_classes; ++p) {
      const uint32_t m = static_cast<uint32_t>(product_offsets[p + 1] - product_offsets[p]);
      size_t k = 0;
      while (k < cls_m.size() &&
             !(cls_m[k] == m && std::memcmp(coefficients + rt::kFrBytes * cls_first[k],
                                            coefficients + rt::kFrBytes * p, rt::kFrBytes) == 0))
        ++k;
      if (k == cls_m.size()) {
        if (k == sc::kMaxClasses || cls_width + m + 1 > sc::kMaxClassWidth) {
          use_classes = false;
          break;
        }
        cls_m.push_back(m);
        cls_first.push_back(static_cast<uint32_t>(p));
        cls_members.emplace_back();
        cls_width += m + 1;
      }
      cls_members[k].push_back(static_cast<uint32_t>(p));
    }
    const size_t C = use_classes ? cls_m.size() : 0;
    const size_t base_bytes = cb + ptr_bytes + off_bytes + idx_bytes + own_bytes;
    const size_t cls_base = (base_bytes + 31) / 32 * 32;  // class coefficients | table | members
    const size_t meta_bytes =
        C ? cls_base + C * rt::kFrBytes + (4 * C + product_count) * sizeof(uint32_t) : base_bytes;
    s->meta = c.allocate(meta_bytes);
    std::vector<uint8_t> host(meta_bytes, 0);
    Fr k5;
    std::memcpy(k5.v, f29::FrConsts::K5, sizeof k5.v);
    if (C) {
      uint32_t* tab = reinterpret_cast<uint32_t*>(host.data() + cls_base + C * rt::kFrBytes);
      uint32_t* mem = tab + 4 * C;
      uint32_t off = 0, at = 0;
      for (size_t k = 0; k < C; ++k) {
        Fr v;
        std::memcpy(v.v, coefficients + rt::kFrBytes * cls_first[k], rt::kFrBytes);
        v = to_mont(v);
        // times 2^(5 (m - 1)): a product of m stored values over m - 1 Fr29
        // products is prod(a) 2^(256 - 5 (m - 1))
        for (uint32_t e = 1; e < cls_m[k]; ++e) v = v * k5;
        std::memcpy(host.data() + cls_base + rt::kFrBytes * k, v.v, rt::kFrBytes);
        tab[4 * k] = cls_m[k];
        tab[4 * k + 1] = off;
        tab[4 * k + 2] = at;
        for (uint32_t p : cls_members[k]) mem[at++] = p;
        tab[4 * k + 3] = at;
        off += cls_m[k] + 1;
      }
    }
    for (size_t p = 0; p < product_count; ++p) {
      Fr v;
      std::memcpy(v.v, coefficients + rt::kFrBytes * p, rt::kFrBytes);
      v = to_mont(v);
      std::memcpy(host.data() + rt::kFrBytes * p, v.v, rt::kFrBytes);
      // times 2^(5 m) for the m Fr29 products of the product's chain
      for (size_t e = product_offsets[p]; e < product_offsets[p + 1]; ++e) v = v * k5;
      std::memcpy(host.data() + coeff_bytes + rt::kFrBytes * p, v.v, rt::kFrBytes);
    }
    uint8_t* dev = static_cast<uint8_t*>(s->meta->ptr);
    std::vector<const void*> ptrs(3 * std::max<size_t>(F, 1), nullptr);
    const size_t a_len = num_vars >= 1 ? (size_t(1) << (num_vars - 1)) : 0;
    for (size_t f = 0; f < F; ++f) {
      ptrs[f] = distinct[f]->ptr;
      if (s->slot_base[0]) {
        ptrs[F + f] = s->slot_base[0] + f * a_len;
        ptrs[2 * F + f] = s->slot_base[1] + f * (a_len / 2);
      }
    }
    std::memcpy(host.data() + cb, ptrs.data(), ptr_bytes);
    uint32_t* offs = reinterpret_cast<uint32_t*>(host.data() + cb + ptr_bytes);
    for (size_t p = 0; p <= product_count; ++p) offs[p] = static_cast<uint32_t>(product_offsets[p]);
    uint32_t* idx = offs + (product_count + 1);
    for (size_t i = 0; i < index_count; ++i) idx[i] = slot_of[factor_indices[i]];
    // the first occurrence of each factor among all product entries writes its fold
    uint8_t* own = host.data() + cb + ptr_bytes + off_bytes + idx_bytes;
    std::vector<uint8_t> taken(std::max<size_t>(F, 1), 0);
    for (size_t i = 0; i < index_count; ++i) {
      own[i] = taken[idx[i]] ? 0 : 1;
      taken[idx[i]] = 1;
    }
    uint8_t* stage = c.staging(meta_bytes);
    std::memcpy(stage, host.data(), meta_bytes);
    rt::hip_check(hipMemcpyAsync(dev, stage, meta_bytes, hipMemcpyHostToDevice, c.stream));
    c.staging_used();
    s->prods.coeffs = reinterpret_cast<const Fr*>(dev);
    s->coeffs29 = reinterpret_cast<const Fr*>(dev + coeff_bytes);
    s->prods.offsets = reinterpret_cast<const uint32_t*>(dev + cb + ptr_bytes);
    s->prods.slots = s->prods.offsets + (product_count + 1);
    s->prods.count = static_cast<uint32_t>(product_count);
    const Fr** dptrs = reinterpret_cast<const Fr**>(dev + cb);
    s->table_ptrs[0] = dptrs;
    s->table_ptrs[1] = dptrs + F;
    s->table_ptrs[2] = dptrs + 2 * F;
    s->owner = dev + cb + ptr_bytes + off_bytes + idx_bytes;
    if (C) {
      s->class_k = reinterpret_cast<const Fr*>(dev + cls_base);
      s->classes = reinterpret_cast<const uint32_t*>(dev + cls_base + C * rt::kFrBytes);
      s->class_count = static_cast<uint32_t>(C);
      s->class_width = cls_width;
    }
    const size_t width = max_degree + 1;
    unsigned max_blocks = message_blocks(size_t(1) << (num_vars - 1));
    for (size_t k = 0; k < num_vars; ++k) {
      uint32_t mode = 0;
      max_blocks = std::max(max_blocks, job_blocks(*s, k, &mode));
    }
    s->partial = c.allocate(size_t(max_blocks) * std::max<size_t>(width, s->class_width) * rt::kFrBytes);
    s->result = c.allocate(width * rt::kFrBytes);
    // the ticket, then the class finish sums, cleared by one memset
    constexpr size_t kSumsOffset = 256, kSumsBytes = sc::kMaxClassWidth * 8 * sizeof(uint64_t);
    s->ticket = c.allocate(kSumsOffset + kSumsBytes);
    s->class_sums = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(s->ticket->ptr) + kSumsOffset);
    hipLaunchKernelGGL(sc::k_zero_fr, dim3(1), dim3(64), 0, c.stream,
                       static_cast<Fr*>(s->result->ptr), width);
    rt::check_launch();
    rt::hip_check(hipMemsetAsync(s->ticket->ptr, 0, kSumsOffset + kSumsBytes, c.stream));
    c.sync();
    auto* handle = new HpMi355xSumCheckState{ctx, s};
    c.resources[handle] = {rt::Context::Kind::sumcheck_state, s};
    *out = handle;
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_sumcheck_prover_state_round(
    HpMi355xContext* ctx, HpMi355xSumCheckState* state, const uint8_t* challenge_or_null,
    uint8_t* out_evaluations, size_t evaluation_count) {
  return hp_guarded<false>(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    const HpMi355xStatus st =
        check_round(ctx, c, state, challenge_or_null, out_evaluations, evaluation_count);
    if (st.kind != HP_MI355X_STATUS_SUCCESS) return st;
    sc::State& s = *state->impl;
    try {
      const size_t width = s.max_degree + 1;
      const Fr ch = challenge_of(s, challenge_or_null);
      if ((tail_ok(s, false) || wide_ok(s)) && tail_rounds(c, {&s}, {ch}, {out_evaluations}))
        return rt::ok();
      quiesce_tails(c);
      if (s.max_degree <= kMaxJobDegree && round_graphs_enabled()) {
        std::vector<sc::RoundJob> jobs{make_job(s, ch, nullptr)};
        std::vector<uint8_t> out(width * rt::kFrBytes);
        run_jobs_graph(c, jobs, {0}, out.size(), out.data());
        if (!rt::canonical_fr_array(out.data(), width)) rt::fail_state(-4);
        std::memcpy(out_evaluations, out.data(), out.size());
        s.round += 1;
        return rt::ok();
      }
      if (s.max_degree <= kMaxJobDegree) {
        // the last block writes the message straight into pinned host memory
        Fr* host = reinterpret_cast<Fr*>(c.results(width * rt::kFrBytes));
        std::vector<sc::RoundJob> jobs{make_job(s, ch, host)};
        launch_jobs(c, jobs);
        c.sync();
        // a class finish whose sums never arrived writes non-canonical words
        if (!rt::canonical_fr_array(reinterpret_cast<const uint8_t*>(host), width)) rt::fail_state(-4);
        std::memcpy(out_evaluations, host, width * rt::kFrBytes);
      } else {
        generic_round(c, s, ch);
        std::vector<uint8_t> out(width * rt::kFrBytes);
        rt::download_bytes(c, s.result->ptr, out.size(), out.data());
        std::memcpy(out_evaluations, out.data(), out.size());
      }
      s.round += 1;
      return rt::ok();
    } catch (const rt::Failure& f) {
      mark_failed(s, f);
      throw;
    }
  });
}

// One round of each of `count` distinct states: the round kernels of all
// states (degree <= 8) share launches and one copy of the messages back.
extern "C" HpMi355xStatus hp_mi355x_v1_sumcheck_prover_state_round_batch(
    HpMi355xContext* ctx, HpMi355xSumCheckState* const* states,
    const uint8_t* const* challenges_or_null, uint8_t* const* out_evaluations,
    const size_t* evaluation_counts, size_t count) {
  return hp_guarded<false>(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (count == 0) return rt::ok();
    if (states == nullptr || out_evaluations == nullptr || evaluation_counts == nullptr)
      return rt::invalid_parameters();
    std::unordered_map<const void*, size_t> seen;
    for (size_t i = 0; i < count; ++i) {
      if (!seen.emplace(states[i], i).second) return rt::invalid_parameters();
      const uint8_t* ch = challenges_or_null ? challenges_or_null[i] : nullptr;
      const HpMi355xStatus st =
          check_round(ctx, c, states[i], ch, out_evaluations[i], evaluation_counts[i]);
      if (st.kind != HP_MI355X_STATUS_SUCCESS) return st;
    }
    std::vector<sc::State*> ss(count);
    std::vector<size_t> offset(count);
    size_t total = 0;
    for (size_t i = 0; i < count; ++i) {
      ss[i] = states[i]->impl.get();
      offset[i] = total;
      total += ss[i]->max_degree + 1;
    }
    try {
      // every block of a loop launch must be resident at once (one block of
      // 512 threads per CU): at most 128 states
      const bool batch = count > 1;
      bool all_tail = count <= kTailMaxStates;
      for (size_t i = 0; i < count; ++i) all_tail = all_tail && tail_ok(*ss[i], batch);
      if (all_tail) {
        std::vector<Fr> chs(count);
        std::vector<uint8_t*> outs(count);
        for (size_t i = 0; i < count; ++i) {
          chs[i] = challenge_of(*ss[i], challenges_or_null ? challenges_or_null[i] : nullptr);
          outs[i] = out_evaluations[i];
        }
        if (tail_rounds(c, ss, chs, outs)) return rt::ok();
      }
      quiesce_tails(c);
      bool all_jobs = true;
      for (size_t i = 0; i < count; ++i) all_jobs = all_jobs && ss[i]->max_degree <= kMaxJobDegree;
      if (all_jobs && round_graphs_enabled()) {
        std::vector<sc::RoundJob> jobs;
        std::vector<size_t> offs;
        for (size_t i = 0; i < count; ++i) {
          sc::State& s = *ss[i];
          const Fr ch = challenge_of(s, challenges_or_null ? challenges_or_null[i] : nullptr);
          jobs.push_back(make_job(s, ch, nullptr));
          offs.push_back(offset[i]);
        }
        std::vector<uint8_t> host(total * rt::kFrBytes);
        run_jobs_graph(c, jobs, offs, host.size(), host.data());
        // a class finish whose sums never arrived writes non-canonical words
        if (!rt::canonical_fr_array(host.data(), total)) rt::fail_state(-4);
        for (size_t i = 0; i < count; ++i) {
          std::memcpy(out_evaluations[i], host.data() + offset[i] * rt::kFrBytes,
                      (ss[i]->max_degree + 1) * rt::kFrBytes);
          ss[i]->round += 1;
        }
        return rt::ok();
      }
      // messages land in pinned host memory: the job kernels' last blocks
      // write them there, the generic rounds copy them there
      Fr* out_base = reinterpret_cast<Fr*>(c.results(total * rt::kFrBytes));
      std::vector<sc::RoundJob> jobs;
      for (size_t i = 0; i < count; ++i) {
        sc::State& s = *ss[i];
        const Fr ch = challenge_of(s, challenges_or_null ? challenges_or_null[i] : nullptr);
        if (s.max_degree <= kMaxJobDegree) {
          jobs.push_back(make_job(s, ch, out_base + offset[i]));
        } else {
          generic_round(c, s, ch);
          rt::hip_check(hipMemcpyAsync(out_base + offset[i], s.result->ptr,
                                       (s.max_degree + 1) * rt::kFrBytes, hipMemcpyDefault,
                                       c.stream));
        }
      }
      if (!jobs.empty()) launch_jobs(c, jobs);
      c.sync();
      // a class finish whose sums never arrived writes non-canonical words
      if (!rt::canonical_fr_array(reinterpret_cast<const uint8_t*>(out_base), total)) rt::fail_state(-4);
      for (size_t i = 0; i < count; ++i) {
        std::memcpy(out_evaluations[i], out_base + offset[i], (ss[i]->max_degree + 1) * rt::kFrBytes);
        ss[i]->round += 1;
      }
      return rt::ok();
    } catch (const rt::Failure& f) {
      for (sc::State* s : ss) mark_failed(*s, f);
      throw;
    }
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_sumcheck_prover_state_destroy(HpMi355xContext* ctx,
                                                                     HpMi355xSumCheckState* state) {
  return hp_guarded_release(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (state == nullptr || !c.holds(state, rt::Context::Kind::sumcheck_state))
      return rt::invalid_parameters();
    if (state->ctx != ctx) return rt::invalid_parameters();
    sc::State& s = *state->impl;
    if (s.tail_active) {
      // its block may still wait on the host: stop the loop before the
      // state's memory and mailbox are recycled
      try {
        quiesce_tails(c);
      } catch (const rt::Failure& f) {
        c.poison(f.status.native_code);
      }
    }
    tail_return_mailbox(c, s);
    c.resources.erase(state);
    delete state;
    return rt::ok();
  });
}
