// SumCheck prover state (SumCheckProver, backend/src/sum_check.rs:17-38).
//
// Round k of a state over n variables and maximum degree d:
//   k > 0: fold every distinct factor with the previous challenge c,
//          new[b] = old[2b] + c (old[2b+1] - old[2b])  (cpu/sum_check.rs:103-118);
//   then   g(t) = sum_p coeff_p sum_b prod_{f in p} (f[2b] + t (f[2b+1] - f[2b]))
//          for t = 0..d, evaluated directly at every t (the CPU backend
//          evaluates t <= arity and extrapolates, cpu/sum_check.rs:125-177:
//          the same polynomial values).
// Tables are HBM-resident Montgomery arrays; a round folds every distinct
// factor into its ping-pong slot (one launch for all factors), evaluates the
// message with per-block partial sums, and reduces the partials in a third
// kernel; (d+1) canonical values come back to the host.  Products are
// (coefficient, factor slots) with repeats kept; the coefficient is applied to
// the first factor's line once per pair, not once per t.
//
// Validation and statuses follow the reference bridge
// (backend/src/ascend910b3/native/bridge_sumcheck.cpp:210-331 create,
// :519-550 round).
#include <algorithm>
#include <unordered_map>
#include <vector>

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

// Message values for t in [0, D] over pairs b < pairs, tables tbl[slot].
// MAXA bounds the arity (= D); lines are kept in registers per product.
template <int D>
__global__ void __launch_bounds__(kThreads) k_message(const Fr* const* tbl, Products prods,
                                                      size_t pairs, Fr* partial) {
  Fr acc[D + 1];
#pragma unroll
  for (int t = 0; t <= D; ++t) acc[t] = Fr::zero();
  for (size_t b = blockIdx.x * (size_t)blockDim.x + threadIdx.x; b < pairs;
       b += (size_t)gridDim.x * blockDim.x) {
    for (uint32_t p = 0; p < prods.count; ++p) {
      const uint32_t begin = prods.offsets[p];
      const uint32_t arity = prods.offsets[p + 1] - begin;
      Fr val[D], step[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        if (k < (int)arity) {
          const Fr* t = tbl[prods.slots[begin + k]];
          const Fr lo = load_fr(t, 2 * b), hi = load_fr(t, 2 * b + 1);
          val[k] = lo;
          step[k] = hi - lo;
        }
      }
      const Fr coeff = load_fr(prods.coeffs, p);
      val[0] = coeff * val[0];
      step[0] = coeff * step[0];
#pragma unroll
      for (int t = 0; t <= D; ++t) {
        Fr term = val[0];
#pragma unroll
        for (int k = 1; k < D; ++k)
          if (k < (int)arity) term = term * val[k];
        acc[t] = acc[t] + term;
        if (t < D) {
#pragma unroll
          for (int k = 0; k < D; ++k)
            if (k < (int)arity) val[k] = val[k] + step[k];
        }
      }
    }
  }
  block_reduce_store<D + 1>(acc, D + 1, partial, D + 1);
}

// One round in one launch, for max degree D <= 6.  With fold set, thread b
// first folds its two new pairs of every distinct factor from the previous
// tables with challenge c (new[2b], new[2b+1] from old[4b..4b+3]) and writes
// them; it then evaluates every product on the pair (re-read from cache).
// Each block leaves its D+1 partial sums; the last block to finish (ticket
// counter) adds all partials, writes the canonical message and resets the
// counter for the next round.
template <int D>
__global__ void __launch_bounds__(kThreads)
    k_round(const Fr* const* src, Fr* const* dst, uint32_t factors, Fr c, int fold,
            Products prods, size_t pairs, Fr* partial, unsigned* ticket, Fr* out) {
  Fr acc[D + 1];
#pragma unroll
  for (int t = 0; t <= D; ++t) acc[t] = Fr::zero();
  const Fr* const* tbl = fold ? dst : src;
  for (size_t b = blockIdx.x * (size_t)blockDim.x + threadIdx.x; b < pairs;
       b += (size_t)gridDim.x * blockDim.x) {
    if (fold) {
      for (uint32_t f = 0; f < factors; ++f) {
        const Fr* o = src[f];
        const Fr a0 = load_fr(o, 4 * b), a1 = load_fr(o, 4 * b + 1);
        const Fr a2 = load_fr(o, 4 * b + 2), a3 = load_fr(o, 4 * b + 3);
        store_fr(dst[f], 2 * b, a0 + c * (a1 - a0));
        store_fr(dst[f], 2 * b + 1, a2 + c * (a3 - a2));
      }
    }
    for (uint32_t p = 0; p < prods.count; ++p) {
      const uint32_t begin = prods.offsets[p];
      const uint32_t arity = prods.offsets[p + 1] - begin;
      Fr val[D], step[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        if (k < (int)arity) {
          const Fr* t = tbl[prods.slots[begin + k]];
          const Fr lo = load_fr(t, 2 * b), hi = load_fr(t, 2 * b + 1);
          val[k] = lo;
          step[k] = hi - lo;
        }
      }
      const Fr coeff = load_fr(prods.coeffs, p);
      val[0] = coeff * val[0];
      step[0] = coeff * step[0];
#pragma unroll
      for (int t = 0; t <= D; ++t) {
        Fr term = val[0];
#pragma unroll
        for (int k = 1; k < D; ++k)
          if (k < (int)arity) term = term * val[k];
        acc[t] = acc[t] + term;
        if (t < D) {
#pragma unroll
          for (int k = 0; k < D; ++k)
            if (k < (int)arity) val[k] = val[k] + step[k];
        }
      }
    }
  }
  block_reduce_store<D + 1>(acc, D + 1, partial, D + 1);
  // last block: sum all block partials
  __shared__ unsigned last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (threadIdx.x <= (unsigned)D) {
    Fr s = Fr::zero();
    for (unsigned blk = 0; blk < gridDim.x; ++blk) s = s + load_fr(partial, blk * (D + 1) + threadIdx.x);
    store_fr(out, threadIdx.x, from_mont(s));
  }
  if (threadIdx.x == 0) *ticket = 0;
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
  rt::BufferRef meta;                 // coeffs | offsets | slots | table pointers
  Products prods{};
  const Fr** table_ptrs[3] = {nullptr, nullptr, nullptr};  // inputs, slot A, slot B
  size_t distinct = 0, product_count = 0;
  rt::BufferRef partial, result, ticket;
  Fr* slot_base[2] = {nullptr, nullptr};
};

}  // namespace sc
}  // namespace hp

struct HpMi355xSumCheckState {
  HpMi355xContext* ctx;
  std::shared_ptr<hp::sc::State> impl;
};

using namespace hp;

namespace {

unsigned message_blocks(size_t pairs) {
  size_t b = (pairs + sc::kThreads - 1) / sc::kThreads;
  return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(b, sc::kMaxBlocks)));
}

// Launch the message for the current tables (pairs entries) of state s.
void launch_message(rt::Context& c, sc::State& s, const Fr* const* tables, size_t pairs) {
  const unsigned blocks = message_blocks(pairs);
  const size_t width = s.max_degree + 1;
  Fr* partial = static_cast<Fr*>(s.partial->ptr);
  switch (s.max_degree) {
#define HP_MSG_CASE(D)                                                                     \
  case D:                                                                                  \
    hipLaunchKernelGGL(sc::k_message<D>, dim3(blocks), dim3(sc::kThreads), 0, c.stream, \
                       tables, s.prods, pairs, partial);                                   \
    break;
    HP_MSG_CASE(1)
    HP_MSG_CASE(2)
    HP_MSG_CASE(3)
    HP_MSG_CASE(4)
    HP_MSG_CASE(5)
    HP_MSG_CASE(6)
#undef HP_MSG_CASE
    default:
      for (size_t t0 = 0; t0 < width; t0 += 8) {
        const uint32_t tc = static_cast<uint32_t>(std::min<size_t>(8, width - t0));
        hipLaunchKernelGGL(sc::k_message_chunk, dim3(blocks), dim3(sc::kThreads), 0, c.stream,
                           tables, s.prods, pairs, static_cast<uint32_t>(t0), tc, partial,
                           width);
        rt::check_launch();
      }
  }
  rt::check_launch();
  hipLaunchKernelGGL(sc::k_reduce_partials, dim3(1), dim3(64), 0, c.stream, partial,
                     (size_t)blocks, width, static_cast<Fr*>(s.result->ptr));
  rt::check_launch();
}

void launch_round(rt::Context& c, sc::State& s, const Fr* const* src, Fr* const* dst,
                  const Fr& ch, bool fold, size_t pairs) {
  const unsigned blocks = message_blocks(pairs);
  Fr* partial = static_cast<Fr*>(s.partial->ptr);
  unsigned* ticket = static_cast<unsigned*>(s.ticket->ptr);
  Fr* out = static_cast<Fr*>(s.result->ptr);
  const uint32_t F = static_cast<uint32_t>(s.distinct);
  switch (s.max_degree) {
#define HP_ROUND_CASE(D)                                                                      \
  case D:                                                                                     \
    hipLaunchKernelGGL(sc::k_round<D>, dim3(blocks), dim3(sc::kThreads), 0, c.stream, src, \
                       dst, F, ch, fold ? 1 : 0, s.prods, pairs, partial, ticket, out);       \
    break;
    HP_ROUND_CASE(1)
    HP_ROUND_CASE(2)
    HP_ROUND_CASE(3)
    HP_ROUND_CASE(4)
    HP_ROUND_CASE(5)
    HP_ROUND_CASE(6)
#undef HP_ROUND_CASE
    default:
      rt::fail_state(-3);
  }
  rt::check_launch();
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
    const size_t F = distinct.size();
    if (F > 0 && num_vars >= 2) {
      const size_t a = size_t(1) << (num_vars - 1), b = size_t(1) << (num_vars - 2);
      s->slots = c.allocate(F * (a + b) * rt::kFrBytes);
      s->slot_base[0] = static_cast<Fr*>(s->slots->ptr);
      s->slot_base[1] = s->slot_base[0] + F * a;
    }
    // metadata block: coeffs (Fr) | table ptrs x3 (F each) | offsets | slots
    const size_t coeff_bytes = std::max<size_t>(product_count, 1) * rt::kFrBytes;
    const size_t ptr_bytes = 3 * std::max<size_t>(F, 1) * sizeof(void*);
    const size_t off_bytes = (product_count + 1) * sizeof(uint32_t);
    const size_t idx_bytes = std::max<size_t>(index_count, 1) * sizeof(uint32_t);
    const size_t meta_bytes = coeff_bytes + ptr_bytes + off_bytes + idx_bytes;
    s->meta = c.allocate(meta_bytes);
    std::vector<uint8_t> host(meta_bytes, 0);
    for (size_t p = 0; p < product_count; ++p) {
      Fr v;
      std::memcpy(v.v, coefficients + rt::kFrBytes * p, rt::kFrBytes);
      v = to_mont(v);
      std::memcpy(host.data() + rt::kFrBytes * p, v.v, rt::kFrBytes);
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
    std::memcpy(host.data() + coeff_bytes, ptrs.data(), ptr_bytes);
    uint32_t* offs = reinterpret_cast<uint32_t*>(host.data() + coeff_bytes + ptr_bytes);
    for (size_t p = 0; p <= product_count; ++p) offs[p] = static_cast<uint32_t>(product_offsets[p]);
    uint32_t* idx = offs + (product_count + 1);
    for (size_t i = 0; i < index_count; ++i) idx[i] = slot_of[factor_indices[i]];
    uint8_t* stage = c.staging(meta_bytes);
    std::memcpy(stage, host.data(), meta_bytes);
    rt::hip_check(hipMemcpyAsync(dev, stage, meta_bytes, hipMemcpyHostToDevice, c.stream));
    c.staging_used();
    s->prods.coeffs = reinterpret_cast<const Fr*>(dev);
    s->prods.offsets = reinterpret_cast<const uint32_t*>(dev + coeff_bytes + ptr_bytes);
    s->prods.slots = s->prods.offsets + (product_count + 1);
    s->prods.count = static_cast<uint32_t>(product_count);
    const Fr** dptrs = reinterpret_cast<const Fr**>(dev + coeff_bytes);
    s->table_ptrs[0] = dptrs;
    s->table_ptrs[1] = dptrs + F;
    s->table_ptrs[2] = dptrs + 2 * F;
    const size_t width = max_degree + 1;
    const size_t pairs0 = size_t(1) << (num_vars - 1);
    s->partial = c.allocate(message_blocks(pairs0) * width * rt::kFrBytes);
    s->result = c.allocate(width * rt::kFrBytes);
    s->ticket = c.allocate(sizeof(unsigned));
    hipLaunchKernelGGL(sc::k_zero_fr, dim3(1), dim3(64), 0, c.stream,
                       static_cast<Fr*>(s->result->ptr), width);
    rt::check_launch();
    rt::hip_check(hipMemsetAsync(s->ticket->ptr, 0, sizeof(unsigned), c.stream));
    c.sync();
    auto* handle = new HpMi355xSumCheckState{ctx, s};
    c.resources[handle] = s;
    *out = handle;
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_sumcheck_prover_state_round(
    HpMi355xContext* ctx, HpMi355xSumCheckState* state, const uint8_t* challenge_or_null,
    uint8_t* out_evaluations, size_t evaluation_count) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (state == nullptr || state->ctx != ctx || !c.resources.count(state))
      return rt::invalid_parameters();
    sc::State& s = *state->impl;
    if (s.failed) return rt::invalid_state(s.failure_code);
    if (out_evaluations == nullptr) return rt::invalid_parameters();
    if (evaluation_count != s.max_degree + 1) return rt::invalid_parameters();
    const size_t k = s.round;
    if (k >= s.num_vars) return rt::invalid_state();
    if (k == 0) {
      if (challenge_or_null != nullptr) return rt::invalid_state();
    } else {
      if (challenge_or_null == nullptr) return rt::invalid_state();
      if (!rt::canonical_fr(challenge_or_null)) return rt::invalid_parameters();
    }
    try {
      const size_t width = s.max_degree + 1;
      const size_t pairs = size_t(1) << (s.num_vars - k - 1);
      if (s.product_count == 0) {
        hipLaunchKernelGGL(sc::k_zero_fr, dim3(1), dim3(64), 0, c.stream,
                           static_cast<Fr*>(s.result->ptr), width);
        rt::check_launch();
      } else {
        Fr ch = Fr::zero();
        if (k > 0) {
          std::memcpy(ch.v, challenge_or_null, rt::kFrBytes);
          ch = to_mont(ch);
        }
        // round k writes slot (k-1)%2 (A holds 2^(n-1), B 2^(n-2) entries) and
        // reads the inputs (k = 1) or slot (k-2)%2
        const Fr* const* src = k <= 1 ? s.table_ptrs[0] : s.table_ptrs[1 + ((k - 2) & 1)];
        Fr* const* dst = k == 0 ? nullptr : const_cast<Fr* const*>(s.table_ptrs[1 + ((k - 1) & 1)]);
        if (s.max_degree <= 6) {
          launch_round(c, s, src, dst, ch, k > 0, pairs);
        } else {
          const Fr* const* tables = src;
          if (k > 0) {
            const size_t half = 2 * pairs;  // new table length
            const size_t total = s.distinct * half;
            unsigned blocks = static_cast<unsigned>(
                std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 65535)));
            hipLaunchKernelGGL(sc::k_fold_factors, dim3(blocks), dim3(256), 0, c.stream, src, dst,
                               s.distinct, ch, half);
            rt::check_launch();
            tables = dst;
          }
          launch_message(c, s, tables, pairs);
        }
      }
      std::vector<uint8_t> out(width * rt::kFrBytes);
      rt::download_bytes(c, s.result->ptr, out.size(), out.data());
      std::memcpy(out_evaluations, out.data(), out.size());
      s.round += 1;
      return rt::ok();
    } catch (const rt::Failure& f) {
      if (f.poison) {
        s.failed = true;
        s.failure_code = f.status.native_code ? f.status.native_code : -1;
      }
      throw;
    }
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_sumcheck_prover_state_destroy(HpMi355xContext* ctx,
                                                                     HpMi355xSumCheckState* state) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (state == nullptr || state->ctx != ctx) return rt::invalid_parameters();
    auto it = c.resources.find(state);
    if (it == c.resources.end()) return rt::invalid_parameters();
    c.resources.erase(it);
    delete state;
    return rt::ok();
  });
}
