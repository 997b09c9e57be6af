// MLE operations of MleBackend (backend/src/multilinear_polynomial.rs:33-121)
// as streaming HIP kernels over HBM-resident Montgomery Fr tables.
//
// Semantics follow the reference CPU backend
// (backend/src/cpu/multilinear_polynomial.rs:29-254, backend/src/common/eq.rs);
// validation order and status kinds follow the reference bridge
// (backend/src/ascend910b3/native/bridge_mle.cpp:173-733).  Table order is
// low-bit-first: index bit i <-> variable i <-> point[i].
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "field29.hpp"
#include "kernels_common.hpp"
#include "mle.hpp"

namespace hp {
namespace {

constexpr unsigned kThreads = 256;

unsigned grid(size_t n) {
  size_t b = (n + kThreads - 1) / kThreads;
  b = std::min<size_t>(std::max<size_t>(b, 1), 65535);
  return static_cast<unsigned>(b);
}

#define GRID_STRIDE(i, n)                                                        \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); \
       i += (size_t)gridDim.x * blockDim.x)

// out[j] = sum_k coeff[k] * tables[k][j]
__global__ void k_linear_combination(const Fr* const* tables, const Fr* coeffs, size_t count,
                                     Fr* out, size_t n) {
  GRID_STRIDE(j, n) {
    Fr acc = Fr::zero();
    for (size_t k = 0; k < count; ++k) acc = acc + load_fr(coeffs, k) * load_fr(tables[k], j);
    store_fr(out, j, acc);
  }
}

}  // namespace

// ---- fold: dst[b] = src[2b] + r (src[2b+1] - src[2b]) ---------------------
__global__ void k_fold(const Fr* src, Fr* dst, const Fr* r, size_t half) {
  const Fr c = load_fr(r, 0);
  GRID_STRIDE(b, half) {
    const Fr lo = load_fr(src, 2 * b);
    const Fr hi = load_fr(src, 2 * b + 1);
    store_fr(dst, b, lo + c * (hi - lo));
  }
}

// Fold the remaining `levels` variables of a table of 2^levels <= 1024 entries
// inside one block (LDS), writing the single result to out[0].
__global__ void k_fold_tail(const Fr* src, const Fr* point, size_t levels, Fr* out) {
  __shared__ Fr buf[1024];
  const size_t n = size_t(1) << levels;
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) buf[i] = load_fr(src, i);
  __syncthreads();
  size_t len = n;
  for (size_t l = 0; l < levels; ++l) {
    const Fr c = load_fr(point, l);
    const size_t half = len / 2;
    Fr v[4];
    size_t cnt = 0;
    for (size_t b = threadIdx.x; b < half; b += blockDim.x) {
      const Fr lo = buf[2 * b], hi = buf[2 * b + 1];
      v[cnt++] = lo + c * (hi - lo);
    }
    __syncthreads();
    cnt = 0;
    for (size_t b = threadIdx.x; b < half; b += blockDim.x) buf[b] = v[cnt++];
    __syncthreads();
    len = half;
  }
  if (threadIdx.x == 0) store_fr(out, 0, buf[0]);
}

namespace {

struct EqChunks {
  const Fr* table[4];
  uint32_t shift[4];
  uint32_t bits[4];
  uint32_t count;
};

// eq table of chunk blockIdx.x (k <= 10 coordinates r[shift .. shift + k)):
// t[x] = prod_{i<k} (bit_i(x) ? r_i : 1 - r_i), built in LDS by prepending one
// coordinate per level from the highest to the lowest (eq.rs:67-95 order).
// The chunks are independent: one launch, one block each.
__device__ __forceinline__ void eq_chunk_body(const Fr* r, size_t k, Fr* table) {
  __shared__ Fr buf[1024];
  if (threadIdx.x == 0) buf[0] = Fr::one();
  __syncthreads();
  size_t len = 1;
  for (size_t step = 0; step < k; ++step) {
    const Fr ri = load_fr(r, k - 1 - step);
    Fr even[2], odd[2];
    size_t cnt = 0;
    for (size_t b = threadIdx.x; b < len; b += blockDim.x) {
      const Fr v = buf[b];
      const Fr t = ri * v;
      even[cnt] = v - t;
      odd[cnt] = t;
      ++cnt;
    }
    __syncthreads();
    cnt = 0;
    for (size_t b = threadIdx.x; b < len; b += blockDim.x) {
      buf[2 * b] = even[cnt];
      buf[2 * b + 1] = odd[cnt];
      ++cnt;
    }
    __syncthreads();
    len *= 2;
  }
  for (size_t i = threadIdx.x; i < len; i += blockDim.x) store_fr(table, i, buf[i]);
}

__global__ void k_eq_chunk(const Fr* r_all, EqChunks chunks) {
  eq_chunk_body(r_all + chunks.shift[blockIdx.x], chunks.bits[blockIdx.x],
                const_cast<Fr*>(chunks.table[blockIdx.x]));
}

// Several eq tables in two launches: one block per (table, chunk), then the
// outer products, each table over its own range of blocks.
struct EqJob {
  const Fr* r;  // the table's point (device, Montgomery)
  Fr* out;
  uint32_t nv, block_begin, blocks;
  EqChunks chunks;
};
struct EqChunkRef {
  uint32_t job, chunk;
};

__global__ void k_eq_chunks_batch(const EqJob* jobs, const EqChunkRef* refs) {
  const EqChunkRef ref = refs[blockIdx.x];
  const EqJob& job = jobs[ref.job];
  eq_chunk_body(job.r + job.chunks.shift[ref.chunk], job.chunks.bits[ref.chunk],
                const_cast<Fr*>(job.chunks.table[ref.chunk]));
}

__global__ void k_eq_outer_batch(const EqJob* jobs, uint32_t count) {
  uint32_t j = 0;
  while (j + 1 < count && blockIdx.x >= jobs[j + 1].block_begin) ++j;
  const EqJob& job = jobs[j];
  const EqChunks& ch = job.chunks;
  const size_t n = size_t(1) << job.nv, stride = (size_t)job.blocks * blockDim.x;
  for (size_t x = (size_t)(blockIdx.x - job.block_begin) * blockDim.x + threadIdx.x; x < n; x += stride) {
    Fr v = load_fr(ch.table[0], x & ((size_t(1) << ch.bits[0]) - 1));
    for (uint32_t c = 1; c < ch.count; ++c)
      v = v * load_fr(ch.table[c], (x >> ch.shift[c]) & ((size_t(1) << ch.bits[c]) - 1));
    store_fr(job.out, x, v);
  }
}

// out[x] = prod_c chunk_c[(x >> shift_c) & mask_c]
__global__ void k_eq_outer(EqChunks chunks, Fr* out, size_t n) {
  GRID_STRIDE(x, n) {
    Fr v = load_fr(chunks.table[0], x & ((size_t(1) << chunks.bits[0]) - 1));
    for (uint32_t c = 1; c < chunks.count; ++c)
      v = v * load_fr(chunks.table[c], (x >> chunks.shift[c]) & ((size_t(1) << chunks.bits[c]) - 1));
    store_fr(out, x, v);
  }
}

// ---- batched evaluation: f(r) = sum_x f[x] eq(r, x) ----------------------
struct DotJob {
  const Fr* table;
  const Fr* eq;  // null: the zero-variable eq table [1]
  Fr* partial;   // one Fr per block of the job
  unsigned* ticket;
  Fr* out;       // canonical result
  uint32_t len, block_begin, blocks, pad;
};

constexpr unsigned kDotPerThread = 16;

__device__ __forceinline__ Fr wave_sum_fr(Fr v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) {
    Fr o;
#pragma unroll
    for (int i = 0; i < 8; ++i) o.v[i] = __shfl_xor(v.v[i], offset, 64);
    v = v + o;
  }
  return v;
}

__global__ void __launch_bounds__(kThreads) k_eval_dots(const DotJob* jobs, uint32_t count) {
  __shared__ Fr lds[kThreads / 64];
  __shared__ unsigned last;
  uint32_t lo = 0, hi = count;  // last job with block_begin <= blockIdx.x
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (jobs[mid].block_begin <= blockIdx.x) lo = mid; else hi = mid;
  }
  const DotJob job = jobs[lo];
  const uint32_t lblock = blockIdx.x - job.block_begin;
  Fr acc = Fr::zero();
  for (size_t x = (size_t)lblock * kThreads + threadIdx.x; x < job.len;
       x += (size_t)job.blocks * kThreads)
    acc = acc + (job.eq ? load_fr(job.table, x) * load_fr(job.eq, x) : load_fr(job.table, x));
  acc = wave_sum_fr(acc);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (job.blocks == 1) {
    // the block's sum is the value: no partial, ticket or device-scope fence
    // (each fence writes back the L2 on this multi-XCD part)
    if (threadIdx.x == 0) {
      Fr s = lds[0];
      for (unsigned w = 1; w < kThreads / 64; ++w) s = s + lds[w];
      store_fr(job.out, 0, from_mont(s));
      __threadfence_system();  // `out` is pinned host memory read after the sync
    }
    return;
  }
  if (threadIdx.x == 0) {
    Fr s = lds[0];
    for (unsigned w = 1; w < kThreads / 64; ++w) s = s + lds[w];
    store_fr(job.partial, lblock, s);
    __threadfence();
    last = atomicAdd(job.ticket, 1u) == job.blocks - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // the last block sums the block partials with all its threads
  Fr s = Fr::zero();
  for (uint32_t b = threadIdx.x; b < job.blocks; b += kThreads) s = s + load_fr(job.partial, b);
  s = wave_sum_fr(s);
  __syncthreads();  // lds reused
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    Fr total = lds[0];
    for (unsigned w = 1; w < kThreads / 64; ++w) total = total + lds[w];
    store_fr(job.out, 0, from_mont(total));
    __threadfence_system();  // `out` is pinned host memory read after the sync
  }
}

__global__ void k_zero_words(unsigned* p, size_t n) {
  GRID_STRIDE(i, n) p[i] = 0;
}

__global__ void k_zero_u32(unsigned* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// Montgomery batch inversion of d_j = prod_k g_k[j] over all entries, out =
// prod_k f_k / prod_k g_k, in four launches and one host inversion:
//   k_frac_prefix  thread t (of T, a power of two) walks its strided entries:
//                  out[j] := the exclusive prefix product of its d_j, tree[t]
//                  := its total; a zero d_j raises aux->zero;
//   k_frac_up      one block: the product tree over the T totals (levels
//                  stored one after another, the root last, copied to aux);
//   host           inverts the root (one inversion instead of one Fermat
//                  chain per thread, whose ~340 dependent products set the
//                  kernel's time: 154 us at nv 12);
//   k_frac_down    one block: the inverses down the tree, in place
//                  (inv(left) = inv(parent) right, inv(right) = inv(parent)
//                  left), leaving 1 / total_t in tree[t];
//   k_frac_back    thread t walks its entries backwards as before.
struct FracAux {
  unsigned zero;
  unsigned pad[7];
  Fr root;
};

__global__ void k_frac_prefix(const Fr* const* dens, size_t count, Fr* out, size_t n, Fr* tree,
                              FracAux* aux) {
  const size_t threads = (size_t)gridDim.x * blockDim.x;
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (tid >= n) return;
  Fr acc = Fr::one();
  bool saw_zero = false;
  for (size_t j = tid; j < n; j += threads) {
    Fr d = load_fr(dens[0], j);
    for (size_t k = 1; k < count; ++k) d = d * load_fr(dens[k], j);
    saw_zero |= d.is_zero();
    store_fr(out, j, acc);
    acc = acc * d;
  }
  if (saw_zero) atomicOr(&aux->zero, 1u);
  store_fr(tree, tid, acc);
}

__global__ void __launch_bounds__(1024) k_frac_up(Fr* tree, size_t T, FracAux* aux) {
  Fr* level = tree;
  for (size_t len = T; len > 1; len /= 2) {
    for (size_t i = threadIdx.x; i < len / 2; i += blockDim.x)
      store_fr(level + len, i, load_fr(level, 2 * i) * load_fr(level, 2 * i + 1));
    __syncthreads();
    level += len;
  }
  if (threadIdx.x == 0) {
    store_fr(&aux->root, 0, load_fr(level, 0));
    __threadfence_system();
  }
}

__global__ void __launch_bounds__(1024) k_frac_down(Fr* tree, size_t T, Fr root_inverse) {
  // level offsets: level l starts at T + T/2 + ... (l terms) = 2T - 2T/2^l
  size_t levels = 0;
  while ((size_t(1) << levels) < T) ++levels;
  if (threadIdx.x == 0) store_fr(tree, 2 * T - 2, root_inverse);
  __syncthreads();
  for (size_t l = levels; l-- > 0;) {
    const size_t len = T >> l, off = 2 * T - 2 * len;  // level l: len entries at off
    Fr* level = tree + off;
    const Fr* parent = tree + off + len;
    for (size_t p = threadIdx.x; p < len / 2; p += blockDim.x) {
      const Fr a = load_fr(level, 2 * p), b = load_fr(level, 2 * p + 1), ip = load_fr(parent, p);
      store_fr(level, 2 * p, ip * b);
      store_fr(level, 2 * p + 1, ip * a);
    }
    __syncthreads();
  }
}

__global__ void k_frac_back(const Fr* const* nums, const Fr* const* dens, size_t count, Fr* out,
                            size_t n, const Fr* tree) {
  const size_t threads = (size_t)gridDim.x * blockDim.x;
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (tid >= n) return;
  Fr inv = load_fr(tree, tid);
  // backwards: inv_j = inv * prefix_j; inv *= d_j
  size_t last = tid + ((n - 1 - tid) / threads) * threads;
  for (size_t j = last;; j -= threads) {
    Fr d = load_fr(dens[0], j);
    for (size_t k = 1; k < count; ++k) d = d * load_fr(dens[k], j);
    Fr num = load_fr(nums[0], j);
    for (size_t k = 1; k < count; ++k) num = num * load_fr(nums[k], j);
    const Fr dinv = inv * load_fr(out, j);
    inv = inv * d;
    store_fr(out, j, num * dinv);
    if (j < threads) break;
  }
}

// Product tree level: dst[i] = src[2i] * src[2i+1].
__global__ void k_pair_products(const Fr* src, Fr* dst, size_t count) {
  GRID_STRIDE(i, count) { store_fr(dst, i, load_fr(src, 2 * i) * load_fr(src, 2 * i + 1)); }
}

// Remaining product-tree levels once a level has <= 1024 entries, in one block.
// `level` points at the current level (len entries); results follow it in place.
__global__ void k_product_tail(Fr* level, size_t len) {
  while (len > 1) {
    const size_t half = len / 2;
    for (size_t i = threadIdx.x; i < half; i += blockDim.x)
      store_fr(level + len, i, load_fr(level, 2 * i) * load_fr(level, 2 * i + 1));
    __syncthreads();
    level += len;
    len = half;
  }
  if (threadIdx.x == 0) store_fr(level, 1, Fr::zero());  // prod(1,...,1) := 0
}

// num_c[j] = f_c[j] + beta (c 2^n + j) + gamma ; den_c[j] = g_c[j] + beta perm_c[j] + gamma
__global__ void k_nums_denoms(const Fr* const* fxs, const Fr* const* gxs, const Fr* const* perms,
                              const Fr* beta_gamma, Fr* const* nums, Fr* const* dens,
                              size_t count, size_t n) {
  const Fr beta = load_fr(beta_gamma, 0);
  const Fr gamma = load_fr(beta_gamma, 1);
  GRID_STRIDE(j, n * count) {
    const size_t c = j / n, row = j % n;
    const Fr id = from_u64<FrParams>((uint64_t)(c * n + row));
    store_fr(nums[c], row, load_fr(fxs[c], row) + beta * id + gamma);
    store_fr(dens[c], row, load_fr(gxs[c], row) + beta * load_fr(perms[c], row) + gamma);
  }
}

// p1 = [frac even | prod even], p2 = [frac odd | prod odd]
__global__ void k_product_factors(const Fr* frac, const Fr* prod, Fr* p1, Fr* p2, size_t half) {
  GRID_STRIDE(i, 2 * half) {
    const Fr* src = i < half ? frac : prod;
    const size_t k = i < half ? i : i - half;
    store_fr(p1, i, load_fr(src, 2 * k));
    store_fr(p2, i, load_fr(src, 2 * k + 1));
  }
}

}  // namespace

namespace mle {

void fold(rt::Context& c, const Fr* src, Fr* dst, const Fr* r, size_t half) {
  hipLaunchKernelGGL(k_fold, dim3(grid(half)), dim3(kThreads), 0, c.stream, src, dst, r, half);
  rt::check_launch();
}

// Evaluate a table of 2^nv entries at a device-resident point; result in out[0].
// Folds ping-pong between two scratch tables (a fold cannot run in place on
// the device: thread b would overwrite entry b while thread b/2 still reads it).
void evaluate(rt::Context& c, const Fr* table, size_t nv, const Fr* point, Fr* out,
              Fr* fold_scratch) {
  if (nv == 0) {
    rt::hip_check(hipMemcpyAsync(out, table, rt::kFrBytes, hipMemcpyDeviceToDevice, c.stream));
    return;
  }
  const Fr* src = table;
  size_t level = 0;
  rt::BufferRef held;
  if (nv > 10) {
    const size_t a = size_t(1) << (nv - 1), b = size_t(1) << (nv - 2);
    if (fold_scratch == nullptr) {
      held = c.allocate((a + b) * rt::kFrBytes);
      fold_scratch = static_cast<Fr*>(held->ptr);
    }
    Fr* bufs[2] = {fold_scratch, fold_scratch + a};
    while (nv - level > 10) {
      const size_t half = size_t(1) << (nv - level - 1);
      Fr* dst = bufs[level & 1];
      fold(c, src, dst, point + level, half);
      src = dst;
      ++level;
    }
  }
  hipLaunchKernelGGL(k_fold_tail, dim3(1), dim3(256), 0, c.stream, src, point + level, nv - level,
                     out);
  rt::check_launch();
}

// Chunk split of an eq table over nv >= 1 coordinates: <= 4 chunks of <= 10
// bits, balanced, low bits first; the chunk tables at `t` (returns its end).
Fr* eq_chunks(size_t nv, Fr* t, EqChunks* eq) {
  const size_t chunks = (nv + 9) / 10;
  eq->count = static_cast<uint32_t>(chunks);
  size_t off = 0;
  for (size_t k = 0; k < chunks; ++k) {
    const size_t bits = (nv - off + (chunks - k) - 1) / (chunks - k);
    eq->table[k] = t;
    eq->shift[k] = static_cast<uint32_t>(off);
    eq->bits[k] = static_cast<uint32_t>(bits);
    t += size_t(1) << bits;
    off += bits;
  }
  return t;
}

size_t eq_chunk_entries(size_t nv) {
  const size_t chunks = (nv + 9) / 10;
  size_t off = 0, total = 0;
  for (size_t k = 0; k < chunks; ++k) {
    const size_t bits = (nv - off + (chunks - k) - 1) / (chunks - k);
    total += size_t(1) << bits;
    off += bits;
  }
  return total;
}

// eq tables of several points (canonical host bytes, nv[i] >= 1 each) into
// outs[i]: the points in one upload, the chunks in one launch, the outer
// products in another (one launch pair per table cost ~14 us of device time
// each at nv 12, most of it fixed).  `keep` receives the scratch buffers.
void build_eq_batch(rt::Context& c, const std::vector<const uint8_t*>& points,
                    const std::vector<size_t>& nv, const std::vector<Fr*>& outs,
                    std::vector<rt::BufferRef>& keep) {
  const size_t m = points.size();
  if (m == 0) return;
  size_t coords = 0, entries = 0;
  for (size_t i = 0; i < m; ++i) {
    if (nv[i] == 0) rt::fail_params();
    coords += nv[i];
    entries += eq_chunk_entries(nv[i]);
  }
  std::vector<uint8_t> bytes(coords * rt::kFrBytes);
  for (size_t i = 0, at = 0; i < m; at += nv[i], ++i)
    std::memcpy(bytes.data() + at * rt::kFrBytes, points[i], nv[i] * rt::kFrBytes);
  rt::BufferRef dpts = c.allocate(coords * rt::kFrBytes);
  rt::upload_fr_values(c, static_cast<Fr*>(dpts->ptr), bytes.data(), coords);
  rt::BufferRef scratch = c.allocate(entries * rt::kFrBytes);
  std::vector<EqJob> jobs(m);
  std::vector<EqChunkRef> refs;
  Fr* t = static_cast<Fr*>(scratch->ptr);
  uint32_t blocks = 0;
  for (size_t i = 0, at = 0; i < m; at += nv[i], ++i) {
    EqJob& j = jobs[i];
    j.r = static_cast<const Fr*>(dpts->ptr) + at;
    j.out = outs[i];
    j.nv = static_cast<uint32_t>(nv[i]);
    t = eq_chunks(nv[i], t, &j.chunks);
    for (uint32_t k = 0; k < j.chunks.count; ++k) refs.push_back({static_cast<uint32_t>(i), k});
    j.block_begin = blocks;
    j.blocks = grid(size_t(1) << nv[i]);
    blocks += j.blocks;
  }
  const size_t jb = m * sizeof(EqJob), rb = refs.size() * sizeof(EqChunkRef);
  rt::BufferRef desc = c.allocate(jb + rb);
  uint8_t* stage = c.staging(jb + rb);
  std::memcpy(stage, jobs.data(), jb);
  std::memcpy(stage + jb, refs.data(), rb);
  rt::hip_check(hipMemcpyAsync(desc->ptr, stage, jb + rb, hipMemcpyHostToDevice, c.stream));
  c.staging_used();
  const EqJob* djobs = static_cast<const EqJob*>(desc->ptr);
  const EqChunkRef* drefs =
      reinterpret_cast<const EqChunkRef*>(static_cast<const uint8_t*>(desc->ptr) + jb);
  hipLaunchKernelGGL(k_eq_chunks_batch, dim3((unsigned)refs.size()), dim3(256), 0, c.stream, djobs,
                     drefs);
  rt::check_launch();
  hipLaunchKernelGGL(k_eq_outer_batch, dim3(blocks), dim3(kThreads), 0, c.stream, djobs,
                     static_cast<uint32_t>(m));
  rt::check_launch();
  keep.push_back(std::move(dpts));
  keep.push_back(std::move(scratch));
  keep.push_back(std::move(desc));
}

void build_eq(rt::Context& c, const Fr* r, size_t nv, Fr* out) {
  // split the coordinates into <= 4 chunks of <= 10 bits, low bits first
  const size_t chunks = (nv + 9) / 10;
  EqChunks eq{};
  eq.count = static_cast<uint32_t>(chunks);
  size_t off = 0, total = 0;
  size_t bits[4];
  for (size_t k = 0; k < chunks; ++k) {
    bits[k] = (nv - off + (chunks - k) - 1) / (chunks - k);  // balanced split
    total += size_t(1) << bits[k];
    off += bits[k];
  }
  rt::BufferRef tables = c.allocate(total * rt::kFrBytes);
  Fr* t = static_cast<Fr*>(tables->ptr);
  off = 0;
  for (size_t k = 0; k < chunks; ++k) {
    eq.table[k] = t;
    eq.shift[k] = static_cast<uint32_t>(off);
    eq.bits[k] = static_cast<uint32_t>(bits[k]);
    t += size_t(1) << bits[k];
    off += bits[k];
  }
  if (chunks == 0) rt::fail_params();  // callers pass nv >= 1
  hipLaunchKernelGGL(k_eq_chunk, dim3((unsigned)chunks), dim3(256), 0, c.stream, r, eq);
  rt::check_launch();
  const size_t n = size_t(1) << nv;
  hipLaunchKernelGGL(k_eq_outer, dim3(grid(n)), dim3(kThreads), 0, c.stream, eq, out, n);
  rt::check_launch();
}

}  // namespace mle
}  // namespace hp

// ============================ C ABI =========================================
using namespace hp;

namespace {

bool dims_ok(size_t nv) { return nv < rt::kMaxNumVars; }

// Upload a small array of canonical Fr values (already validated).
rt::BufferRef upload_small(rt::Context& c, const uint8_t* bytes, size_t count) {
  rt::BufferRef b = c.allocate(std::max<size_t>(count, 1) * rt::kFrBytes);
  if (count) rt::upload_fr_values(c, static_cast<Fr*>(b->ptr), bytes, count);
  return b;
}

rt::BufferRef upload_pointers(rt::Context& c, const std::vector<const void*>& ptrs) {
  rt::BufferRef b = c.allocate(std::max<size_t>(ptrs.size(), 1) * sizeof(void*));
  if (!ptrs.empty()) {
    uint8_t* stage = c.staging(ptrs.size() * sizeof(void*));
    std::memcpy(stage, ptrs.data(), ptrs.size() * sizeof(void*));
    rt::hip_check(hipMemcpyAsync(b->ptr, stage, ptrs.size() * sizeof(void*),
                                 hipMemcpyHostToDevice, c.stream));
    c.staging_used();
  }
  return b;
}

}  // namespace

extern "C" HpMi355xStatus hp_mi355x_v1_mle_upload(HpMi355xContext* ctx, size_t num_vars,
                                                  const uint8_t* evaluations,
                                                  size_t evaluation_count,
                                                  HpMi355xDeviceAddress* out) {
  if (out) *out = nullptr;
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out == nullptr || !dims_ok(num_vars)) return rt::invalid_parameters();
    if (evaluation_count != (size_t(1) << num_vars) || evaluations == nullptr)
      return rt::invalid_parameters();
    if (!rt::canonical_fr_array(evaluations, evaluation_count)) return rt::invalid_parameters();
    rt::BufferRef buf = rt::upload_fr_table(c, evaluations, evaluation_count, num_vars);
    c.sync();  // host buffer is borrowed only until return
    *out = c.publish(std::move(buf));
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_mle_download(HpMi355xContext* ctx,
                                                    HpMi355xDeviceAddress polynomial,
                                                    size_t num_vars, uint8_t* out_evaluations,
                                                    size_t evaluation_count) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out_evaluations == nullptr || !dims_ok(num_vars)) return rt::invalid_parameters();
    if (evaluation_count != (size_t(1) << num_vars)) return rt::invalid_parameters();
    rt::BufferRef buf = c.lookup(polynomial, num_vars);
    if (!buf) return rt::invalid_parameters();
    rt::download_fr(c, static_cast<const Fr*>(buf->ptr), evaluation_count, out_evaluations);
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_mle_linear_combination(
    HpMi355xContext* ctx, size_t num_vars, const HpMi355xDeviceAddress* polynomials,
    const uint8_t* coefficients, size_t count, HpMi355xDeviceAddress* out) {
  if (out) *out = nullptr;
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out == nullptr || polynomials == nullptr || coefficients == nullptr || count == 0)
      return rt::invalid_parameters();
    if (!dims_ok(num_vars)) return rt::invalid_parameters();
    std::vector<rt::BufferRef> inputs;
    for (size_t k = 0; k < count; ++k) {
      rt::BufferRef b = c.lookup(polynomials[k], num_vars);
      if (!b) return rt::invalid_parameters();
      inputs.push_back(b);
    }
    if (!rt::canonical_fr_array(coefficients, count)) return rt::invalid_parameters();
    // zero coefficients contribute nothing (cpu/multilinear_polynomial.rs:72-74)
    std::vector<const void*> ptrs;
    std::vector<uint8_t> coeffs;
    for (size_t k = 0; k < count; ++k) {
      const uint8_t* cb = coefficients + rt::kFrBytes * k;
      if (std::all_of(cb, cb + rt::kFrBytes, [](uint8_t v) { return v == 0; })) continue;
      ptrs.push_back(inputs[k]->ptr);
      coeffs.insert(coeffs.end(), cb, cb + rt::kFrBytes);
    }
    const size_t n = size_t(1) << num_vars;
    rt::BufferRef result = c.allocate(n * rt::kFrBytes, num_vars);
    rt::BufferRef dptrs = upload_pointers(c, ptrs);
    rt::BufferRef dcoef = upload_small(c, coeffs.data(), ptrs.size());
    hipLaunchKernelGGL(k_linear_combination, dim3(grid(n)), dim3(kThreads), 0, c.stream,
                       static_cast<const Fr* const*>(dptrs->ptr),
                       static_cast<const Fr*>(dcoef->ptr), ptrs.size(),
                       static_cast<Fr*>(result->ptr), n);
    rt::check_launch();
    *out = c.publish(std::move(result));
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_mle_evaluate(HpMi355xContext* ctx,
                                                    HpMi355xDeviceAddress polynomial,
                                                    size_t num_vars, const uint8_t* point,
                                                    uint8_t* out_evaluation) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out_evaluation == nullptr) return rt::invalid_parameters();
    if (point == nullptr && num_vars != 0) return rt::invalid_parameters();
    if (!dims_ok(num_vars)) return rt::invalid_parameters();
    rt::BufferRef table = c.lookup(polynomial, num_vars);
    if (!table) return rt::invalid_parameters();
    if (num_vars && !rt::canonical_fr_array(point, num_vars)) return rt::invalid_parameters();
    rt::BufferRef dpoint = upload_small(c, point, num_vars);
    rt::BufferRef result = c.allocate(rt::kFrBytes);
    mle::evaluate(c, static_cast<const Fr*>(table->ptr), num_vars,
                  static_cast<const Fr*>(dpoint->ptr), static_cast<Fr*>(result->ptr));
    rt::download_fr(c, static_cast<const Fr*>(result->ptr), 1, out_evaluation);
    return rt::ok();
  });
}

// ---- evaluate_mle without allocation -------------------------------------
//
// One file-scope device array per GPU holds everything the evaluation needs
// beyond its input table and output: the point (Montgomery form), the value
// and the two fold tables (2^(nv-1) and 2^(nv-2) entries) for up to
// HP_MI355X_NOALLOC_MAX_NUM_VARS variables.  Every word is written before it
// is read within a call, so nothing has to start at zero.  All calls on one
// GPU share the array: calls that overlap on one GPU (two contexts of the
// device driven from two threads) race.
namespace hp {
namespace {

constexpr size_t kNoAllocMaxNumVars = HP_MI355X_NOALLOC_MAX_NUM_VARS;
constexpr size_t kNoAllocPointFr = 32;  // point coordinates (num_vars <= 31)
constexpr size_t kNoAllocScratchFr = kNoAllocPointFr + 2 + (size_t(1) << (kNoAllocMaxNumVars - 1)) +
                                     (size_t(1) << (kNoAllocMaxNumVars - 2));
static_assert(sizeof(Fr) == 32, "Fr is eight 32-bit words");

// 32-byte words; host code never names this array as a device address: it
// uses the per-device pointer from hipGetSymbolAddress below.
static __device__ __align__(256) uint4 scratch[2 * kNoAllocScratchFr];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mutex;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};
int32_t g_noalloc_error[kNoAllocMaxDevices] = {};
bool g_noalloc_done[kNoAllocMaxDevices] = {};

struct PointArg {
  Fr v[kNoAllocPointFr];
};

// the point (Montgomery form, by value) into the scratch point slots
__global__ void k_store_point(PointArg p, uint32_t count, Fr* dst) {
  if (threadIdx.x < count) store_fr(dst, threadIdx.x, p.v[threadIdx.x]);
}

// canonical value of a Montgomery cell
__global__ void k_canonical_value(const Fr* src, Fr* dst) {
  if (threadIdx.x == 0) store_fr(dst, 0, from_mont(load_fr(src, 0)));
}

}  // namespace

namespace mle {
void noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (g_noalloc_done[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    g_noalloc_error[device] = static_cast<int32_t>(e);
    ptr = nullptr;
  }
  g_noalloc_ptr[device] = ptr;
  g_noalloc_done[device] = true;
}
}  // namespace mle

}  // namespace hp

namespace {
// This device's scratch array, or a failure status.
HpMi355xStatus noalloc_scratch(int device, Fr** out) {
  *out = nullptr;
  if (device < 0 || device >= kNoAllocMaxDevices) return rt::invalid_state();
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_noalloc_done[device]) return rt::invalid_state();
  if (g_noalloc_ptr[device] == nullptr) return rt::invalid_state(g_noalloc_error[device]);
  *out = static_cast<Fr*>(g_noalloc_ptr[device]);
  return rt::ok();
}
}  // namespace

extern "C" HpMi355xStatus hp_mi355x_v1_mle_evaluate_noalloc(HpMi355xContext* ctx,
                                                            HpMi355xDeviceAddress polynomial,
                                                            size_t num_vars, const uint8_t* point,
                                                            uint8_t* out_evaluation) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    // the checks of mle_evaluate, then the size limit of the static scratch
    if (out_evaluation == nullptr) return rt::invalid_parameters();
    if (point == nullptr && num_vars != 0) return rt::invalid_parameters();
    if (!dims_ok(num_vars)) return rt::invalid_parameters();
    rt::BufferRef table = c.lookup(polynomial, num_vars);
    if (!table) return rt::invalid_parameters();
    if (num_vars && !rt::canonical_fr_array(point, num_vars)) return rt::invalid_parameters();
    if (num_vars > kNoAllocMaxNumVars)
      return rt::invalid_parameters(HP_MI355X_NATIVE_EXCEEDS_SCRATCH);
    Fr* base = nullptr;
    const HpMi355xStatus st = noalloc_scratch(c.device, &base);
    if (!rt::is_ok(st)) return st;
    Fr* dpoint = base;
    Fr* value = base + kNoAllocPointFr;       // Montgomery
    Fr* canonical = value + 1;
    Fr* folds = base + kNoAllocPointFr + 2;   // 2^(max-1) + 2^(max-2) entries
    PointArg arg;
    for (size_t i = 0; i < num_vars; ++i) {
      Fr v;
      std::memcpy(v.v, point + rt::kFrBytes * i, rt::kFrBytes);
      arg.v[i] = to_mont(v);
    }
    for (size_t i = num_vars; i < kNoAllocPointFr; ++i) arg.v[i] = Fr::zero();
    if (num_vars) {
      hipLaunchKernelGGL(k_store_point, dim3(1), dim3(64), 0, c.stream, arg,
                         static_cast<uint32_t>(num_vars), dpoint);
      rt::check_launch();
    }
    mle::evaluate(c, static_cast<const Fr*>(table->ptr), num_vars, dpoint, value, folds);
    hipLaunchKernelGGL(k_canonical_value, dim3(1), dim3(64), 0, c.stream, value, canonical);
    rt::check_launch();
    rt::hip_check(hipMemcpyAsync(out_evaluation, canonical, rt::kFrBytes, hipMemcpyDeviceToHost,
                                 c.stream));
    c.sync();
    return rt::ok();
  });
}

// Every (table, point) evaluation in one pass: one eq(point, .) table per
// distinct point, then one launch of dot products (f(r) = sum_x f[x] eq(r, x),
// the same field value as the fold chain) and one copy back.
extern "C" HpMi355xStatus hp_mi355x_v1_mle_evaluate_batch(
    HpMi355xContext* ctx, const HpMi355xDeviceAddress* polynomials, const size_t* num_vars,
    const uint8_t* const* points, uint8_t* out_evaluations, size_t count) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (count == 0) return rt::ok();
    if (polynomials == nullptr || num_vars == nullptr || points == nullptr ||
        out_evaluations == nullptr)
      return rt::invalid_parameters();
    std::vector<rt::BufferRef> tables(count);
    for (size_t i = 0; i < count; ++i) {
      if (!dims_ok(num_vars[i])) return rt::invalid_parameters();
      if (num_vars[i] && (points[i] == nullptr || !rt::canonical_fr_array(points[i], num_vars[i])))
        return rt::invalid_parameters();
      tables[i] = c.lookup(polynomials[i], num_vars[i]);
      if (!tables[i]) return rt::invalid_parameters();
    }
    // distinct points (by value), their eq tables
    std::vector<size_t> point_of(count);
    std::vector<size_t> firsts;
    for (size_t i = 0; i < count; ++i) {
      size_t d = 0;
      for (; d < firsts.size(); ++d) {
        const size_t j = firsts[d];
        if (num_vars[j] == num_vars[i] &&
            std::memcmp(points[j], points[i], num_vars[i] * rt::kFrBytes) == 0)
          break;
      }
      if (d == firsts.size()) firsts.push_back(i);
      point_of[i] = d;
    }
    std::vector<rt::BufferRef> eqs(firsts.size()), keep;
    {
      std::vector<const uint8_t*> pts;
      std::vector<size_t> nvs;
      std::vector<Fr*> outs;
      for (size_t d = 0; d < firsts.size(); ++d) {
        const size_t j = firsts[d];
        if (num_vars[j] == 0) continue;
        eqs[d] = c.allocate((size_t(1) << num_vars[j]) * rt::kFrBytes);
        pts.push_back(points[j]);
        nvs.push_back(num_vars[j]);
        outs.push_back(static_cast<Fr*>(eqs[d]->ptr));
      }
      mle::build_eq_batch(c, pts, nvs, outs, keep);
    }
    std::vector<DotJob> jobs(count);
    uint32_t grid_blocks = 0;
    for (size_t i = 0; i < count; ++i) {
      const size_t len = size_t(1) << num_vars[i];
      const size_t per_block = size_t(kThreads) * kDotPerThread;
      jobs[i].len = static_cast<uint32_t>(len);
      jobs[i].blocks = static_cast<uint32_t>(std::max<size_t>(1, (len + per_block - 1) / per_block));
      jobs[i].block_begin = grid_blocks;
      grid_blocks += jobs[i].blocks;
    }
    rt::BufferRef partial = c.allocate(size_t(grid_blocks) * rt::kFrBytes);
    rt::BufferRef tickets = c.allocate(count * sizeof(unsigned));
    // the values land in pinned host memory
    Fr* results = reinterpret_cast<Fr*>(c.results(count * rt::kFrBytes));
    for (size_t i = 0; i < count; ++i) {
      jobs[i].table = static_cast<const Fr*>(tables[i]->ptr);
      jobs[i].eq = eqs[point_of[i]] ? static_cast<const Fr*>(eqs[point_of[i]]->ptr) : nullptr;
      jobs[i].partial = static_cast<Fr*>(partial->ptr) + jobs[i].block_begin;
      jobs[i].ticket = static_cast<unsigned*>(tickets->ptr) + i;
      jobs[i].out = results + i;
    }
    rt::BufferRef djobs = c.allocate(count * sizeof(DotJob));
    uint8_t* stage = c.staging(count * sizeof(DotJob));
    std::memcpy(stage, jobs.data(), count * sizeof(DotJob));
    rt::hip_check(hipMemcpyAsync(djobs->ptr, stage, count * sizeof(DotJob), hipMemcpyHostToDevice,
                                 c.stream));
    c.staging_used();
    hipLaunchKernelGGL(k_zero_words, dim3(grid(count)), dim3(kThreads), 0, c.stream,
                       static_cast<unsigned*>(tickets->ptr), count);
    rt::check_launch();
    hipLaunchKernelGGL(k_eval_dots, dim3(grid_blocks), dim3(kThreads), 0, c.stream,
                       static_cast<const DotJob*>(djobs->ptr), static_cast<uint32_t>(count));
    rt::check_launch();
    c.sync();
    std::memcpy(out_evaluations, results, count * rt::kFrBytes);
    return rt::ok();
  });
}

// Time `reps` folds of one resident table by its lowest variable with HIP
// events on the context stream (the fold of fix_variables and of every
// sumcheck round: 2 reads and 1 write of 32 bytes per output entry).
extern "C" HpMi355xStatus hp_mi355x_v1_mle_fold_profile(HpMi355xContext* ctx,
                                                       HpMi355xDeviceAddress polynomial,
                                                       size_t num_vars, const uint8_t* r,
                                                       size_t reps, float* ms_per_fold) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (r == nullptr || ms_per_fold == nullptr || reps == 0 || num_vars == 0 || !dims_ok(num_vars))
      return rt::invalid_parameters();
    if (!rt::canonical_fr_array(r, 1)) return rt::invalid_parameters();
    rt::BufferRef table = c.lookup(polynomial, num_vars);
    if (!table) return rt::invalid_parameters();
    const size_t half = size_t(1) << (num_vars - 1);
    rt::BufferRef dr = upload_small(c, r, 1);
    rt::BufferRef out = c.allocate(half * rt::kFrBytes);
    hipEvent_t e0, e1;
    rt::hip_check(hipEventCreate(&e0));
    rt::hip_check(hipEventCreate(&e1));
    mle::fold(c, static_cast<const Fr*>(table->ptr), static_cast<Fr*>(out->ptr),
              static_cast<const Fr*>(dr->ptr), half);  // warm
    rt::hip_check(hipEventRecord(e0, c.stream));
    for (size_t i = 0; i < reps; ++i)
      mle::fold(c, static_cast<const Fr*>(table->ptr), static_cast<Fr*>(out->ptr),
                static_cast<const Fr*>(dr->ptr), half);
    rt::hip_check(hipEventRecord(e1, c.stream));
    rt::hip_check(hipEventSynchronize(e1));
    float ms = 0;
    rt::hip_check(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_fold = ms / static_cast<float>(reps);
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_mle_build_eq_x_r(HpMi355xContext* ctx, const uint8_t* r,
                                                        size_t num_vars,
                                                        HpMi355xDeviceAddress* out) {
  if (out) *out = nullptr;
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out == nullptr || r == nullptr) return rt::invalid_parameters();
    if (num_vars == 0 || !dims_ok(num_vars)) return rt::invalid_parameters();
    if (!rt::canonical_fr_array(r, num_vars)) return rt::invalid_parameters();
    rt::BufferRef dr = upload_small(c, r, num_vars);
    rt::BufferRef result = c.allocate((size_t(1) << num_vars) * rt::kFrBytes, num_vars);
    mle::build_eq(c, static_cast<const Fr*>(dr->ptr), num_vars, static_cast<Fr*>(result->ptr));
    *out = c.publish(std::move(result));
    return rt::ok();
  });
}

// a R -> 1/a R (Montgomery form, R = 2^256; a != 0) on the host: the binary
// extended Euclidean algorithm on the integer A = a R gives A^-1 = 1/a R^-1,
// and one Montgomery product by R^3 mod r returns 1/a R.
Fr host_inverse_fr(const Fr& a) {
  using u64 = uint64_t;
  using u128 = unsigned __int128;
  auto limbs = [](const uint32_t* w, u64* x) {
    for (int i = 0; i < 4; ++i) x[i] = (u64)w[2 * i] | ((u64)w[2 * i + 1] << 32);
  };
  u64 r[4], u[4], v[4], x1[4] = {1, 0, 0, 0}, x2[4] = {0, 0, 0, 0};
  limbs(FrParams::MOD, r);
  limbs(a.v, u);
  std::memcpy(v, r, sizeof v);
  auto is_one = [](const u64* x) { return x[0] == 1 && (x[1] | x[2] | x[3]) == 0; };
  auto geq = [](const u64* x, const u64* y) {
    for (int i = 3; i >= 0; --i)
      if (x[i] != y[i]) return x[i] > y[i];
    return true;
  };
  auto sub = [](u64* x, const u64* y) {  // x -= y, no underflow
    u64 borrow = 0;
    for (int i = 0; i < 4; ++i) {
      const u128 d = (u128)x[i] - y[i] - borrow;
      x[i] = (u64)d;
      borrow = (u64)(d >> 64) & 1;
    }
  };
  auto half_mod = [&](u64* x) {  // x / 2 mod r for x < r (x + r < 2^256)
    if (x[0] & 1) {
      u64 carry = 0;
      for (int i = 0; i < 4; ++i) {
        const u128 t = (u128)x[i] + r[i] + carry;
        x[i] = (u64)t;
        carry = (u64)(t >> 64);
      }
    }
    for (int i = 0; i < 3; ++i) x[i] = (x[i] >> 1) | (x[i + 1] << 63);
    x[3] >>= 1;
  };
  auto half = [](u64* x) {
    for (int i = 0; i < 3; ++i) x[i] = (x[i] >> 1) | (x[i + 1] << 63);
    x[3] >>= 1;
  };
  auto sub_mod = [&](u64* x, const u64* y) {  // x - y mod r, both < r
    if (geq(x, y)) {
      sub(x, y);
    } else {
      u64 t[4];
      std::memcpy(t, r, sizeof t);
      sub(t, y);
      u64 carry = 0;
      for (int i = 0; i < 4; ++i) {
        const u128 s = (u128)x[i] + t[i] + carry;
        x[i] = (u64)s;
        carry = (u64)(s >> 64);
      }
    }
  };
  // invariants: x1 A = u, x2 A = v (mod r); u, v > 0 (A != 0 mod r)
  while (!is_one(u) && !is_one(v)) {
    while ((u[0] & 1) == 0) {
      half(u);
      half_mod(x1);
    }
    while ((v[0] & 1) == 0) {
      half(v);
      half_mod(x2);
    }
    if (geq(u, v)) {
      sub(u, v);
      sub_mod(x1, x2);
    } else {
      sub(v, u);
      sub_mod(x2, x1);
    }
  }
  const u64* x = is_one(u) ? x1 : x2;
  Fr ainv, r2;
  for (int i = 0; i < 4; ++i) {
    ainv.v[2 * i] = (uint32_t)x[i];
    ainv.v[2 * i + 1] = (uint32_t)(x[i] >> 32);
  }
  std::memcpy(r2.v, FrParams::R2, sizeof r2.v);
  const Fr r3 = r2 * r2;  // R^4 / R
  return ainv * r3;
}

extern "C" HpMi355xStatus hp_mi355x_v1_mle_compute_frac_poly(
    HpMi355xContext* ctx, size_t num_vars, const HpMi355xDeviceAddress* numerators,
    const HpMi355xDeviceAddress* denominators, size_t count, HpMi355xDeviceAddress* out) {
  if (out) *out = nullptr;
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out == nullptr || numerators == nullptr || denominators == nullptr || count == 0)
      return rt::invalid_parameters();
    if (!dims_ok(num_vars)) return rt::invalid_parameters();
    std::vector<rt::BufferRef> keep;
    std::vector<const void*> nums, dens;
    for (size_t k = 0; k < count; ++k) {
      rt::BufferRef f = c.lookup(numerators[k], num_vars);
      rt::BufferRef g = c.lookup(denominators[k], num_vars);
      if (!f || !g) return rt::invalid_parameters();
      nums.push_back(f->ptr);
      dens.push_back(g->ptr);
      keep.push_back(f);
      keep.push_back(g);
    }
    const size_t n = size_t(1) << num_vars;
    rt::BufferRef result = c.allocate(n * rt::kFrBytes, num_vars);
    rt::BufferRef dn = upload_pointers(c, nums);
    rt::BufferRef dd = upload_pointers(c, dens);
    rt::BufferRef aux_buf = c.allocate(sizeof(FracAux));
    FracAux* aux = static_cast<FracAux*>(aux_buf->ptr);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c.stream, &aux->zero);
    rt::check_launch();
    // T threads (a power of two): small tables one entry per thread, large
    // ones ~16 entries per thread with 16384 to 65536 threads
    size_t T = 1;
    while (T < std::min<size_t>(n, std::max<size_t>(16384, std::min<size_t>(n / 16, 65536)))) T *= 2;
    T = std::min(T, n);
    rt::BufferRef tree_buf = c.allocate((2 * T - 1) * rt::kFrBytes);
    Fr* tree = static_cast<Fr*>(tree_buf->ptr);
    const auto* dnp = static_cast<const Fr* const*>(dn->ptr);
    const auto* ddp = static_cast<const Fr* const*>(dd->ptr);
    Fr* res = static_cast<Fr*>(result->ptr);
    hipLaunchKernelGGL(k_frac_prefix, dim3(rt::blocks_for(T, 64)), dim3(64), 0, c.stream, ddp, count,
                       res, n, tree, aux);
    rt::check_launch();
    hipLaunchKernelGGL(k_frac_up, dim3(1), dim3(1024), 0, c.stream, tree, T, aux);
    rt::check_launch();
    FracAux host{};
    rt::download_bytes(c, aux, sizeof(FracAux), reinterpret_cast<uint8_t*>(&host));
    if (host.zero) return rt::invalid_parameters();  // result buffer is released unpublished
    hipLaunchKernelGGL(k_frac_down, dim3(1), dim3(1024), 0, c.stream, tree, T,
                       host_inverse_fr(host.root));
    rt::check_launch();
    hipLaunchKernelGGL(k_frac_back, dim3(rt::blocks_for(T, 64)), dim3(64), 0, c.stream, dnp, ddp,
                       count, res, n, static_cast<const Fr*>(tree));
    rt::check_launch();
    *out = c.publish(std::move(result));
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_mle_compute_product_poly(HpMi355xContext* ctx,
                                                                HpMi355xDeviceAddress frac,
                                                                size_t num_vars,
                                                                HpMi355xDeviceAddress* out) {
  if (out) *out = nullptr;
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out == nullptr) return rt::invalid_parameters();
    if (num_vars == 0 || !dims_ok(num_vars)) return rt::invalid_parameters();
    rt::BufferRef f = c.lookup(frac, num_vars);
    if (!f) return rt::invalid_parameters();
    const size_t n = size_t(1) << num_vars;
    rt::BufferRef result = c.allocate(n * rt::kFrBytes, num_vars);
    Fr* o = static_cast<Fr*>(result->ptr);
    // level 1 from frac, then level by level until <= 1024 entries remain
    size_t len = n / 2;
    hipLaunchKernelGGL(k_pair_products, dim3(grid(len)), dim3(kThreads), 0, c.stream,
                       static_cast<const Fr*>(f->ptr), o, len);
    rt::check_launch();
    Fr* level = o;
    while (len > 1024) {
      hipLaunchKernelGGL(k_pair_products, dim3(grid(len / 2)), dim3(kThreads), 0, c.stream, level,
                         level + len, len / 2);
      rt::check_launch();
      level += len;
      len /= 2;
    }
    hipLaunchKernelGGL(k_product_tail, dim3(1), dim3(256), 0, c.stream, level, len);
    rt::check_launch();
    *out = c.publish(std::move(result));
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_mle_compute_nums_and_denoms(
    HpMi355xContext* ctx, size_t num_vars, const uint8_t* beta, const uint8_t* gamma,
    const HpMi355xDeviceAddress* fxs, const HpMi355xDeviceAddress* gxs,
    const HpMi355xDeviceAddress* permutations, size_t count,
    HpMi355xDeviceAddress* out_numerators, HpMi355xDeviceAddress* out_denominators) {
  if (out_numerators && out_denominators)
    for (size_t i = 0; i < count; ++i) out_numerators[i] = out_denominators[i] = nullptr;
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (beta == nullptr || gamma == nullptr || fxs == nullptr || gxs == nullptr ||
        permutations == nullptr || out_numerators == nullptr || out_denominators == nullptr ||
        count == 0)
      return rt::invalid_parameters();
    if (!dims_ok(num_vars)) return rt::invalid_parameters();
    std::vector<rt::BufferRef> keep;
    std::vector<const void*> f, g, p;
    for (size_t k = 0; k < count; ++k) {
      rt::BufferRef a = c.lookup(fxs[k], num_vars), b = c.lookup(gxs[k], num_vars),
                    d = c.lookup(permutations[k], num_vars);
      if (!a || !b || !d) return rt::invalid_parameters();
      f.push_back(a->ptr);
      g.push_back(b->ptr);
      p.push_back(d->ptr);
      keep.insert(keep.end(), {a, b, d});
    }
    if (!rt::canonical_fr(beta) || !rt::canonical_fr(gamma)) return rt::invalid_parameters();
    if (count > (size_t(1) << 20)) return rt::invalid_parameters();
    const size_t n = size_t(1) << num_vars;
    std::vector<rt::BufferRef> nums, dens;
    std::vector<const void*> np, dp;
    for (size_t k = 0; k < count; ++k) {
      nums.push_back(c.allocate(n * rt::kFrBytes, num_vars));
      dens.push_back(c.allocate(n * rt::kFrBytes, num_vars));
      np.push_back(nums.back()->ptr);
      dp.push_back(dens.back()->ptr);
    }
    uint8_t bg[64];
    std::memcpy(bg, beta, 32);
    std::memcpy(bg + 32, gamma, 32);
    rt::BufferRef dbg = upload_small(c, bg, 2);
    rt::BufferRef df = upload_pointers(c, f), dgp = upload_pointers(c, g),
                  dpp = upload_pointers(c, p), dn = upload_pointers(c, np),
                  dd = upload_pointers(c, dp);
    hipLaunchKernelGGL(k_nums_denoms, dim3(grid(n * count)), dim3(kThreads), 0, c.stream,
                       static_cast<const Fr* const*>(df->ptr),
                       static_cast<const Fr* const*>(dgp->ptr),
                       static_cast<const Fr* const*>(dpp->ptr), static_cast<const Fr*>(dbg->ptr),
                       static_cast<Fr* const*>(dn->ptr), static_cast<Fr* const*>(dd->ptr), count,
                       n);
    rt::check_launch();
    for (size_t k = 0; k < count; ++k) {
      out_numerators[k] = c.publish(nums[k]);
      out_denominators[k] = c.publish(dens[k]);
    }
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_mle_compute_product_factors(
    HpMi355xContext* ctx, HpMi355xDeviceAddress frac, HpMi355xDeviceAddress product,
    size_t num_vars, HpMi355xDeviceAddress* out_p1, HpMi355xDeviceAddress* out_p2) {
  if (out_p1) *out_p1 = nullptr;
  if (out_p2) *out_p2 = nullptr;
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out_p1 == nullptr || out_p2 == nullptr) return rt::invalid_parameters();
    if (num_vars == 0 || !dims_ok(num_vars)) return rt::invalid_parameters();
    rt::BufferRef f = c.lookup(frac, num_vars), p = c.lookup(product, num_vars);
    if (!f || !p) return rt::invalid_parameters();
    const size_t n = size_t(1) << num_vars;
    rt::BufferRef a = c.allocate(n * rt::kFrBytes, num_vars);
    rt::BufferRef b = c.allocate(n * rt::kFrBytes, num_vars);
    hipLaunchKernelGGL(k_product_factors, dim3(grid(n)), dim3(kThreads), 0, c.stream,
                       static_cast<const Fr*>(f->ptr), static_cast<const Fr*>(p->ptr),
                       static_cast<Fr*>(a->ptr), static_cast<Fr*>(b->ptr), n / 2);
    rt::check_launch();
    *out_p1 = c.publish(std::move(a));
    *out_p2 = c.publish(std::move(b));
    return rt::ok();
  });
}
