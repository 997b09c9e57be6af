// Multilinear KZG (MultilinearKzgBackend, backend/src/multilinear_kzg.rs:20-55)
// over a device-resident prepared SRS.
//
//   commit(f)       = MSM(layer[N - nv], f)                (cpu/multilinear_kzg.rs:40-59)
//   multi_commit    = independent commits, order and duplicates kept; all MSMs
//                     are queued back to back and synchronised once
//                     (the reference bridge syncs after each, bridge_kzg.cpp:555-575)
//   open(f, x)      = for j < nv: q[b] = f[2b+1] - f[2b], f[b] = f[2b] + x_j q[b],
//                     pi_j = MSM(layer[N - nv + 1 + j], q)  (cpu/multilinear_kzg.rs:79-137)
// Prepared layers hold window-shifted copies of every point (see msm.hip).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kernels_common.hpp"
#include "mle.hpp"
#include "msm.hpp"

namespace hp {
namespace kzg {

struct Params {
  size_t num_vars = 0;
  std::vector<msm::Layer> layers;  // N + 1
  std::vector<bool> ready;
  uint32_t window_override = 0;    // 0 = plan by size
};

namespace {

__global__ void k_fold_quotient(const Fr* src, Fr* dst, Fr* q, const Fr* x, size_t half) {
  const Fr c = load_fr(x, 0);
  for (size_t b = blockIdx.x * (size_t)blockDim.x + threadIdx.x; b < half;
       b += (size_t)gridDim.x * blockDim.x) {
    const Fr lo = load_fr(src, 2 * b), hi = load_fr(src, 2 * b + 1);
    const Fr d = hi - lo;
    store_fr(q, b, d);
    store_fr(dst, b, lo + c * d);
  }
}

// The last `levels` fold levels of an open in one block, from a table of
// 2^levels <= 1024 entries: each level's quotients to q (consecutive, as the
// level launches write them) and the final value to out[0]; the same field
// operations as k_fold_quotient, without a launch per level.
constexpr uint32_t kFoldTailMax = 10;
__global__ void __launch_bounds__(512) k_fold_quotient_tail(const Fr* src, Fr* q, const Fr* x,
                                                            uint32_t levels, Fr* out) {
  __shared__ Fr buf[size_t(1) << kFoldTailMax];
  size_t len = size_t(1) << levels;
  for (size_t i = threadIdx.x; i < len; i += blockDim.x) buf[i] = load_fr(src, i);
  __syncthreads();
  for (uint32_t l = 0; l < levels; ++l) {
    const Fr c = load_fr(x, l);
    const size_t half = len / 2;  // <= 512 = blockDim.x
    Fr v = Fr::zero();
    if (threadIdx.x < half) {
      const Fr lo = buf[2 * threadIdx.x], hi = buf[2 * threadIdx.x + 1];
      const Fr d = hi - lo;
      store_fr(q, threadIdx.x, d);
      v = lo + c * d;
    }
    __syncthreads();
    if (threadIdx.x < half) buf[threadIdx.x] = v;
    __syncthreads();
    q += half;
    len = half;
  }
  if (threadIdx.x == 0) store_fr(out, 0, buf[0]);
}

unsigned grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(b, 65535)));
}

}  // namespace
}  // namespace kzg
}  // namespace hp

struct HpMi355xKzgParams {
  HpMi355xContext* ctx;
  std::shared_ptr<hp::kzg::Params> impl;
};

namespace hp {
namespace kzg {
// An asynchronous batch: its work runs on the context's side stream; every
// buffer it reads or writes is held here until the wait.
struct Ticket {
  std::shared_ptr<Params> params;  // the layers the batch reads
  std::vector<rt::BufferRef> hold;
  msm::Batch batch;
  G1Xyzz* slots_host = nullptr;  // pinned: level-1 results of every MSM
  Fr* values_host = nullptr;     // pinned: opened values, stored form
  std::vector<uint8_t*> packet_outs, value_outs;
  hipEvent_t done = nullptr;
  ~Ticket() {
    if (slots_host) (void)hipHostFree(slots_host);
    if (values_host) (void)hipHostFree(values_host);
    if (done) (void)hipEventDestroy(done);
  }
};
}  // namespace kzg
}  // namespace hp

struct HpMi355xKzgTicket {
  HpMi355xContext* ctx;
  std::shared_ptr<hp::kzg::Ticket> impl;
};

using namespace hp;

namespace {

kzg::Params* bind(HpMi355xContext* ctx, rt::Context& c, const HpMi355xKzgParams* params) {
  // membership first: nothing is read through a handle this context does not hold
  if (params == nullptr || !c.holds(params, rt::Context::Kind::kzg_params)) return nullptr;
  if (params->ctx != ctx) return nullptr;
  return params->impl.get();
}

static_assert(sizeof(G1Xyzz) % 32 == 0, "Fr values after the result slots stay 32-byte aligned");

// Run a batch of MSMs and write one 97-byte packet per MSM; `extra` Fr
// values on the device come back canonical into extra_out under the same
// synchronisation.
void run_batch(rt::Context& c, const msm::Batch& batch, uint8_t* out97, const Fr* extra = nullptr,
               size_t extra_count = 0, uint8_t* extra_out = nullptr) {
  const size_t slots = batch.result_slots();
  // the level-1 reduction writes its points straight into pinned host memory
  // (G1Xyzz is a multiple of 32 bytes: the extra values stay aligned)
  const size_t slot_bytes = std::max<size_t>(slots, 1) * sizeof(G1Xyzz);
  uint8_t* base = c.results(slot_bytes + extra_count * rt::kFrBytes);
  G1Xyzz* host = reinterpret_cast<G1Xyzz*>(base);
  if (extra_count) rt::enqueue_from_mont(c, extra, extra_count, reinterpret_cast<Fr*>(base + slot_bytes));
  batch.enqueue(c, host);
  c.sync();
  batch.finish_packets(host, out97);
  if (extra_count) std::memcpy(extra_out, base + slot_bytes, extra_count * rt::kFrBytes);
}

}  // namespace

extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_create(HpMi355xContext* ctx,
                                                                     size_t num_vars,
                                                                     HpMi355xKzgParams** out) {
  if (out) *out = nullptr;
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out == nullptr || num_vars > 30) return rt::invalid_parameters();
    auto p = std::make_shared<kzg::Params>();
    p->num_vars = num_vars;
    p->layers.resize(num_vars + 1);
    p->ready.assign(num_vars + 1, false);
    if (const char* env = std::getenv("HP_MI355X_MSM_WINDOW")) {
      const int v = std::atoi(env);
      if (v >= 4 && v <= 20) p->window_override = static_cast<uint32_t>(v);
    }
    auto* handle = new HpMi355xKzgParams{ctx, p};
    c.resources[handle] = {rt::Context::Kind::kzg_params, p};
    *out = handle;
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_upload_layer(
    HpMi355xContext* ctx, HpMi355xKzgParams* params, size_t layer_index, const uint8_t* points,
    size_t point_count) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    kzg::Params* p = bind(ctx, c, params);
    if (!p || points == nullptr || layer_index > p->num_vars) return rt::invalid_parameters();
    if (point_count != (size_t(1) << (p->num_vars - layer_index))) return rt::invalid_parameters();
    msm::Layer layer;
    if (!msm::upload_layer(c, layer, points, point_count, p->window_override))
      return rt::invalid_parameters();
    c.sync();
    p->layers[layer_index] = std::move(layer);
    p->ready[layer_index] = true;
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_generate(HpMi355xContext* ctx,
                                                                       HpMi355xKzgParams* params,
                                                                       const uint8_t* t) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    kzg::Params* p = bind(ctx, c, params);
    if (!p) return rt::invalid_parameters();
    const size_t N = p->num_vars;
    if (N > 0 && (t == nullptr || !rt::canonical_fr_array(t, N))) return rt::invalid_parameters();
    rt::BufferRef tr = c.allocate(std::max<size_t>(N, 1) * rt::kFrBytes);
    if (N) rt::upload_fr_values(c, static_cast<Fr*>(tr->ptr), t, N);
    rt::BufferRef table = msm::make_fixed_table(c);
    for (size_t i = 0; i <= N; ++i) {
      const size_t len = N - i, n = size_t(1) << len;
      rt::BufferRef scal = c.allocate(n * rt::kFrBytes);
      if (len == 0) {
        uint8_t one[32] = {1};
        rt::upload_fr_values(c, static_cast<Fr*>(scal->ptr), one, 1);
      } else {
        mle::build_eq(c, static_cast<const Fr*>(tr->ptr) + i, len, static_cast<Fr*>(scal->ptr));
      }
      rt::BufferRef pts = c.allocate(n * sizeof(G1Affine));
      msm::generate_points(c, static_cast<const Fr*>(scal->ptr), n,
                           static_cast<G1Affine*>(pts->ptr),
                           static_cast<const G1Affine*>(table->ptr));
      msm::Layer layer;
      msm::prepare_layer_from_device(c, layer, static_cast<const G1Affine*>(pts->ptr), n,
                                     p->window_override);
      p->layers[i] = std::move(layer);
      p->ready[i] = true;
    }
    c.sync();
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_download_layer(
    HpMi355xContext* ctx, const HpMi355xKzgParams* params, size_t layer_index, uint8_t* points,
    size_t point_count) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    kzg::Params* p = bind(ctx, c, params);
    if (!p || points == nullptr || layer_index > p->num_vars) return rt::invalid_parameters();
    if (point_count != (size_t(1) << (p->num_vars - layer_index))) return rt::invalid_parameters();
    if (!p->ready[layer_index]) return rt::invalid_state();
    msm::download_layer(c, p->layers[layer_index], points);
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_upload_layer_compressed(
    HpMi355xContext* ctx, HpMi355xKzgParams* params, size_t layer_index, const uint8_t* compressed,
    size_t point_count) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    kzg::Params* p = bind(ctx, c, params);
    if (!p || compressed == nullptr || layer_index > p->num_vars) return rt::invalid_parameters();
    if (point_count != (size_t(1) << (p->num_vars - layer_index))) return rt::invalid_parameters();
    msm::Layer layer;
    if (!msm::upload_layer_compressed(c, layer, compressed, point_count, p->window_override))
      return rt::invalid_parameters();
    c.sync();
    p->layers[layer_index] = std::move(layer);
    p->ready[layer_index] = true;
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_download_layer_compressed(
    HpMi355xContext* ctx, const HpMi355xKzgParams* params, size_t layer_index, uint8_t* compressed,
    size_t point_count) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    kzg::Params* p = bind(ctx, c, params);
    if (!p || compressed == nullptr || layer_index > p->num_vars) return rt::invalid_parameters();
    if (point_count != (size_t(1) << (p->num_vars - layer_index))) return rt::invalid_parameters();
    if (!p->ready[layer_index]) return rt::invalid_state();
    msm::download_layer_compressed(c, p->layers[layer_index], compressed);
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_params_destroy(HpMi355xContext* ctx,
                                                                      HpMi355xKzgParams* params) {
  return hp_guarded_release(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (!bind(ctx, c, params)) return rt::invalid_parameters();
    c.resources.erase(params);
    delete params;
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_commit(HpMi355xContext* ctx,
                                                              const HpMi355xKzgParams* params,
                                                              HpMi355xDeviceAddress polynomial,
                                                              size_t num_vars,
                                                              uint8_t* out_commitment) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    kzg::Params* p = bind(ctx, c, params);
    if (!p || out_commitment == nullptr) return rt::invalid_parameters();
    if (num_vars > p->num_vars) return rt::invalid_parameters();
    rt::BufferRef f = c.lookup(polynomial, num_vars);
    if (!f) return rt::invalid_parameters();
    const size_t layer = p->num_vars - num_vars;
    if (!p->ready[layer]) return rt::invalid_state();
    msm::Batch batch;
    batch.add(p->layers[layer], static_cast<const Fr*>(f->ptr));
    run_batch(c, batch, out_commitment);
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_multi_commit(
    HpMi355xContext* ctx, const HpMi355xKzgParams* params, const HpMi355xDeviceAddress* polynomials,
    const size_t* num_vars, size_t count, uint8_t* out_commitments) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    kzg::Params* p = bind(ctx, c, params);
    if (!p) return rt::invalid_parameters();
    if (count == 0) return rt::ok();
    if (polynomials == nullptr || num_vars == nullptr || out_commitments == nullptr)
      return rt::invalid_parameters();
    // validate the whole batch before any side effect
    std::vector<rt::BufferRef> tables(count);
    for (size_t i = 0; i < count; ++i) {
      if (num_vars[i] > p->num_vars) return rt::invalid_parameters();
      tables[i] = c.lookup(polynomials[i], num_vars[i]);
      if (!tables[i]) return rt::invalid_parameters();
    }
    for (size_t i = 0; i < count; ++i)
      if (!p->ready[p->num_vars - num_vars[i]]) return rt::invalid_state();
    msm::Batch batch;
    for (size_t i = 0; i < count; ++i)
      batch.add(p->layers[p->num_vars - num_vars[i]], static_cast<const Fr*>(tables[i]->ptr));
    run_batch(c, batch, out_commitments);
    return rt::ok();
  });
}

namespace {

// Folds of one open, queued on the stream: quotient j (2^(nv-1-j) entries)
// lands at q + (n - 2^(nv-j)) and its MSM against layer N - nv + 1 + j joins
// `batch`; *value points at the device cell holding f(point) afterwards.
// The buffers stay alive through `keep`.
void enqueue_open(rt::Context& c, const kzg::Params& p, const Fr* table, size_t num_vars,
                  const uint8_t* point, msm::Batch& batch, std::vector<rt::BufferRef>& keep,
                  const Fr** value) {
  if (num_vars == 0) {
    *value = table;
    return;
  }
  const size_t N = p.num_vars;
  const size_t n = size_t(1) << num_vars;
  rt::BufferRef dpoint = c.allocate(num_vars * rt::kFrBytes);
  rt::upload_fr_values(c, static_cast<Fr*>(dpoint->ptr), point, num_vars);
  // folded tables ping-pong in fa/fb; the quotients are kept so that all
  // MSMs run as one batch afterwards
  rt::BufferRef work = c.allocate((n / 2 + n / 4 + 1 + n) * rt::kFrBytes);
  Fr* fa = static_cast<Fr*>(work->ptr);
  Fr* fb = fa + n / 2;
  Fr* q = fb + n / 4 + 1;
  const Fr* src = table;
  size_t qoff = 0;
  for (size_t j = 0; j < num_vars; ++j) {
    const size_t half = n >> (j + 1);
    if (num_vars - j <= kzg::kFoldTailMax) {
      // the remaining levels in one block; the value lands where the level
      // launches would have left it
      Fr* last = ((num_vars - 1) & 1) ? fb : fa;
      hipLaunchKernelGGL(kzg::k_fold_quotient_tail, dim3(1), dim3(512), 0, c.stream, src, q + qoff,
                         static_cast<const Fr*>(dpoint->ptr) + j, static_cast<uint32_t>(num_vars - j), last);
      rt::check_launch();
      for (size_t jj = j; jj < num_vars; ++jj) {
        batch.add(p.layers[N - num_vars + 1 + jj], q + qoff);
        qoff += n >> (jj + 1);
      }
      src = last;
      break;
    }
    Fr* dst = (j & 1) ? fb : fa;
    hipLaunchKernelGGL(kzg::k_fold_quotient, dim3(kzg::grid_for(half)), dim3(256), 0, c.stream,
                       src, dst, q + qoff, static_cast<const Fr*>(dpoint->ptr) + j, half);
    rt::check_launch();
    batch.add(p.layers[N - num_vars + 1 + j], q + qoff);
    qoff += half;
    src = dst;
  }
  *value = src;
  keep.push_back(std::move(dpoint));
  keep.push_back(std::move(work));
}

// f(point) cells of several opens -> canonical bytes, one copy back
__global__ void k_gather_values(const Fr* const* cells, size_t count, Fr* out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    store_fr(out, i, load_fr(cells[i], 0));
}

void download_values(rt::Context& c, const std::vector<const Fr*>& cells, uint8_t* const* outs) {
  if (cells.empty()) return;
  const size_t k = cells.size();
  rt::BufferRef ptrs = c.allocate(k * sizeof(void*));
  uint8_t* stage = c.staging(k * sizeof(void*));
  std::memcpy(stage, cells.data(), k * sizeof(void*));
  rt::hip_check(hipMemcpyAsync(ptrs->ptr, stage, k * sizeof(void*), hipMemcpyHostToDevice, c.stream));
  c.staging_used();
  rt::BufferRef vals = c.allocate(k * rt::kFrBytes);
  hipLaunchKernelGGL(k_gather_values, dim3(1), dim3(256), 0, c.stream,
                     static_cast<const Fr* const*>(ptrs->ptr), k, static_cast<Fr*>(vals->ptr));
  rt::check_launch();
  std::vector<uint8_t> host(k * rt::kFrBytes);
  rt::download_fr(c, static_cast<const Fr*>(vals->ptr), k, host.data());
  for (size_t i = 0; i < k; ++i) std::memcpy(outs[i], host.data() + i * rt::kFrBytes, rt::kFrBytes);
}

}  // namespace

extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_open(
    HpMi355xContext* ctx, const HpMi355xKzgParams* params, HpMi355xDeviceAddress polynomial,
    size_t num_vars, const uint8_t* point, uint8_t* out_quotients, uint8_t* out_evaluation) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    kzg::Params* p = bind(ctx, c, params);
    if (!p || out_evaluation == nullptr) return rt::invalid_parameters();
    if (num_vars > p->num_vars) return rt::invalid_parameters();
    if (num_vars > 0 && (point == nullptr || out_quotients == nullptr))
      return rt::invalid_parameters();
    rt::BufferRef f = c.lookup(polynomial, num_vars);
    if (!f) return rt::invalid_parameters();
    if (num_vars && !rt::canonical_fr_array(point, num_vars)) return rt::invalid_parameters();
    const size_t N = p->num_vars;
    for (size_t j = 0; j < num_vars; ++j)
      if (!p->ready[N - num_vars + 1 + j]) return rt::invalid_state();
    msm::Batch batch;
    std::vector<rt::BufferRef> keep;
    const Fr* value = nullptr;
    enqueue_open(c, *p, static_cast<const Fr*>(f->ptr), num_vars, point, batch, keep, &value);
    if (batch.size())
      run_batch(c, batch, out_quotients, value, 1, out_evaluation);
    else
      rt::download_fr(c, value, 1, out_evaluation);
    return rt::ok();
  });
}

// Every request of a batch, before any work: arguments, then layer readiness.
static HpMi355xStatus check_requests(rt::Context& c, const kzg::Params& p,
                                     const HpMi355xKzgRequest* requests, size_t count,
                                     std::vector<rt::BufferRef>& tables) {
  const size_t N = p.num_vars;
  for (size_t i = 0; i < count; ++i) {
    const HpMi355xKzgRequest& r = requests[i];
    if (r.kind != HP_MI355X_KZG_COMMIT && r.kind != HP_MI355X_KZG_OPEN)
      return rt::invalid_parameters();
    if (r.num_vars > N) return rt::invalid_parameters();
    tables[i] = c.lookup(r.polynomial, r.num_vars);
    if (!tables[i]) return rt::invalid_parameters();
    if (r.kind == HP_MI355X_KZG_COMMIT) {
      if (r.out_points == nullptr) return rt::invalid_parameters();
    } else {
      if (r.out_evaluation == nullptr) return rt::invalid_parameters();
      if (r.num_vars > 0 && (r.point == nullptr || r.out_points == nullptr))
        return rt::invalid_parameters();
      if (r.num_vars && !rt::canonical_fr_array(r.point, r.num_vars)) return rt::invalid_parameters();
    }
  }
  for (size_t i = 0; i < count; ++i) {
    const HpMi355xKzgRequest& r = requests[i];
    if (r.kind == HP_MI355X_KZG_COMMIT) {
      if (!p.ready[N - r.num_vars]) return rt::invalid_state();
    } else {
      for (size_t j = 0; j < r.num_vars; ++j)
        if (!p.ready[N - r.num_vars + 1 + j]) return rt::invalid_state();
    }
  }
  return rt::ok();
}

// Commits and opens of many polynomials with all their MSMs in one batch.
extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_batch(HpMi355xContext* ctx,
                                                             const HpMi355xKzgParams* params,
                                                             const HpMi355xKzgRequest* requests,
                                                             size_t count) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    kzg::Params* p = bind(ctx, c, params);
    if (!p) return rt::invalid_parameters();
    if (count == 0) return rt::ok();
    if (requests == nullptr) return rt::invalid_parameters();
    const size_t N = p->num_vars;
    std::vector<rt::BufferRef> tables(count);
    const HpMi355xStatus st = check_requests(c, *p, requests, count, tables);
    if (!rt::is_ok(st)) return st;
    msm::Batch batch;
    std::vector<rt::BufferRef> keep;
    std::vector<const Fr*> cells;
    std::vector<uint8_t*> value_outs;
    std::vector<uint8_t*> packet_outs;  // one per MSM, in batch order
    for (size_t i = 0; i < count; ++i) {
      const HpMi355xKzgRequest& r = requests[i];
      const Fr* table = static_cast<const Fr*>(tables[i]->ptr);
      if (r.kind == HP_MI355X_KZG_COMMIT) {
        batch.add(p->layers[N - r.num_vars], table);
        packet_outs.push_back(r.out_points);
      } else {
        const Fr* value = nullptr;
        enqueue_open(c, *p, table, r.num_vars, r.point, batch, keep, &value);
        for (size_t j = 0; j < r.num_vars; ++j) packet_outs.push_back(r.out_points + rt::kG1Bytes * j);
        cells.push_back(value);
        value_outs.push_back(r.out_evaluation);
      }
    }
    if (batch.size()) {
      std::vector<uint8_t> packets(batch.size() * rt::kG1Bytes);
      run_batch(c, batch, packets.data());
      for (size_t j = 0; j < packet_outs.size(); ++j)
        std::memcpy(packet_outs[j], packets.data() + rt::kG1Bytes * j, rt::kG1Bytes);
    }
    download_values(c, cells, value_outs.data());
    return rt::ok();
  });
}

// The requests of multilinear_kzg_batch enqueued on the context's side stream
// behind everything already queued on the context stream; the call returns
// at once and the results land in the requests' outputs at the wait, so
// other work (sumcheck rounds) overlaps the MSMs on the device.
extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_batch_submit(
    HpMi355xContext* ctx, const HpMi355xKzgParams* params, const HpMi355xKzgRequest* requests,
    size_t count, HpMi355xKzgTicket** out) {
  if (out) *out = nullptr;
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (out == nullptr) return rt::invalid_parameters();
    kzg::Params* p = bind(ctx, c, params);
    if (!p) return rt::invalid_parameters();
    if (count > 0 && requests == nullptr) return rt::invalid_parameters();
    std::vector<rt::BufferRef> tables(count);
    const HpMi355xStatus st = check_requests(c, *p, requests, count, tables);
    if (!rt::is_ok(st)) return st;
    auto t = std::make_shared<kzg::Ticket>();
    t->params = std::static_pointer_cast<kzg::Params>(c.resources.at(params).object);
    t->hold = tables;
    const size_t N = p->num_vars;
    hipStream_t side = c.side_stream();
    {
      hipEvent_t ready = nullptr;
      rt::hip_check(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
      const hipError_t r1 = hipEventRecord(ready, c.stream);
      const hipError_t r2 = r1 == hipSuccess ? hipStreamWaitEvent(side, ready, 0) : r1;
      (void)hipEventDestroy(ready);
      rt::hip_check(r2);
    }
    rt::StreamSwap swap(c, side);
    std::vector<const Fr*> cells;
    for (size_t i = 0; i < count; ++i) {
      const HpMi355xKzgRequest& r = requests[i];
      const Fr* table = static_cast<const Fr*>(tables[i]->ptr);
      if (r.kind == HP_MI355X_KZG_COMMIT) {
        t->batch.add(p->layers[N - r.num_vars], table);
        t->packet_outs.push_back(r.out_points);
      } else {
        const Fr* value = nullptr;
        enqueue_open(c, *p, table, r.num_vars, r.point, t->batch, t->hold, &value);
        for (size_t j = 0; j < r.num_vars; ++j) t->packet_outs.push_back(r.out_points + rt::kG1Bytes * j);
        cells.push_back(value);
        t->value_outs.push_back(r.out_evaluation);
      }
    }
    if (t->batch.size()) {
      const size_t slots = std::max<size_t>(t->batch.result_slots(), 1);
      rt::hip_check(hipHostMalloc(reinterpret_cast<void**>(&t->slots_host), slots * sizeof(G1Xyzz),
                                  hipHostMallocDefault));
      t->batch.enqueue(c, t->slots_host, nullptr, &t->hold);
    }
    if (!cells.empty()) {
      const size_t k = cells.size();
      rt::BufferRef ptrs = c.allocate(k * sizeof(void*));
      uint8_t* stage = c.staging(k * sizeof(void*));
      std::memcpy(stage, cells.data(), k * sizeof(void*));
      rt::hip_check(hipMemcpyAsync(ptrs->ptr, stage, k * sizeof(void*), hipMemcpyHostToDevice, c.stream));
      c.staging_used();
      rt::BufferRef vals = c.allocate(k * rt::kFrBytes);
      hipLaunchKernelGGL(k_gather_values, dim3(1), dim3(256), 0, c.stream,
                         static_cast<const Fr* const*>(ptrs->ptr), k, static_cast<Fr*>(vals->ptr));
      rt::check_launch();
      rt::hip_check(hipHostMalloc(reinterpret_cast<void**>(&t->values_host), k * rt::kFrBytes,
                                  hipHostMallocDefault));
      rt::hip_check(hipMemcpyAsync(t->values_host, vals->ptr, k * rt::kFrBytes, hipMemcpyDeviceToHost,
                                   c.stream));
      t->hold.push_back(std::move(ptrs));
      t->hold.push_back(std::move(vals));
    }
    rt::hip_check(hipEventCreateWithFlags(&t->done, hipEventDisableTiming));
    rt::hip_check(hipEventRecord(t->done, c.stream));
    auto* handle = new HpMi355xKzgTicket{ctx, t};
    c.resources[handle] = {rt::Context::Kind::kzg_ticket, t};
    *out = handle;
    return rt::ok();
  });
}

// Waits for a submitted batch and writes its results (once per ticket; the
// ticket is released).
extern "C" HpMi355xStatus hp_mi355x_v1_multilinear_kzg_batch_wait(HpMi355xContext* ctx,
                                                                  HpMi355xKzgTicket* ticket) {
  return hp_guarded<false>(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (ticket == nullptr || !c.holds(ticket, rt::Context::Kind::kzg_ticket))
      return rt::invalid_parameters();
    if (ticket->ctx != ctx) return rt::invalid_parameters();
    kzg::Ticket& t = *ticket->impl;
    rt::hip_check(hipEventSynchronize(t.done));
    if (t.batch.size()) {
      std::vector<uint8_t> packets(t.batch.size() * rt::kG1Bytes);
      t.batch.finish_packets(t.slots_host, packets.data());
      for (size_t j = 0; j < t.packet_outs.size(); ++j)
        std::memcpy(t.packet_outs[j], packets.data() + rt::kG1Bytes * j, rt::kG1Bytes);
    }
    for (size_t i = 0; i < t.value_outs.size(); ++i) {
      const Fr v = from_mont(t.values_host[i]);
      std::memcpy(t.value_outs[i], v.v, rt::kFrBytes);
    }
    c.resources.erase(ticket);
    delete ticket;
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_msm_profile(HpMi355xContext* ctx,
                                                   const HpMi355xKzgParams* params,
                                                   HpMi355xDeviceAddress polynomial,
                                                   size_t num_vars, float* stage_ms,
                                                   size_t stage_count, uint32_t* plan_info) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    kzg::Params* p = bind(ctx, c, params);
    if (!p || stage_ms == nullptr || stage_count < 7 || plan_info == nullptr)
      return rt::invalid_parameters();
    if (num_vars > p->num_vars) return rt::invalid_parameters();
    rt::BufferRef f = c.lookup(polynomial, num_vars);
    if (!f) return rt::invalid_parameters();
    const size_t layer = p->num_vars - num_vars;
    if (!p->ready[layer]) return rt::invalid_state();
    msm::StageEvents ev;
    for (auto& e : ev.marks) rt::hip_check(hipEventCreate(&e));
    msm::Batch batch;
    batch.add(p->layers[layer], static_cast<const Fr*>(f->ptr));
    const size_t slots = batch.result_slots();
    rt::BufferRef res = c.allocate(slots * sizeof(G1Xyzz) + sizeof(uint32_t));
    ev.entries_out = reinterpret_cast<uint32_t*>(static_cast<G1Xyzz*>(res->ptr) + slots);
    batch.enqueue(c, static_cast<G1Xyzz*>(res->ptr), &ev);
    c.sync();
    for (int i = 0; i < 6; ++i)
      rt::hip_check(hipEventElapsedTime(&stage_ms[i], ev.marks[i], ev.marks[i + 1]));
    rt::hip_check(hipEventElapsedTime(&stage_ms[6], ev.marks[0], ev.marks[6]));
    for (auto& e : ev.marks) (void)hipEventDestroy(e);
    uint32_t entries = 0;
    rt::download_bytes(c, ev.entries_out, sizeof(uint32_t), reinterpret_cast<uint8_t*>(&entries));
    const msm::Plan& plan = p->layers[layer].plan;
    plan_info[0] = plan.c;
    plan_info[1] = plan.windows;
    plan_info[2] = plan.buckets;
    plan_info[3] = entries;
    return rt::ok();
  });
}

// Autotune window bits and lane task cap per layer size on this device and
// install them in the context (layers prepared afterwards use them).
extern "C" HpMi355xStatus hp_mi355x_v1_msm_autotune(HpMi355xContext* ctx, const size_t* log_sizes,
                                                    size_t count, uint32_t* out_window,
                                                    uint32_t* out_cap, float* out_ms) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (count == 0) return rt::ok();
    if (log_sizes == nullptr || out_window == nullptr || out_cap == nullptr || out_ms == nullptr)
      return rt::invalid_parameters();
    for (size_t i = 0; i < count; ++i)
      if (log_sizes[i] < 1 || log_sizes[i] > 24) return rt::invalid_parameters();
    for (size_t i = 0; i < count; ++i)
      msm::autotune(c, log_sizes[i], &out_window[i], &out_cap[i], &out_ms[i]);
    return rt::ok();
  });
}

// Install a saved tuning table (window 0 / cap 0: the built-in rule).
extern "C" HpMi355xStatus hp_mi355x_v1_msm_set_tuning(HpMi355xContext* ctx, const size_t* log_sizes,
                                                      const uint32_t* window, const uint32_t* cap,
                                                      size_t count) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (count == 0) return rt::ok();
    if (log_sizes == nullptr || window == nullptr || cap == nullptr) return rt::invalid_parameters();
    for (size_t i = 0; i < count; ++i) {
      if (log_sizes[i] > 40) return rt::invalid_parameters();
      if (window[i] != 0 && (window[i] < 4 || window[i] > 20)) return rt::invalid_parameters();
      if (cap[i] > 255) return rt::invalid_parameters();
    }
    for (size_t i = 0; i < count; ++i) {
      c.tuned_window[log_sizes[i]] = static_cast<uint8_t>(window[i]);
      c.tuned_cap[log_sizes[i]] = static_cast<uint8_t>(cap[i]);
    }
    return rt::ok();
  });
}
