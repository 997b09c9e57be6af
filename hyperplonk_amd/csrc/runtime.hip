// Runtime: contexts, caching allocator, pinned staging, handle registry and
// the runtime entry points of the C ABI (include/hp_mi355x.h).
//
// Mirrors the lifetime contract of the reference runtime
// (backend/src/ascend910b3/native/bridge_runtime.h:1-41): contexts are host
// metadata, buffers are released to a stream-ordered reuse queue rather than
// freed while work may still read them, and context_destroy drains the stream
// before returning memory to the driver.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_set>

#include "kernels_common.hpp"
#include "mle.hpp"
#include "runtime.hpp"

namespace hp {
namespace rt {

thread_local Context* g_call_context = nullptr;

namespace {
std::mutex g_contexts_mutex;
std::unordered_set<HpMi355xContext*> g_contexts;

template <int N>
bool canonical_limbs(const uint8_t* bytes, const uint32_t* mod) {
  // compare from the most significant limb down
  for (int i = N - 1; i >= 0; --i) {
    uint32_t w;
    std::memcpy(&w, bytes + 4 * i, 4);
    if (w < mod[i]) return true;
    if (w > mod[i]) return false;
  }
  return false;  // equal to the modulus
}
}  // namespace

bool canonical_fr(const uint8_t* bytes) {
  return canonical_limbs<8>(bytes, FrParams::MOD);
}
bool canonical_fq(const uint8_t* bytes) {
  return canonical_limbs<12>(bytes, FqParams::MOD);
}
bool canonical_fr_array(const uint8_t* bytes, size_t count) {
  for (size_t i = 0; i < count; ++i)
    if (!canonical_fr(bytes + kFrBytes * i)) return false;
  return true;
}

size_t Allocator::round_up(size_t bytes) {
  if (bytes <= 256) return 256;
  if (bytes <= (size_t(1) << 20)) {
    size_t r = 256;
    while (r < bytes) r <<= 1;
    return r;
  }
  const size_t granule = size_t(2) << 20;
  return (bytes + granule - 1) / granule * granule;
}

Allocator::Allocator(hipStream_t stream) : stream_(stream) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0)
    limit_ = std::min<size_t>(size_t(16) << 30, std::max<size_t>(size_t(1) << 30, total_b / 8));
  else
    (void)hipGetLastError();
}

void* Allocator::acquire(size_t bytes, size_t* capacity) {
  const size_t size = round_up(bytes);
  auto it = free_.lower_bound(size);
  while (it != free_.end() && it->second.empty()) ++it;
  if (it != free_.end() && it->first <= 2 * size) {
    void* p = it->second.back();
    it->second.pop_back();
    cached_ -= it->first;
    *capacity = it->first;
    live_ += it->first;
    peak_ = std::max(peak_, live_);
    return p;
  }
  if (cached_ > limit_) trim();
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, size);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    e = hipMalloc(&p, size);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail_oom();
    }
  }
  *capacity = size;
  live_ += size;
  peak_ = std::max(peak_, live_);
  return p;
}

void Allocator::release(void* ptr, size_t capacity) {
  if (!ptr) return;
  free_[capacity].push_back(ptr);
  cached_ += capacity;
  live_ -= std::min(live_, capacity);
}

void Allocator::trim() {
  (void)hipStreamSynchronize(stream_);
  for (auto& kv : free_)
    for (void* p : kv.second) (void)hipFree(p);
  free_.clear();
  cached_ = 0;
}

Allocator::~Allocator() { trim(); }

Buffer::~Buffer() {
  if (ctx && ptr && ctx->alloc) ctx->alloc->release(ptr, capacity);
}

BufferRef Context::allocate(size_t bytes, size_t nv) {
  auto b = std::make_shared<Buffer>();
  b->ctx = this;
  b->bytes = bytes;
  b->num_vars = nv;
  b->ptr = alloc->acquire(bytes == 0 ? 1 : bytes, &b->capacity);
  return b;
}

void* Context::publish(BufferRef buffer) {
  void* p = buffer->ptr;
  registry[p] = std::move(buffer);
  return p;
}

BufferRef Context::lookup(void* handle, size_t nv) const {
  if (handle == nullptr) return nullptr;
  auto it = registry.find(handle);
  if (it == registry.end()) return nullptr;
  if (it->second->num_vars != nv) return nullptr;
  return it->second;
}

namespace {
// Waits until the copy that last used a staging area has completed.
void wait_staging(hipEvent_t e) {
  if (e == nullptr) return;
  const hipError_t q = hipEventQuery(e);
  if (q == hipErrorNotReady) {
    (void)hipGetLastError();
    hip_check(hipEventSynchronize(e));
  } else {
    hip_check(q);
  }
}
}  // namespace

uint8_t* Context::staging(size_t bytes) {
  if (bytes <= kRingSlotBytes) {
    if (ring == nullptr)
      hip_check(hipHostMalloc(reinterpret_cast<void**>(&ring), kRingSlots * kRingSlotBytes,
                              hipHostMallocDefault));
    const size_t s = ring_next;
    ring_next = (ring_next + 1) % kRingSlots;
    wait_staging(ring_free[s]);
    staging_slot = static_cast<int>(s);
    return ring + s * kRingSlotBytes;
  }
  staging_slot = -1;
  // the previous copy out of the staging buffer has usually completed (its
  // call synchronised): wait only when it has not
  wait_staging(pinned_free);
  if (bytes > pinned_bytes) {
    if (pinned) hip_check(hipHostFree(pinned));
    pinned = nullptr;
    pinned_bytes = 0;
    size_t want = std::max(bytes, size_t(1) << 20);
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&pinned), want, hipHostMallocDefault));
    pinned_bytes = want;
  }
  return pinned;
}

uint8_t* Context::results(size_t bytes) {
  if (bytes > results_bytes) {
    hip_check(hipStreamSynchronize(stream));  // nothing still writes the old block
    if (results_host) hip_check(hipHostFree(results_host));
    results_host = nullptr;
    results_bytes = 0;
    const size_t want = std::max(bytes, size_t(64) << 10);
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&results_host), want, hipHostMallocDefault));
    results_bytes = want;
  }
  return results_host;
}

void Context::staging_used() {
  hipEvent_t& e = staging_slot >= 0 ? ring_free[staging_slot] : pinned_free;
  if (!e) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  hip_check(hipEventRecord(e, stream));
}

void Context::sync() { hip_check(hipStreamSynchronize(stream)); }

hipStream_t Context::side_stream() {
  if (side == nullptr) hip_check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
  return side;
}

// ---- boundary conversions ----------------------------------------------
namespace {
__global__ void k_fr_to_mont(Fr* data, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    store_fr(data, i, to_mont(load_fr(data, i)));
}

__global__ void k_fr_from_mont(const Fr* src, Fr* dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    store_fr(dst, i, from_mont(load_fr(src, i)));
  __threadfence_system();  // dst may be pinned host memory read after the sync
}

unsigned grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b == 0) b = 1;
  return static_cast<unsigned>(b);
}

constexpr size_t kStagingChunk = size_t(64) << 20;
}  // namespace

void upload_fr_values(Context& c, Fr* dst, const uint8_t* host, size_t count) {
  // a handful of values (points, challenges, coefficients): converted on the
  // host (~0.1 us each), which saves the conversion launch
  constexpr size_t kHostConvert = 64;
  if (count > 0 && count <= kHostConvert) {
    uint8_t* stage = c.staging(count * kFrBytes);
    for (size_t i = 0; i < count; ++i) {
      Fr v;
      std::memcpy(v.v, host + i * kFrBytes, kFrBytes);
      v = to_mont(v);
      std::memcpy(stage + i * kFrBytes, v.v, kFrBytes);
    }
    hip_check(hipMemcpyAsync(dst, stage, count * kFrBytes, hipMemcpyHostToDevice, c.stream));
    c.staging_used();
    return;
  }
  size_t done = 0;
  while (done < count) {
    const size_t n = std::min(count - done, kStagingChunk / kFrBytes);
    uint8_t* stage = c.staging(n * kFrBytes);
    std::memcpy(stage, host + done * kFrBytes, n * kFrBytes);
    hip_check(hipMemcpyAsync(dst + done, stage, n * kFrBytes, hipMemcpyHostToDevice, c.stream));
    c.staging_used();
    done += n;
  }
  if (count) {
    hipLaunchKernelGGL(k_fr_to_mont, dim3(grid_for(count)), dim3(256), 0, c.stream, dst, count);
    check_launch();
  }
}

BufferRef upload_fr_table(Context& c, const uint8_t* host, size_t count, size_t num_vars) {
  BufferRef buf = c.allocate(count * kFrBytes, num_vars);
  upload_fr_values(c, static_cast<Fr*>(buf->ptr), host, count);
  return buf;
}

void download_bytes(Context& c, const void* src, size_t bytes, uint8_t* host) {
  size_t done = 0;
  while (done < bytes) {
    const size_t n = std::min(bytes - done, kStagingChunk);
    uint8_t* stage = c.staging(n);
    hip_check(hipMemcpyAsync(stage, static_cast<const uint8_t*>(src) + done, n,
                             hipMemcpyDeviceToHost, c.stream));
    c.staging_used();
    hip_check(hipStreamSynchronize(c.stream));
    std::memcpy(host + done, stage, n);
    done += n;
  }
}

void enqueue_from_mont(Context& c, const Fr* src, size_t count, Fr* dst) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_fr_from_mont, dim3(grid_for(count)), dim3(256), 0, c.stream, src, dst, count);
  check_launch();
}

void download_fr(Context& c, const Fr* src, size_t count, uint8_t* host) {
  if (count == 0) return;
  if (count * kFrBytes <= kStagingChunk) {
    // canonical values written straight into pinned host memory
    uint8_t* out = c.results(count * kFrBytes);
    hipLaunchKernelGGL(k_fr_from_mont, dim3(grid_for(count)), dim3(256), 0, c.stream, src,
                       reinterpret_cast<Fr*>(out), count);
    check_launch();
    c.sync();
    std::memcpy(host, out, count * kFrBytes);
    return;
  }
  BufferRef tmp = c.allocate(count * kFrBytes);
  hipLaunchKernelGGL(k_fr_from_mont, dim3(grid_for(count)), dim3(256), 0, c.stream, src,
                     static_cast<Fr*>(tmp->ptr), count);
  check_launch();
  download_bytes(c, tmp->ptr, count * kFrBytes, host);
}

namespace {
// A batch of items: workers and the caller claim indices until none is left.
// Polls `ready` for at most `limit` (a pause between polls); returns whether
// it became true.
template <class F>
bool spin_until(F ready, std::chrono::microseconds limit) {
  const auto until = std::chrono::steady_clock::now() + limit;
  for (uint32_t k = 0;; ++k) {
    if (ready()) return true;
    __builtin_ia32_pause();
    if ((k & 63) == 63 && std::chrono::steady_clock::now() >= until) return false;
  }
}

struct ForkJoin {
  const std::function<void(size_t)>* body;
  size_t n;
  std::atomic<size_t> next{0}, done{0};
  std::mutex mu;
  std::condition_variable cv;
  void run() {
    for (size_t i; (i = next.fetch_add(1)) < n;) {
      (*body)(i);
      if (done.fetch_add(1) + 1 == n) {
        std::lock_guard<std::mutex> g(mu);
        cv.notify_all();
      }
    }
  }
};

class Pool {
 public:
  static Pool& get() {
    static Pool* pool = new Pool();  // workers live until the process ends
    return *pool;
  }
  size_t workers() const { return threads_.size(); }
  void submit(const std::shared_ptr<ForkJoin>& job, size_t helpers) {
    {
      std::lock_guard<std::mutex> g(mu_);
      for (size_t k = 0; k < helpers; ++k) queue_.push_back(job);
      pending_.fetch_add(helpers, std::memory_order_release);
    }
    cv_.notify_all();
  }

 private:
  Pool() {
    // up to 15 workers and the caller: the 16 host threads one GPU's share of
    // a machine provides
    const unsigned hw = std::max(2u, std::thread::hardware_concurrency());
    const unsigned n = std::min(15u, hw - 1);
    for (unsigned k = 0; k < n; ++k) threads_.emplace_back([this] { loop(); });
    for (auto& t : threads_) t.detach();
  }
  void loop() {
    for (;;) {
      std::shared_ptr<ForkJoin> job;
      // fork-joins come in bursts (an MSM finish runs three in a row): look
      // for the next one for a short while before sleeping, so it does not
      // wait for a wake-up
      spin_until([&] { return pending_.load(std::memory_order_acquire) != 0; }, kWorkerSpin);
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !queue_.empty(); });
        job = std::move(queue_.front());
        queue_.pop_front();
        pending_.fetch_sub(1, std::memory_order_relaxed);
      }
      job->run();
    }
  }
  static constexpr std::chrono::microseconds kWorkerSpin{200};
  std::atomic<size_t> pending_{0};  // queued entries, read without the lock
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::shared_ptr<ForkJoin>> queue_;
  std::vector<std::thread> threads_;
};
}  // namespace

void parallel_for(size_t n, const std::function<void(size_t)>& body) {
  if (n == 0) return;
  if (n == 1) {
    body(0);
    return;
  }
  auto job = std::make_shared<ForkJoin>();
  job->body = &body;
  job->n = n;
  Pool& pool = Pool::get();
  pool.submit(job, std::min(n - 1, pool.workers()));
  job->run();
  // the helpers finish their last items within microseconds of the caller:
  // poll before sleeping on the condition variable
  if (spin_until([&] { return job->done.load(std::memory_order_acquire) == n; },
                 std::chrono::microseconds(2000)))
    return;
  std::unique_lock<std::mutex> lk(job->mu);
  job->cv.wait(lk, [&] { return job->done.load() == n; });
}

bool context_live(HpMi355xContext* ctx) {
  std::lock_guard<std::mutex> lock(g_contexts_mutex);
  return g_contexts.count(ctx) != 0;
}

}  // namespace rt
}  // namespace hp

using namespace hp;

extern "C" HpMi355xStatus hp_mi355x_v1_context_create(int32_t device_id, HpMi355xContext** out) {
  if (out == nullptr) return rt::invalid_parameters();
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) {
    (void)hipGetLastError();
    return rt::invalid_state();
  }
  if (device_id < 0 || device_id >= count) return rt::invalid_parameters();
  if (hipSetDevice(device_id) != hipSuccess) return rt::invalid_state();
  HpMi355xContext* ctx = nullptr;
  try {
    ctx = new HpMi355xContext();
    ctx->impl.device = device_id;
    rt::hip_check(hipStreamCreateWithFlags(&ctx->impl.stream, hipStreamNonBlocking));
    ctx->impl.alloc = std::make_unique<rt::Allocator>(ctx->impl.stream);
    mle::noalloc_scratch_setup(device_id);  // once per device, this device current
  } catch (const rt::Failure& f) {
    if (ctx) {
      if (ctx->impl.stream) (void)hipStreamDestroy(ctx->impl.stream);
      delete ctx;
    }
    return f.status;
  } catch (...) {
    delete ctx;
    return rt::invalid_state(-2);
  }
  {
    std::lock_guard<std::mutex> lock(rt::g_contexts_mutex);
    rt::g_contexts.insert(ctx);
  }
  *out = ctx;
  return rt::ok();
}

extern "C" HpMi355xStatus hp_mi355x_v1_context_destroy(HpMi355xContext* ctx) {
  if (ctx == nullptr) return rt::invalid_parameters();
  {
    std::lock_guard<std::mutex> lock(rt::g_contexts_mutex);
    if (!rt::g_contexts.erase(ctx)) return rt::invalid_parameters();
  }
  rt::Context& c = ctx->impl;
  (void)hipSetDevice(c.device);
  if (c.quiesce != nullptr) {
    try {
      c.quiesce(c);  // stops persistent round loops (they wait on the host)
    } catch (...) {
    }
  }
  hipError_t drained = hipStreamSynchronize(c.stream);
  if (c.side != nullptr) {
    const hipError_t e = hipStreamSynchronize(c.side);
    if (drained == hipSuccess) drained = e;
  }
  c.resources.clear();
  if (c.release_tails != nullptr) c.release_tails(c);
  c.registry.clear();
  c.alloc.reset();  // frees every cached block (after the drain)
  if (c.pinned) (void)hipHostFree(c.pinned);
  if (c.results_host) (void)hipHostFree(c.results_host);
  if (c.pinned_free) (void)hipEventDestroy(c.pinned_free);
  if (c.ring) (void)hipHostFree(c.ring);
  for (hipEvent_t e : c.ring_free)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c.stream);
  if (c.side != nullptr) (void)hipStreamDestroy(c.side);
  const bool poisoned = c.poisoned;
  const int32_t first_error = c.first_error;
  delete ctx;
  if (drained != hipSuccess) return rt::invalid_state(static_cast<int32_t>(drained));
  // everything is released; a context that failed earlier still reports its
  // first error (bridge_runtime.cpp:1317)
  if (poisoned) return rt::invalid_state(first_error);
  return rt::ok();
}

extern "C" HpMi355xStatus hp_mi355x_v1_memory_info(HpMi355xContext* ctx, size_t* device_free,
                                                   size_t* device_total, size_t* context_live,
                                                   size_t* context_peak, int32_t reset_peak) {
  return hp_guarded<false>(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (device_free == nullptr || device_total == nullptr || context_live == nullptr ||
        context_peak == nullptr)
      return rt::invalid_parameters();
    rt::hip_check(hipMemGetInfo(device_free, device_total));
    // blocks in the context's cache are free for the context's own use
    *device_free += c.alloc->cached_bytes();
    *context_live = c.alloc->live_bytes();
    *context_peak = c.alloc->peak_bytes();
    if (reset_peak) c.alloc->reset_peak();
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_debug_inject_failure(HpMi355xContext* ctx, uint32_t after,
                                                            int32_t native_code) {
  return hp_guarded(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    if (after != 0 && native_code == 0) return rt::invalid_parameters();
    c.inject_after = after;
    c.inject_code = native_code;
    return rt::ok();
  });
}

extern "C" HpMi355xStatus hp_mi355x_v1_buffer_retire(HpMi355xContext* ctx,
                                                     HpMi355xDeviceAddress address) {
  return hp_guarded_release(ctx, [&](rt::Context& c) -> HpMi355xStatus {
    auto it = c.registry.find(address);
    if (address == nullptr || it == c.registry.end()) return rt::invalid_parameters();
    c.registry.erase(it);  // states holding references keep the memory alive
    return rt::ok();
  });
}
