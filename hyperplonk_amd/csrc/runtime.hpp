// Context, device memory and status plumbing shared by every entry point.
//
// One context = one device + one HIP stream, driven by one host thread at a
// time (the single-thread contract of the reference bridge,
// backend/src/ascend910b3/native/bridge.h:16-21).  Device memory comes from a
// per-context caching allocator: freed blocks go to size-keyed free lists and
// are reused in stream order, so a handle's memory is recycled only after the
// work already queued on the stream that touched it (all work of a context is
// on its stream).  MLE handles are the base addresses of registered buffers;
// prover states hold shared references so a retired handle stays alive for as
// long as a state still reads it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/hp_mi355x.h"

struct HpMi355xContext;

namespace hp {
namespace rt {

constexpr size_t kFrBytes = 32;
constexpr size_t kG1Bytes = 97;
// Tables of at most 2^31 entries (num_vars < 32): device-side lengths, pair
// counts and grid sizes are 32-bit.
constexpr size_t kMaxNumVars = 32;

inline HpMi355xStatus ok() { return {HP_MI355X_STATUS_SUCCESS, 0}; }
inline HpMi355xStatus invalid_parameters(int32_t code = 0) {
  return {HP_MI355X_STATUS_INVALID_PARAMETERS, code};
}
inline HpMi355xStatus invalid_state(int32_t code = 0) {
  return {HP_MI355X_STATUS_INVALID_STATE, code};
}
inline bool is_ok(const HpMi355xStatus& s) { return s.kind == 0 && s.native_code == 0; }

// Thrown inside the library only; converted to a status at the ABI.
// `poison` marks device-side failures, after which the context is unusable;
// validation failures and exhausted memory leave the context intact.
struct Failure {
  HpMi355xStatus status;
  bool poison;
};

[[noreturn]] inline void fail_state(int32_t code) { throw Failure{invalid_state(code), true}; }
[[noreturn]] inline void fail_params() { throw Failure{invalid_parameters(), false}; }
[[noreturn]] inline void fail_oom() {
  throw Failure{invalid_state(static_cast<int32_t>(hipErrorOutOfMemory)), false};
}


// Canonical little-endian field element check (value < modulus).
bool canonical_fr(const uint8_t* bytes);
bool canonical_fq(const uint8_t* bytes);
bool canonical_fr_array(const uint8_t* bytes, size_t count);

// Caching device allocator.  Freed blocks are kept by size class and reused
// best-fit (a cached block up to twice the request); the cache is bounded
// (`limit`, an eighth of the device's memory, at most 16 GiB) so that blocks
// of sizes that never recur (batch-dependent scratch) cannot fill the device
// and starve the runtime's own allocations (kernel scratch, queues).
class Allocator {
 public:
  explicit Allocator(hipStream_t stream);
  ~Allocator();
  // returns a block of at least `bytes`; *capacity gets its size class
  void* acquire(size_t bytes, size_t* capacity);
  void release(void* ptr, size_t capacity);
  void trim();  // synchronises, then frees every cached block
  size_t cached_bytes() const { return cached_; }
  // bytes handed out and not yet released, and their high-water mark
  size_t live_bytes() const { return live_; }
  size_t peak_bytes() const { return peak_; }
  void reset_peak() { peak_ = live_; }

 private:
  static size_t round_up(size_t bytes);
  hipStream_t stream_;
  std::map<size_t, std::vector<void*>> free_;
  size_t cached_ = 0;
  size_t live_ = 0, peak_ = 0;
  size_t limit_ = size_t(4) << 30;
};

struct Context;

// A device allocation owned by a context.  Releasing the last reference
// returns the block to the context allocator.
struct Buffer {
  Context* ctx = nullptr;
  void* ptr = nullptr;
  size_t bytes = 0;
  size_t capacity = 0;  // size class of the block
  size_t num_vars = 0;  // for MLE tables
  ~Buffer();
};
using BufferRef = std::shared_ptr<Buffer>;

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::unique_ptr<Allocator> alloc;
  std::unordered_map<void*, BufferRef> registry;  // live MLE handles
  // Live prover states and KZG parameters (released before memory on destroy),
  // keyed by handle address with the handle's kind, so that a handle is
  // checked by lookup here before anything is read through it, and a freed
  // handle whose address was reused by a handle of another kind is refused.
  enum class Kind : uint8_t { kzg_params = 1, sumcheck_state = 2, internal = 3, kzg_ticket = 4 };
  struct Resource {
    Kind kind;
    std::shared_ptr<void> object;
  };
  std::unordered_map<const void*, Resource> resources;
  bool holds(const void* handle, Kind kind) const {
    auto it = resources.find(handle);
    return it != resources.end() && it->second.kind == kind;
  }
  bool poisoned = false;
  int32_t first_error = 0;
  // MSM tuning by log2 of the layer size (0 = the built-in rule): window bits
  // for layers prepared afterwards, lane task cap for batches
  uint8_t tuned_window[64] = {};
  uint8_t tuned_cap[64] = {};
  // Pinned host staging for uploads and small results.
  uint8_t* pinned = nullptr;
  size_t pinned_bytes = 0;
  hipEvent_t pinned_free = nullptr;  // last copy that read/wrote `pinned`
  // Uses of at most kRingSlotBytes take the slots of a ring in turn, each with
  // its own event, so the small uploads of one call do not wait for each
  // other's copies (a copy waits behind all earlier work on the stream).
  static constexpr size_t kRingSlots = 16, kRingSlotBytes = 64 * 1024;
  uint8_t* ring = nullptr;
  hipEvent_t ring_free[kRingSlots] = {};
  size_t ring_next = 0;
  int staging_slot = -1;  // slot of the current staging use; -1: `pinned`

  // Allocate a new device buffer (not registered).
  BufferRef allocate(size_t bytes, size_t num_vars = 0);
  // Register an MLE buffer as a handle.
  void* publish(BufferRef buffer);
  // Look up a live MLE handle with the expected dimension.
  BufferRef lookup(void* handle, size_t num_vars) const;
  // Pinned host bytes of at least `bytes`, after previous users completed.
  uint8_t* staging(size_t bytes);
  void staging_used();  // record that queued work reads/writes staging
  // Pinned host memory the device writes small results into directly (no
  // copy back): valid after the stream is synchronised; reused by the next
  // call, so callers copy out before returning.
  uint8_t* results(size_t bytes);
  uint8_t* results_host = nullptr;
  size_t results_bytes = 0;
  void sync();
  // A second stream for work that overlaps the context stream (asynchronous
  // KZG batches), created on first use; context_destroy drains both.
  hipStream_t side = nullptr;
  hipStream_t side_stream();
  void poison(int32_t code) {
    if (!poisoned) first_error = code ? code : -1;
    poisoned = true;
  }
  // Persistent sumcheck round loops (sumcheck.hip) running on `stream`: while
  // any is active, work enqueued by another entry point would wait behind it,
  // so hp_guarded calls `quiesce` first (it stops them and drains the stream).
  void (*quiesce)(Context&) = nullptr;
  void* tails = nullptr;  // owned by sumcheck.hip (released by `release_tails`)
  void (*release_tails)(Context&) = nullptr;
  // Failure injection (hp_mi355x_v1_debug_inject_failure): the inject_after-th
  // checked HIP call of this context from now on reports inject_code as its
  // error, after the call itself completed normally on the device.
  uint32_t inject_after = 0;
  int32_t inject_code = 0;
};

// The context of the ABI call running on this thread (set by hp_guarded).
extern thread_local Context* g_call_context;

inline void hip_check(hipError_t e) {
  Context* c = g_call_context;
  if (e == hipSuccess && c != nullptr && c->inject_after != 0 && --c->inject_after == 0)
    e = static_cast<hipError_t>(c->inject_code);
  if (e != hipSuccess) fail_state(static_cast<int32_t>(e));
}

// Points the context's stream at `s` for one scope: everything the library
// enqueues meanwhile (copies, kernels, staging events) goes to `s`.
struct StreamSwap {
  Context& c;
  hipStream_t saved;
  StreamSwap(Context& ctx, hipStream_t s) : c(ctx), saved(ctx.stream) { c.stream = s; }
  ~StreamSwap() { c.stream = saved; }
};

// Sets g_call_context for the lifetime of one ABI call.
struct CallScope {
  Context* saved;
  explicit CallScope(Context* c) : saved(g_call_context) { g_call_context = c; }
  ~CallScope() { g_call_context = saved; }
};

bool context_live(HpMi355xContext* ctx);

// Host fork-join over n independent items on a process-wide pool of worker
// threads (the calling thread works too); returns when every item is done.
// Used for the host ends of MSMs (combining bucket-reduction groups).
void parallel_for(size_t n, const std::function<void(size_t)>& body);

// Launch-error check after kernel launches.
inline void check_launch() { hip_check(hipGetLastError()); }

// Grid helpers.
inline unsigned blocks_for(size_t n, unsigned threads) {
  size_t b = (n + threads - 1) / threads;
  if (b == 0) b = 1;
  return static_cast<unsigned>(b);
}

}  // namespace rt
}  // namespace hp

struct HpMi355xContext {
  hp::rt::Context impl;
};

// Run `body` as an ABI call: validate the context, select its device, map
// library failures to statuses and poison the context on device errors.
// kQuiesce: stop the context's persistent round loops before the body
// enqueues work (every entry point but the sumcheck rounds, which manage them).
template <bool kQuiesce = true, class F>
HpMi355xStatus hp_guarded(HpMi355xContext* ctx, F&& body) noexcept {
  if (ctx == nullptr || !hp::rt::context_live(ctx)) return hp::rt::invalid_parameters();
  hp::rt::Context& c = ctx->impl;
  if (c.poisoned) return hp::rt::invalid_state(c.first_error);
  hp::rt::CallScope scope(&c);
  try {
    if (hipSetDevice(c.device) != hipSuccess) return hp::rt::invalid_state();
    if (kQuiesce && c.quiesce != nullptr) c.quiesce(c);
    return body(c);
  } catch (const hp::rt::Failure& f) {
    if (f.poison) c.poison(f.status.native_code);
    return f.status;
  } catch (...) {
    c.poison(-2);
    return hp::rt::invalid_state(-2);
  }
}

// Run a release entry point (buffer_retire, params and state destroy): as
// hp_guarded, but also on a poisoned context, since releasing host-side
// ownership must succeed after a failure was reported (the owner's drop path
// cannot recover from a refused release).  Device memory goes back to the
// context's cache and is freed by context_destroy after its drain.
template <class F>
HpMi355xStatus hp_guarded_release(HpMi355xContext* ctx, F&& body) noexcept {
  if (ctx == nullptr || !hp::rt::context_live(ctx)) return hp::rt::invalid_parameters();
  hp::rt::Context& c = ctx->impl;
  try {
    return body(c);
  } catch (...) {
    return hp::rt::invalid_state(-2);
  }
}
