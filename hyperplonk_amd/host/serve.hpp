// Prover service on one MI355X: continuous batching of independent proofs.
//
// Every admitted request runs the ordinary protocol (snark.hpp: preprocess +
// prove) on a worker thread of its own against BatchingBackend, which has the
// backend interface of Mi355xBackend:
//   * every call is handed to the Scheduler, whose device thread gathers what
//     every active request is waiting for (until all of them wait, or a short
//     window passes) and runs one tick: the table-producing calls (uploads,
//     linear combinations, eq tables, frac, product, nums/denoms, factors,
//     prover_init), which only queue kernels, one after another, then each
//     kind of result-returning call as ONE batched call:
//     multilinear_kzg_batch (all MSMs of all requests in one launch sequence),
//     sumcheck round_batch (all rounds in shared launches), evaluate_batch.
// Requests arrive on a trace; a request is admitted when a slot is free and
// its device footprint (measured per request class on the device before the
// trace) fits the HBM budget (measured free device memory, or a set value),
// so new requests join the batches of those already on the GPU.  A proof depends only on its own
// request, so it is byte-identical alone and in any mix.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "circuit.hpp"
#include "mi355x_backend.hpp"

namespace hyperplonk {
namespace serve {

using host::Fr;
using mi355x::Mi355xBackend;

class Scheduler {
 public:
  // window: how long a tick waits for requests still on the host;
  // kzg_hold: how long commits/opens may be held back to merge with those of
  // other requests while the others still have cheaper work (rounds, tables)
  Scheduler(const Mi355xBackend& be, std::chrono::microseconds window,
            std::chrono::microseconds kzg_hold = std::chrono::microseconds(0))
      : be_(be), window_(window), kzg_hold_(kzg_hold), thread_([this] { loop(); }) {}
  ~Scheduler() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_submit_.notify_all();
    thread_.join();
  }

  void join() {
    std::lock_guard<std::mutex> g(mu_);
    ++active_;
  }
  void leave() {
    {
      std::lock_guard<std::mutex> g(mu_);
      --active_;
    }
    cv_submit_.notify_all();
  }

  std::vector<FrBytes> round(mi355x::ProverState& st, const std::optional<FrBytes>& ch) {
    Pending p;
    p.kind = Kind::Round;
    p.state = &st;
    p.challenge = ch;
    submit(p);
    return std::move(p.messages);
  }
  std::vector<Mi355xBackend::KzgResult> kzg(const mi355x::PreparedProverParam& pp,
                                            std::vector<Mi355xBackend::KzgOp> ops) {
    Pending p;
    p.kind = Kind::Kzg;
    p.pp = &pp;
    p.kzg_ops = std::move(ops);
    submit(p);
    return std::move(p.kzg_results);
  }
  std::vector<FrBytes> evaluate(std::vector<mi355x::Mle> mles, std::vector<std::vector<FrBytes>> points) {
    Pending p;
    p.kind = Kind::Eval;
    p.mles = std::move(mles);
    p.points = std::move(points);
    submit(p);
    return std::move(p.messages);
  }

  // any other backend call, run on the device thread in the next tick
  template <class F>
  auto call(F&& f) -> decltype(f()) {
    std::optional<decltype(f())> result;
    Pending p;
    p.kind = Kind::Generic;
    p.fn = [&] { result.emplace(f()); };
    submit(p);
    return std::move(*result);
  }

  struct Stats {
    size_t ticks = 0, ops = 0, kzg_calls = 0, round_calls = 0, eval_calls = 0, kzg_held = 0;
    double gather_ms = 0, generic_ms = 0, kzg_ms = 0, round_ms = 0, eval_ms = 0;  // device thread
    double waiting_ms = 0;  // request threads, summed: from submit to result
  };
  Stats stats() const {
    std::lock_guard<std::mutex> g(mu_);
    return stats_;
  }

 private:
  enum class Kind { Round, Kzg, Eval, Generic };
  struct Pending {
    Kind kind;
    std::function<void()> fn;
    mi355x::ProverState* state = nullptr;
    std::optional<FrBytes> challenge;
    const mi355x::PreparedProverParam* pp = nullptr;
    std::vector<Mi355xBackend::KzgOp> kzg_ops;
    std::vector<mi355x::Mle> mles;
    std::vector<std::vector<FrBytes>> points;
    std::vector<FrBytes> messages;
    std::vector<Mi355xBackend::KzgResult> kzg_results;
    std::exception_ptr error;
    bool done = false;
    std::chrono::steady_clock::time_point since;
  };

  static double ms(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  }

  void submit(Pending& p) {
    const auto t = std::chrono::steady_clock::now();
    std::unique_lock<std::mutex> lk(mu_);
    p.since = t;
    queue_.push_back(&p);
    cv_submit_.notify_all();
    cv_done_.wait(lk, [&] { return p.done; });
    stats_.waiting_ms += ms(t, std::chrono::steady_clock::now());
    if (p.error) std::rethrow_exception(p.error);
  }

  void loop() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_submit_.wait(lk, [&] { return stop_ || !queue_.empty(); });
      if (queue_.empty()) return;  // stop requested and nothing pending
      // gather: until every active request waits here, or the window passes
      const auto g0 = std::chrono::steady_clock::now();
      cv_submit_.wait_until(lk, g0 + window_, [&] { return queue_.size() >= active_; });
      stats_.gather_ms += ms(g0, std::chrono::steady_clock::now());
      std::vector<Pending*> batch;
      hold_back_kzg(&batch);
      ++stats_.ticks;
      stats_.ops += batch.size();
      lk.unlock();
      execute(batch);
      lk.lock();
      for (Pending* p : batch) p->done = true;
      cv_done_.notify_all();
    }
  }

  // Take the queue into *batch, except commits/opens that are worth holding
  // back: fewer than half of the active requests wait on one, others have
  // other work in this tick, and none has waited kzg_hold yet.
  void hold_back_kzg(std::vector<Pending*>* batch) {
    size_t kzg = 0;
    auto oldest = std::chrono::steady_clock::time_point::max();
    for (Pending* p : queue_)
      if (p->kind == Kind::Kzg) {
        ++kzg;
        oldest = std::min(oldest, p->since);
      }
    const bool hold = kzg_hold_.count() > 0 && kzg > 0 && kzg < queue_.size() &&
                      2 * kzg < active_ &&
                      std::chrono::steady_clock::now() - oldest < kzg_hold_;
    std::deque<Pending*> rest;
    for (Pending* p : queue_) {
      if (hold && p->kind == Kind::Kzg)
        rest.push_back(p);
      else
        batch->push_back(p);
    }
    queue_.swap(rest);
    if (hold) ++stats_.kzg_held;
  }

  template <class F>
  static void guarded(const std::vector<Pending*>& group, F&& f) {
    try {
      f();
    } catch (...) {
      for (Pending* p : group) p->error = std::current_exception();
    }
  }

  void execute(const std::vector<Pending*>& batch) {
    std::vector<Pending*> rounds, evals;
    std::map<const mi355x::PreparedProverParam*, std::vector<Pending*>> kzgs;
    for (Pending* p : batch) {
      if (p->kind == Kind::Round) rounds.push_back(p);
      else if (p->kind == Kind::Eval) evals.push_back(p);
      else if (p->kind == Kind::Kzg) kzgs[p->pp].push_back(p);
    }
    // calls that only queue device work first, then the batched groups
    auto t = std::chrono::steady_clock::now();
    auto lap = [&](double Stats::*field) {
      const auto now = std::chrono::steady_clock::now();
      std::lock_guard<std::mutex> g(mu_);
      stats_.*field += ms(t, now);
      t = now;
    };
    for (Pending* p : batch)
      if (p->kind == Kind::Generic) guarded({p}, [&] { p->fn(); });
    lap(&Stats::generic_ms);
    for (auto& [pp, group] : kzgs) {
      guarded(group, [&] {
        std::vector<Mi355xBackend::KzgOp> ops;
        for (Pending* p : group) ops.insert(ops.end(), p->kzg_ops.begin(), p->kzg_ops.end());
        auto res = be_.kzg_batch(*pp, ops);
        size_t k = 0;
        for (Pending* p : group)
          for (size_t i = 0; i < p->kzg_ops.size(); ++i) p->kzg_results.push_back(std::move(res[k++]));
      });
      lap(&Stats::kzg_ms);
      std::lock_guard<std::mutex> g(mu_);
      ++stats_.kzg_calls;
    }
    if (!rounds.empty()) {
      guarded(rounds, [&] {
        std::vector<mi355x::ProverState*> states;
        std::vector<std::optional<FrBytes>> chs;
        for (Pending* p : rounds) {
          states.push_back(p->state);
          chs.push_back(p->challenge);
        }
        auto res = be_.prove_rounds_batch(states, chs);
        for (size_t i = 0; i < rounds.size(); ++i) rounds[i]->messages = std::move(res[i]);
      });
      lap(&Stats::round_ms);
      std::lock_guard<std::mutex> g(mu_);
      ++stats_.round_calls;
    }
    if (!evals.empty()) {
      guarded(evals, [&] {
        std::vector<mi355x::Mle> mles;
        std::vector<std::vector<FrBytes>> pts;
        for (Pending* p : evals) {
          mles.insert(mles.end(), p->mles.begin(), p->mles.end());
          pts.insert(pts.end(), p->points.begin(), p->points.end());
        }
        auto res = be_.evaluate_mles(mles, pts);
        size_t k = 0;
        for (Pending* p : evals)
          for (size_t i = 0; i < p->mles.size(); ++i) p->messages.push_back(res[k++]);
      });
      lap(&Stats::eval_ms);
      std::lock_guard<std::mutex> g(mu_);
      ++stats_.eval_calls;
    }
  }

  const Mi355xBackend& be_;
  std::chrono::microseconds window_, kzg_hold_;
  mutable std::mutex mu_;
  std::condition_variable cv_submit_, cv_done_;
  std::deque<Pending*> queue_;
  size_t active_ = 0;
  bool stop_ = false;
  Stats stats_;
  std::thread thread_;
};

// The backend interface of Mi355xBackend, with the waiting calls batched
// across requests through a Scheduler.
class BatchingBackend {
 public:
  using Mle = mi355x::Mle;
  using VirtualPolynomial = mi355x::VirtualPolynomial;
  using PreparedProverParam = mi355x::PreparedProverParam;
  using ProverState = mi355x::ProverState;

  BatchingBackend(const Mi355xBackend& be, Scheduler& s) : be_(be), s_(s) {}

  Mle mle_from_evaluations(size_t nv, const std::vector<FrBytes>& e) const {
    return s_.call([&] { return be_.mle_from_evaluations(nv, e); });
  }
  Mle linear_combination(const std::vector<Mle>& p, const std::vector<FrBytes>& c) const {
    return s_.call([&] { return be_.linear_combination(p, c); });
  }
  Mle build_eq_x_r(const std::vector<FrBytes>& r) const {
    return s_.call([&] { return be_.build_eq_x_r(r); });
  }
  Mle compute_frac_poly(const std::vector<Mle>& f, const std::vector<Mle>& g) const {
    return s_.call([&] { return be_.compute_frac_poly(f, g); });
  }
  Mle compute_product_poly(const Mle& f) const {
    return s_.call([&] { return be_.compute_product_poly(f); });
  }
  std::pair<std::vector<Mle>, std::vector<Mle>> compute_nums_and_denoms(
      const FrBytes& beta, const FrBytes& gamma, const std::vector<Mle>& f, const std::vector<Mle>& g,
      const std::vector<Mle>& p) const {
    return s_.call([&] { return be_.compute_nums_and_denoms(beta, gamma, f, g, p); });
  }
  std::array<Mle, 2> compute_product_factors(const Mle& f, const Mle& p) const {
    return s_.call([&] { return be_.compute_product_factors(f, p); });
  }
  ProverState prover_init(const VirtualPolynomial& vp) const {
    return s_.call([&] { return be_.prover_init(vp); });
  }

  std::vector<FrBytes> prove_round_and_update_state(ProverState& st, const std::optional<FrBytes>& ch) const {
    return s_.round(st, ch);
  }
  FrBytes evaluate_mle(const Mle& m, const std::vector<FrBytes>& point) const {
    return s_.evaluate({m}, {point}).at(0);
  }
  std::vector<FrBytes> evaluate_mles(const std::vector<Mle>& ms,
                                     const std::vector<std::vector<FrBytes>>& points) const {
    return s_.evaluate(ms, points);
  }
  G1Packet commit(const PreparedProverParam& pp, const Mle& m) const {
    return s_.kzg(pp, {{false, m, {}}}).at(0).points.at(0);
  }
  std::vector<G1Packet> multi_commit(const PreparedProverParam& pp, const std::vector<Mle>& ms) const {
    if (ms.empty()) return {};
    std::vector<Mi355xBackend::KzgOp> ops;
    for (const auto& m : ms) ops.push_back({false, m, {}});
    std::vector<G1Packet> out;
    for (auto& r : s_.kzg(pp, std::move(ops))) out.push_back(r.points.at(0));
    return out;
  }
  std::pair<std::vector<G1Packet>, FrBytes> open(const PreparedProverParam& pp, const Mle& m,
                                                  const std::vector<FrBytes>& point) const {
    auto r = s_.kzg(pp, {{true, m, point}});
    return {std::move(r.at(0).points), r.at(0).value};
  }

 private:
  const Mi355xBackend& be_;
  Scheduler& s_;
};

// ------------------------------------------------------------------ traces
struct Request {
  size_t id = 0;
  bool jellyfish = false;
  size_t nv = 0;
  uint64_t seed = 0;
  double arrival_s = 0;  // offset from the start of the run
};

// Poisson arrivals at `rate` per second (rate <= 0: everything at t = 0),
// gate jellyfish with probability jelly_frac, nv uniform in [nv_min, nv_max].
inline std::vector<Request> make_trace(size_t count, double rate, size_t nv_min, size_t nv_max,
                                       double jelly_frac, uint64_t seed) {
  host::SplitMix64 rng(seed);
  std::vector<Request> out;
  double t = 0;
  for (size_t i = 0; i < count; ++i) {
    Request r;
    r.id = i;
    const double u = (rng.next() >> 11) * (1.0 / 9007199254740992.0);
    r.jellyfish = u < jelly_frac;
    r.nv = nv_min + rng.below(nv_max - nv_min + 1);
    r.seed = rng.next();
    // drawn for every rate, so the request mix does not depend on the rate
    const double e = ((rng.next() >> 11) + 1) * (1.0 / 9007199254740993.0);
    if (rate > 0) t += -std::log(e) / rate;
    r.arrival_s = t;
    out.push_back(r);
  }
  return out;
}

// Request payloads (witness + index) come from a pool of `variants` seeded
// circuits per (gate, nv): request r carries variant r.seed % variants.  The
// pool is built before any clock starts, so serving figures measure the
// prover and not the synthesis of test inputs (an nv-20 mock circuit takes
// seconds of host time to solve).
inline uint64_t circuit_seed(const Request& r, size_t variants) {
  const uint64_t v = r.seed % std::max<size_t>(variants, 1);
  return ((uint64_t)r.nv << 40) ^ ((uint64_t)r.jellyfish << 32) ^ (v * 0x9E3779B97F4A7C15ULL) ^ 0x7061796cULL;
}

inline host::MockCircuit circuit_of(const Request& r, size_t variants) {
  const uint64_t seed = circuit_seed(r, variants);
  return r.jellyfish ? host::jellyfish_circuit(r.nv, seed) : host::vanilla_circuit(r.nv, seed);
}

// Device bytes a request may hold at its peak when nothing was measured for
// its class: its tables (witnesses, selectors, permutation oracles,
// nums/denoms, frac/prod/factors, eq tables), the sumcheck fold slots and MSM
// working memory, ~64 tables of 2^nv Fr.
inline size_t estimated_footprint(const Request& r) { return size_t(64) * 32 * (size_t(1) << r.nv); }

// Measured footprints by request class (jellyfish, nv): the context's peak
// live device bytes over one serial preprocess + prove of the class, above
// what was live before, with a 25 % margin (batched MSM work memory of a tick
// is shared between the requests in it).
using Footprints = std::map<std::pair<bool, size_t>, size_t>;

class CircuitPool {
 public:
  CircuitPool(const std::vector<Request>& trace, size_t variants, size_t threads)
      : variants_(variants) {
    std::vector<const Request*> todo;
    for (const Request& r : trace)
      if (map_.emplace(key(r), nullptr).second) todo.push_back(&r);
    // largest first, so the slow nv-20 solves overlap
    std::sort(todo.begin(), todo.end(), [](const Request* a, const Request* b) { return a->nv > b->nv; });
    std::atomic<size_t> next{0};
    std::mutex mu;
    auto work = [&] {
      for (size_t i; (i = next.fetch_add(1)) < todo.size();) {
        auto c = std::make_shared<const host::MockCircuit>(circuit_of(*todo[i], variants_));
        std::lock_guard<std::mutex> g(mu);
        map_[key(*todo[i])] = std::move(c);
      }
    };
    std::vector<std::thread> pool;
    for (size_t t = 0; t < std::max<size_t>(threads, 1); ++t) pool.emplace_back(work);
    for (auto& t : pool) t.join();
  }
  std::shared_ptr<const host::MockCircuit> get(const Request& r) const { return map_.at(key(r)); }
  size_t size() const { return map_.size(); }
  size_t variants() const { return variants_; }

 private:
  std::tuple<bool, size_t, uint64_t> key(const Request& r) const {
    return {r.jellyfish, r.nv, r.seed % std::max<size_t>(variants_, 1)};
  }
  size_t variants_;
  std::map<std::tuple<bool, size_t, uint64_t>, std::shared_ptr<const host::MockCircuit>> map_;
};

struct Outcome {
  double start_s = 0, done_s = 0;  // offsets from the run start
  std::vector<uint8_t> proof;      // serialized proof
  std::string error;
};

struct ServeConfig {
  size_t slots = 16;
  // device bytes the admitted requests may hold together; 0: measured, 90 %
  // of the device memory free when the trace starts (the SRS already resident)
  size_t hbm_budget = 0;
  Footprints footprints;  // per class; classes missing here use estimated_footprint
  size_t gen_threads = 8;  // threads that build the payload pool
  size_t variants = 2;     // pooled circuits per (gate, nv)
  std::chrono::microseconds window{300};
  std::chrono::microseconds kzg_hold{3000};
  bool batched = true;  // false: serial FIFO, one request at a time, plain backend
  size_t progress_every = 0;  // > 0: a line on stderr every that many completions
};

struct ServeReport {
  std::vector<Outcome> outcomes;
  double span_s = 0;  // first arrival to last completion
  Scheduler::Stats stats;
  size_t max_concurrent = 0;
  size_t hbm_budget = 0;      // the budget admission used
  size_t budget_waits = 0;    // admissions that waited for memory with a slot free
  size_t peak_admitted = 0;   // largest sum of admitted footprints
  size_t device_peak = 0;     // the context's peak live bytes over the run
};

inline size_t footprint_of(const ServeConfig& cfg, const Request& r) {
  auto it = cfg.footprints.find({r.jellyfish, r.nv});
  return it != cfg.footprints.end() ? it->second : estimated_footprint(r);
}

inline Footprints measure_footprints(const Mi355xBackend& be, const mi355x::PreparedProverParam& pp,
                                     const host::MultilinearVerifierParam& vk,
                                     const std::vector<Request>& trace, const CircuitPool& payloads) {
  Footprints out;
  for (const Request& r : trace) {
    const std::pair<bool, size_t> key{r.jellyfish, r.nv};
    if (out.count(key)) continue;
    auto c = payloads.get(r);
    const size_t before = be.memory_info(true).live;
    {
      auto [pk, vkey] = host::preprocess(be, c->index, pp, vk);
      (void)host::prove(be, pk, c->public_inputs, c->witnesses);
    }
    const size_t peak = be.memory_info(false).peak;
    const size_t used = peak > before ? peak - before : 0;
    out[key] = used + used / 4;
  }
  return out;
}

// Run the trace.  Batched: admission by slots and HBM budget, every admitted
// request on its own thread through the Scheduler.  Serial: FIFO, one at a
// time on the plain backend.
inline ServeReport run_trace(const Mi355xBackend& be, const mi355x::PreparedProverParam& pp,
                             const host::MultilinearVerifierParam& vk, const std::vector<Request>& trace,
                             const CircuitPool& payloads, const ServeConfig& cfg) {
  using Clock = std::chrono::steady_clock;
  ServeReport rep;
  rep.outcomes.resize(trace.size());
  const auto t0 = Clock::now();
  auto since = [&](Clock::time_point t) { return std::chrono::duration<double>(t - t0).count(); };
  std::atomic<size_t> completed{0};
  auto note_done = [&] {
    const size_t k = ++completed;
    if (cfg.progress_every && (k % cfg.progress_every == 0 || k == trace.size())) {
      std::fprintf(stderr, "serve %s: %zu/%zu done at %.1f s\n", cfg.batched ? "batched" : "serial", k,
                   trace.size(), since(Clock::now()));
      std::fflush(stderr);
    }
  };
  auto wait_arrival = [&](const Request& r) {
    const auto at = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(r.arrival_s));
    std::this_thread::sleep_until(at);
  };
  if (!cfg.batched) {
    for (const Request& r : trace) {
      auto c = payloads.get(r);
      wait_arrival(r);
      Outcome& o = rep.outcomes[r.id];
      o.start_s = since(Clock::now());
      try {
        auto [pk, vkey] = host::preprocess(be, c->index, pp, vk);
        o.proof = host::serialize_proof(host::prove(be, pk, c->public_inputs, c->witnesses));
      } catch (const std::exception& e) {
        o.error = e.what();
      }
      o.done_s = since(Clock::now());
      note_done();
    }
  } else {
    size_t budget = cfg.hbm_budget;
    if (budget == 0) budget = be.memory_info().device_free / 10 * 9;
    rep.hbm_budget = budget;
    Scheduler sched(be, cfg.window, cfg.kzg_hold);
    BatchingBackend bb(be, sched);
    std::mutex mu;
    std::condition_variable cv;
    size_t running = 0, used = 0;
    bool closed = false;
    std::deque<std::pair<const Request*, std::shared_ptr<const host::MockCircuit>>> admitted;
    auto worker = [&] {
      for (;;) {
        std::pair<const Request*, std::shared_ptr<const host::MockCircuit>> job;
        {
          std::unique_lock<std::mutex> lk(mu);
          cv.wait(lk, [&] { return closed || !admitted.empty(); });
          if (admitted.empty()) return;
          job = admitted.front();
          admitted.pop_front();
        }
        const Request& r = *job.first;
        Outcome& o = rep.outcomes[r.id];
        try {
          auto [pk, vkey] = host::preprocess(bb, job.second->index, pp, vk);
          o.proof = host::serialize_proof(host::prove(bb, pk, job.second->public_inputs, job.second->witnesses));
        } catch (const std::exception& e) {
          o.error = e.what();
        }
        job.second.reset();
        o.done_s = since(Clock::now());
        note_done();
        sched.leave();
        {
          std::lock_guard<std::mutex> g(mu);
          --running;
          used -= std::min(footprint_of(cfg, r), budget);
        }
        cv.notify_all();
      }
    };
    (void)be.memory_info(true);  // peak from here
    std::vector<std::thread> pool;
    for (size_t i = 0; i < cfg.slots; ++i) pool.emplace_back(worker);
    for (const Request& r : trace) {
      auto c = payloads.get(r);
      wait_arrival(r);
      const size_t need = std::min(footprint_of(cfg, r), budget);
      std::unique_lock<std::mutex> lk(mu);
      if (running < cfg.slots && used + need > budget) ++rep.budget_waits;
      cv.wait(lk, [&] { return running < cfg.slots && used + need <= budget; });
      ++running;
      used += need;
      rep.peak_admitted = std::max(rep.peak_admitted, used);
      rep.max_concurrent = std::max(rep.max_concurrent, running);
      sched.join();
      rep.outcomes[r.id].start_s = since(Clock::now());
      admitted.emplace_back(&r, std::move(c));
      lk.unlock();
      cv.notify_all();
    }
    {
      std::lock_guard<std::mutex> g(mu);
      closed = true;
    }
    cv.notify_all();
    for (auto& t : pool) t.join();
    rep.stats = sched.stats();
    rep.device_peak = be.memory_info().peak;
  }
  double last = 0;
  for (const auto& o : rep.outcomes) last = std::max(last, o.done_s);
  rep.span_s = last - (trace.empty() ? 0 : trace.front().arrival_s);
  return rep;
}

}  // namespace serve
}  // namespace hyperplonk
