// Prover benchmark on one MI355X (hyperplonk/examples/full_prove_bench.rs and
// subroutines/examples/sumcheck_bench.rs timing boundaries):
//   prove     <nv> <vanilla|jellyfish> <proofs> <warmup> <seed>
//             preprocess once, then time each complete `prove` call; one proof is
//             verified on the host (pairings) and its time reported separately
//   sumcheck  <nv> <2|5> <reps> <seed>
//             whole sumcheck proof with host Fiat-Shamir: from before
//             prover_init through every round to the state drop
//   batch     <reps> <seed>
//             13 MLEs with nv drawn from {14,16,18,20,22} committed against an
//             nv-22 SRS, plus 64 independent sumchecks (nv in [10,18], degree
//             2..5, products (c0, f0..f_{d-1}) + (c1, f_d f0), challenges from
//             each state's own transcript): serially (one commit / one state at
//             a time), batched (one multi_commit; every round of all 64
//             states in shared launches) and overlapped (the commit batch
//             submitted to the side stream while the 64 sumchecks run);
//             outputs must be bit-identical
//   serve     <requests> <rate|0|auto|cap> <nv_min> <nv_max> <jelly_frac> <slots> <seed>
//             [<budget bytes|auto> [<srs nv> [single]]]
//             a request trace (Poisson arrivals at `rate` per second; 0 = all at
//             once; auto = 0.8 x the serial capacity, cap = 0.8 x the batched
//             capacity; both capacities measured on a closed-loop probe of the
//             same request mix), then (unless `single`) again at a high load
//             (0.9 x the batched capacity, above the serial one), each served
//             with continuous batching and again serially (FIFO, one proof at
//             a time); latency = completion - arrival; every proof of the two
//             runs must be identical.  Admission by slots and an HBM budget
//             (bytes, or auto: 90 % of the free device memory) against
//             per-class footprints measured on the device first; one SRS of
//             max(nv_max, srs nv) variables serves every request
//   srs       <nv> <path>
//             device SRS generation from a seeded trapdoor, compression on the
//             device and the canonical file write (load_or_generate_srs), then
//             the file read back and its layers decompressed and prepared on
//             the device; the prepared points must equal the generated ones
// Prints one JSON object per run on stdout.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "circuit.hpp"
#include "mi355x_backend.hpp"
#include "serve.hpp"
#include "srs_file.hpp"

using namespace hyperplonk;
using namespace hyperplonk::host;
using Clock = std::chrono::steady_clock;

static double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

static double percentile(std::vector<double> v, double p) {
  std::sort(v.begin(), v.end());
  const size_t i = std::min(v.size() - 1, static_cast<size_t>(p * (v.size() - 1) + 0.5));
  return v[i];
}

static std::string hex8(const std::vector<uint8_t>& b) {
  const auto d = sha3_256(b.data(), b.size());
  char s[17];
  for (int i = 0; i < 8; ++i) std::snprintf(s + 2 * i, 3, "%02x", d[i]);
  return s;
}

static int run_prove(size_t nv, const std::string& gate, int proofs, int warmup, uint64_t seed) {
  mi355x::Mi355xBackend be(0);
  const MockCircuit c = gate == "vanilla" ? vanilla_circuit(nv, seed) : jellyfish_circuit(nv, seed);
  SplitMix64 trng(seed ^ 0x5352535f54415553ULL);
  std::vector<Fr> t;
  for (size_t i = 0; i < nv; ++i) t.push_back(trng.field());
  auto t0 = Clock::now();
  const auto pp = be.generate_prover_param(nv, to_wire(t));
  const double srs_ms = ms_since(t0);
  const auto full_vk = MultilinearVerifierParam::from_trapdoor(t);
  t0 = Clock::now();
  auto [pk, vk] = preprocess(be, c.index, pp, full_vk);
  const double pre_ms = ms_since(t0);
  HyperPlonkProof first;
  for (int i = 0; i < warmup; ++i) first = prove(be, pk, c.public_inputs, c.witnesses);
  std::vector<double> lat;
  const auto start = Clock::now();
  for (int i = 0; i < proofs; ++i) {
    const auto p0 = Clock::now();
    HyperPlonkProof p = prove(be, pk, c.public_inputs, c.witnesses);
    lat.push_back(ms_since(p0));
    if (i == 0) first = std::move(p);
  }
  const double total_ms = ms_since(start);
  t0 = Clock::now();
  const bool ok = verify(vk, c.public_inputs, first);
  const double verify_ms = ms_since(t0);
  std::printf(
      "{\"mode\": \"prove\", \"nv\": %zu, \"gate\": \"%s\", \"proofs\": %d, \"proofs_per_s\": %.3f, "
      "\"ms_per_proof\": %.4f, \"p50_ms\": %.4f, \"p99_ms\": %.4f, \"verify_ms\": %.3f, \"verified\": %s, "
      "\"srs_ms\": %.3f, \"preprocess_ms\": %.3f, \"proof_sha3_8\": \"%s\"}\n",
      nv, gate.c_str(), proofs, 1e3 * proofs / total_ms, total_ms / proofs, percentile(lat, 0.5),
      percentile(lat, 0.99), verify_ms, ok ? "true" : "false", srs_ms, pre_ms,
      hex8(serialize_proof(first)).c_str());
  return ok ? 0 : 1;
}

// degree 2: c0 f0 f1 + c1 f2 f3 + c2 f4 f5; degree 5: the 5-wire jellyfish terms
// with a scalar-coefficient w1^5 term (14 MLEs)
static int run_sumcheck(size_t nv, int degree, int reps, uint64_t seed) {
  mi355x::Mi355xBackend be(0);
  SplitMix64 rng(seed);
  const size_t k = degree == 2 ? 6 : 14;
  std::vector<mi355x::Mle> m;
  for (size_t i = 0; i < k; ++i) {
    std::vector<FrBytes> tab(size_t(1) << nv);
    for (auto& x : tab) x = rng.field().bytes();
    m.push_back(be.mle_from_evaluations(nv, tab));
  }
  mi355x::VirtualPolynomial vp(nv);
  if (degree == 2) {
    for (int i = 0; i < 3; ++i) vp.add_mle_list({m[2 * i], m[2 * i + 1]}, rng.field().bytes());
  } else {
    const FrBytes one = Fr::one().bytes(), neg = (-Fr::one()).bytes();
    vp.add_mle_list({m[0], m[9]}, one);
    vp.add_mle_list({m[1], m[10]}, one);
    vp.add_mle_list({m[2], m[11]}, one);
    vp.add_mle_list({m[3], m[12]}, one);
    vp.add_mle_list({m[4], m[9], m[10]}, one);
    vp.add_mle_list({m[5], m[11], m[12]}, one);
    vp.add_mle_list({m[6], m[9], m[10], m[11], m[12]}, one);
    vp.add_mle_list({m[7], m[13]}, neg);
    vp.add_mle_list({m[8]}, one);
    vp.add_mle_list({m[9], m[9], m[9], m[9], m[9]}, rng.field().bytes());
  }
  std::vector<double> ms;
  std::string digest;
  bool identical = true;
  for (int r = 0; r < reps + 1; ++r) {
    const auto t0 = Clock::now();
    IOPTranscript tr("sumcheck bench");
    tr.append_message("seed", reinterpret_cast<const uint8_t*>(&seed), 8);
    const IOPProof proof = sumcheck_prove(be, vp, tr);
    const double t = ms_since(t0);
    Writer w;
    for (const auto& msg : proof.proofs) w.frs(msg);
    const std::string d = hex8(w.buf);
    identical = identical && (digest.empty() || d == digest);
    digest = d;  // the last run's messages (runs after the first may replay cached launches)
    if (r == 0) continue;  // first run is warmup
    ms.push_back(t);
  }
  std::printf(
      "{\"mode\": \"sumcheck\", \"nv\": %zu, \"degree\": %d, \"reps\": %d, \"median_ms\": %.4f, "
      "\"min_ms\": %.4f, \"messages_sha3_8\": \"%s\", \"runs_identical\": %s}\n",
      nv, degree, reps, percentile(ms, 0.5), percentile(ms, 0.0), digest.c_str(),
      identical ? "true" : "false");
  return identical ? 0 : 1;
}

struct BatchCase {
  std::vector<mi355x::Mle> commit_mles;
  std::vector<mi355x::VirtualPolynomial> vps;
};

static BatchCase make_batch_case(const mi355x::Mi355xBackend& be, uint64_t seed) {
  SplitMix64 rng(seed);
  BatchCase bc;
  const size_t choices[5] = {14, 16, 18, 20, 22};
  for (int i = 0; i < 13; ++i) {
    const size_t nv = choices[rng.below(5)];
    std::vector<FrBytes> tab(size_t(1) << nv);
    for (auto& x : tab) x = rng.field().bytes();
    bc.commit_mles.push_back(be.mle_from_evaluations(nv, tab));
  }
  for (int i = 0; i < 64; ++i) {
    const size_t nv = 10 + rng.below(9);
    const size_t d = 2 + rng.below(4);
    std::vector<mi355x::Mle> f;
    for (size_t k = 0; k <= d; ++k) {
      std::vector<FrBytes> tab(size_t(1) << nv);
      for (auto& x : tab) x = rng.field().bytes();
      f.push_back(be.mle_from_evaluations(nv, tab));
    }
    mi355x::VirtualPolynomial vp(nv);
    vp.add_mle_list(std::vector<mi355x::Mle>(f.begin(), f.begin() + d), rng.field().bytes());
    vp.add_mle_list({f[d], f[0]}, rng.field().bytes());
    bc.vps.push_back(std::move(vp));
  }
  return bc;
}

// all messages of the 64 sumchecks; batched: rounds of every live state per call
static std::vector<std::vector<std::vector<Fr>>> run_sumchecks(const mi355x::Mi355xBackend& be,
                                                               const BatchCase& bc, bool batched) {
  const size_t n = bc.vps.size();
  std::vector<std::vector<std::vector<Fr>>> msgs(n);
  std::vector<IOPTranscript> tr;
  for (size_t i = 0; i < n; ++i) {
    tr.emplace_back("batched sumcheck");
    const uint64_t id = i;
    tr.back().append_message("state", reinterpret_cast<const uint8_t*>(&id), 8);
    tr.back().append_message("aux info", ser_aux(bc.vps[i].aux_info));
  }
  if (!batched) {
    for (size_t i = 0; i < n; ++i) {
      auto st = be.prover_init(bc.vps[i]);
      std::optional<FrBytes> ch;
      for (size_t r = 0; r < bc.vps[i].aux_info.num_variables; ++r) {
        msgs[i].push_back(from_wire(be.prove_round_and_update_state(st, ch)));
        tr[i].append_message("prover msg", ser_frs(msgs[i].back()));
        ch = tr[i].get_and_append_challenge("Internal round").bytes();
      }
    }
    return msgs;
  }
  std::vector<mi355x::ProverState> states;
  for (size_t i = 0; i < n; ++i) states.push_back(be.prover_init(bc.vps[i]));
  std::vector<std::optional<FrBytes>> ch(n);
  for (size_t r = 0;; ++r) {
    std::vector<mi355x::ProverState*> live;
    std::vector<std::optional<FrBytes>> lch;
    std::vector<size_t> idx;
    for (size_t i = 0; i < n; ++i)
      if (r < bc.vps[i].aux_info.num_variables) {
        live.push_back(&states[i]);
        lch.push_back(ch[i]);
        idx.push_back(i);
      }
    if (live.empty()) break;
    const auto out = be.prove_rounds_batch(live, lch);
    for (size_t k = 0; k < idx.size(); ++k) {
      const size_t i = idx[k];
      msgs[i].push_back(from_wire(out[k]));
      tr[i].append_message("prover msg", ser_frs(msgs[i].back()));
      ch[i] = tr[i].get_and_append_challenge("Internal round").bytes();
    }
  }
  return msgs;
}

static int run_batch(int reps, uint64_t seed) {
  mi355x::Mi355xBackend be(0);
  const size_t N = 22;
  SplitMix64 trng(seed ^ 0x5352535f54415553ULL);
  std::vector<Fr> t;
  for (size_t i = 0; i < N; ++i) t.push_back(trng.field());
  auto t0 = Clock::now();
  const auto pp = be.generate_prover_param(N, to_wire(t));
  const double srs_ms = ms_since(t0);
  const BatchCase bc = make_batch_case(be, seed);
  size_t commit_points = 0;
  for (const auto& m : bc.commit_mles) commit_points += size_t(1) << m.num_vars();
  std::vector<double> sc_serial, sc_batched, cm_serial, cm_batched, overlapped;
  bool same = true;
  std::vector<mi355x::Mi355xBackend::KzgOp> commit_ops;
  for (const auto& m : bc.commit_mles) commit_ops.push_back({false, m, {}});
  for (int rep = 0; rep < reps + 1; ++rep) {
    t0 = Clock::now();
    std::vector<G1Packet> serial_c;
    for (const auto& m : bc.commit_mles) serial_c.push_back(be.commit(pp, m));
    const double a = ms_since(t0);
    t0 = Clock::now();
    const auto batched_c = be.multi_commit(pp, bc.commit_mles);
    const double b = ms_since(t0);
    t0 = Clock::now();
    const auto ms = run_sumchecks(be, bc, false);
    const double c = ms_since(t0);
    t0 = Clock::now();
    const auto mb = run_sumchecks(be, bc, true);
    const double d = ms_since(t0);
    // overlapped: the commit batch runs on the side stream while the 64
    // sumchecks' rounds run on the context stream
    t0 = Clock::now();
    auto pending = be.kzg_batch_submit(pp, commit_ops);
    const auto mo = run_sumchecks(be, bc, true);
    const auto co = pending.wait();
    const double e = ms_since(t0);
    same = same && serial_c == batched_c && ms == mb && mo == ms;
    for (size_t i = 0; i < co.size(); ++i) same = same && co[i].points[0] == serial_c[i];
    if (rep == 0) continue;  // warmup
    cm_serial.push_back(a);
    cm_batched.push_back(b);
    sc_serial.push_back(c);
    sc_batched.push_back(d);
    overlapped.push_back(e);
  }
  // standalone MSMs (config 2 sizes) against the same SRS
  std::string sizes;
  SplitMix64 mrng(seed ^ 0x4d534dULL);
  for (size_t lg : {16, 18, 20, 22}) {
    std::vector<FrBytes> tab(size_t(1) << lg);
    for (auto& x : tab) x = mrng.field().bytes();
    const auto m = be.mle_from_evaluations(lg, tab);
    be.commit(pp, m);
    std::vector<double> v;
    for (int r = 0; r < 5; ++r) {
      t0 = Clock::now();
      be.commit(pp, m);
      v.push_back(ms_since(t0));
    }
    const double med = percentile(v, 0.5);
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s\"2^%zu\": {\"ms\": %.3f, \"pts_per_s\": %.0f}", sizes.empty() ? "" : ", ",
                  lg, med, (double)(size_t(1) << lg) / (med * 1e-3));
    sizes += buf;
  }
  const double cs = percentile(cm_serial, 0.5), cb = percentile(cm_batched, 0.5);
  const double ss = percentile(sc_serial, 0.5), sb = percentile(sc_batched, 0.5);
  const double ov = percentile(overlapped, 0.5);
  std::printf(
      "{\"mode\": \"batch\", \"reps\": %d, \"commit_points\": %zu, \"srs_ms\": %.1f, "
      "\"commit_serial_ms\": %.3f, \"commit_batched_ms\": %.3f, \"sumcheck64_serial_ms\": %.3f, "
      "\"sumcheck64_batched_ms\": %.3f, \"total_serial_ms\": %.3f, \"total_batched_ms\": %.3f, "
      "\"total_overlapped_ms\": %.3f, \"speedup\": %.3f, \"speedup_overlapped\": %.3f, "
      "\"overlapped_over_serial\": %.3f, \"bit_identical\": %s, \"msm_sizes\": {%s}}\n",
      reps, commit_points, srs_ms, cs, cb, ss, sb, cs + ss, cb + sb, ov, (cs + ss) / (cb + sb),
      (cs + ss) / ov, ov / (cs + ss), same ? "true" : "false", sizes.c_str());
  return same ? 0 : 1;
}

struct Summary {
  double proofs_per_s = 0, p50 = 0, p99 = 0, mean = 0;
  size_t errors = 0;
};

static Summary summarize(const std::vector<serve::Request>& trace, const serve::ServeReport& rep) {
  Summary s;
  std::vector<double> lat;
  for (size_t i = 0; i < trace.size(); ++i) {
    const auto& o = rep.outcomes[i];
    if (!o.error.empty()) ++s.errors;
    lat.push_back(1e3 * (o.done_s - trace[i].arrival_s));
  }
  s.proofs_per_s = trace.size() / rep.span_s;
  s.p50 = percentile(lat, 0.5);
  s.p99 = percentile(lat, 0.99);
  for (double x : lat) s.mean += x / lat.size();
  return s;
}

static double busy_ms(const serve::ServeReport& rep) {
  double t = 0;
  for (const auto& o : rep.outcomes) t += 1e3 * (o.done_s - o.start_s);
  return t;
}

struct ServeOptions {
  size_t budget = 0;   // 0: measured (90 % of the free device memory)
  size_t srs_nv = 0;   // SRS size (at least nv_max): one resident SRS serves every nv (trim)
  bool single = false; // only the first rate (no high-load pair)
};

static int run_serve(size_t count, const std::string& rate_arg, size_t nv_min, size_t nv_max, double jelly,
                     size_t slots, uint64_t seed, const ServeOptions& opt) {
  mi355x::Mi355xBackend be(0);
  const size_t N = std::max(nv_max, opt.srs_nv);
  SplitMix64 trng(seed ^ 0x5352535f54415553ULL);
  std::vector<Fr> t;
  for (size_t i = 0; i < N; ++i) t.push_back(trng.field());
  auto t0 = Clock::now();
  const auto pp = be.generate_prover_param(N, to_wire(t));
  const double srs_ms = ms_since(t0);
  const auto vk = MultilinearVerifierParam::from_trapdoor(t);
  serve::ServeConfig cfg;
  cfg.slots = slots;
  if (const char* e = std::getenv("HP_SERVE_KZG_HOLD_US")) cfg.kzg_hold = std::chrono::microseconds(std::atol(e));
  if (const char* e = std::getenv("HP_SERVE_WINDOW_US")) cfg.window = std::chrono::microseconds(std::atol(e));
  cfg.gen_threads = std::max<unsigned>(2, std::min<unsigned>(16, std::thread::hardware_concurrency()) - 2);
  cfg.progress_every = count >= 500 ? 100 : 0;
  double rate = 0;
  double capacity = 0;
  // the request mix (arrival times do not change it) and the probe's mix share one payload pool
  const size_t probe = std::min<size_t>(count, std::max<size_t>(8 * slots, 128));
  const auto ptrace = serve::make_trace(probe, 0, nv_min, nv_max, jelly, seed ^ 0x70726f6265ULL);
  auto trace = serve::make_trace(count, 1.0, nv_min, nv_max, jelly, seed);
  std::vector<serve::Request> all = ptrace;
  all.insert(all.end(), trace.begin(), trace.end());
  t0 = Clock::now();
  const serve::CircuitPool payloads(all, cfg.variants, cfg.gen_threads);
  const double pool_ms = ms_since(t0);
  // admission: per-class device footprints measured on this device, and the budget
  t0 = Clock::now();
  cfg.footprints = serve::measure_footprints(be, pp, vk, all, payloads);
  const double footprint_ms = ms_since(t0);
  cfg.hbm_budget = opt.budget;
  size_t max_footprint = 0;
  for (const auto& kv : cfg.footprints) max_footprint = std::max(max_footprint, kv.second);
  // closed-loop probes (everything at t = 0): the capacity of each mode
  double capacity_serial = 0;
  {
    const auto prep = serve::run_trace(be, pp, vk, ptrace, payloads, cfg);
    capacity = probe / prep.span_s;
    serve::ServeConfig serial = cfg;
    serial.batched = false;
    const auto srep = serve::run_trace(be, pp, vk, ptrace, payloads, serial);
    capacity_serial = probe / srep.span_s;
  }
  // auto: a fixed rate the serial prover sustains (0.8 x its capacity), so
  // both modes keep up and the latencies compare like for like; cap: 0.8 x
  // the batched capacity (the prover-service stream's definition)
  rate = rate_arg == "auto"  ? 0.8 * capacity_serial
         : rate_arg == "cap" ? 0.8 * capacity
                             : std::atof(rate_arg.c_str());
  trace = serve::make_trace(count, rate, nv_min, nv_max, jelly, seed);
  serve::ServeConfig batched = cfg;
  const auto brep = serve::run_trace(be, pp, vk, trace, payloads, batched);
  serve::ServeConfig serial = cfg;
  serial.batched = false;
  const auto srep = serve::run_trace(be, pp, vk, trace, payloads, serial);
  size_t identical = 0;
  for (size_t i = 0; i < trace.size(); ++i)
    identical += brep.outcomes[i].error.empty() && brep.outcomes[i].proof == srep.outcomes[i].proof;
  // verify the proofs of the first few requests on the host
  size_t verified = 0, checked = 0;
  for (size_t i = 0; i < trace.size() && checked < 3; ++i) {
    if (trace[i].nv > 14) continue;
    const auto c = serve::circuit_of(trace[i], cfg.variants);
    auto [pk, vkey] = preprocess(be, c.index, pp, vk);
    verified += verify(vkey, c.public_inputs, prove(be, pk, c.public_inputs, c.witnesses));
    ++checked;
  }
  const Summary b = summarize(trace, brep), s = summarize(trace, srep);
  // high load: 0.9 x the batched capacity, above what the serial prover sustains
  const double rate_hi = 0.9 * capacity;
  std::vector<serve::Request> trace_hi;
  serve::ServeReport brep_hi, srep_hi;
  size_t identical_hi = 0;
  Summary bh, sh;
  if (!opt.single) {
    trace_hi = serve::make_trace(count, rate_hi, nv_min, nv_max, jelly, seed);
    brep_hi = serve::run_trace(be, pp, vk, trace_hi, payloads, batched);
    srep_hi = serve::run_trace(be, pp, vk, trace_hi, payloads, serial);
    for (size_t i = 0; i < trace_hi.size(); ++i)
      identical_hi += brep_hi.outcomes[i].error.empty() &&
                      brep_hi.outcomes[i].proof == srep_hi.outcomes[i].proof &&
                      brep_hi.outcomes[i].proof == brep.outcomes[i].proof;
    bh = summarize(trace_hi, brep_hi);
    sh = summarize(trace_hi, srep_hi);
  }
  size_t jelly_n = 0;
  for (const auto& r : trace) jelly_n += r.jellyfish;
  std::printf(
      "{\"mode\": \"serve\", \"requests\": %zu, \"jellyfish\": %zu, \"nv\": [%zu, %zu], \"rate_per_s\": %.3f, "
      "\"capacity_probe_per_s\": %.3f, \"capacity_serial_probe_per_s\": %.3f, \"probe_requests\": %zu, "
      "\"payload_circuits\": %zu, "
      "\"payload_pool_ms\": %.0f, \"slots\": %zu, \"srs_num_vars\": %zu, \"srs_ms\": %.1f, "
      "\"batched\": {\"proofs_per_s\": %.3f, \"p50_ms\": %.2f, \"p99_ms\": %.2f, \"mean_ms\": %.2f, "
      "\"errors\": %zu, \"max_concurrent\": %zu, \"ticks\": %zu, \"ops\": %zu, \"kzg_calls\": %zu, "
      "\"round_calls\": %zu, \"eval_calls\": %zu, \"kzg_held\": %zu, \"kzg_hold_us\": %ld, \"device_ms\": {\"gather\": %.1f, \"generic\": %.1f, "
      "\"kzg\": %.1f, \"rounds\": %.1f, \"evals\": %.1f}, \"request_wait_ms\": %.1f, \"request_total_ms\": %.1f}, "
      "\"serial\": {\"proofs_per_s\": %.3f, \"p50_ms\": %.2f, \"p99_ms\": %.2f, \"mean_ms\": %.2f, "
      "\"errors\": %zu}, \"identical_proofs\": %zu, "
      "\"high_load\": {\"rate_per_s\": %.3f, \"batched\": {\"proofs_per_s\": %.3f, \"p50_ms\": %.2f, "
      "\"p99_ms\": %.2f, \"mean_ms\": %.2f, \"errors\": %zu}, \"serial\": {\"proofs_per_s\": %.3f, "
      "\"p50_ms\": %.2f, \"p99_ms\": %.2f, \"mean_ms\": %.2f, \"errors\": %zu}, \"identical_proofs\": %zu}, "
      "\"verified\": %zu, \"verify_checked\": %zu, \"rate_rule\": \"%s\", \"hbm_budget_bytes\": %zu, "
      "\"budget_waits\": %zu, \"peak_admitted_bytes\": %zu, \"device_peak_bytes\": %zu, "
      "\"footprint_classes\": %zu, \"max_footprint_bytes\": %zu, \"footprint_measure_ms\": %.0f, "
      "\"high_load_run\": %s}\n",
      trace.size(), jelly_n, nv_min, nv_max, rate, capacity, capacity_serial, probe, payloads.size(),
      pool_ms, slots, N,
      srs_ms, b.proofs_per_s, b.p50, b.p99,
      b.mean, b.errors, brep.max_concurrent, brep.stats.ticks, brep.stats.ops, brep.stats.kzg_calls,
      brep.stats.round_calls, brep.stats.eval_calls, brep.stats.kzg_held, (long)cfg.kzg_hold.count(),
      brep.stats.gather_ms, brep.stats.generic_ms,
      brep.stats.kzg_ms, brep.stats.round_ms, brep.stats.eval_ms, brep.stats.waiting_ms, busy_ms(brep), s.proofs_per_s, s.p50, s.p99, s.mean, s.errors, identical,
      rate_hi, bh.proofs_per_s, bh.p50, bh.p99, bh.mean, bh.errors, sh.proofs_per_s, sh.p50, sh.p99,
      sh.mean, sh.errors, identical_hi, verified, checked,
      rate_arg == "auto" ? "0.8 x serial capacity" : rate_arg == "cap" ? "0.8 x batched capacity" : "given",
      brep.hbm_budget, brep.budget_waits, brep.peak_admitted, brep.device_peak, cfg.footprints.size(),
      max_footprint, footprint_ms, opt.single ? "false" : "true");
  return identical == trace.size() && identical_hi == trace_hi.size() && verified == checked ? 0 : 1;
}

static int run_srs(size_t nv, const std::string& path) {
  mi355x::Mi355xBackend be(0);
  SplitMix64 trng(0x5352535f54415553ULL ^ nv);
  std::vector<Fr> t;
  for (size_t i = 0; i < nv; ++i) t.push_back(trng.field());
  std::remove(path.c_str());
  // warm the device (context, kernels) outside the timed setups
  (void)be.generate_prover_param(2, to_wire(std::vector<Fr>(t.begin(), t.begin() + 2)));
  auto t0 = Clock::now();
  const auto gpp = be.generate_prover_param(nv, to_wire(t));
  const double gen_device_ms = ms_since(t0);
  const SrsSetup made = load_or_generate_srs(be, nv, t, path);
  const SrsSetup loaded = load_or_generate_srs(be, nv, t, path);
  t0 = Clock::now();
  const auto lpp = be.prepare_prover_param_compressed(nv, loaded.srs.prover_layers(nv));
  const double prepare_ms = ms_since(t0);
  const bool same = be.download_prover_param(lpp, true) == be.download_prover_param(gpp, true);
  const size_t bytes = serialize_srs(loaded.srs).size();
  std::remove(path.c_str());
  std::printf(
      "{\"mode\": \"srs\", \"num_vars\": %zu, \"points\": %zu, \"file_bytes\": %zu, "
      "\"generate_device_ms\": %.1f, \"generate_and_write_ms\": %.1f, \"read_ms\": %.1f, "
      "\"prepare_from_file_ms\": %.1f, \"load_total_ms\": %.1f, \"identical\": %s}\n",
      nv, (size_t(2) << nv) - 1, bytes, gen_device_ms, made.setup_ms, loaded.load_ms, prepare_ms,
      loaded.load_ms + prepare_ms, same ? "true" : "false");
  return same ? 0 : 1;
}

int main(int argc, char** argv) {
  try {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "prove" && argc == 7)
      return run_prove(std::strtoul(argv[2], nullptr, 10), argv[3], std::atoi(argv[4]), std::atoi(argv[5]),
                       std::strtoull(argv[6], nullptr, 0));
    if (mode == "batch" && argc == 4)
      return run_batch(std::atoi(argv[2]), std::strtoull(argv[3], nullptr, 0));
    if (mode == "serve" && argc >= 9 && argc <= 12) {
      ServeOptions opt;
      if (argc > 9 && std::string(argv[9]) != "auto") opt.budget = std::strtoull(argv[9], nullptr, 0);
      if (argc > 10) opt.srs_nv = std::strtoul(argv[10], nullptr, 10);
      if (argc > 11) opt.single = std::string(argv[11]) == "single";
      return run_serve(std::strtoul(argv[2], nullptr, 10), argv[3], std::strtoul(argv[4], nullptr, 10),
                       std::strtoul(argv[5], nullptr, 10), std::atof(argv[6]), std::strtoul(argv[7], nullptr, 10),
                       std::strtoull(argv[8], nullptr, 0), opt);
    }
    if (mode == "srs" && argc == 4) return run_srs(std::strtoul(argv[2], nullptr, 10), argv[3]);
    if (mode == "sumcheck" && argc == 6)
      return run_sumcheck(std::strtoul(argv[2], nullptr, 10), std::atoi(argv[3]), std::atoi(argv[4]),
                          std::strtoull(argv[5], nullptr, 0));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "prove_bench: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr,
               "usage: %s prove <nv> <vanilla|jellyfish> <proofs> <warmup> <seed>\n"
               "       %s sumcheck <nv> <2|5> <reps> <seed>\n"
               "       %s batch <reps> <seed>\n"
               "       %s serve <requests> <rate|0|auto|cap> <nv_min> <nv_max> <jelly_frac> <slots> <seed>"
               " [<budget bytes|auto> [<srs nv> [single]]]\n"
               "       %s srs <nv> <path>\n",
               argv[0], argv[0], argv[0], argv[0], argv[0]);
  return 2;
}
