// The prover service (hyperplonk_amd/host/serve.hpp) under an HBM budget that
// binds: a closed-loop trace of mixed vanilla / jellyfish requests is served
// with continuous batching while admission holds back requests whose measured
// footprint does not fit, then serially; every served proof must equal the
// serial one, and sampled served proofs must equal the proofs the CPU oracle
// backend makes of the same circuits with the same SRS trapdoor (and verify).
// Usage: serve_oracle_test <requests> <nv_min> <nv_max> <slots> <budget_in_footprints> <seed>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../hyperplonk_amd/host/serve.hpp"
#include "oracle_backend.hpp"

using namespace hyperplonk;
using namespace hyperplonk::host;

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s <requests> <nv_min> <nv_max> <slots> <budget_in_footprints> <seed>\n",
                 argv[0]);
    return 2;
  }
  const size_t count = std::strtoul(argv[1], nullptr, 10);
  const size_t nv_min = std::strtoul(argv[2], nullptr, 10), nv_max = std::strtoul(argv[3], nullptr, 10);
  const size_t slots = std::strtoul(argv[4], nullptr, 10);
  const double budget_units = std::atof(argv[5]);
  const uint64_t seed = std::strtoull(argv[6], nullptr, 0);
  mi355x::Mi355xBackend gpu(0);
  oracle_cpu::OracleBackend cpu(8);
  SplitMix64 trng(seed ^ 0x5352535f54415553ULL);
  std::vector<Fr> t;
  for (size_t i = 0; i < nv_max; ++i) t.push_back(trng.field());
  const auto gpp = gpu.generate_prover_param(nv_max, to_wire(t));
  const auto vk = MultilinearVerifierParam::from_trapdoor(t);
  const auto trace = serve::make_trace(count, 0, nv_min, nv_max, 0.3, seed);
  serve::ServeConfig cfg;
  cfg.slots = slots;
  const serve::CircuitPool payloads(trace, cfg.variants, 4);
  cfg.footprints = serve::measure_footprints(gpu, gpp, vk, trace, payloads);
  size_t largest = 0;
  for (const auto& kv : cfg.footprints) largest = std::max(largest, kv.second);
  cfg.hbm_budget = static_cast<size_t>(budget_units * static_cast<double>(largest));
  std::printf("classes %zu largest footprint %zu budget %zu\n", cfg.footprints.size(), largest, cfg.hbm_budget);
  std::fflush(stdout);
  const auto brep = serve::run_trace(gpu, gpp, vk, trace, payloads, cfg);
  serve::ServeConfig serial = cfg;
  serial.batched = false;
  const auto srep = serve::run_trace(gpu, gpp, vk, trace, payloads, serial);
  int failures = 0;
  auto check = [&](bool ok, const char* what) {
    if (!ok) {
      std::printf("  FAIL %s\n", what);
      ++failures;
    }
  };
  size_t identical = 0;
  for (size_t i = 0; i < trace.size(); ++i)
    identical += brep.outcomes[i].error.empty() && brep.outcomes[i].proof == srep.outcomes[i].proof;
  std::printf("identical %zu/%zu max_concurrent %zu budget_waits %zu peak_admitted %zu device_peak %zu\n",
              identical, trace.size(), brep.max_concurrent, brep.budget_waits, brep.peak_admitted,
              brep.device_peak);
  check(identical == trace.size(), "served proofs equal serial proofs");
  check(brep.budget_waits > 0, "the budget held requests back");
  check(brep.peak_admitted <= cfg.hbm_budget, "admitted footprints within the budget");
  check(brep.max_concurrent > 1, "requests overlapped");
  // sampled proofs against the oracle backend: one vanilla, one jellyfish, the largest
  const auto cpp = cpu.generate_prover_param(nv_max, to_wire(t));
  std::vector<size_t> sample;
  for (size_t i = 0; i < trace.size() && sample.size() < 1; ++i)
    if (!trace[i].jellyfish) sample.push_back(i);
  for (size_t i = 0; i < trace.size() && sample.size() < 2; ++i)
    if (trace[i].jellyfish) sample.push_back(i);
  size_t big = 0;
  for (size_t i = 0; i < trace.size(); ++i)
    if (trace[i].nv > trace[big].nv) big = i;
  sample.push_back(big);
  for (size_t i : sample) {
    const auto c = payloads.get(trace[i]);
    auto [pk, cvk] = preprocess(cpu, c->index, cpp, vk);
    const HyperPlonkProof cp = prove(cpu, pk, c->public_inputs, c->witnesses);
    const auto cb = serialize_proof(cp);
    const bool same = cb == brep.outcomes[i].proof;
    const bool ok = verify(cvk, c->public_inputs, cp);
    std::printf("  request %zu nv %zu %s: served proof %s oracle proof, verify %s\n", i, trace[i].nv,
                trace[i].jellyfish ? "jellyfish" : "vanilla", same ? "==" : "!=", ok ? "accept" : "REJECT");
    check(same && ok, "served proof equals the oracle backend's");
  }
  std::printf("SERVE_ORACLE %s\n", failures ? "FAIL" : "PASS");
  return failures ? 1 : 0;
}
