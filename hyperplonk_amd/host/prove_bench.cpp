// Prover benchmark on one MI355X (hyperplonk/examples/full_prove_bench.rs and
// subroutines/examples/sumcheck_bench.rs timing boundaries):
//   prove     <nv> <vanilla|jellyfish> <proofs> <warmup> <seed>
//             preprocess once, then time each complete `prove` call; one proof is
//             verified on the host (pairings) and its time reported separately
//   sumcheck  <nv> <2|5> <reps> <seed>
//             whole sumcheck proof with host Fiat-Shamir: from before
//             prover_init through every round to the state drop
// Prints one JSON object per run on stdout.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "circuit.hpp"
#include "mi355x_backend.hpp"

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
  for (int r = 0; r < reps + 1; ++r) {
    const auto t0 = Clock::now();
    IOPTranscript tr("sumcheck bench");
    tr.append_message("seed", reinterpret_cast<const uint8_t*>(&seed), 8);
    const IOPProof proof = sumcheck_prove(be, vp, tr);
    const double t = ms_since(t0);
    if (r == 0) {  // first run is warmup
      Writer w;
      for (const auto& msg : proof.proofs) w.frs(msg);
      digest = hex8(w.buf);
      continue;
    }
    ms.push_back(t);
  }
  std::printf(
      "{\"mode\": \"sumcheck\", \"nv\": %zu, \"degree\": %d, \"reps\": %d, \"median_ms\": %.4f, "
      "\"min_ms\": %.4f, \"messages_sha3_8\": \"%s\"}\n",
      nv, degree, reps, percentile(ms, 0.5), percentile(ms, 0.0), digest.c_str());
  return 0;
}

int main(int argc, char** argv) {
  try {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "prove" && argc == 7)
      return run_prove(std::strtoul(argv[2], nullptr, 10), argv[3], std::atoi(argv[4]), std::atoi(argv[5]),
                       std::strtoull(argv[6], nullptr, 0));
    if (mode == "sumcheck" && argc == 6)
      return run_sumcheck(std::strtoul(argv[2], nullptr, 10), std::atoi(argv[3]), std::atoi(argv[4]),
                          std::strtoull(argv[5], nullptr, 0));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "prove_bench: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr,
               "usage: %s prove <nv> <vanilla|jellyfish> <proofs> <warmup> <seed>\n"
               "       %s sumcheck <nv> <2|5> <reps> <seed>\n",
               argv[0], argv[0]);
  return 2;
}
