// HyperPlonk proofs on the MI355X backend against the same protocol run on the
// CPU oracle backend: identical seeded circuit and SRS trapdoor, proofs compared
// byte for byte, the GPU proof verified with pairings.
// Usage: protocol_gpu_test <nv> <vanilla|jellyfish> <seed> [<nv> <gate> <seed> ...]
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../hyperplonk_amd/host/circuit.hpp"
#include "../../hyperplonk_amd/host/mi355x_backend.hpp"
#include "oracle_backend.hpp"

using namespace hyperplonk;
using namespace hyperplonk::host;

static std::string hex(const uint8_t* p, size_t n) {
  std::string s;
  char b[3];
  for (size_t i = 0; i < n; ++i) {
    std::snprintf(b, 3, "%02x", p[i]);
    s += b;
  }
  return s;
}

int main(int argc, char** argv) {
  if (argc < 4 || (argc - 1) % 3) {
    std::fprintf(stderr, "usage: %s <nv> <vanilla|jellyfish> <seed> ...\n", argv[0]);
    return 2;
  }
  mi355x::Mi355xBackend gpu(0);
  oracle_cpu::OracleBackend cpu(8);
  int failures = 0;
  for (int a = 1; a + 2 < argc; a += 3) {
    const size_t nv = std::strtoul(argv[a], nullptr, 10);
    const std::string gate = argv[a + 1];
    const uint64_t seed = std::strtoull(argv[a + 2], nullptr, 0);
    std::printf("case nv=%zu gate=%s seed=%llu\n", nv, gate.c_str(), (unsigned long long)seed);
    std::fflush(stdout);
    const MockCircuit c = gate == "vanilla" ? vanilla_circuit(nv, seed) : jellyfish_circuit(nv, seed);
    SplitMix64 trng(seed ^ 0x5352535f54415553ULL);
    std::vector<Fr> t;
    for (size_t i = 0; i < nv; ++i) t.push_back(trng.field());
    const auto full_vk = MultilinearVerifierParam::from_trapdoor(t);
    const auto gpp = gpu.generate_prover_param(nv, to_wire(t));
    const auto cpp = cpu.generate_prover_param(nv, to_wire(t));
    auto [gpk, gvk] = preprocess(gpu, c.index, gpp, full_vk);
    auto [cpk, cvk] = preprocess(cpu, c.index, cpp, full_vk);
    const bool keys_equal = gvk.selector_commitments == cvk.selector_commitments &&
                            gvk.perm_commitments == cvk.perm_commitments;
    const HyperPlonkProof gp = prove(gpu, gpk, c.public_inputs, c.witnesses);
    const HyperPlonkProof cp = prove(cpu, cpk, c.public_inputs, c.witnesses);
    const auto gb = serialize_proof(gp), cb = serialize_proof(cp);
    const auto gd = sha3_256(gb.data(), gb.size()), cd = sha3_256(cb.data(), cb.size());
    const bool same = gb == cb;
    const bool ok = verify(gvk, c.public_inputs, gp);
    std::printf("  keys %s proof gpu %s oracle %s %s verify %s\n", keys_equal ? "equal" : "DIFFER",
                hex(gd.data(), 8).c_str(), hex(cd.data(), 8).c_str(), same ? "MATCH" : "MISMATCH",
                ok ? "accept" : "REJECT");
    if (!keys_equal || !same || !ok) ++failures;
  }
  std::printf("PROTOCOL_GPU %s\n", failures ? "FAIL" : "PASS");
  return failures ? 1 : 0;
}
