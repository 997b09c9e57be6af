// HyperPlonk proofs on the MI355X backend against the same protocol run on the
// CPU oracle backend: identical seeded circuit and SRS trapdoor, proofs compared
// byte for byte, the GPU proof verified with pairings.
// Usage: protocol_gpu_test <nv> <vanilla|jellyfish> <seed> [<nv> <gate> <seed> ...]
//        protocol_gpu_test srs <nv> <dir>
//   srs: device-generated SRS compressed on the device and on the host (same
//   bytes), written as a canonical SRS file, read back and decompressed on the
//   device (same points as generated, also trimmed), a proof from the loaded
//   SRS identical to one from the generated SRS and accepted by the verifier
//   key read from the file; malformed layers rejected.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../hyperplonk_amd/host/circuit.hpp"
#include "../../hyperplonk_amd/host/mi355x_backend.hpp"
#include "../../hyperplonk_amd/host/srs_file.hpp"
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


static int srs_case(size_t nv, const std::string& dir) {
  mi355x::Mi355xBackend gpu(0);
  int failures = 0;
  auto check = [&](bool ok, const char* what) {
    if (!ok) {
      std::printf("  FAIL %s\n", what);
      ++failures;
    }
  };
  std::printf("case srs nv=%zu\n", nv);
  std::fflush(stdout);
  SplitMix64 trng(0x5352535f54415553ULL ^ nv);
  std::vector<Fr> t;
  for (size_t i = 0; i < nv; ++i) t.push_back(trng.field());
  const std::string path = dir + "/srs_nv" + std::to_string(nv) + ".bin";
  std::remove(path.c_str());
  const SrsSetup made = load_or_generate_srs(gpu, nv, t, path);
  check(made.source == "generated_and_written_canonical_file", "first setup generates and writes");
  const SrsSetup loaded = load_or_generate_srs(gpu, nv, t, path);
  check(loaded.source == "loaded_canonical_file", "second setup loads");
  check(loaded.srs.layers == made.srs.layers, "loaded layers");
  const auto gpp = gpu.generate_prover_param(nv, to_wire(t));
  const auto packets = gpu.download_prover_param(gpp, false);
  // device compression against host compression of the device's points
  bool host_same = true;
  for (size_t i = 0; i <= nv; ++i)
    for (size_t x = 0; x < packets[i].size() / 97; ++x) {
      G1Packet p;
      std::copy(packets[i].begin() + 97 * x, packets[i].begin() + 97 * (x + 1), p.begin());
      const auto c = g1_compressed(p);
      host_same = host_same && std::equal(c.begin(), c.end(), made.srs.layers[i].begin() + 48 * x);
    }
  check(host_same, "device compression equals host compression");
  // host decompression of a sample equals the device points
  bool sample_same = true;
  for (size_t i = 0; i <= nv; ++i) {
    const size_t n = packets[i].size() / 97;
    for (size_t x : {size_t(0), n / 3, n - 1}) {
      const G1Packet p = g1_to_packet(g1_decompress(loaded.srs.layers[i].data() + 48 * x));
      sample_same = sample_same && std::equal(p.begin(), p.end(), packets[i].begin() + 97 * x);
    }
  }
  check(sample_same, "host decompression of sampled points");
  // device decompression: the full parameters and a trimmed set
  const auto lpp = gpu.prepare_prover_param_compressed(nv, loaded.srs.prover_layers(nv));
  check(gpu.download_prover_param(lpp, false) == packets, "device decompression (full)");
  if (nv >= 2) {
    const auto tpp = gpu.prepare_prover_param_compressed(nv - 2, loaded.srs.prover_layers(nv - 2));
    const auto tp = gpu.download_prover_param(tpp, false);
    check(std::equal(tp.begin(), tp.end(), packets.begin() + 2), "device decompression (trimmed)");
  }
  // the verifier key read from the file
  const auto full_vk = MultilinearVerifierParam::from_trapdoor(t);
  const auto fvk = loaded.srs.verifier_param(nv);
  bool vk_same = fvk.h_mask.size() == full_vk.h_mask.size();
  for (size_t i = 0; vk_same && i < fvk.h_mask.size(); ++i)
    vk_same = fvk.h_mask[i].x == full_vk.h_mask[i].x && fvk.h_mask[i].y == full_vk.h_mask[i].y;
  check(vk_same, "verifier masks");
  // a proof from the loaded SRS equals one from the generated SRS and verifies
  const MockCircuit c = vanilla_circuit(nv, 5);
  auto [gpk, gvk] = preprocess(gpu, c.index, gpp, full_vk);
  auto [lpk, lvk] = preprocess(gpu, c.index, lpp, fvk);
  const auto gb = serialize_proof(prove(gpu, gpk, c.public_inputs, c.witnesses));
  const HyperPlonkProof lp = prove(gpu, lpk, c.public_inputs, c.witnesses);
  check(serialize_proof(lp) == gb, "proof from loaded SRS");
  check(verify(lvk, c.public_inputs, lp), "verify with file key");
  // malformed layers are refused on the device; the encodings of infinity and
  // of -P decompress to the identity and to -P
  auto refused = [&](std::vector<uint8_t> layer0) {
    std::vector<std::vector<uint8_t>> layers(loaded.srs.layers.begin(), loaded.srs.layers.end());
    layers[0] = std::move(layer0);
    std::vector<const std::vector<uint8_t>*> ptrs;
    for (const auto& l : layers) ptrs.push_back(&l);
    try {
      (void)gpu.prepare_prover_param_compressed(nv, ptrs);
    } catch (const BackendError&) {
      return true;
    }
    return false;
  };
  {
    auto l = loaded.srs.layers[0];
    l[48 * 1] &= 0x7f;
    check(refused(l), "missing compression flag refused");
  }
  {
    auto l = loaded.srs.layers[0];
    std::fill(l.begin() + 48 * 2 + 1, l.begin() + 48 * 3, 0xff);
    l[48 * 2] = 0x9f;  // x >= q
    check(refused(l), "x >= q refused");
  }
  {
    auto l = loaded.srs.layers[0];
    bool placed = false;
    for (uint8_t v = 1; v < 200 && !placed; ++v) {
      l[48 * 3 + 47] = v;
      const Fq x = srs_detail::get_be(l.data() + 48 * 3);
      Fq y;
      placed = !srs_detail::fq_sqrt(x.square() * x + Fq::from_u64(4), &y);
    }
    check(placed && refused(l), "x off the curve refused");
  }
  {
    auto l = loaded.srs.layers[0];
    l[48 * 4 + 5] = 1;
    l[48 * 4] = 0xC0;  // infinity with stray bits
    check(refused(l), "bad infinity encoding refused");
  }
  {
    auto l = loaded.srs.layers[0];
    std::fill(l.begin() + 48 * 5, l.begin() + 48 * 6, 0);
    l[48 * 5] = 0xC0;
    l[48 * 6] ^= 0x20;
    std::vector<std::vector<uint8_t>> layers(loaded.srs.layers.begin(), loaded.srs.layers.end());
    layers[0] = l;
    std::vector<const std::vector<uint8_t>*> ptrs;
    for (const auto& x : layers) ptrs.push_back(&x);
    const auto pp = gpu.prepare_prover_param_compressed(nv, ptrs);
    const auto got = gpu.download_prover_param(pp, false);
    check(got[0][97 * 5] == 1, "infinity decompresses to the identity");
    G1Packet p6;
    std::copy(packets[0].begin() + 97 * 6, packets[0].begin() + 97 * 7, p6.begin());
    const G1Packet n6 = g1_to_packet(to_aff(jac_neg(to_jac(g1_from_packet(p6)))));
    check(std::equal(n6.begin(), n6.end(), got[0].begin() + 97 * 6), "flipped sign flag gives -P");
  }
  std::printf("  srs generation %.1f ms load %.1f ms, %zu bytes\n", made.generation_ms, loaded.load_ms,
              serialize_srs(loaded.srs).size());
  std::remove(path.c_str());
  std::printf("SRS_GPU %s\n", failures ? "FAIL" : "PASS");
  return failures ? 1 : 0;
}

int main(int argc, char** argv) {
  if (argc == 4 && std::string(argv[1]) == "srs")
    return srs_case(std::strtoul(argv[2], nullptr, 10), argv[3]);
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
