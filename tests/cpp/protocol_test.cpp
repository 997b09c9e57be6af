// Host protocol layer checks (CPU only; the CPU oracle is the backend):
//   kat      field, transcript, serialization and pairing known answers and
//            algebraic laws; prints SHA3-256 digests that the Python test
//            recomputes with hashlib
//   prove    HyperPlonk prove + verify of a seeded mock circuit over the oracle
//            backend, then verification of tampered proofs
//   cpu-prove  timed proofs on the oracle backend (CPU baseline)
// Usage: protocol_test kat | prove <nv> <vanilla|jellyfish> <seed> |
//        cpu-prove <nv> <vanilla|jellyfish> <reps> <threads>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../hyperplonk_amd/host/circuit.hpp"
#include "oracle_backend.hpp"

using namespace hyperplonk;
using namespace hyperplonk::host;

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);      \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static std::string hex(const uint8_t* p, size_t n) {
  std::string s;
  char b[3];
  for (size_t i = 0; i < n; ++i) {
    std::snprintf(b, 3, "%02x", p[i]);
    s += b;
  }
  return s;
}

static void kat() {
  // SHA3-256 over lengths around the 136-byte rate; the test recomputes them
  for (size_t len : {0, 1, 3, 135, 136, 137, 272, 300}) {
    std::vector<uint8_t> m(len);
    for (size_t i = 0; i < len; ++i) m[i] = static_cast<uint8_t>(i * 7 + 3);
    const auto d = sha3_256(m.data(), m.size());
    std::printf("sha3 %zu %s\n", len, hex(d.data(), 32).c_str());
  }
  // Merlin known answer (merlin's simple transcript vector)
  {
    MerlinTranscript t("test protocol");
    t.append_message("some label", reinterpret_cast<const uint8_t*>("some data"), 9);
    uint8_t c[32];
    t.challenge_bytes("challenge", c, 32);
    std::printf("merlin %s\n", hex(c, 32).c_str());
  }
  // IOPTranscript challenge stream (recomputed by the Python test's own STROBE)
  {
    IOPTranscript t("hyperplonk");
    bool threw = false;
    try {
      t.get_and_append_challenge("x");
    } catch (const TranscriptError&) {
      threw = true;
    }
    CHECK(threw);  // an empty transcript refuses challenges
    t.append_field_element("w", Fr::from_u64(5));
    const Fr c0 = t.get_and_append_challenge("beta");
    const auto v = t.get_and_append_challenge_vectors("0check r", 2);
    std::printf("iop %s %s %s\n", c0.hex().c_str(), v[0].hex().c_str(), v[1].hex().c_str());
  }
  // field laws
  SplitMix64 rng(7);
  for (int i = 0; i < 50; ++i) {
    const Fr a = rng.field(), b = rng.field(), c = rng.field();
    CHECK((a + b) * c == a * c + b * c);
    CHECK(a * a.inverse() == Fr::one() || a.is_zero());
    CHECK(a - b + b == a);
  }
  CHECK(Fr::from_i64(-3) + Fr::from_u64(3) == Fr::zero());
  {
    // from_le_bytes_mod_order: 2^256 mod r, and a 64-byte all-ones string
    uint8_t b[64] = {};
    b[32] = 1;
    std::printf("modorder %s\n", Fr::from_le_bytes_mod_order(b, 64).hex().c_str());
    std::memset(b, 0xff, 64);
    std::printf("modorder %s\n", Fr::from_le_bytes_mod_order(b, 64).hex().c_str());
  }
  // curves
  const G1Aff g1 = g1_generator();
  const G2Aff g2 = g2_generator();
  CHECK(g1_on_curve(g1));
  CHECK(g2_on_curve(g2));
  const Limbs<4> r = FrSpec::MOD;
  CHECK(jac_mul(to_jac(g1), r).is_infinity());
  CHECK(jac_mul(to_jac(g2), r).is_infinity());
  const auto gc = g1_compressed(g1_to_packet(g1));
  std::printf("g1gen %s\n", hex(gc.data(), 48).c_str());
  const auto gn = g1_compressed(g1_to_packet(to_aff(jac_neg(to_jac(g1)))));
  std::printf("g1neg %s\n", hex(gn.data(), 48).c_str());
  G1Packet inf{};
  inf[0] = 1;
  std::printf("g1inf %s\n", hex(g1_compressed(inf).data(), 48).c_str());
  // pairing: non-degenerate, order r, bilinear
  const Fq12 e = pairing(g1, g2);
  CHECK(e != Fq12::one());
  CHECK(e.pow(r) == Fq12::one());
  const Fr a = rng.field(), b = rng.field();
  const G1Aff pa = to_aff(jac_mul(to_jac(g1), a.canonical()));
  const G2Aff qb = to_aff(jac_mul(to_jac(g2), b.canonical()));
  CHECK(pairing(pa, qb) == e.pow((a * b).canonical()));
  CHECK(pairing(pa, g2) == pairing(g1, to_aff(jac_mul(to_jac(g2), a.canonical()))));
  const G1Aff neg = to_aff(jac_neg(to_jac(pa)));
  CHECK(final_exponentiation(multi_miller_loop({{pa, qb}, {neg, qb}})) == Fq12::one());
  CHECK(final_exponentiation(multi_miller_loop({{pa, qb}, {pa, qb}})) != Fq12::one());
  // eq / interpolation
  {
    const std::vector<Fr> x{rng.field(), rng.field(), rng.field()};
    const auto tab = eq_table(x);
    for (size_t i = 0; i < 8; ++i) {
      std::vector<Fr> bits;
      for (int k = 0; k < 3; ++k) bits.push_back((i >> k) & 1 ? Fr::one() : Fr::zero());
      CHECK(tab[i] == eq_eval(bits, x));
    }
    // a cubic through 4 points, evaluated elsewhere and at a node
    auto cubic = [](const Fr& t) { return t * t * t * Fr::from_u64(3) + t * Fr::from_u64(5) + Fr::from_u64(11); };
    std::vector<Fr> p;
    for (int i = 0; i < 4; ++i) p.push_back(cubic(Fr::from_u64(i)));
    const Fr z = rng.field();
    CHECK(interpolate_uni_poly(p, z) == cubic(z));
    CHECK(interpolate_uni_poly(p, Fr::from_u64(2)) == p[2]);
  }
  std::printf("KAT %s\n", failures ? "FAIL" : "PASS");
}

static int prove_case(size_t nv, const std::string& gate, uint64_t seed) {
  const MockCircuit c = gate == "vanilla" ? vanilla_circuit(nv, seed) : jellyfish_circuit(nv, seed);
  CHECK(is_satisfied(c));
  SplitMix64 trng(seed ^ 0x5352535f54415553ULL);
  std::vector<Fr> t;
  for (size_t i = 0; i < nv; ++i) t.push_back(trng.field());
  oracle_cpu::OracleBackend be(8);
  const auto pp = be.generate_prover_param(nv, to_wire(t));
  const auto full_vk = MultilinearVerifierParam::from_trapdoor(t);
  auto [pk, vk] = preprocess(be, c.index, pp, full_vk);
  const HyperPlonkProof proof = prove(be, pk, c.public_inputs, c.witnesses);
  const std::vector<uint8_t> bytes = serialize_proof(proof);
  const auto d = sha3_256(bytes.data(), bytes.size());
  std::printf("proof %zu bytes sha3 %s\n", bytes.size(), hex(d.data(), 32).c_str());
  CHECK(proof.batch_openings.f_i_eval_at_point_i.size() == 3 * c.witnesses.size() + c.index.selectors.size() + 8);
  CHECK(verify(vk, c.public_inputs, proof));
  // a second prove is deterministic
  CHECK(serialize_proof(prove(be, pk, c.public_inputs, c.witnesses)) == bytes);
  // tampering: an evaluation, a round message, an opening proof, a public input
  auto rejected = [&](const HyperPlonkProof& p, const std::vector<Fr>& pub) {
    try {
      return !verify(vk, pub, p);
    } catch (const std::exception&) {
      return true;
    }
  };
  {
    HyperPlonkProof p = proof;
    p.batch_openings.f_i_eval_at_point_i[5] += Fr::one();
    CHECK(rejected(p, c.public_inputs));
  }
  {
    HyperPlonkProof p = proof;
    p.zero_check_proof.proofs[1][2] += Fr::one();
    CHECK(rejected(p, c.public_inputs));
  }
  {
    HyperPlonkProof p = proof;
    std::swap(p.batch_openings.g_prime_proof[0], p.batch_openings.g_prime_proof[1]);
    CHECK(rejected(p, c.public_inputs));
  }
  {
    std::vector<Fr> pub = c.public_inputs;
    pub[0] += Fr::one();
    CHECK(rejected(proof, pub));
  }
  {
    // a witness that breaks one gate row cannot produce an accepted proof
    std::vector<Column> w = c.witnesses;
    w[1][3] = (from_wire(w[1][3]) + Fr::one()).bytes();
    bool ok_proof = true;
    try {
      ok_proof = verify(vk, c.public_inputs, prove(be, pk, c.public_inputs, w));
    } catch (const std::exception&) {
      ok_proof = false;
    }
    CHECK(!ok_proof);
  }
  std::printf("PROVE %s\n", failures ? "FAIL" : "PASS");
  return failures ? 1 : 0;
}

// Timed prove on the oracle backend: the CPU baseline of the proof benchmark
static int cpu_prove(size_t nv, const std::string& gate, int reps, int threads) {
  const MockCircuit c = gate == "vanilla" ? vanilla_circuit(nv, 1) : jellyfish_circuit(nv, 2);
  SplitMix64 trng(0x5352535f54415553ULL);
  std::vector<Fr> t;
  for (size_t i = 0; i < nv; ++i) t.push_back(trng.field());
  oracle_cpu::OracleBackend be(threads);
  const auto pp = be.generate_prover_param(nv, to_wire(t));
  auto [pk, vk] = preprocess(be, c.index, pp, MultilinearVerifierParam::from_trapdoor(t));
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < reps; ++i) (void)prove(be, pk, c.public_inputs, c.witnesses);
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("{\"nv\": %zu, \"gate\": \"%s\", \"reps\": %d, \"threads\": %d, \"proofs_per_s\": %.4f, "
              "\"ms_per_proof\": %.2f}\n",
              nv, gate.c_str(), reps, threads, reps / s, 1e3 * s / reps);
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::string(argv[1]) == "kat") {
    kat();
    return failures ? 1 : 0;
  }
  if (argc >= 6 && std::string(argv[1]) == "cpu-prove")
    return cpu_prove(std::strtoul(argv[2], nullptr, 10), argv[3], std::atoi(argv[4]), std::atoi(argv[5]));
  if (argc >= 5 && std::string(argv[1]) == "prove")
    return prove_case(std::strtoul(argv[2], nullptr, 10), argv[3], std::strtoull(argv[4], nullptr, 0));
  std::fprintf(stderr, "usage: %s kat | prove <nv> <vanilla|jellyfish> <seed>\n", argv[0]);
  return 2;
}
