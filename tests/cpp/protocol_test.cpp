// Host protocol layer checks (CPU only; the CPU oracle is the backend):
//   kat      field, transcript, serialization and pairing known answers and
//            algebraic laws; prints SHA3-256 digests that the Python test
//            recomputes with hashlib
//   prove    HyperPlonk prove + verify of a seeded mock circuit over the oracle
//            backend, then verification of tampered proofs
//   cpu-prove  timed proofs on the oracle backend (CPU baseline)
//   srs      SRS file encoding: compressed G1/G2 known answers, square roots
//            against Euler's criterion, a small universal SRS written, read
//            back and decompressed point by point, malformed files rejected
// Usage: protocol_test kat | prove <nv> <vanilla|jellyfish> <seed> |
//        cpu-prove <nv> <vanilla|jellyfish> <reps> <threads> | srs <dir>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../hyperplonk_amd/host/circuit.hpp"
#include "../../hyperplonk_amd/host/srs_file.hpp"
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
    // random strings of lengths around the chunk sizes (32 bytes for Fr, 48
    // for Fq): "modorderx <field> <input hex, byte order as given> <result>"
    SplitMix64 brng(11);
    const size_t lens[] = {0, 1, 7, 31, 32, 33, 47, 48, 49, 64, 65, 96, 97, 130};
    for (size_t len : lens) {
      uint8_t in[130];
      for (size_t i = 0; i < len; ++i) in[i] = static_cast<uint8_t>(brng.next() >> 56);
      if (len == 96) std::memset(in, 0xff, len);
      std::string hx;
      for (size_t i = 0; i < len; ++i) {
        char h[3];
        std::snprintf(h, 3, "%02x", in[i]);
        hx += h;
      }
      if (hx.empty()) hx = "-";
      std::printf("modorderx r %s %s\n", hx.c_str(), Fr::from_le_bytes_mod_order(in, len).hex().c_str());
      std::printf("modorderx q %s %s\n", hx.c_str(), Fq::from_le_bytes_mod_order(in, len).hex().c_str());
    }
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

template <class F>
static bool throws(F&& f) {
  try {
    f();
  } catch (const SrsFileError&) {
    return true;
  }
  return false;
}

static void srs_checks(const std::string& dir) {
  const G1Aff g1 = g1_generator();
  const G2Aff g2 = g2_generator();
  // compressed generators (the encodings the BLS12-381 serialization
  // standard publishes for them)
  const auto c1 = g1_compressed(g1_to_packet(g1));
  const auto c2 = g2_compressed(g2);
  CHECK(hex(c1.data(), 48) ==
        "97f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb");
  CHECK(hex(c2.data(), 96) ==
        "93e02b6052719f607dacd3a088274f65596bd0d09920b61ab5da61bbdc7f5049334cf11213945d57e5ac7d055d042b7e"
        "024aa2b2f08f0a91260805272dc51051c6e47ad4fa403b02b4510b647ae3d1770bac0326a805bbefd48056c8c121bdb8");
  std::printf("g2gen %s\n", hex(c2.data(), 96).c_str());
  auto same1 = [](const G1Aff& a, const G1Aff& b) {
    return a.infinity == b.infinity && (a.infinity || (a.x == b.x && a.y == b.y));
  };
  auto same2 = [](const G2Aff& a, const G2Aff& b) {
    return a.infinity == b.infinity && (a.infinity || (a.x == b.x && a.y == b.y));
  };
  CHECK(same1(g1_decompress(c1.data()), g1));
  CHECK(same2(g2_decompress(c2.data()), g2));
  // round trips of random multiples and their negations; the sign flag is
  // the only difference between P and -P
  SplitMix64 rng(99);
  for (int i = 0; i < 8; ++i) {
    const Fr k = rng.field();
    const G1Aff p = to_aff(jac_mul(to_jac(g1), k.canonical()));
    const G1Aff n = to_aff(jac_neg(to_jac(p)));
    const auto cp = g1_compressed(g1_to_packet(p)), cn = g1_compressed(g1_to_packet(n));
    CHECK(same1(g1_decompress(cp.data()), p));
    CHECK(same1(g1_decompress(cn.data()), n));
    CHECK((cp[0] ^ cn[0]) == 0x20);
    CHECK(std::equal(cp.begin() + 1, cp.end(), cn.begin() + 1));
    const G2Aff q = to_aff(jac_mul(to_jac(g2), k.canonical()));
    const G2Aff qn = to_aff(jac_neg(to_jac(q)));
    const auto cq = g2_compressed(q), cqn = g2_compressed(qn);
    CHECK(same2(g2_decompress(cq.data()), q));
    CHECK(same2(g2_decompress(cqn.data()), qn));
    CHECK((cq[0] ^ cqn[0]) == 0x20);
  }
  // square roots against Euler's criterion: a in Fq2 is a square iff its
  // norm c0^2 + c1^2 is a square in Fq
  const Limbs<6> half = srs_detail::shifted_q(-1, 1);
  int squares = 0, non_squares = 0;
  for (int i = 0; i < 24; ++i) {
    uint8_t b[96];
    for (auto& v : b) v = static_cast<uint8_t>(rng.next());
    const Fq2 a{Fq::from_le_bytes_mod_order(b, 48), Fq::from_le_bytes_mod_order(b + 48, 48)};
    const bool euler = (a.c0.square() + a.c1.square()).pow(half) == Fq::one();
    Fq2 r;
    const bool found = srs_detail::fq2_sqrt(a, &r);
    CHECK(found == euler);
    if (found) CHECK(r.square() == a);
    const Fq x = a.c0;
    Fq s;
    const bool fq_found = srs_detail::fq_sqrt(x, &s);
    CHECK(fq_found == (x.pow(half) == Fq::one()));
    if (fq_found) CHECK(s.square() == x);
    (found ? squares : non_squares)++;
  }
  CHECK(squares > 0 && non_squares > 0);
  // an x off the curve: x^3 + 4 not a square
  std::array<uint8_t, 48> off = c1;
  for (uint8_t v = 1;; ++v) {
    off[47] = v;
    const Fq x = srs_detail::get_be(off.data());
    Fq y;
    if (!srs_detail::fq_sqrt(x.square() * x + Fq::from_u64(4), &y)) break;
  }
  CHECK(throws([&] { g1_decompress(off.data()); }));
  std::array<uint8_t, 48> unflagged = c1;
  unflagged[0] &= 0x7f;
  CHECK(throws([&] { g1_decompress(unflagged.data()); }));
  std::array<uint8_t, 48> inf{};
  inf[0] = 0xC0;
  CHECK(g1_decompress(inf.data()).infinity);
  inf[5] = 1;
  CHECK(throws([&] { g1_decompress(inf.data()); }));

  // a small universal SRS: layer i = { g * eq(t[i..], x) }, last layer [g]
  const size_t nv = 3;
  std::vector<Fr> t;
  for (size_t i = 0; i < nv; ++i) t.push_back(rng.field());
  SrsFile s;
  s.num_vars = nv;
  std::vector<std::vector<G1Aff>> points;
  for (size_t i = 0; i <= nv; ++i) {
    const std::vector<Fr> ti(t.begin() + i, t.end());
    std::vector<G1Aff> layer;
    std::vector<uint8_t> bytes;
    for (size_t x = 0; x < (size_t(1) << (nv - i)); ++x) {
      std::vector<Fr> bits;
      for (size_t k = 0; k < nv - i; ++k) bits.push_back((x >> k) & 1 ? Fr::one() : Fr::zero());
      const Fr e = ti.empty() ? Fr::one() : eq_eval(bits, ti);
      layer.push_back(to_aff(jac_mul(to_jac(g1), e.canonical())));
      const auto c = g1_compressed(g1_to_packet(layer.back()));
      bytes.insert(bytes.end(), c.begin(), c.end());
    }
    points.push_back(layer);
    s.layers.push_back(bytes);
  }
  const MultilinearVerifierParam vp = MultilinearVerifierParam::from_trapdoor(t);
  s.g = vp.g;
  s.h = vp.h;
  s.h_mask = vp.h_mask;
  const std::vector<uint8_t> bytes = serialize_srs(s);
  // 8 + 8 + sum(8 + 48 n) + 48 + 96 + 8 + 96 nv
  CHECK(bytes.size() == 16 + (8 * 4 + 48 * 15) + 48 + 96 + 8 + 96 * nv);
  std::printf("srs_nv3 %zu %s\n", bytes.size(), hex(sha3_256(bytes.data(), bytes.size()).data(), 8).c_str());
  const SrsFile r = deserialize_srs(bytes);
  CHECK(r.num_vars == nv && r.layers == s.layers && same1(r.g, s.g) && same2(r.h, s.h));
  for (size_t i = 0; i < nv; ++i) CHECK(same2(r.h_mask[i], s.h_mask[i]));
  for (size_t i = 0; i <= nv; ++i)
    for (size_t x = 0; x < points[i].size(); ++x)
      CHECK(same1(g1_decompress(r.layers[i].data() + 48 * x), points[i][x]));
  CHECK(same1(points[nv][0], g1));
  // the trimmed verifier key keeps the last masks; prover layers start at N - nv
  CHECK(r.verifier_param(2).h_mask.size() == 2 && same2(r.verifier_param(2).h_mask[0], s.h_mask[1]));
  CHECK(r.prover_layers(2).size() == 3 && r.prover_layers(2)[0] == &r.layers[1]);
  // file round trip; no overwrite
  const std::string path = dir + "/srs_nv3.bin";
  std::remove(path.c_str());
  write_srs_file(path, s);
  CHECK(throws([&] { write_srs_file(path, s); }));
  CHECK(serialize_srs(read_srs_file(path)) == bytes);
  std::remove(path.c_str());
  // malformed files
  auto bad = [&](std::vector<uint8_t> b) { return throws([&] { deserialize_srs(b); }); };
  CHECK(bad(std::vector<uint8_t>(bytes.begin(), bytes.end() - 1)));
  {
    auto b = bytes;
    b.push_back(0);
    CHECK(bad(b));
  }
  {
    auto b = bytes;
    b[8] = 3;  // layer count != num_vars + 1
    CHECK(bad(b));
  }
  {
    auto b = bytes;
    b[16] = 7;  // first layer length
    CHECK(bad(b));
  }
  {
    auto b = bytes;
    const size_t g_at = 16 + (8 * 4 + 48 * 15);
    b[g_at] &= 0x7f;  // g not flagged compressed
    CHECK(bad(b));
  }
  {
    // first mask with an x whose x^3 + 4 (1 + u) has a non-square norm
    auto b = bytes;
    const size_t mask_at = 16 + (8 * 4 + 48 * 15) + 48 + 96 + 8;
    bool placed = false;
    for (uint8_t v = 1; v < 200 && !placed; ++v) {
      b[mask_at + 95] = v;
      const Fq2 x{srs_detail::get_be(b.data() + mask_at + 48), srs_detail::get_be(b.data() + mask_at)};
      const Fq2 rhs = x.square() * x + Fq2{Fq::from_u64(4), Fq::from_u64(4)};
      placed = (rhs.c0.square() + rhs.c1.square()).pow(half) != Fq::one();
    }
    CHECK(placed);
    CHECK(bad(b));
  }
  std::printf("SRS %s\n", failures ? "FAIL" : "PASS");
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::string(argv[1]) == "kat") {
    kat();
    return failures ? 1 : 0;
  }
  if (argc >= 6 && std::string(argv[1]) == "cpu-prove")
    return cpu_prove(std::strtoul(argv[2], nullptr, 10), argv[3], std::atoi(argv[4]), std::atoi(argv[5]));
  if (argc >= 3 && std::string(argv[1]) == "srs") {
    srs_checks(argv[2]);
    return failures ? 1 : 0;
  }
  if (argc >= 5 && std::string(argv[1]) == "prove")
    return prove_case(std::strtoul(argv[2], nullptr, 10), argv[3], std::strtoull(argv[4], nullptr, 0));
  std::fprintf(stderr, "usage: %s kat | prove <nv> <vanilla|jellyfish> <seed> | srs <dir>\n", argv[0]);
  return 2;
}
