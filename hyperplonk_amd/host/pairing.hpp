// Host verifier arithmetic: BLS12-381 G1/G2 and the optimal ate pairing, used
// by the multilinear KZG verifier (subroutines/src/pcs/multilinear_kzg/mod.rs:
// 169-221) and the batch verifier (batching.rs:196-292).
//
// Tower: Fq2 = Fq[u]/(u^2 + 1), Fq6 = Fq2[v]/(v^3 - xi) with xi = 1 + u,
// Fq12 = Fq6[w]/(w^2 - v).  G2 lives on the M-type sextic twist
// y^2 = x^3 + 4 xi, mapped into E(Fq12) by (x, y) -> (x / w^2, y / w^3).
// Miller loop over |x| = 0xd201000000010000 (x < 0: the result is conjugated),
// with the twist point T kept affine over Fq2.  A line through T with twist
// slope L, evaluated at P = (xP, yP) and scaled by w^3 (a factor the final
// exponentiation removes), is (L xT - yT) - L xP v + yP v w.
// Final exponentiation f^((q^12 - 1) / r) = ((f^(q^6 - 1))^(q^2 + 1))^((q^4 - q^2 + 1) / r),
// the first factor by conjugation and inversion, the other two by plain
// square-and-multiply over their (precomputed) exponents.
#pragma once
#include <vector>

#include "ff.hpp"
#include "pcs.hpp"

namespace hyperplonk {
namespace host {

// ------------------------------------------------------------------ towers
struct Fq2 {
  Fq c0, c1;
  static Fq2 zero() { return {}; }
  static Fq2 one() { return {Fq::one(), Fq::zero()}; }
  bool is_zero() const { return c0.is_zero() && c1.is_zero(); }
  bool operator==(const Fq2& o) const { return c0 == o.c0 && c1 == o.c1; }
  bool operator!=(const Fq2& o) const { return !(*this == o); }
  Fq2 operator+(const Fq2& o) const { return {c0 + o.c0, c1 + o.c1}; }
  Fq2 operator-(const Fq2& o) const { return {c0 - o.c0, c1 - o.c1}; }
  Fq2 operator-() const { return {-c0, -c1}; }
  Fq2 operator*(const Fq2& o) const {
    return {c0 * o.c0 - c1 * o.c1, c0 * o.c1 + c1 * o.c0};
  }
  Fq2 scale(const Fq& s) const { return {c0 * s, c1 * s}; }
  Fq2 square() const { return *this * *this; }
  Fq2 mul_by_xi() const { return {c0 - c1, c0 + c1}; }
  Fq2 inverse() const {
    const Fq t = (c0.square() + c1.square()).inverse();
    return {c0 * t, -(c1 * t)};
  }
};

struct Fq6 {
  Fq2 c0, c1, c2;
  static Fq6 zero() { return {}; }
  static Fq6 one() { return {Fq2::one(), Fq2::zero(), Fq2::zero()}; }
  bool operator==(const Fq6& o) const { return c0 == o.c0 && c1 == o.c1 && c2 == o.c2; }
  Fq6 operator+(const Fq6& o) const { return {c0 + o.c0, c1 + o.c1, c2 + o.c2}; }
  Fq6 operator-(const Fq6& o) const { return {c0 - o.c0, c1 - o.c1, c2 - o.c2}; }
  Fq6 operator-() const { return {-c0, -c1, -c2}; }
  Fq6 operator*(const Fq6& o) const {
    return {c0 * o.c0 + (c1 * o.c2 + c2 * o.c1).mul_by_xi(),
            c0 * o.c1 + c1 * o.c0 + (c2 * o.c2).mul_by_xi(),
            c0 * o.c2 + c1 * o.c1 + c2 * o.c0};
  }
  Fq6 mul_by_v() const { return {c2.mul_by_xi(), c0, c1}; }
  Fq6 inverse() const {
    const Fq2 a = c0.square() - (c1 * c2).mul_by_xi();
    const Fq2 b = c2.square().mul_by_xi() - c0 * c1;
    const Fq2 c = c1.square() - c0 * c2;
    const Fq2 t = (c0 * a + (c2 * b + c1 * c).mul_by_xi()).inverse();
    return {a * t, b * t, c * t};
  }
};

struct Fq12 {
  Fq6 c0, c1;
  static Fq12 one() { return {Fq6::one(), Fq6::zero()}; }
  bool operator==(const Fq12& o) const { return c0 == o.c0 && c1 == o.c1; }
  bool operator!=(const Fq12& o) const { return !(*this == o); }
  Fq12 operator*(const Fq12& o) const {
    return {c0 * o.c0 + (c1 * o.c1).mul_by_v(), c0 * o.c1 + c1 * o.c0};
  }
  Fq12 square() const { return *this * *this; }
  Fq12 conjugate() const { return {c0, -c1}; }
  Fq12 inverse() const {
    const Fq6 t = (c0 * c0 - (c1 * c1).mul_by_v()).inverse();
    return {c0 * t, -(c1 * t)};
  }
  template <size_t M>
  Fq12 pow(const Limbs<M>& e) const {
    Fq12 r = one();
    bool started = false;
    for (size_t i = M; i-- > 0;)
      for (int b = 63; b >= 0; --b) {
        if (started) r = r.square();
        if ((e[i] >> b) & 1) {
          r = started ? r * *this : *this;
          started = true;
        }
      }
    return r;
  }
};

// ------------------------------------------------------------------ curves
// Jacobian points over a field F with y^2 = x^3 + b; infinity has Z = 0.
template <class F>
struct Jac {
  F x, y, z;
  static Jac infinity() { return {F::one(), F::one(), F::zero()}; }
  bool is_infinity() const { return z.is_zero(); }
};

template <class F>
struct Aff {
  F x, y;
  bool infinity = true;
};

template <class F>
Jac<F> jac_dbl(const Jac<F>& p) {
  if (p.is_infinity()) return p;
  const F a = p.x.square(), b = p.y.square(), c = b.square();
  const F d = ((p.x + b).square() - a - c) + ((p.x + b).square() - a - c);
  const F e = a + a + a;
  const F f = e.square();
  Jac<F> r;
  r.x = f - d - d;
  const F c8 = c + c + c + c + c + c + c + c;
  r.y = e * (d - r.x) - c8;
  r.z = (p.y * p.z) + (p.y * p.z);
  return r;
}

template <class F>
Jac<F> jac_add(const Jac<F>& p, const Jac<F>& q) {
  if (p.is_infinity()) return q;
  if (q.is_infinity()) return p;
  const F z1z1 = p.z.square(), z2z2 = q.z.square();
  const F u1 = p.x * z2z2, u2 = q.x * z1z1;
  const F s1 = p.y * q.z * z2z2, s2 = q.y * p.z * z1z1;
  if (u1 == u2) return s1 == s2 ? jac_dbl(p) : Jac<F>::infinity();
  const F h = u2 - u1, rr = s2 - s1;
  const F hh = h.square(), hhh = h * hh, v = u1 * hh;
  Jac<F> out;
  out.x = rr.square() - hhh - v - v;
  out.y = rr * (v - out.x) - s1 * hhh;
  out.z = p.z * q.z * h;
  return out;
}

template <class F>
Jac<F> jac_neg(const Jac<F>& p) {
  return {p.x, -p.y, p.z};
}

template <class F>
Jac<F> to_jac(const Aff<F>& a) {
  return a.infinity ? Jac<F>::infinity() : Jac<F>{a.x, a.y, F::one()};
}

template <class F>
Aff<F> to_aff(const Jac<F>& p) {
  if (p.is_infinity()) return {};
  const F zi = p.z.inverse(), zi2 = zi.square();
  return {p.x * zi2, p.y * zi2 * zi, false};
}

// scalar: canonical little-endian limbs
template <class F>
Jac<F> jac_mul(const Jac<F>& p, const Limbs<4>& k) {
  Jac<F> r = Jac<F>::infinity();
  for (size_t i = 4; i-- > 0;)
    for (int b = 63; b >= 0; --b) {
      r = jac_dbl(r);
      if ((k[i] >> b) & 1) r = jac_add(r, p);
    }
  return r;
}

template <class F>
bool jac_eq(const Jac<F>& p, const Jac<F>& q) {
  if (p.is_infinity() || q.is_infinity()) return p.is_infinity() && q.is_infinity();
  const F z1z1 = p.z.square(), z2z2 = q.z.square();
  return p.x * z2z2 == q.x * z1z1 && p.y * q.z * z2z2 == q.y * p.z * z1z1;
}

using G1Jac = Jac<Fq>;
using G1Aff = Aff<Fq>;
using G2Jac = Jac<Fq2>;
using G2Aff = Aff<Fq2>;

inline Fq fq_hex(const char* h) {
  Limbs<6> c{};
  const size_t n = std::strlen(h);
  for (size_t i = 0; i < n; ++i) {
    const char ch = h[n - 1 - i];
    const u64 d = ch <= '9' ? ch - '0' : (ch | 32) - 'a' + 10;
    c[i / 16] |= d << (4 * (i % 16));
  }
  return Fq::from_canonical(c);
}

inline G1Aff g1_generator() {
  return {fq_hex("17f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb"),
          fq_hex("08b3f481e3aaa0f1a09e30ed741d8ae4fcf5e095d5d00af600db18cb2c04b3edd03cc744a2888ae40caa232946c5e7e1"),
          false};
}

inline G2Aff g2_generator() {
  return {{fq_hex("024aa2b2f08f0a91260805272dc51051c6e47ad4fa403b02b4510b647ae3d1770bac0326a805bbefd48056c8c121bdb8"),
           fq_hex("13e02b6052719f607dacd3a088274f65596bd0d09920b61ab5da61bbdc7f5049334cf11213945d57e5ac7d055d042b7e")},
          {fq_hex("0ce5d527727d6e118cc9cdc6da2e351aadfd9baa8cbdd3a76d429a695160d12c923ac9cc3baca289e193548608b82801"),
           fq_hex("0606c4a02ea734cc32acd2b02bc28b99cb3e287e85a763af267492ab572e99ab3f370d275cec1da1aaa9075ff05f79be")},
          false};
}

inline bool g1_on_curve(const G1Aff& p) {
  return p.infinity || p.y.square() == p.x.square() * p.x + Fq::from_u64(4);
}
inline bool g2_on_curve(const G2Aff& p) {
  const Fq2 b = Fq2{Fq::from_u64(4), Fq::from_u64(4)};  // 4 (1 + u)
  return p.infinity || p.y.square() == p.x.square() * p.x + b;
}

// 97-byte packet <-> affine G1
inline G1Aff g1_from_packet(const G1Packet& p) {
  if (p[0] == 1) return {};
  G1Aff a;
  a.infinity = false;
  if (!Fq::from_bytes(p.data() + 1, &a.x) || !Fq::from_bytes(p.data() + 49, &a.y))
    throw std::invalid_argument("non-canonical G1 packet");
  return a;
}
inline G1Packet g1_to_packet(const G1Aff& a) {
  G1Packet p{};
  if (a.infinity) {
    p[0] = 1;
    return p;
  }
  a.x.to_bytes(p.data() + 1);
  a.y.to_bytes(p.data() + 49);
  return p;
}

// ----------------------------------------------------------------- pairing
namespace detail {
constexpr u64 kAteLoop = 0xd201000000010000ULL;
constexpr Limbs<12> kExpQ2Plus1 = {0x26aa00001c718e3aULL, 0x7ced6b1d76382eabULL, 0x162c338362113cfdULL,
                                   0x66bf91ed3e71b743ULL, 0x292e85a87091a049ULL, 0x1d68619c86185c7bULL,
                                   0xf53149330978ef01ULL, 0x50a62cfd16ddca6eULL, 0x66e59e49349e8bd0ULL,
                                   0xe2dc90e50e7046b4ULL, 0x4bd278eaa22f25e9ULL, 0x02a437a4b8c35fc7ULL};
constexpr Limbs<20> kExpHard = {0xe516c3f438e3ba79ULL, 0xfa9912aae208ccf1ULL, 0x905ce937335d5b68ULL,
                                0xc71a2629b0dea236ULL, 0x83774940996754c8ULL, 0x21d160aeb6a1e799ULL,
                                0x2ed0b283ed237db4ULL, 0x915c97f36c6f1821ULL, 0x67f17fcbde783765ULL,
                                0x2378b9039096d1b7ULL, 0x7988f8761bdc51dcULL, 0x2076995003fc77a1ULL,
                                0x827eca0ba621315bULL, 0xe5a72bce8d63cb9fULL, 0xf68f7764c28b6f8aULL,
                                0x2f230063cf081517ULL, 0x94506632528d6a9aULL, 0xd3cde88eeb996ca3ULL,
                                0xc0bd38c3195c899eULL, 0x000f686b3d807d01ULL};

inline Fq12 line(const Fq2& slope, const Fq2& xt, const Fq2& yt, const G1Aff& p) {
  Fq12 l;
  l.c0 = {slope * xt - yt, -slope.scale(p.x), Fq2::zero()};
  l.c1 = {Fq2::zero(), Fq2{p.y, Fq::zero()}, Fq2::zero()};
  return l;
}
}  // namespace detail

// product of Miller loops of the pairs, before the final exponentiation
inline Fq12 multi_miller_loop(const std::vector<std::pair<G1Aff, G2Aff>>& pairs) {
  struct Item {
    G1Aff p;
    G2Aff q;
    Fq2 tx, ty;
  };
  std::vector<Item> items;
  for (const auto& [p, q] : pairs)
    if (!p.infinity && !q.infinity) items.push_back({p, q, q.x, q.y});
  Fq12 f = Fq12::one();
  const Fq2 three{Fq::from_u64(3), Fq::zero()};
  for (int bit = 62; bit >= 0; --bit) {
    f = f.square();
    for (auto& it : items) {
      const Fq2 slope = (it.tx.square() * three) * (it.ty + it.ty).inverse();
      f = f * detail::line(slope, it.tx, it.ty, it.p);
      const Fq2 x3 = slope.square() - it.tx - it.tx;
      it.ty = slope * (it.tx - x3) - it.ty;
      it.tx = x3;
    }
    if ((detail::kAteLoop >> bit) & 1) {
      for (auto& it : items) {
        const Fq2 slope = (it.q.y - it.ty) * (it.q.x - it.tx).inverse();
        f = f * detail::line(slope, it.tx, it.ty, it.p);
        const Fq2 x3 = slope.square() - it.tx - it.q.x;
        it.ty = slope * (it.tx - x3) - it.ty;
        it.tx = x3;
      }
    }
  }
  return f.conjugate();
}

inline Fq12 final_exponentiation(const Fq12& f) {
  const Fq12 easy = f.conjugate() * f.inverse();
  return easy.pow(detail::kExpQ2Plus1).pow(detail::kExpHard);
}

inline Fq12 pairing(const G1Aff& p, const G2Aff& q) {
  return final_exponentiation(multi_miller_loop({{p, q}}));
}

// ----------------------------------------------------------- KZG verifier
// MultilinearVerifierParam (srs.rs:103-114): h_mask[i] = h * t_i
struct MultilinearVerifierParam {
  size_t num_vars = 0;
  G1Aff g;
  G2Aff h;
  std::vector<G2Aff> h_mask;

  static MultilinearVerifierParam from_trapdoor(const std::vector<Fr>& t) {
    MultilinearVerifierParam vp;
    vp.num_vars = t.size();
    vp.g = g1_generator();
    vp.h = g2_generator();
    for (const Fr& ti : t) vp.h_mask.push_back(to_aff(jac_mul(to_jac(vp.h), ti.canonical())));
    return vp;
  }
  // srs.rs:126-153: keep the last nv masks
  MultilinearVerifierParam trim(size_t nv) const {
    if (nv > num_vars) throw PCSError("SRS does not support target number of vars");
    MultilinearVerifierParam out = *this;
    out.num_vars = nv;
    out.h_mask.assign(h_mask.begin() + (num_vars - nv), h_mask.end());
    return out;
  }
};

// verify_internal (multilinear_kzg/mod.rs:169-221)
inline bool kzg_verify(const MultilinearVerifierParam& vp, const G1Packet& commitment,
                       const std::vector<Fr>& point, const Fr& value,
                       const std::vector<G1Packet>& proofs) {
  const size_t nv = point.size();
  if (nv > vp.num_vars) throw PCSError("point length exceeds param limit");
  const size_t ignored = vp.num_vars - nv;
  std::vector<std::pair<G1Aff, G2Aff>> pairs;
  const G2Jac h = to_jac(vp.h);
  for (size_t i = 0; i < nv && i < proofs.size(); ++i) {
    const G2Jac hv = jac_add(to_jac(vp.h_mask[ignored + i]), jac_neg(jac_mul(h, point[i].canonical())));
    pairs.push_back({g1_from_packet(proofs[i]), to_aff(hv)});
  }
  const G1Jac lhs =
      jac_add(jac_mul(to_jac(vp.g), value.canonical()), jac_neg(to_jac(g1_from_packet(commitment))));
  pairs.push_back({to_aff(lhs), vp.h});
  return final_exponentiation(multi_miller_loop(pairs)) == Fq12::one();
}

// batch_verify_internal (batching.rs:196-292)
inline bool batch_verify(const MultilinearVerifierParam& vp, const std::vector<G1Packet>& comms,
                         const std::vector<std::vector<Fr>>& points, const BatchProof& proof,
                         IOPTranscript& transcript) {
  const size_t k = comms.size();
  if (k == 0 || points.size() != k || proof.f_i_eval_at_point_i.size() != k)
    throw PCSError("batch commitments, points, and evaluations must have equal nonzero lengths");
  const size_t nv = points[0].size();
  for (const auto& p : points)
    if (nv == 0 || p.size() != nv) throw PCSError("batch points must have the same positive dimension");
  for (const auto& p : points) transcript.append_message("eval_point", ser_frs(p));
  for (const auto& e : proof.f_i_eval_at_point_i) transcript.append_field_element("eval", e);
  const std::vector<Fr> t = transcript.get_and_append_challenge_vectors("t", ceil_log2(k));
  const std::vector<Fr> w = batch_weights(t);
  Fr sum = Fr::zero();
  for (size_t i = 0; i < k; ++i) sum += w[i] * proof.f_i_eval_at_point_i[i];
  SumCheckSubClaim sub;
  try {
    sub = sumcheck_verify(sum, proof.sum_check_proof, {2, nv}, transcript);
  } catch (const PolyIOPError&) {
    throw PCSError("Sumcheck in batch verification failed");
  }
  G1Jac acc = G1Jac::infinity();
  for (size_t i = 0; i < k; ++i) {
    const Fr s = eq_eval(sub.point, points[i]) * w[i];
    acc = jac_add(acc, jac_mul(to_jac(g1_from_packet(comms[i])), s.canonical()));
  }
  return kzg_verify(vp, g1_to_packet(to_aff(acc)), sub.point, sub.expected_evaluation,
                    proof.g_prime_proof);
}

}  // namespace host
}  // namespace hyperplonk
