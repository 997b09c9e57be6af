// Host-side G1 arithmetic on 6 x 64-bit limbs for the MSM's last steps.
//
// The device keeps Fq in Montgomery form with R = 2^384 on 12 x 32-bit limbs;
// the same little-endian bytes read as 6 x 64-bit limbs are the same
// Montgomery representation (R = 2^384 either way), so device XYZZ results
// are used here without conversion.  The host finishes what is inherently
// sequential and short (combining the last bucket-reduction groups, the
// final affine normalisation), where one CPU core with 64 x 64 -> 128-bit
// products beats a single GPU wavefront by an order of magnitude.
#pragma once
#include <cstdint>
#include <cstring>

#include "curve.hpp"

namespace hp {
namespace host {

struct Fq64 {
  uint64_t v[6];
};

constexpr uint64_t kQ[6] = {0xb9feffffffffaaabull, 0x1eabfffeb153ffffull, 0x6730d2a0f6b0f624ull,
                            0x64774b84f38512bfull, 0x4b1ba7b6434bacd7ull, 0x1a0111ea397fe69aull};
constexpr uint64_t kQInv = 0x89f3fffcfffcfffdull;  // -q^{-1} mod 2^64

inline bool is_zero(const Fq64& a) {
  return (a.v[0] | a.v[1] | a.v[2] | a.v[3] | a.v[4] | a.v[5]) == 0;
}

inline void cond_sub_q(uint64_t t[6]) {
  uint64_t d[6];
  unsigned __int128 borrow = 0;
  for (int i = 0; i < 6; ++i) {
    unsigned __int128 x = (unsigned __int128)t[i] - kQ[i] - (uint64_t)borrow;
    d[i] = (uint64_t)x;
    borrow = (x >> 64) & 1;
  }
  if (!borrow) std::memcpy(t, d, sizeof d);
}

inline Fq64 add(const Fq64& a, const Fq64& b) {
  Fq64 r;
  unsigned __int128 c = 0;
  for (int i = 0; i < 6; ++i) {
    c += (unsigned __int128)a.v[i] + b.v[i];
    r.v[i] = (uint64_t)c;
    c >>= 64;
  }
  cond_sub_q(r.v);
  return r;
}

inline Fq64 sub(const Fq64& a, const Fq64& b) {
  Fq64 r;
  unsigned __int128 borrow = 0;
  for (int i = 0; i < 6; ++i) {
    unsigned __int128 x = (unsigned __int128)a.v[i] - b.v[i] - (uint64_t)borrow;
    r.v[i] = (uint64_t)x;
    borrow = (x >> 64) & 1;
  }
  if (borrow) {
    unsigned __int128 c = 0;
    for (int i = 0; i < 6; ++i) {
      c += (unsigned __int128)r.v[i] + kQ[i];
      r.v[i] = (uint64_t)c;
      c >>= 64;
    }
  }
  return r;
}

// CIOS Montgomery product in the "no final carry" form: q's top limb is below
// 2^63 - 1, so every intermediate fits in six words (the carries A and C of
// the product and reduction rows never overflow a word).  Rows written out so
// the compiler keeps t in registers (about half the time of the looped form).
inline Fq64 mul(const Fq64& a, const Fq64& b) {
  using u128 = unsigned __int128;
  uint64_t t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0;
#define HP_CIOS_ROW(i)                                                               \
  {                                                                                  \
    const uint64_t bi = b.v[i];                                                      \
    u128 x = (u128)a.v[0] * bi + t0;                                                 \
    uint64_t A = (uint64_t)(x >> 64);                                                \
    const uint64_t m = (uint64_t)x * kQInv;                                          \
    u128 y = (u128)m * kQ[0] + (uint64_t)x;                                          \
    uint64_t C = (uint64_t)(y >> 64);                                                \
    x = (u128)a.v[1] * bi + t1 + A;                                                  \
    A = (uint64_t)(x >> 64);                                                         \
    y = (u128)m * kQ[1] + (uint64_t)x + C;                                           \
    C = (uint64_t)(y >> 64);                                                         \
    t0 = (uint64_t)y;                                                                \
    x = (u128)a.v[2] * bi + t2 + A;                                                  \
    A = (uint64_t)(x >> 64);                                                         \
    y = (u128)m * kQ[2] + (uint64_t)x + C;                                           \
    C = (uint64_t)(y >> 64);                                                         \
    t1 = (uint64_t)y;                                                                \
    x = (u128)a.v[3] * bi + t3 + A;                                                  \
    A = (uint64_t)(x >> 64);                                                         \
    y = (u128)m * kQ[3] + (uint64_t)x + C;                                           \
    C = (uint64_t)(y >> 64);                                                         \
    t2 = (uint64_t)y;                                                                \
    x = (u128)a.v[4] * bi + t4 + A;                                                  \
    A = (uint64_t)(x >> 64);                                                         \
    y = (u128)m * kQ[4] + (uint64_t)x + C;                                           \
    C = (uint64_t)(y >> 64);                                                         \
    t3 = (uint64_t)y;                                                                \
    x = (u128)a.v[5] * bi + t5 + A;                                                  \
    A = (uint64_t)(x >> 64);                                                         \
    y = (u128)m * kQ[5] + (uint64_t)x + C;                                           \
    C = (uint64_t)(y >> 64);                                                         \
    t4 = (uint64_t)y;                                                                \
    t5 = C + A;                                                                      \
  }
  HP_CIOS_ROW(0) HP_CIOS_ROW(1) HP_CIOS_ROW(2) HP_CIOS_ROW(3) HP_CIOS_ROW(4) HP_CIOS_ROW(5)
#undef HP_CIOS_ROW
  // t < 2q: one conditional subtraction, branch-free
  const uint64_t t[6] = {t0, t1, t2, t3, t4, t5};
  uint64_t d[6], borrow = 0;
  for (int i = 0; i < 6; ++i) {
    const u128 z = (u128)t[i] - kQ[i] - borrow;
    d[i] = (uint64_t)z;
    borrow = (uint64_t)(z >> 64) & 1;
  }
  const uint64_t keep = 0 - borrow;  // borrow: t < q, keep t
  Fq64 r;
  for (int i = 0; i < 6; ++i) r.v[i] = (t[i] & keep) | (d[i] & ~keep);
  return r;
}

namespace detail {
inline bool is_one_int(const uint64_t x[6]) {
  return x[0] == 1 && (x[1] | x[2] | x[3] | x[4] | x[5]) == 0;
}
inline bool geq_int(const uint64_t x[6], const uint64_t y[6]) {
  for (int i = 5; i >= 0; --i)
    if (x[i] != y[i]) return x[i] > y[i];
  return true;
}
inline void sub_int(uint64_t x[6], const uint64_t y[6]) {  // x -= y, x >= y
  unsigned __int128 borrow = 0;
  for (int i = 0; i < 6; ++i) {
    const unsigned __int128 d = (unsigned __int128)x[i] - y[i] - (uint64_t)borrow;
    x[i] = (uint64_t)d;
    borrow = (d >> 64) & 1;
  }
}
inline void half_int(uint64_t x[6]) {  // x /= 2, x even
  for (int i = 0; i < 5; ++i) x[i] = (x[i] >> 1) | (x[i + 1] << 63);
  x[5] >>= 1;
}
// x / 2 mod q for x < q: (x + q) / 2 when x is odd (x + q < 2^382, no overflow)
inline void half_mod(uint64_t x[6]) {
  if (x[0] & 1) {
    unsigned __int128 c = 0;
    for (int i = 0; i < 6; ++i) {
      c += (unsigned __int128)x[i] + kQ[i];
      x[i] = (uint64_t)c;
      c >>= 64;
    }
  }
  half_int(x);
}
inline void sub_mod(uint64_t x[6], const uint64_t y[6]) {  // x - y mod q, both < q
  if (geq_int(x, y)) {
    sub_int(x, y);
  } else {
    uint64_t t[6];
    for (int i = 0; i < 6; ++i) t[i] = kQ[i];
    sub_int(t, y);
    unsigned __int128 c = 0;
    for (int i = 0; i < 6; ++i) {
      c += (unsigned __int128)x[i] + t[i];
      x[i] = (uint64_t)c;
      c >>= 64;
    }
  }
}
}  // namespace detail

// Montgomery-form inverse (zero maps to zero) by the binary extended
// Euclidean algorithm on the integer A = a R: it gives A^-1 = a^-1 R^-1, and one
// product by R^3 mod q returns a^-1 R.  About 10x faster on one core than the
// exponentiation a^(q - 2).
inline Fq64 inverse(const Fq64& a) {
  using namespace detail;
  if (is_zero(a)) return a;
  uint64_t u[6], v[6], x1[6] = {1, 0, 0, 0, 0, 0}, x2[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 6; ++i) {
    u[i] = a.v[i];
    v[i] = kQ[i];
  }
  // invariants: x1 A = u, x2 A = v (mod q); u, v > 0
  while (!is_one_int(u) && !is_one_int(v)) {
    while ((u[0] & 1) == 0) {
      half_int(u);
      half_mod(x1);
    }
    while ((v[0] & 1) == 0) {
      half_int(v);
      half_mod(x2);
    }
    if (geq_int(u, v)) {
      sub_int(u, v);
      sub_mod(x1, x2);
    } else {
      sub_int(v, u);
      sub_mod(x2, x1);
    }
  }
  Fq64 r;
  for (int i = 0; i < 6; ++i) r.v[i] = is_one_int(u) ? x1[i] : x2[i];
  static const Fq64 kR3{{0xed48ac6bd94ca1e0ull, 0x315f831e03a7adf8ull, 0x9a53352a615e29ddull,
                         0x34c04e5e921e1761ull, 0x2512d43565724728ull, 0x0aa6346091755d4dull}};
  return mul(r, kR3);
}

inline Fq64 from_mont(const Fq64& a) {
  Fq64 one{{1, 0, 0, 0, 0, 0}};
  return mul(a, one);
}

struct Xyzz64 {
  Fq64 X, Y, ZZ, ZZZ;
  bool is_identity() const { return is_zero(ZZ); }
};

inline Xyzz64 from_device(const G1Xyzz& p) {
  Xyzz64 r;
  std::memcpy(&r, &p, sizeof r);
  return r;
}

inline G1Xyzz to_device(const Xyzz64& p) {
  G1Xyzz r;
  std::memcpy(&r, &p, sizeof r);
  return r;
}

inline Xyzz64 neg(const Xyzz64& p) { return {p.X, sub(Fq64{}, p.Y), p.ZZ, p.ZZZ}; }

inline Xyzz64 dbl(const Xyzz64& p) {
  if (p.is_identity() || is_zero(p.Y)) return Xyzz64{{}, {}, {}, {}};
  const Fq64 U = add(p.Y, p.Y), V = mul(U, U), W = mul(U, V), S = mul(p.X, V);
  const Fq64 xx = mul(p.X, p.X), M = add(add(xx, xx), xx);
  const Fq64 X3 = sub(sub(mul(M, M), S), S);
  const Fq64 Y3 = sub(mul(M, sub(S, X3)), mul(W, p.Y));
  return {X3, Y3, mul(V, p.ZZ), mul(W, p.ZZZ)};
}

inline Xyzz64 add(const Xyzz64& p, const Xyzz64& q) {
  if (q.is_identity()) return p;
  if (p.is_identity()) return q;
  const Fq64 U1 = mul(p.X, q.ZZ), U2 = mul(q.X, p.ZZ);
  const Fq64 S1 = mul(p.Y, q.ZZZ), S2 = mul(q.Y, p.ZZZ);
  const Fq64 P = sub(U2, U1), R = sub(S2, S1);
  if (is_zero(P)) {
    if (is_zero(R)) return dbl(p);
    return Xyzz64{{}, {}, {}, {}};
  }
  const Fq64 PP = mul(P, P), PPP = mul(P, PP), Q = mul(U1, PP);
  const Fq64 X3 = sub(sub(sub(mul(R, R), PPP), Q), Q);
  const Fq64 Y3 = sub(mul(R, sub(Q, X3)), mul(S1, PPP));
  return {X3, Y3, mul(mul(p.ZZ, q.ZZ), PP), mul(mul(p.ZZZ, q.ZZZ), PPP)};
}

// Canonical 97-byte packet of an XYZZ point (identity -> tag 1).
inline void to_packet(const Xyzz64& p, uint8_t* out) {
  std::memset(out, 0, 97);
  if (p.is_identity()) {
    out[0] = 1;
    return;
  }
  const Fq64 inv = inverse(mul(p.ZZ, p.ZZZ));
  const Fq64 x = from_mont(mul(p.X, mul(inv, p.ZZZ)));
  const Fq64 y = from_mont(mul(p.Y, mul(inv, p.ZZ)));
  std::memcpy(out + 1, x.v, 48);
  std::memcpy(out + 49, y.v, 48);
}

}  // namespace host
}  // namespace hp
