// BLS12-381 prime fields on limbs of 29 bits, Montgomery form with R = 2^(29 L),
// lazy reduction: Fq29 (14 limbs, R = 2^406) is the arithmetic of the MSM
// bucket stages, Fr29 (9 limbs, R = 2^261) that of the sumcheck rounds.
//
// Why this radix: a 29 x 29-bit product is below 2^58, so a column of the
// Montgomery product (at most 2 L products plus the carry) fits a 64-bit
// accumulator and every limb product is ONE v_mad_u64_u32 with no carry
// bookkeeping, against a multiply-add plus a carry add on 32-bit limbs
// (2 * 14^2 = 392 products for Fq against 2 x 288 instructions).
//
// Lazy reduction: (a b + m p) / R < a b / R + p, and R / p is 2^25.3 (Fq) or
// 70.7 (Fr), so products of operands a few times p land just above p: no
// final subtraction; additions only propagate carries; a subtraction adds a
// multiple k p of the modulus that bounds the subtrahend.  Values stay well
// below R with normalised limbs (< 2^29, the top limb holds the rest); equality
// with zero needs the exact test is_zero_mod.
//
// Storage: Fq29 prepared SRS points hold x 2^406 mod p (canonical, < p) packed
// in twelve 32-bit words, the same 96 bytes as the 32-bit-limb form.  Fr29 is
// applied to values stored in the 32-bit-limb Montgomery form A = a 2^256
// (read as integers): mul(c 2^261, A) = c A, so a prescaled operand keeps
// results in the stored form (see the sumcheck kernels).
#pragma once
#include "field.hpp"

namespace hp {
namespace f29 {

constexpr uint32_t kBits = 29;
constexpr uint32_t kMask = (1u << kBits) - 1u;

struct FqP {
  static constexpr int L = 14;
  static constexpr int Words = 12;              // packed form
  static constexpr uint32_t INV = 0x1ffcfffdu;  // -p^-1 mod 2^29
  // p, 2p, 4p, 8p, 16p, normalised limbs
  static constexpr uint32_t P[L] = {0x1fffaaab, 0xff7ffff, 0x14ffffee, 0x17fffd62, 0xf6241ea,
                                    0x9507b58,  0xafd9cc3, 0x109e70a2, 0x1764774b, 0x121a5d66,
                                    0x12c6e9ed, 0x12ffcd34, 0x111ea3,  0xd};
  static constexpr uint32_t P2[L] = {0x1fff5556, 0x1fefffff, 0x9ffffdc,  0xffffac5,  0x1ec483d5,
                                     0x12a0f6b0, 0x15fb3986, 0x13ce144,  0xec8ee97,  0x434bacd,
                                     0x58dd3db,  0x5ff9a69,  0x223d47,   0x1a};
  static constexpr uint32_t P4[L] = {0x1ffeaaac, 0x1fdfffff, 0x13ffffb9, 0x1ffff58a, 0x1d8907aa,
                                     0x541ed61,  0xbf6730d,  0x279c289,  0x1d91dd2e, 0x869759a,
                                     0xb1ba7b6,  0xbff34d2,  0x447a8e,   0x34};
  static constexpr uint32_t P8[L] = {0x1ffd5558, 0x1fbfffff, 0x7ffff73,  0x1fffeb15, 0x1b120f55,
                                     0xa83dac3,  0x17ece61a, 0x4f38512,  0x1b23ba5c, 0x10d2eb35,
                                     0x16374f6c, 0x17fe69a4, 0x88f51c,   0x68};
  static constexpr uint32_t P16[L] = {0x1ffaaab0, 0x1f7fffff, 0xffffee7,  0x1fffd62a, 0x16241eab,
                                      0x1507b587, 0xfd9cc34,  0x9e70a25,  0x164774b8, 0x1a5d66b,
                                      0xc6e9ed9,  0xffcd349,  0x111ea39,  0xd0};
  static constexpr uint64_t kTop = 6980443811ull;  // p >> (29 (L - 2)) = p >> 348
};

struct FrP {
  static constexpr int L = 9;
  static constexpr int Words = 8;
  static constexpr uint32_t INV = 0x1fffffffu;  // -r^-1 mod 2^29 (r = 1 mod 2^32)
  static constexpr uint32_t P[L] = {0x1,       0x1ffffff8, 0x1f96ffbf, 0x1b4805ff, 0x1d80553b,
                                    0xc0404d0, 0x1520cce7, 0xa6533af,  0x73eda7};
  static constexpr uint32_t P2[L] = {0x2,        0x1ffffff0, 0x1f2dff7f, 0x16900bff, 0x1b00aa77,
                                     0x180809a1, 0xa4199ce,  0x14ca675f, 0xe7db4e};
  static constexpr uint32_t P4[L] = {0x4,        0x1fffffe0, 0x1e5bfeff, 0xd2017ff,  0x160154ef,
                                     0x10101343, 0x1483339d, 0x994cebe,  0x1cfb69d};
  static constexpr uint32_t P8[L] = {0x8,      0x1fffffc0, 0x1cb7fdff, 0x1a402fff, 0xc02a9de,
                                     0x202687, 0x906673b,  0x13299d7d, 0x39f6d3a};
  static constexpr uint32_t P16[L] = {0x10,     0x1fffff80, 0x196ffbff, 0x14805fff, 0x180553bd,
                                      0x404d0e, 0x120cce76, 0x6533afa,  0x73eda75};
  static constexpr uint64_t kTop = 4078865654035375ull;  // r >> (29 (L - 2)) = r >> 203
};

template <class P>
struct F29 {
  uint32_t l[P::L];
};
using Fq29 = F29<FqP>;
using Fr29 = F29<FrP>;

template <class P>
HP_HD F29<P> zero_t() {
  F29<P> r;
#pragma unroll
  for (int i = 0; i < P::L; ++i) r.l[i] = 0;
  return r;
}

template <class P>
HP_HD bool all_zero(const F29<P>& a) {
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < P::L; ++i) acc |= a.l[i];
  return acc == 0;
}

// (a b + m p) / R by product scanning: normalised limbs in, normalised out.
template <class P>
HP_HD F29<P> mul(const F29<P>& a, const F29<P>& b) {
  constexpr int L = P::L;
  uint32_t m[L];
  F29<P> t;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < L; ++k) {
#pragma unroll
    for (int i = 0; i < k; ++i) {
      acc += (uint64_t)a.l[i] * b.l[k - i];
      acc += (uint64_t)m[i] * P::P[k - i];
    }
    acc += (uint64_t)a.l[k] * b.l[0];
    m[k] = ((uint32_t)acc * P::INV) & kMask;
    acc += (uint64_t)m[k] * P::P[0];
    acc >>= kBits;  // the low 29 bits are zero
  }
#pragma unroll
  for (int k = L; k < 2 * L - 1; ++k) {
#pragma unroll
    for (int i = k - L + 1; i < L; ++i) {
      acc += (uint64_t)a.l[i] * b.l[k - i];
      acc += (uint64_t)m[i] * P::P[k - i];
    }
    t.l[k - L] = (uint32_t)acc & kMask;
    acc >>= kBits;
  }
  t.l[L - 1] = (uint32_t)acc;
  return t;
}

// a^2 / R: every cross product once, against the doubled operand (2 a_j is
// below 2^30); the columns are those of mul(a, a), so the result is the same
// integer.  Column bound: 7 cross products below 2^59, one square and 14
// m p products below 2^58: 29 * 2^58 < 2^63.
template <class P>
HP_HD F29<P> sqr(const F29<P>& a) {
  constexpr int L = P::L;
  uint32_t d[L], m[L];
#pragma unroll
  for (int i = 0; i < L; ++i) d[i] = a.l[i] << 1;
  F29<P> t;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < L; ++k) {
#pragma unroll
    for (int i = 0; i < k - i; ++i) acc += (uint64_t)a.l[i] * d[k - i];
    if ((k & 1) == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
    for (int i = 0; i < k; ++i) acc += (uint64_t)m[i] * P::P[k - i];
    m[k] = ((uint32_t)acc * P::INV) & kMask;
    acc += (uint64_t)m[k] * P::P[0];
    acc >>= kBits;
  }
#pragma unroll
  for (int k = L; k < 2 * L - 1; ++k) {
#pragma unroll
    for (int i = k - L + 1; i < k - i; ++i) acc += (uint64_t)a.l[i] * d[k - i];
    if ((k & 1) == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
    for (int i = k - L + 1; i < L; ++i) acc += (uint64_t)m[i] * P::P[k - i];
    t.l[k - L] = (uint32_t)acc & kMask;
    acc >>= kBits;
  }
  t.l[L - 1] = (uint32_t)acc;
  return t;
}

// (a b + c d) / R with one Montgomery reduction.  A column holds at most
// 3 L = 42 products below 2^58 (< 2^63.4) plus the carry; the result is below
// (a b + c d) / R + p.
template <class P>
HP_HD F29<P> mul_sum(const F29<P>& a, const F29<P>& b, const F29<P>& c, const F29<P>& d) {
  constexpr int L = P::L;
  uint32_t m[L];
  F29<P> t;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < L; ++k) {
#pragma unroll
    for (int i = 0; i < k; ++i) {
      acc += (uint64_t)a.l[i] * b.l[k - i];
      acc += (uint64_t)c.l[i] * d.l[k - i];
      acc += (uint64_t)m[i] * P::P[k - i];
    }
    acc += (uint64_t)a.l[k] * b.l[0];
    acc += (uint64_t)c.l[k] * d.l[0];
    m[k] = ((uint32_t)acc * P::INV) & kMask;
    acc += (uint64_t)m[k] * P::P[0];
    acc >>= kBits;
  }
#pragma unroll
  for (int k = L; k < 2 * L - 1; ++k) {
#pragma unroll
    for (int i = k - L + 1; i < L; ++i) {
      acc += (uint64_t)a.l[i] * b.l[k - i];
      acc += (uint64_t)c.l[i] * d.l[k - i];
      acc += (uint64_t)m[i] * P::P[k - i];
    }
    t.l[k - L] = (uint32_t)acc & kMask;
    acc >>= kBits;
  }
  t.l[L - 1] = (uint32_t)acc;
  return t;
}

template <class P>
HP_HD F29<P> add(const F29<P>& a, const F29<P>& b) {
  constexpr int L = P::L;
  F29<P> r;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < L - 1; ++i) {
    const uint32_t s = a.l[i] + b.l[i] + c;
    r.l[i] = s & kMask;
    c = s >> kBits;
  }
  r.l[L - 1] = a.l[L - 1] + b.l[L - 1] + c;
  return r;
}

// a + b + b
template <class P>
HP_HD F29<P> add_twice(const F29<P>& a, const F29<P>& b) {
  constexpr int L = P::L;
  F29<P> r;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < L - 1; ++i) {
    const uint32_t s = a.l[i] + 2u * b.l[i] + c;
    r.l[i] = s & kMask;
    c = s >> kBits;
  }
  r.l[L - 1] = a.l[L - 1] + 2u * b.l[L - 1] + c;
  return r;
}

// a + k p - b for the constant limbs kp of k p >= b (signed carries)
template <class P>
HP_HD F29<P> sub_kp(const F29<P>& a, const F29<P>& b, const uint32_t* kp) {
  constexpr int L = P::L;
  F29<P> r;
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < L - 1; ++i) {
    const int32_t s = (int32_t)(a.l[i] + kp[i]) - (int32_t)b.l[i] + c;
    r.l[i] = (uint32_t)s & kMask;
    c = s >> kBits;  // arithmetic shift
  }
  r.l[L - 1] = (uint32_t)((int32_t)(a.l[L - 1] + kp[L - 1]) - (int32_t)b.l[L - 1] + c);
  return r;
}

template <class P>
HP_HD F29<P> sub1p(const F29<P>& a, const F29<P>& b) {
  return sub_kp(a, b, P::P);
}
template <class P>
HP_HD F29<P> sub2p(const F29<P>& a, const F29<P>& b) {
  return sub_kp(a, b, P::P2);
}
template <class P>
HP_HD F29<P> sub4p(const F29<P>& a, const F29<P>& b) {
  return sub_kp(a, b, P::P4);
}
template <class P>
HP_HD F29<P> sub8p(const F29<P>& a, const F29<P>& b) {
  return sub_kp(a, b, P::P8);
}
template <class P>
HP_HD F29<P> sub16p(const F29<P>& a, const F29<P>& b) {
  return sub_kp(a, b, P::P16);
}
// p - a for a canonical a (< p)
template <class P>
HP_HD F29<P> neg_canonical(const F29<P>& a) {
  return sub_kp(zero_t<P>(), a, P::P);
}

// a - k p when a >= k p, else a
template <class P>
HP_HD F29<P> cond_sub(const F29<P>& a, const uint32_t* kp) {
  constexpr int L = P::L;
  F29<P> d;
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int32_t x = (int32_t)a.l[i] - (int32_t)kp[i] + c;
    d.l[i] = i < L - 1 ? ((uint32_t)x & kMask) : (uint32_t)x;
    c = i < L - 1 ? (x >> kBits) : (x >> 31);
  }
  return c == 0 ? d : a;
}

// a mod p in [0, p) for a < 32 p
template <class P>
HP_HD F29<P> canonical(const F29<P>& a) {
  F29<P> v = cond_sub(a, P::P16);
  v = cond_sub(v, P::P8);
  v = cond_sub(v, P::P4);
  v = cond_sub(v, P::P2);
  return cond_sub(v, P::P);
}

// a - k p for k = floor(a / p) - 1 (estimated from the top 58 bits), clamped
// at 0: a representative below 3 p of a value below 2^(29 L) with normalised
// limbs (for Fr, values up to 70 r; the sumcheck accumulators stay below 41 r)
template <class P>
HP_HD F29<P> reduce_est(const F29<P>& a) {
  constexpr int L = P::L;
  const uint64_t top = ((uint64_t)a.l[L - 1] << kBits) | a.l[L - 2];
  const double q = (double)top * (1.0 / (double)P::kTop);
  const uint32_t k = q >= 1.0 ? (uint32_t)q - 1u : 0u;
  F29<P> r;
  int64_t c = 0;
#pragma unroll
  for (int i = 0; i < L - 1; ++i) {
    const int64_t x = (int64_t)a.l[i] - (int64_t)k * P::P[i] + c;
    r.l[i] = (uint32_t)x & kMask;
    c = x >> kBits;
  }
  r.l[L - 1] = (uint32_t)((int64_t)a.l[L - 1] - (int64_t)k * P::P[L - 1] + c);
  return r;
}

// a == k p exactly (limbs normalised, k < 64)
template <class P>
HP_HD bool equals_kp(const F29<P>& a, uint32_t k) {
  constexpr int L = P::L;
  uint64_t c = 0;
  uint32_t diff = 0;
#pragma unroll
  for (int i = 0; i < L - 1; ++i) {
    c += (uint64_t)k * P::P[i];
    diff |= ((uint32_t)c & kMask) ^ a.l[i];
    c >>= kBits;
  }
  c += (uint64_t)k * P::P[L - 1];
  diff |= (uint32_t)c ^ a.l[L - 1];
  return diff == 0;
}

// a = 0 (mod p) for a < 32 p: the only candidate multiple is k = round(a / p),
// read off the top 58 bits; the low limb rules out all but ~2^-29 of the
// non-multiples before the full comparison.
template <class P>
HP_HD bool is_zero_mod(const F29<P>& a) {
  constexpr int L = P::L;
  const uint64_t top = ((uint64_t)a.l[L - 1] << kBits) | a.l[L - 2];
  const uint32_t k = (uint32_t)((double)top * (1.0 / (double)P::kTop) + 0.5);
  if ((((uint64_t)k * P::P[0]) & kMask) != a.l[0]) return false;
  return equals_kp(a, k);
}

// Packed value (P::Words 32-bit words, < 2^(32 Words)) <-> limbs.
template <class P>
HP_HD F29<P> unpack_t(const uint32_t* w) {
  F29<P> r;
#pragma unroll
  for (int i = 0; i < P::L; ++i) {
    const int bit = (int)kBits * i, j = bit >> 5, sh = bit & 31;
    uint64_t x = w[j];
    if (j + 1 < P::Words) x |= (uint64_t)w[j + 1] << 32;
    r.l[i] = (uint32_t)(x >> sh) & kMask;
  }
  return r;
}

template <class P>
HP_HD void pack_t(const F29<P>& a, uint32_t* w) {
#pragma unroll
  for (int j = 0; j < P::Words; ++j) w[j] = 0;
#pragma unroll
  for (int i = 0; i < P::L; ++i) {
    const int bit = (int)kBits * i, j = bit >> 5, sh = bit & 31;
    const uint64_t x = (uint64_t)a.l[i] << sh;
    w[j] |= (uint32_t)x;
    if (j + 1 < P::Words) w[j + 1] |= (uint32_t)(x >> 32);
  }
}

// ---- Fq29 names used by the MSM -------------------------------------------
constexpr int L = FqP::L;

// R mod p (Montgomery one) and the factors that move a value between
// R = 2^384 (Fq) and R = 2^406 (Fq29) with one Fq product: 2^406 mod p, 2^362
struct Consts {
  static constexpr uint32_t ONE[L] = {0x3a9fb84,  0xba00690,  0x71288f1,  0xf59bcc5,  0x126cb614,
                                      0x585bf36,  0x1b85ac3d, 0x1cf856fa, 0x1891ecbd, 0x1a7eec05,
                                      0x155a88f0, 0x741ac6d,  0x1317c30f, 0x9};
  static constexpr uint32_t TO29[12] = {0x3a9fb84,  0xc57400d2, 0x629c4a23, 0x6147acde,
                                        0x7e6d26cb, 0x6b0f4b0b, 0xc2b7d6e1, 0x91ecbde7,
                                        0x4fdd80b8, 0xd56a23c3, 0xf3a0d636, 0x13317c30};
  static constexpr uint32_t FROM29[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0x400};
};

HP_HD Fq29 zero() { return zero_t<FqP>(); }

HP_HD Fq29 one() {
  Fq29 r;
#pragma unroll
  for (int i = 0; i < L; ++i) r.l[i] = Consts::ONE[i];
  return r;
}

HP_HD Fq29 unpack(const uint32_t* w) { return unpack_t<FqP>(w); }
HP_HD void pack(const Fq29& a, uint32_t* w) { pack_t(a, w); }

// Fq (Montgomery, R = 2^384) -> packed canonical Fq29 value and back.
HP_HD Fq to29_packed(const Fq& x) {
  Fq c;
#pragma unroll
  for (int i = 0; i < 12; ++i) c.v[i] = Consts::TO29[i];
  return x * c;  // x R384 * 2^406 / R384 = x 2^406, canonical
}

// packed canonical x 2^406 -> x R384
HP_HD Fq from29_packed(const Fq& packed) {
  Fq c;
#pragma unroll
  for (int i = 0; i < 12; ++i) c.v[i] = Consts::FROM29[i];
  return packed * c;
}

HP_HD Fq from29(const Fq29& a) {
  Fq packed;
  pack(canonical(a), packed.v);
  return from29_packed(packed);  // x 2^406 * 2^362 / 2^384 = x R384
}

// ---- Fr29 names used by the sumcheck ----------------------------------------
// 2^261 mod r as 32-bit words: the Fr (R = 2^256) Montgomery form of 2^5, so
// c (R = 2^256) * K5 = c 2^261, the prescaled operand of Fr29 products.
struct FrConsts {
  static constexpr uint32_t K5[8] = {0xffffffba, 0x45,       0x72d846,   0x1a25272e,
                                     0x5dbeee8b, 0xfe2eedcd, 0x9eefbe41, 0x4d043f42};
};

HP_HD Fr29 unpack_fr(const Fr& a) { return unpack_t<FrP>(a.v); }

// canonical packed Fr (the stored form) from a value below 32 r
HP_HD Fr pack_fr(const Fr29& a) {
  Fr r;
  pack_t(canonical(a), r.v);
  return r;
}

// canonical packed Fr from a value below 4 r (a reduced value)
HP_HD Fr pack_fr_reduced(const Fr29& a) {
  Fr r;
  pack_t(cond_sub(cond_sub(a, FrP::P2), FrP::P), r.v);
  return r;
}

// packed Fr of a value already below r
HP_HD Fr pack_fr_canonical(const Fr29& a) {
  Fr r;
  pack_t(a, r.v);
  return r;
}

// c (stored form, c 2^256) -> c 2^261 in Fr29 limbs
HP_HD Fr29 prescale_fr(const Fr& c) {
  Fr k;
#pragma unroll
  for (int i = 0; i < 8; ++i) k.v[i] = FrConsts::K5[i];
  return unpack_fr(c * k);
}

// a^-1 in the stored form (zero maps to zero): a^(r - 2) with 4-bit windows
// on Fr29 (R = 2^261) from the prescaled a, then one product by the stored
// form of 1 (2^256) to return to the stored form.  Half the
// instructions of the 32-bit-limb exponentiation on the latency-bound paths
// (one inversion per thread in batch inversions).
HP_HD Fr inverse_fr(const Fr& a) {
  uint32_t exp[8];  // r - 2
  uint32_t borrow = 2;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint64_t d = (uint64_t)FrParams::MOD[i] - borrow;
    exp[i] = (uint32_t)d;
    borrow = (uint32_t)(d >> 63);
  }
  // fixed 4-bit windows: x^1 .. x^15, then 4 squarings and at most one
  // product per nibble (252 squarings and <= 77 products, against 255 and
  // ~130 bit by bit); the top nibble of r - 2 (7) starts the chain
  Fr29 tab[16];
  tab[1] = prescale_fr(a);
#pragma unroll
  for (int k = 2; k < 16; ++k) tab[k] = mul(tab[k - 1], tab[1]);
  Fr29 acc = tab[(exp[7] >> 28) & 15u];
  for (int nib = 62; nib >= 0; --nib) {
    acc = sqr(sqr(sqr(sqr(acc))));
    const uint32_t d = (exp[nib >> 3] >> (4 * (nib & 7))) & 15u;
    if (d) acc = mul(acc, tab[d]);
  }
  return pack_fr(mul(acc, unpack_fr(Fr::one())));
}

}  // namespace f29
}  // namespace hp
