// BLS12-381 Fq on 14 limbs of 29 bits, Montgomery form with R = 2^406, lazy
// reduction: the arithmetic of the MSM bucket stages.
//
// Why this radix: a 29 x 29-bit product is below 2^58, so a column of the
// Montgomery product (at most 28 products plus the carry) fits a 64-bit
// accumulator and every limb product is ONE v_mad_u64_u32 with no carry
// bookkeeping, against a multiply-add plus a carry add on 32-bit limbs.
// 14^2 + 14^2 = 392 products cost fewer VALU instructions than 2 x 288.
//
// Lazy reduction: R / p > 2^25, so for inputs below 16 p the product
// (a b + m p) / R is below p (1 + 2^-17): multiplication never needs a final
// subtraction, additions only propagate carries, and a subtraction adds a
// multiple k p of the modulus that bounds the subtrahend.  Values therefore
// live in [0, 32 p) with normalised limbs (< 2^29, the top limb holds the
// rest); only equality with zero needs the exact test is_zero_mod.
//
// Storage: prepared SRS points hold x 2^406 mod p (canonical, < p) packed in
// twelve 32-bit words, the same 96 bytes as the 32-bit-limb form.
#pragma once
#include "field.hpp"

namespace hp {
namespace f29 {

constexpr int L = 14;
constexpr uint32_t kBits = 29;
constexpr uint32_t kMask = (1u << kBits) - 1u;
constexpr uint32_t kInv = 0x1ffcfffdu;  // -p^-1 mod 2^29

struct Fq29 {
  uint32_t l[L];
};

// p, 2p, 4p, 8p, 16p and R mod p (Montgomery one), normalised limbs
struct Consts {
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
  static constexpr uint32_t ONE[L] = {0x3a9fb84,  0xba00690,  0x71288f1,  0xf59bcc5,  0x126cb614,
                                      0x585bf36,  0x1b85ac3d, 0x1cf856fa, 0x1891ecbd, 0x1a7eec05,
                                      0x155a88f0, 0x741ac6d,  0x1317c30f, 0x9};
  // 2^406 mod p and 2^362 as 32-bit words: the Montgomery factors that move
  // a value between R = 2^384 (Fq) and R = 2^406 (Fq29) with one Fq product
  static constexpr uint32_t TO29[12] = {0x3a9fb84,  0xc57400d2, 0x629c4a23, 0x6147acde,
                                        0x7e6d26cb, 0x6b0f4b0b, 0xc2b7d6e1, 0x91ecbde7,
                                        0x4fdd80b8, 0xd56a23c3, 0xf3a0d636, 0x13317c30};
  static constexpr uint32_t FROM29[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0x400};
  static constexpr uint64_t kPTop = 6980443811ull;  // p >> 348
};

HP_HD Fq29 zero() {
  Fq29 r;
#pragma unroll
  for (int i = 0; i < L; ++i) r.l[i] = 0;
  return r;
}

HP_HD Fq29 one() {
  Fq29 r;
#pragma unroll
  for (int i = 0; i < L; ++i) r.l[i] = Consts::ONE[i];
  return r;
}

HP_HD bool all_zero(const Fq29& a) {
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < L; ++i) acc |= a.l[i];
  return acc == 0;
}

// (a b + m p) / 2^406 by product scanning; inputs below 16 p with normalised
// limbs give an output below p (1 + 2^-17), normalised.
HP_HD Fq29 mul(const Fq29& a, const Fq29& b) {
  uint32_t m[L];
  Fq29 t;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < L; ++k) {
#pragma unroll
    for (int i = 0; i < k; ++i) {
      acc += (uint64_t)a.l[i] * b.l[k - i];
      acc += (uint64_t)m[i] * Consts::P[k - i];
    }
    acc += (uint64_t)a.l[k] * b.l[0];
    m[k] = ((uint32_t)acc * kInv) & kMask;
    acc += (uint64_t)m[k] * Consts::P[0];
    acc >>= kBits;  // the low 29 bits are zero
  }
#pragma unroll
  for (int k = L; k < 2 * L - 1; ++k) {
#pragma unroll
    for (int i = k - L + 1; i < L; ++i) {
      acc += (uint64_t)a.l[i] * b.l[k - i];
      acc += (uint64_t)m[i] * Consts::P[k - i];
    }
    t.l[k - L] = (uint32_t)acc & kMask;
    acc >>= kBits;
  }
  t.l[L - 1] = (uint32_t)acc;
  return t;
}

HP_HD Fq29 sqr(const Fq29& a) { return mul(a, a); }

HP_HD Fq29 add(const Fq29& a, const Fq29& b) {
  Fq29 r;
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
HP_HD Fq29 add_twice(const Fq29& a, const Fq29& b) {
  Fq29 r;
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
HP_HD Fq29 sub_kp(const Fq29& a, const Fq29& b, const uint32_t* kp) {
  Fq29 r;
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

HP_HD Fq29 sub2p(const Fq29& a, const Fq29& b) { return sub_kp(a, b, Consts::P2); }
HP_HD Fq29 sub4p(const Fq29& a, const Fq29& b) { return sub_kp(a, b, Consts::P4); }
HP_HD Fq29 sub8p(const Fq29& a, const Fq29& b) { return sub_kp(a, b, Consts::P8); }
HP_HD Fq29 sub16p(const Fq29& a, const Fq29& b) { return sub_kp(a, b, Consts::P16); }
// p - a for a canonical a (< p)
HP_HD Fq29 neg_canonical(const Fq29& a) { return sub_kp(zero(), a, Consts::P); }

// a == k p exactly (limbs normalised, k < 64)
HP_HD bool equals_kp(const Fq29& a, uint32_t k) {
  uint64_t c = 0;
  uint32_t diff = 0;
#pragma unroll
  for (int i = 0; i < L - 1; ++i) {
    c += (uint64_t)k * Consts::P[i];
    diff |= ((uint32_t)c & kMask) ^ a.l[i];
    c >>= kBits;
  }
  c += (uint64_t)k * Consts::P[L - 1];
  diff |= (uint32_t)c ^ a.l[L - 1];
  return diff == 0;
}

// a = 0 (mod p) for a < 32 p: the only candidate multiple is k = round(a / p),
// read off the top 58 bits; the low limb rules out all but ~2^-29 of the
// non-multiples before the full comparison.
HP_HD bool is_zero_mod(const Fq29& a) {
  const uint64_t top = ((uint64_t)a.l[L - 1] << kBits) | a.l[L - 2];  // a >> 348
  const uint32_t k = (uint32_t)((double)top * (1.0 / (double)Consts::kPTop) + 0.5);
  if ((((uint64_t)k * Consts::P[0]) & kMask) != a.l[0]) return false;
  return equals_kp(a, k);
}

// a mod p in [0, p) for a < 32 p (conversions only)
HP_HD Fq29 canonical(const Fq29& a) {
  Fq29 v = a;
  const uint32_t* steps[5] = {Consts::P16, Consts::P8, Consts::P4, Consts::P2, Consts::P};
#pragma unroll
  for (int s = 0; s < 5; ++s) {
    Fq29 d;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const int32_t x = (int32_t)v.l[i] - (int32_t)steps[s][i] + c;
      d.l[i] = i < L - 1 ? ((uint32_t)x & kMask) : (uint32_t)x;
      c = i < L - 1 ? (x >> kBits) : (x >> 31);
    }
    if (c == 0) v = d;  // no borrow: v >= k p
  }
  return v;
}

// Packed canonical value (twelve 32-bit words, < 2^384) <-> limbs.
HP_HD Fq29 unpack(const uint32_t* w) {
  Fq29 r;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int bit = (int)kBits * i, j = bit >> 5, sh = bit & 31;
    uint64_t x = w[j];
    if (j + 1 < 12) x |= (uint64_t)w[j + 1] << 32;
    r.l[i] = (uint32_t)(x >> sh) & kMask;
  }
  return r;
}

HP_HD void pack(const Fq29& a, uint32_t* w) {
#pragma unroll
  for (int j = 0; j < 12; ++j) w[j] = 0;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int bit = (int)kBits * i, j = bit >> 5, sh = bit & 31;
    const uint64_t x = (uint64_t)a.l[i] << sh;
    w[j] |= (uint32_t)x;
    if (j + 1 < 12) w[j + 1] |= (uint32_t)(x >> 32);
  }
}

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
  Fq packed, c;
  pack(canonical(a), packed.v);
#pragma unroll
  for (int i = 0; i < 12; ++i) c.v[i] = Consts::FROM29[i];
  return packed * c;  // x 2^406 * 2^362 / 2^384 = x R384
}

}  // namespace f29
}  // namespace hp
