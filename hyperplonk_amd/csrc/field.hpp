// BLS12-381 prime fields on 32-bit limbs, Montgomery form (R = 2^(32 N)).
//
// Fr (N = 8) is the scalar field of every MLE table, challenge and sumcheck
// message; Fq (N = 12) is the base field of G1.  Tables live in HBM in
// Montgomery form; the canonical little-endian form exists only at the
// host boundary (see convert kernels).  The same code compiles for host and
// device so host-side set-up (fixed-base tables, constants) cannot drift from
// what the kernels compute.
//
// Multiplication is CIOS with u32 x u32 -> u64 multiply-adds (v_mad_u64_u32
// on gfx950, full VALU rate).  Both moduli leave their top limb below
// 2^31 - 1, so the "no final carry word" form of CIOS applies and the result
// of every step stays below 2p before the single conditional subtraction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HP_HD __host__ __device__ __forceinline__

namespace hp {

struct FrParams {
  static constexpr int N = 8;
  static constexpr uint32_t MOD[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                                      0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
  static constexpr uint32_t INV = 0xffffffffu;  // -p^{-1} mod 2^32
  static constexpr uint32_t ONE[8] = {0xfffffffeu, 0x00000001u, 0x00034802u, 0x5884b7fau,
                                      0xecbc4ff5u, 0x998c4fefu, 0xacc5056fu, 0x1824b159u};
  static constexpr uint32_t R2[8] = {0xf3f29c6du, 0xc999e990u, 0x87925c23u, 0x2b6cedcbu,
                                     0x7254398fu, 0x05d31496u, 0x9f59ff11u, 0x0748d9d9u};
};

struct FqParams {
  static constexpr int N = 12;
  static constexpr uint32_t MOD[12] = {0xffffaaabu, 0xb9feffffu, 0xb153ffffu, 0x1eabfffeu,
                                       0xf6b0f624u, 0x6730d2a0u, 0xf38512bfu, 0x64774b84u,
                                       0x434bacd7u, 0x4b1ba7b6u, 0x397fe69au, 0x1a0111eau};
  static constexpr uint32_t INV = 0xfffcfffdu;
  static constexpr uint32_t ONE[12] = {0x0002fffdu, 0x76090000u, 0xc40c0002u, 0xebf4000bu,
                                       0x53c758bau, 0x5f489857u, 0x70525745u, 0x77ce5853u,
                                       0xa256ec6du, 0x5c071a97u, 0xfa80e493u, 0x15f65ec3u};
  static constexpr uint32_t R2[12] = {0x1c341746u, 0xf4df1f34u, 0x09d104f1u, 0x0a76e6a6u,
                                      0x4c95b6d5u, 0x8de5476cu, 0x939d83c0u, 0x67eb88a9u,
                                      0xb519952du, 0x9a793e85u, 0x92cae3aau, 0x11988fe5u};
};

template <class P>
struct Fe {
  static constexpr int N = P::N;
  uint32_t v[P::N];

  HP_HD static Fe zero() {
    Fe r;
#pragma unroll
    for (int i = 0; i < N; ++i) r.v[i] = 0;
    return r;
  }
  HP_HD static Fe one() {
    Fe r;
#pragma unroll
    for (int i = 0; i < N; ++i) r.v[i] = P::ONE[i];
    return r;
  }
  HP_HD static Fe modulus() {
    Fe r;
#pragma unroll
    for (int i = 0; i < N; ++i) r.v[i] = P::MOD[i];
    return r;
  }
  HP_HD bool is_zero() const {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) acc |= v[i];
    return acc == 0;
  }
  HP_HD bool operator==(const Fe& o) const {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) acc |= v[i] ^ o.v[i];
    return acc == 0;
  }
  HP_HD bool operator!=(const Fe& o) const { return !(*this == o); }
};

using Fr = Fe<FrParams>;
using Fq = Fe<FqParams>;

// a >= p (as integers)
template <class P>
HP_HD bool geq_modulus(const uint32_t* a) {
  // borrow of a - p
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < P::N; ++i) {
    uint64_t d = (uint64_t)a[i] - P::MOD[i] - borrow;
    borrow = (uint32_t)(d >> 63);
  }
  return borrow == 0;
}

// Conditionally subtract p from a (a < 2p): returns a mod p.
// Device code chains the hardware carry (v_add_co / v_addc, v_sub_co / v_subb:
// one instruction per limb) where the portable form takes several 64-bit
// operations per limb.
template <class P>
HP_HD void reduce_once(uint32_t* a) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t t[P::N];
  unsigned borrow = 0;
#pragma unroll
  for (int i = 0; i < P::N; ++i) t[i] = __builtin_subc(a[i], P::MOD[i], borrow, &borrow);
#pragma unroll
  for (int i = 0; i < P::N; ++i) a[i] = borrow ? a[i] : t[i];
#else
  uint32_t t[P::N];
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < P::N; ++i) {
    uint64_t d = (uint64_t)a[i] - P::MOD[i] - borrow;
    t[i] = (uint32_t)d;
    borrow = (uint32_t)(d >> 63);
  }
  // borrow == 0  <=>  a >= p  -> take t
#pragma unroll
  for (int i = 0; i < P::N; ++i) a[i] = borrow ? a[i] : t[i];
#endif
}

template <class P>
HP_HD Fe<P> operator+(const Fe<P>& a, const Fe<P>& b) {
  Fe<P> r;
#if defined(__HIP_DEVICE_COMPILE__)
  {
    unsigned c = 0;
#pragma unroll
    for (int i = 0; i < P::N; ++i) r.v[i] = __builtin_addc(a.v[i], b.v[i], c, &c);
  }
#else
  uint32_t carry = 0;
#pragma unroll
  for (int i = 0; i < P::N; ++i) {
    uint64_t s = (uint64_t)a.v[i] + b.v[i] + carry;
    r.v[i] = (uint32_t)s;
    carry = (uint32_t)(s >> 32);
  }
#endif
  // both moduli are below 2^(32N - 1): the sum never carries out.
  reduce_once<P>(r.v);
  return r;
}

template <class P>
HP_HD Fe<P> operator-(const Fe<P>& a, const Fe<P>& b) {
  Fe<P> r;
#if defined(__HIP_DEVICE_COMPILE__)
  {
    unsigned bo = 0;
#pragma unroll
    for (int i = 0; i < P::N; ++i) r.v[i] = __builtin_subc(a.v[i], b.v[i], bo, &bo);
    const uint32_t mask = 0u - bo;
    unsigned c = 0;
#pragma unroll
    for (int i = 0; i < P::N; ++i) r.v[i] = __builtin_addc(r.v[i], P::MOD[i] & mask, c, &c);
  }
#else
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < P::N; ++i) {
    uint64_t d = (uint64_t)a.v[i] - b.v[i] - borrow;
    r.v[i] = (uint32_t)d;
    borrow = (uint32_t)(d >> 63);
  }
  // add p back when the subtraction borrowed
  uint32_t mask = 0u - borrow;
  uint32_t carry = 0;
#pragma unroll
  for (int i = 0; i < P::N; ++i) {
    uint64_t s = (uint64_t)r.v[i] + (P::MOD[i] & mask) + carry;
    r.v[i] = (uint32_t)s;
    carry = (uint32_t)(s >> 32);
  }
#endif
  return r;
}

template <class P>
HP_HD Fe<P> neg(const Fe<P>& a) {
  return Fe<P>::zero() - a;
}

template <class P>
HP_HD Fe<P> dbl(const Fe<P>& a) {
  return a + a;
}

#if defined(__HIP_DEVICE_COMPILE__)
// acc (64 bit) += a * b with the carry out of bit 64 counted into hi:
// v_mad_u64_u32 leaves its carry in an SGPR pair, v_addc_co_u32 folds it in.
// Two VALU instructions per limb product and no zero-extension moves.
__device__ __forceinline__ void mac_vv(uint64_t& acc, uint32_t& hi, uint32_t a, uint32_t b) {
  uint64_t cc;
  asm("v_mad_u64_u32 %0, %1, %3, %4, %0\n\t"
      "v_addc_co_u32_e64 %2, %1, %2, 0, %1"
      : "+v"(acc), "=&s"(cc), "+v"(hi)
      : "v"(a), "v"(b));
}
// Same with the second factor a wave-uniform constant held in an SGPR.
__device__ __forceinline__ void mac_vs(uint64_t& acc, uint32_t& hi, uint32_t a, uint32_t b) {
  uint64_t cc;
  asm("v_mad_u64_u32 %0, %1, %3, %4, %0\n\t"
      "v_addc_co_u32_e64 %2, %1, %2, 0, %1"
      : "+v"(acc), "=&s"(cc), "+v"(hi)
      : "v"(a), "s"(b));
}

// Montgomery product by finely integrated product scanning: column k of
// a * b and of m * p is accumulated in one (64 + 32)-bit register triple,
// m_k is chosen to clear the column's low word, and the column shifts out.
template <class P>
__device__ __forceinline__ Fe<P> mul_fips(const Fe<P>& a, const Fe<P>& b) {
  constexpr int N = P::N;
  uint32_t m[N], t[N];
  uint64_t acc = 0;
  uint32_t hi = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
#pragma unroll
    for (int i = 0; i < k; ++i) {
      mac_vv(acc, hi, a.v[i], b.v[k - i]);
      mac_vs(acc, hi, m[i], P::MOD[k - i]);
    }
    mac_vv(acc, hi, a.v[k], b.v[0]);
    m[k] = (uint32_t)acc * P::INV;
    mac_vs(acc, hi, m[k], P::MOD[0]);
    acc = (acc >> 32) | ((uint64_t)hi << 32);
    hi = 0;
  }
#pragma unroll
  for (int k = N; k < 2 * N - 1; ++k) {
#pragma unroll
    for (int i = k - N + 1; i < N; ++i) {
      mac_vv(acc, hi, a.v[i], b.v[k - i]);
      mac_vs(acc, hi, m[i], P::MOD[k - i]);
    }
    t[k - N] = (uint32_t)acc;
    acc = (acc >> 32) | ((uint64_t)hi << 32);
    hi = 0;
  }
  t[N - 1] = (uint32_t)acc;
  Fe<P> r;
#pragma unroll
  for (int i = 0; i < N; ++i) r.v[i] = t[i];
  reduce_once<P>(r.v);
  return r;
}
#endif

// Montgomery product a * b * 2^(-32N) mod p.  Device code uses the product
// scanning form above; host code (set-up tables, final normalisation) uses
// CIOS in the no-final-carry variant.
template <class P>
HP_HD Fe<P> operator*(const Fe<P>& a, const Fe<P>& b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return mul_fips(a, b);
#else
  constexpr int N = P::N;
  uint32_t t[N];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    // t += a * b[i]
    uint64_t s = (uint64_t)a.v[0] * b.v[i] + t[0];
    uint32_t A = (uint32_t)(s >> 32);
    uint32_t t0 = (uint32_t)s;
    uint32_t m = t0 * P::INV;
    uint64_t u = (uint64_t)m * P::MOD[0] + t0;
    uint32_t C = (uint32_t)(u >> 32);
#pragma unroll
    for (int j = 1; j < N; ++j) {
      s = (uint64_t)a.v[j] * b.v[i] + t[j] + A;
      A = (uint32_t)(s >> 32);
      u = (uint64_t)m * P::MOD[j] + (uint32_t)s + C;
      C = (uint32_t)(u >> 32);
      t[j - 1] = (uint32_t)u;
    }
    t[N - 1] = C + A;
  }
  Fe<P> r;
#pragma unroll
  for (int i = 0; i < N; ++i) r.v[i] = t[i];
  reduce_once<P>(r.v);
  return r;
#endif
}

template <class P>
HP_HD Fe<P> sqr(const Fe<P>& a) {
  return a * a;
}

// Montgomery form of a canonical integer already known to be < p.
template <class P>
HP_HD Fe<P> to_mont(const Fe<P>& canonical) {
  Fe<P> r2;
#pragma unroll
  for (int i = 0; i < P::N; ++i) r2.v[i] = P::R2[i];
  return canonical * r2;
}

template <class P>
HP_HD Fe<P> from_mont(const Fe<P>& a) {
  Fe<P> one_raw = Fe<P>::zero();
  one_raw.v[0] = 1;
  return a * one_raw;
}

// a^(p-2) by left-to-right square-and-multiply over the constant exponent;
// zero maps to zero (callers reject zero denominators before inverting).
template <class P>
HP_HD Fe<P> inverse(const Fe<P>& a) {
  uint32_t exp[P::N];  // p - 2 (folded at compile time)
  uint32_t borrow = 2;
#pragma unroll
  for (int i = 0; i < P::N; ++i) {
    uint64_t d = (uint64_t)P::MOD[i] - borrow;
    exp[i] = (uint32_t)d;
    borrow = (uint32_t)(d >> 63);
  }
  Fe<P> acc = Fe<P>::one();
  for (int i = P::N - 1; i >= 0; --i) {
    const uint32_t e = exp[i];
    for (int b = 31; b >= 0; --b) {
      acc = sqr(acc);
      if ((e >> b) & 1u) acc = acc * a;
    }
  }
  return acc;
}

template <class P>
HP_HD Fe<P> from_u64(uint64_t x) {
  Fe<P> r = Fe<P>::zero();
  r.v[0] = (uint32_t)x;
  if (P::N > 1) r.v[1] = (uint32_t)(x >> 32);
  return to_mont(r);
}

}  // namespace hp
