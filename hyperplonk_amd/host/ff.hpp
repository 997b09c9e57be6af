// Host prime fields of BLS12-381 for the protocol layer (transcript
// challenges, verifier arithmetic, interpolation): 64-bit limbs in Montgomery
// form, R = 2^(64 N), CIOS multiplication over unsigned __int128.
//
// Canonical encodings follow ark-ff as the reference uses it:
//   * an element serialises as its canonical integer, little-endian, in
//     ceil(bits / 8) bytes (transcript/src/lib.rs:64-66 -> serialize_compressed);
//   * from_le_bytes_mod_order reads any byte string as a little-endian integer
//     and reduces it mod p (transcript/src/lib.rs:91).
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>

namespace hyperplonk {
namespace host {

using u64 = uint64_t;
using u128 = unsigned __int128;

template <size_t N>
using Limbs = std::array<u64, N>;

// r = 0x73eda753 299d7d48 3339d808 09a1d805 53bda402 fffe5bfe ffffffff 00000001
struct FrSpec {
  static constexpr size_t N = 4;
  static constexpr size_t BYTES = 32;
  static constexpr Limbs<4> MOD = {0xffffffff00000001ULL, 0x53bda402fffe5bfeULL,
                                   0x3339d80809a1d805ULL, 0x73eda753299d7d48ULL};
  static constexpr u64 INV = 0xfffffffeffffffffULL;  // -MOD^-1 mod 2^64
  static constexpr Limbs<4> R1 = {0x00000001fffffffeULL, 0x5884b7fa00034802ULL,
                                  0x998c4fefecbc4ff5ULL, 0x1824b159acc5056fULL};
  static constexpr Limbs<4> R2 = {0xc999e990f3f29c6dULL, 0x2b6cedcb87925c23ULL,
                                  0x05d314967254398fULL, 0x0748d9d99f59ff11ULL};
};

// q: the 381-bit base field modulus
struct FqSpec {
  static constexpr size_t N = 6;
  static constexpr size_t BYTES = 48;
  static constexpr Limbs<6> MOD = {0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL,
                                   0x6730d2a0f6b0f624ULL, 0x64774b84f38512bfULL,
                                   0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};
  static constexpr u64 INV = 0x89f3fffcfffcfffdULL;
  static constexpr Limbs<6> R1 = {0x760900000002fffdULL, 0xebf4000bc40c0002ULL,
                                  0x5f48985753c758baULL, 0x77ce585370525745ULL,
                                  0x5c071a97a256ec6dULL, 0x15f65ec3fa80e493ULL};
  static constexpr Limbs<6> R2 = {0xf4df1f341c341746ULL, 0x0a76e6a609d104f1ULL,
                                  0x8de5476c4c95b6d5ULL, 0x67eb88a9939d83c0ULL,
                                  0x9a793e85b519952dULL, 0x11988fe592cae3aaULL};
};

template <size_t N>
inline bool limbs_less(const Limbs<N>& a, const Limbs<N>& b) {
  for (size_t i = N; i-- > 0;)
    if (a[i] != b[i]) return a[i] < b[i];
  return false;
}

// a -= b, returns the borrow
template <size_t N>
inline u64 limbs_sub(Limbs<N>& a, const Limbs<N>& b) {
  u64 borrow = 0;
  for (size_t i = 0; i < N; ++i) {
    const u128 d = (u128)a[i] - b[i] - borrow;
    a[i] = (u64)d;
    borrow = (u64)(d >> 64) & 1;
  }
  return borrow;
}

template <class S>
struct Fp {
  static constexpr size_t N = S::N;
  Limbs<N> v{};  // Montgomery representation

  static Fp zero() { return Fp{}; }
  static Fp one() {
    Fp r;
    r.v = S::R1;
    return r;
  }
  static Fp from_u64(u64 x) {
    Limbs<N> c{};
    c[0] = x;
    return from_limbs_reduce(c);
  }
  static Fp from_i64(int64_t x) {
    return x < 0 ? -from_u64(static_cast<u64>(-(x + 1)) + 1) : from_u64(static_cast<u64>(x));
  }
  // canonical (< MOD) limbs -> element
  static Fp from_canonical(const Limbs<N>& c) {
    Fp a;
    a.v = c;
    Fp r2;
    r2.v = S::R2;
    return a * r2;
  }
  // any limbs -> element (reduced by conditional subtraction loop)
  static Fp from_limbs_reduce(Limbs<N> c) {
    while (!limbs_less(c, S::MOD)) limbs_sub(c, S::MOD);
    return from_canonical(c);
  }

  Limbs<N> canonical() const {
    Fp one_raw;
    one_raw.v = Limbs<N>{};
    one_raw.v[0] = 1;
    return (*this * one_raw).v;
  }

  bool is_zero() const {
    for (u64 x : v)
      if (x) return false;
    return true;
  }
  bool operator==(const Fp& o) const { return v == o.v; }
  bool operator!=(const Fp& o) const { return v != o.v; }

  Fp operator+(const Fp& o) const {
    Fp r;
    u64 carry = 0;
    for (size_t i = 0; i < N; ++i) {
      const u128 s = (u128)v[i] + o.v[i] + carry;
      r.v[i] = (u64)s;
      carry = (u64)(s >> 64);
    }
    if (carry || !limbs_less(r.v, S::MOD)) limbs_sub(r.v, S::MOD);
    return r;
  }
  Fp operator-(const Fp& o) const {
    Fp r = *this;
    if (limbs_sub(r.v, o.v)) {
      u64 carry = 0;
      for (size_t i = 0; i < N; ++i) {
        const u128 s = (u128)r.v[i] + S::MOD[i] + carry;
        r.v[i] = (u64)s;
        carry = (u64)(s >> 64);
      }
    }
    return r;
  }
  Fp operator-() const { return Fp{} - *this; }
  Fp dbl() const { return *this + *this; }

  // CIOS Montgomery product (both moduli leave a spare top bit, so one
  // conditional subtraction suffices).
  Fp operator*(const Fp& o) const {
    u64 t[N + 2] = {};
    for (size_t i = 0; i < N; ++i) {
      u64 carry = 0;
      for (size_t j = 0; j < N; ++j) {
        const u128 s = (u128)v[j] * o.v[i] + t[j] + carry;
        t[j] = (u64)s;
        carry = (u64)(s >> 64);
      }
      u128 s = (u128)t[N] + carry;
      t[N] = (u64)s;
      t[N + 1] = (u64)(s >> 64);
      const u64 m = t[0] * S::INV;
      s = (u128)m * S::MOD[0] + t[0];
      carry = (u64)(s >> 64);
      for (size_t j = 1; j < N; ++j) {
        s = (u128)m * S::MOD[j] + t[j] + carry;
        t[j - 1] = (u64)s;
        carry = (u64)(s >> 64);
      }
      s = (u128)t[N] + carry;
      t[N - 1] = (u64)s;
      t[N] = t[N + 1] + (u64)(s >> 64);
    }
    Fp r;
    for (size_t i = 0; i < N; ++i) r.v[i] = t[i];
    if (t[N] || !limbs_less(r.v, S::MOD)) limbs_sub(r.v, S::MOD);
    return r;
  }
  Fp& operator+=(const Fp& o) { return *this = *this + o; }
  Fp& operator-=(const Fp& o) { return *this = *this - o; }
  Fp& operator*=(const Fp& o) { return *this = *this * o; }
  Fp square() const { return *this * *this; }

  // exponent given as little-endian 64-bit limbs
  template <size_t M>
  Fp pow(const Limbs<M>& e) const {
    Fp r = one();
    for (size_t i = M; i-- > 0;)
      for (int b = 63; b >= 0; --b) {
        r = r.square();
        if ((e[i] >> b) & 1) r = r * *this;
      }
    return r;
  }
  // Fermat: a^(p-2); zero maps to zero
  Fp inverse() const {
    Limbs<N> e = S::MOD;
    Limbs<N> two{};
    two[0] = 2;
    limbs_sub(e, two);
    return pow(e);
  }

  // canonical little-endian bytes (S::BYTES)
  void to_bytes(uint8_t* out) const {
    const Limbs<N> c = canonical();
    for (size_t i = 0; i < N; ++i)
      for (size_t b = 0; b < 8; ++b) out[8 * i + b] = static_cast<uint8_t>(c[i] >> (8 * b));
  }
  std::array<uint8_t, S::BYTES> bytes() const {
    std::array<uint8_t, S::BYTES> out{};
    to_bytes(out.data());
    return out;
  }
  // canonical bytes -> element; false if the integer is >= MOD
  static bool from_bytes(const uint8_t* in, Fp* out) {
    Limbs<N> c{};
    for (size_t i = 0; i < N; ++i)
      for (size_t b = 0; b < 8; ++b) c[i] |= static_cast<u64>(in[8 * i + b]) << (8 * b);
    if (!limbs_less(c, S::MOD)) return false;
    *out = from_canonical(c);
    return true;
  }
  // little-endian integer of any length, reduced mod p: Horner over chunks of
  // N limbs from the most significant end, base 2^(64N) (= S::R1 mod p), so a
  // 64-byte Fr challenge costs three Montgomery products instead of 128
  static Fp from_le_bytes_mod_order(const uint8_t* in, size_t len) {
    constexpr size_t kChunk = 8 * N;
    const Fp base = from_canonical(S::R1);
    Fp acc = zero();
    size_t end = len;
    while (end > 0) {
      const size_t begin = (end - 1) / kChunk * kChunk;  // the top chunk may be partial
      Limbs<N> c{};
      for (size_t i = begin; i < end; ++i) c[(i - begin) / 8] |= static_cast<u64>(in[i]) << (8 * ((i - begin) % 8));
      acc = acc * base + from_limbs_reduce(c);
      end = begin;
    }
    return acc;
  }
  // arkworks' "y > -y" on canonical integers
  bool lexicographically_largest() const {
    const Limbs<N> a = canonical(), b = (-*this).canonical();
    return limbs_less(b, a);
  }
  std::string hex() const {  // big-endian hex of the canonical integer
    static const char* d = "0123456789abcdef";
    std::string s;
    const Limbs<N> c = canonical();
    for (size_t i = N; i-- > 0;)
      for (int b = 60; b >= 0; b -= 4) s.push_back(d[(c[i] >> b) & 15]);
    return s;
  }
};

using Fr = Fp<FrSpec>;
using Fq = Fp<FqSpec>;

}  // namespace host
}  // namespace hyperplonk
