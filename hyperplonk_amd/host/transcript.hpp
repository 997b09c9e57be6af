// Fiat-Shamir transcript of the protocol layer (transcript/src/lib.rs:30-121):
// a Merlin transcript (third-party crate merlin 3.0.0: STROBE-128 over
// Keccak-f[1600]) wrapped as IOPTranscript<Fr>.
//
// Keccak-f[1600] follows FIPS 202 section 3: the round constants come from
// the degree-8 LFSR of its algorithm 5 and the rho offsets and pi lane walk
// are generated from (x, y) -> (y, 2x + 3y), rather than tabulated.
// The STROBE subset is what Merlin v1.0 uses: meta-AD, AD and PRF operations
// with rate 166 bytes, initialised from the "STROBEv1.0.2" domain block.
// Merlin frames each message as meta-AD(label) || meta-AD(u32 LE length)
// followed by AD(message) or PRF(length bytes).
#pragma once
#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "ff.hpp"

namespace hyperplonk {
namespace host {

// FIPS 202 algorithm 5
constexpr int keccak_rc_bit(int t) {
  if (t % 255 == 0) return 1;
  unsigned r = 1;
  for (int i = 1; i <= t % 255; ++i) {
    r <<= 1;
    if (r & 0x100) r ^= 0x171;
  }
  return r & 1;
}

struct KeccakTables {
  uint64_t rc[24];
  int rho[25];
};

// evaluated by the compiler, so the permutation's rotations are constants
constexpr KeccakTables make_keccak_tables() {
  KeccakTables t{};
  for (int ir = 0; ir < 24; ++ir) {
    uint64_t rc = 0;
    for (int j = 0; j <= 6; ++j)
      if (keccak_rc_bit(j + 7 * ir)) rc |= uint64_t(1) << ((1 << j) - 1);
    t.rc[ir] = rc;
  }
  t.rho[0] = 0;
  int x = 1, y = 0;
  for (int k = 0; k < 24; ++k) {
    t.rho[x + 5 * y] = ((k + 1) * (k + 2) / 2) % 64;
    const int nx = y, ny = (2 * x + 3 * y) % 5;
    x = nx;
    y = ny;
  }
  return t;
}

class Keccak {
 public:
  static void permute(uint8_t state[200]) {
    static constexpr KeccakTables t = make_keccak_tables();
    uint64_t a[25];
    std::memcpy(a, state, 200);  // lanes are little-endian 64-bit words
    for (int round = 0; round < 24; ++round) {
      uint64_t c[5], d[5], b[25];
#pragma GCC unroll 5
      for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
#pragma GCC unroll 5
      for (int x = 0; x < 5; ++x) d[x] = c[(x + 4) % 5] ^ rotl(c[(x + 1) % 5], 1);
#pragma GCC unroll 5
      for (int x = 0; x < 5; ++x)
#pragma GCC unroll 5
        for (int y = 0; y < 5; ++y)
          b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl(a[x + 5 * y] ^ d[x], t.rho[x + 5 * y]);
#pragma GCC unroll 5
      for (int y = 0; y < 5; ++y)
#pragma GCC unroll 5
        for (int x = 0; x < 5; ++x)
          a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
      a[0] ^= t.rc[round];
    }
    std::memcpy(state, a, 200);
  }

 private:
  static uint64_t rotl(uint64_t v, int n) { return n ? (v << n) | (v >> (64 - n)) : v; }
};

// SHA3-256 (FIPS 202: rate 136, domain bits 01 then pad10*1); used by the
// tests to pin the permutation against an independent implementation.
inline std::array<uint8_t, 32> sha3_256(const uint8_t* data, size_t len) {
  uint8_t st[200] = {};
  const size_t rate = 136;
  size_t pos = 0;
  for (size_t i = 0; i < len; ++i) {
    st[pos++] ^= data[i];
    if (pos == rate) {
      Keccak::permute(st);
      pos = 0;
    }
  }
  st[pos] ^= 0x06;
  st[rate - 1] ^= 0x80;
  Keccak::permute(st);
  std::array<uint8_t, 32> out{};
  std::memcpy(out.data(), st, 32);
  return out;
}

class Strobe128 {
 public:
  explicit Strobe128(const std::string& protocol_label) {
    st_[0] = 1;
    st_[1] = kRate + 2;
    st_[2] = 1;
    st_[3] = 0;
    st_[4] = 1;
    st_[5] = 96;
    std::memcpy(st_ + 6, "STROBEv1.0.2", 12);
    Keccak::permute(st_);
    meta_ad(reinterpret_cast<const uint8_t*>(protocol_label.data()), protocol_label.size(), false);
  }
  void meta_ad(const uint8_t* data, size_t len, bool more) {
    begin_op(kFlagM | kFlagA, more);
    absorb(data, len);
  }
  void ad(const uint8_t* data, size_t len, bool more) {
    begin_op(kFlagA, more);
    absorb(data, len);
  }
  void prf(uint8_t* out, size_t len, bool more) {
    begin_op(kFlagI | kFlagA | kFlagC, more);
    squeeze(out, len);
  }

 private:
  static constexpr uint8_t kRate = 166;
  static constexpr uint8_t kFlagI = 1, kFlagA = 2, kFlagC = 4, kFlagT = 8, kFlagM = 16,
                           kFlagK = 32;

  void run_f() {
    st_[pos_] ^= pos_begin_;
    st_[pos_ + 1] ^= 0x04;
    st_[kRate + 1] ^= 0x80;
    Keccak::permute(st_);
    pos_ = 0;
    pos_begin_ = 0;
  }
  void absorb(const uint8_t* data, size_t len) {
    for (size_t i = 0; i < len; ++i) {
      st_[pos_++] ^= data[i];
      if (pos_ == kRate) run_f();
    }
  }
  void squeeze(uint8_t* out, size_t len) {
    for (size_t i = 0; i < len; ++i) {
      out[i] = st_[pos_];
      st_[pos_++] = 0;
      if (pos_ == kRate) run_f();
    }
  }
  void begin_op(uint8_t flags, bool more) {
    if (more) {
      if (flags != cur_flags_) throw std::logic_error("strobe: continued operation changed flags");
      return;
    }
    if (flags & kFlagT) throw std::logic_error("strobe: transport operations are not supported");
    const uint8_t old_begin = pos_begin_;
    pos_begin_ = static_cast<uint8_t>(pos_ + 1);
    cur_flags_ = flags;
    const uint8_t frame[2] = {old_begin, flags};
    absorb(frame, 2);
    if ((flags & (kFlagC | kFlagK)) && pos_ != 0) run_f();
  }

  uint8_t st_[200] = {};
  uint8_t pos_ = 0, pos_begin_ = 0, cur_flags_ = 0;
};

class MerlinTranscript {
 public:
  explicit MerlinTranscript(const std::string& label) : strobe_("Merlin v1.0") {
    append_message("dom-sep", reinterpret_cast<const uint8_t*>(label.data()), label.size());
  }
  void append_message(const std::string& label, const uint8_t* msg, size_t len) {
    const uint8_t n[4] = {static_cast<uint8_t>(len), static_cast<uint8_t>(len >> 8),
                          static_cast<uint8_t>(len >> 16), static_cast<uint8_t>(len >> 24)};
    strobe_.meta_ad(reinterpret_cast<const uint8_t*>(label.data()), label.size(), false);
    strobe_.meta_ad(n, 4, true);
    strobe_.ad(msg, len, false);
  }
  void challenge_bytes(const std::string& label, uint8_t* out, size_t len) {
    const uint8_t n[4] = {static_cast<uint8_t>(len), static_cast<uint8_t>(len >> 8),
                          static_cast<uint8_t>(len >> 16), static_cast<uint8_t>(len >> 24)};
    strobe_.meta_ad(reinterpret_cast<const uint8_t*>(label.data()), label.size(), false);
    strobe_.meta_ad(n, 4, true);
    strobe_.prf(out, len, false);
  }

 private:
  Strobe128 strobe_;
};

struct TranscriptError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// transcript/src/lib.rs:30-121
class IOPTranscript {
 public:
  explicit IOPTranscript(const std::string& label) : t_(label) {}

  void append_message(const std::string& label, const std::vector<uint8_t>& msg) {
    t_.append_message(label, msg.data(), msg.size());
    empty_ = false;
  }
  void append_message(const std::string& label, const uint8_t* msg, size_t len) {
    t_.append_message(label, msg, len);
    empty_ = false;
  }
  // ark-serialize of a field element: canonical LE bytes
  void append_field_element(const std::string& label, const Fr& f) {
    const auto b = f.bytes();
    append_message(label, b.data(), b.size());
  }
  // 64 PRF bytes -> from_le_bytes_mod_order -> appended back under the same label
  Fr get_and_append_challenge(const std::string& label) {
    if (empty_) throw TranscriptError("transcript is empty");
    uint8_t buf[64];
    t_.challenge_bytes(label, buf, 64);
    const Fr c = Fr::from_le_bytes_mod_order(buf, 64);
    append_field_element(label, c);
    return c;
  }
  std::vector<Fr> get_and_append_challenge_vectors(const std::string& label, size_t len) {
    if (empty_) throw TranscriptError("transcript is empty");
    std::vector<Fr> out;
    out.reserve(len);
    for (size_t i = 0; i < len; ++i) out.push_back(get_and_append_challenge(label));
    return out;
  }

 private:
  MerlinTranscript t_;
  bool empty_ = true;
};

}  // namespace host
}  // namespace hyperplonk
