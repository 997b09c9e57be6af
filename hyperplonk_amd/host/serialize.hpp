// ark-serialize compressed encodings of the values the protocol appends to
// its transcript (transcript/src/lib.rs:54-70, to_bytes! :125-131):
//   usize          u64 little-endian
//   Vec<T>         u64 length, then the elements
//   Fr             32 bytes, canonical little-endian
//   G1Affine       48 bytes: x big-endian with flags in the top three bits of
//                  the first byte (0x80 compressed, 0x40 point at infinity,
//                  0x20 y is the larger of y and -y); infinity is 0xC0 then zeros
//                  (the zcash BLS12-381 format ark-bls12-381 0.6 implements)
//   Commitment     its G1Affine (a one-field struct)
//   VPAuxInfo      max_degree, num_variables (usize each; PhantomData is empty)
//   IOPProverMessage  its evaluations as Vec<Fr>
#pragma once
#include <cstdint>
#include <cstring>
#include <vector>

#include "ff.hpp"
#include "virtual_polynomial.hpp"

namespace hyperplonk {
namespace host {

// packet: tag (0 finite, 1 identity) | x LE48 | y LE48
inline std::array<uint8_t, 48> g1_compressed(const G1Packet& p) {
  std::array<uint8_t, 48> out{};
  if (p[0] == 1) {
    out[0] = 0xC0;
    return out;
  }
  for (int i = 0; i < 48; ++i) out[i] = p[1 + 47 - i];
  Fq y;
  if (!Fq::from_bytes(p.data() + 49, &y)) throw std::invalid_argument("non-canonical G1 y");
  out[0] |= 0x80;
  if (y.lexicographically_largest()) out[0] |= 0x20;
  return out;
}

class Writer {
 public:
  std::vector<uint8_t> buf;
  void u64(uint64_t v) {
    for (int i = 0; i < 8; ++i) buf.push_back(static_cast<uint8_t>(v >> (8 * i)));
  }
  void fr(const Fr& f) {
    const auto b = f.bytes();
    buf.insert(buf.end(), b.begin(), b.end());
  }
  void frs(const std::vector<Fr>& v) {
    u64(v.size());
    for (const auto& f : v) fr(f);
  }
  void g1(const G1Packet& p) {
    const auto c = g1_compressed(p);
    buf.insert(buf.end(), c.begin(), c.end());
  }
  void g1s(const std::vector<G1Packet>& v) {
    u64(v.size());
    for (const auto& p : v) g1(p);
  }
  void aux(const VPAuxInfo& a) {
    u64(a.max_degree);
    u64(a.num_variables);
  }
};

inline std::vector<uint8_t> ser_frs(const std::vector<Fr>& v) {
  Writer w;
  w.frs(v);
  return w.buf;
}
inline std::vector<uint8_t> ser_g1(const G1Packet& p) {
  Writer w;
  w.g1(p);
  return w.buf;
}
inline std::vector<uint8_t> ser_aux(const VPAuxInfo& a) {
  Writer w;
  w.aux(a);
  return w.buf;
}

inline FrBytes to_wire(const Fr& f) { return f.bytes(); }
inline std::vector<FrBytes> to_wire(const std::vector<Fr>& v) {
  std::vector<FrBytes> out;
  out.reserve(v.size());
  for (const auto& f : v) out.push_back(f.bytes());
  return out;
}
inline Fr from_wire(const FrBytes& b) {
  Fr f;
  if (!Fr::from_bytes(b.data(), &f)) throw std::invalid_argument("non-canonical Fr from backend");
  return f;
}
inline std::vector<Fr> from_wire(const std::vector<FrBytes>& v) {
  std::vector<Fr> out;
  out.reserve(v.size());
  for (const auto& b : v) out.push_back(from_wire(b));
  return out;
}

}  // namespace host
}  // namespace hyperplonk
