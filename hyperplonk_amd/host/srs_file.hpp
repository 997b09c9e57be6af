// Canonical SRS files: MultilinearUniversalParams serialized with
// ark-serialize, compressed (subroutines/src/pcs/multilinear_kzg/srs.rs:62-76;
// read and written by hyperplonk/examples/full_prove_bench.rs:665-760 and
// hyperplonk/benches/bench.rs:53-61):
//   num_vars u64
//   powers_of_g: u64 count (num_vars + 1), then per layer u64 len + len G1
//                (layer i holds 2^(num_vars - i) points, the last one [g])
//   g: G1 | h: G2 | h_mask: u64 len (num_vars) + len G2
// Compressed G1 is 48 bytes (x big-endian, flags 0x80 compressed, 0x40
// infinity, 0x20 y > -y on the first byte); compressed G2 is 96 bytes
// (x.c1 then x.c0, big-endian, the same flags, and y > -y compares y.c1
// first, then y.c0).  The big layers stay compressed on the host: they are
// decompressed on the device when the prover parameters are prepared
// (multilinear_kzg_params_upload_layer_compressed).  Reading performs no
// subgroup check, as deserialize_compressed_unchecked.
#pragma once
#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "pairing.hpp"
#include "serialize.hpp"

namespace hyperplonk {
namespace host {

struct SrsFileError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

namespace srs_detail {

// (q + add) >> shift for add in {+1, -1, -3}, little-endian limbs
inline Limbs<6> shifted_q(int add, int shift) {
  Limbs<6> e = FqSpec::MOD;
  if (add > 0) {
    Limbs<6> a{};
    a[0] = static_cast<uint64_t>(add);
    u64 carry = 0;
    for (int i = 0; i < 6; ++i) {
      const unsigned __int128 s = (unsigned __int128)e[i] + a[i] + carry;
      e[i] = static_cast<uint64_t>(s);
      carry = static_cast<uint64_t>(s >> 64);
    }
  } else {
    Limbs<6> s{};
    s[0] = static_cast<uint64_t>(-add);
    limbs_sub(e, s);
  }
  for (int k = 0; k < shift; ++k)
    for (int i = 0; i < 6; ++i) e[i] = (e[i] >> 1) | (i + 1 < 6 ? e[i + 1] << 63 : 0);
  return e;
}

// q = 3 mod 4: a^((q + 1) / 4) is a root whenever one exists
inline bool fq_sqrt(const Fq& a, Fq* out) {
  static const Limbs<6> e = shifted_q(1, 2);
  const Fq r = a.pow(e);
  if (r.square() != a) return false;
  *out = r;
  return true;
}

inline Fq2 fq2_pow(const Fq2& a, const Limbs<6>& e) {
  Fq2 r = Fq2::one();
  for (size_t i = 6; i-- > 0;)
    for (int b = 63; b >= 0; --b) {
      r = r.square();
      if ((e[i] >> b) & 1) r = r * a;
    }
  return r;
}

// Square root in Fq2 = Fq[u]/(u^2 + 1) for q = 3 mod 4: Adj and
// Rodriguez-Henriquez, "Square root computation over even extension fields",
// algorithm 9.
inline bool fq2_sqrt(const Fq2& a, Fq2* out) {
  if (a.is_zero()) {
    *out = a;
    return true;
  }
  static const Limbs<6> e1 = shifted_q(-3, 2), e2 = shifted_q(-1, 1);
  const Fq2 a1 = fq2_pow(a, e1);
  const Fq2 alpha = a1.square() * a;
  const Fq2 a0 = Fq2{alpha.c0, -alpha.c1} * alpha;  // alpha^q * alpha
  const Fq2 minus_one{-Fq::one(), Fq::zero()};
  if (a0 == minus_one) return false;
  const Fq2 x0 = a1 * a;
  const Fq2 x = alpha == minus_one ? Fq2{-x0.c1, x0.c0}  // u * x0
                                   : fq2_pow(Fq2::one() + alpha, e2) * x0;
  if (x.square() != a) return false;
  *out = x;
  return true;
}

inline bool fq2_largest(const Fq2& y) {
  const Fq2 n = -y;
  const Limbs<6> y1 = y.c1.canonical(), n1 = n.c1.canonical();
  if (y1 != n1) return limbs_less(n1, y1);
  return limbs_less(n.c0.canonical(), y.c0.canonical());
}

inline void put_be(const Fq& f, uint8_t* out) {
  uint8_t le[48];
  f.to_bytes(le);
  for (int i = 0; i < 48; ++i) out[i] = le[47 - i];
}

inline Fq get_be(const uint8_t* in) {
  uint8_t le[48];
  for (int i = 0; i < 48; ++i) le[i] = in[47 - i];
  le[47] &= 0x1f;  // flag bits
  Fq f;
  if (!Fq::from_bytes(le, &f)) throw SrsFileError("non-canonical coordinate");
  return f;
}

inline bool infinity_bits_clear(const uint8_t* in, size_t len) {
  if (in[0] != 0xC0) return false;
  for (size_t i = 1; i < len; ++i)
    if (in[i]) return false;
  return true;
}

}  // namespace srs_detail

inline std::array<uint8_t, 96> g2_compressed(const G2Aff& p) {
  std::array<uint8_t, 96> out{};
  if (p.infinity) {
    out[0] = 0xC0;
    return out;
  }
  srs_detail::put_be(p.x.c1, out.data());
  srs_detail::put_be(p.x.c0, out.data() + 48);
  out[0] |= srs_detail::fq2_largest(p.y) ? 0xA0 : 0x80;
  return out;
}

inline G1Aff g1_decompress(const uint8_t* in) {
  const uint8_t flags = in[0];
  if (!(flags & 0x80)) throw SrsFileError("G1 point is not compressed");
  if (flags & 0x40) {
    if (!srs_detail::infinity_bits_clear(in, 48)) throw SrsFileError("bad G1 infinity encoding");
    return {};
  }
  const Fq x = srs_detail::get_be(in);
  Fq y;
  if (!srs_detail::fq_sqrt(x.square() * x + Fq::from_u64(4), &y))
    throw SrsFileError("G1 x is not on the curve");
  if (y.lexicographically_largest() != bool(flags & 0x20)) y = -y;
  return {x, y, false};
}

inline G2Aff g2_decompress(const uint8_t* in) {
  const uint8_t flags = in[0];
  if (!(flags & 0x80)) throw SrsFileError("G2 point is not compressed");
  if (flags & 0x40) {
    if (!srs_detail::infinity_bits_clear(in, 96)) throw SrsFileError("bad G2 infinity encoding");
    return {};
  }
  const Fq2 x{srs_detail::get_be(in + 48), srs_detail::get_be(in)};
  const Fq2 b{Fq::from_u64(4), Fq::from_u64(4)};  // 4 (1 + u)
  Fq2 y;
  if (!srs_detail::fq2_sqrt(x.square() * x + b, &y)) throw SrsFileError("G2 x is not on the curve");
  if (srs_detail::fq2_largest(y) != bool(flags & 0x20)) y = -y;
  return {x, y, false};
}

// MultilinearUniversalParams with its G1 layers kept in compressed form.
struct SrsFile {
  size_t num_vars = 0;
  std::vector<std::vector<uint8_t>> layers;  // 48 bytes a point
  G1Aff g;
  G2Aff h;
  std::vector<G2Aff> h_mask;  // h * t_i

  // MultilinearUniversalParams::trim's verifier half (srs.rs:126-153)
  MultilinearVerifierParam verifier_param(size_t nv) const {
    MultilinearVerifierParam vp;
    vp.num_vars = num_vars;
    vp.g = g;
    vp.h = h;
    vp.h_mask = h_mask;
    return vp.trim(nv);
  }
  // the prover half: layers num_vars - nv .. num_vars
  std::vector<const std::vector<uint8_t>*> prover_layers(size_t nv) const {
    if (nv > num_vars) throw PCSError("SRS does not support target number of vars");
    std::vector<const std::vector<uint8_t>*> out;
    for (size_t i = num_vars - nv; i <= num_vars; ++i) out.push_back(&layers[i]);
    return out;
  }
};

inline std::vector<uint8_t> serialize_srs(const SrsFile& s) {
  Writer w;
  w.u64(s.num_vars);
  w.u64(s.layers.size());
  for (const auto& layer : s.layers) {
    w.u64(layer.size() / 48);
    w.buf.insert(w.buf.end(), layer.begin(), layer.end());
  }
  w.g1(g1_to_packet(s.g));
  const auto h = g2_compressed(s.h);
  w.buf.insert(w.buf.end(), h.begin(), h.end());
  w.u64(s.h_mask.size());
  for (const auto& m : s.h_mask) {
    const auto c = g2_compressed(m);
    w.buf.insert(w.buf.end(), c.begin(), c.end());
  }
  return std::move(w.buf);
}

// Parse and check the shape full_prove_bench.rs:627-663 validates: num_vars
// + 1 layers of 2^(num_vars - i) points and num_vars masks; g, h and the
// masks are decompressed here, the layers on the device.
inline SrsFile deserialize_srs(const std::vector<uint8_t>& buf) {
  size_t pos = 0;
  auto need = [&](size_t n) {
    if (n > buf.size() - pos) throw SrsFileError("truncated SRS file");
  };
  auto u64v = [&] {
    need(8);
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= static_cast<uint64_t>(buf[pos + i]) << (8 * i);
    pos += 8;
    return v;
  };
  SrsFile s;
  s.num_vars = u64v();
  if (s.num_vars > 30) throw SrsFileError("SRS num_vars out of range");
  if (u64v() != s.num_vars + 1) throw SrsFileError("SRS layer count is not num_vars + 1");
  for (size_t l = 0; l <= s.num_vars; ++l) {
    const uint64_t n = u64v();
    if (n != (uint64_t(1) << (s.num_vars - l))) throw SrsFileError("SRS layer has the wrong length");
    need(48 * n);
    s.layers.emplace_back(buf.begin() + pos, buf.begin() + pos + 48 * n);
    pos += 48 * n;
  }
  need(48);
  s.g = g1_decompress(buf.data() + pos);
  pos += 48;
  need(96);
  s.h = g2_decompress(buf.data() + pos);
  pos += 96;
  if (u64v() != s.num_vars) throw SrsFileError("SRS h_mask length is not num_vars");
  for (size_t i = 0; i < s.num_vars; ++i) {
    need(96);
    s.h_mask.push_back(g2_decompress(buf.data() + pos));
    pos += 96;
  }
  if (pos != buf.size()) throw SrsFileError("trailing bytes after the SRS");
  return s;
}

inline void write_srs_file(const std::string& path, const SrsFile& s) {
  const std::vector<uint8_t> bytes = serialize_srs(s);
  // never overwrite: "x" fails if the path exists (create_new)
  FILE* f = std::fopen(path.c_str(), "wbx");
  if (!f) throw SrsFileError("create canonical SRS file " + path + " without overwriting failed");
  const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
  const bool closed = std::fclose(f) == 0;
  if (!ok || !closed) throw SrsFileError("write canonical SRS to " + path + " failed");
}

inline SrsFile read_srs_file(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw SrsFileError("open SRS file " + path + " failed");
  std::vector<uint8_t> buf;
  if (std::fseek(f, 0, SEEK_END) == 0) {
    const long size = std::ftell(f);
    if (size > 0) buf.resize(static_cast<size_t>(size));
    std::rewind(f);
  }
  const bool ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  std::fclose(f);
  if (!ok) throw SrsFileError("read SRS file " + path + " failed");
  return deserialize_srs(buf);
}

// load_or_generate_srs (full_prove_bench.rs:665-760): load the file at
// `path` when it exists; otherwise generate from the trapdoor t on the
// device (g and h the standard generators) and, with a path, write the file
// without overwriting.
struct SrsSetup {
  SrsFile srs;
  std::string source;  // loaded_canonical_file | generated_and_written_canonical_file | generated_in_memory
  double setup_ms = 0, generation_ms = 0, load_ms = 0;
};

template <class B>
SrsSetup load_or_generate_srs(const B& backend, size_t num_vars, const std::vector<Fr>& t,
                                 const std::string& path) {
  using Clock = std::chrono::steady_clock;
  auto ms = [](Clock::time_point a) {
    return std::chrono::duration<double, std::milli>(Clock::now() - a).count();
  };
  const auto start = Clock::now();
  SrsSetup out;
  struct stat st;
  if (!path.empty() && ::stat(path.c_str(), &st) == 0) {
    out.srs = read_srs_file(path);
    out.load_ms = ms(start);
    if (out.srs.num_vars < num_vars)
      throw SrsFileError("SRS " + path + " supports fewer variables than requested");
    out.source = "loaded_canonical_file";
  } else {
    if (t.size() != num_vars) throw SrsFileError("trapdoor length must equal num_vars");
    // refuse to generate when the file could never be stored
    const size_t slash = path.rfind('/');
    if (!path.empty() && slash != std::string::npos && slash > 0) {
      const std::string parent = path.substr(0, slash);
      if (::stat(parent.c_str(), &st) != 0 || !S_ISDIR(st.st_mode))
        throw SrsFileError("SRS parent " + parent + " is not a directory");
    }
    const auto pp = backend.generate_prover_param(num_vars, to_wire(t));
    out.srs.num_vars = num_vars;
    out.srs.layers = backend.download_prover_param(pp, true);
    const MultilinearVerifierParam vp = MultilinearVerifierParam::from_trapdoor(t);
    out.srs.g = vp.g;
    out.srs.h = vp.h;
    out.srs.h_mask = vp.h_mask;
    out.generation_ms = ms(start);
    if (path.empty()) {
      out.source = "generated_in_memory";
    } else {
      write_srs_file(path, out.srs);
      out.source = "generated_and_written_canonical_file";
    }
  }
  out.setup_ms = ms(start);
  return out;
}

}  // namespace host
}  // namespace hyperplonk
