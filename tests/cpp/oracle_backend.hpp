// Test infrastructure: the CPU oracle (oracle/liboracle.so) behind the same
// backend interface the protocol layer drives (Mle / VirtualPolynomial /
// PreparedProverParam / ProverState and the MleBackend, MultilinearKzgBackend
// and SumCheckProver methods), so one protocol run on the oracle and one on
// the MI355X backend can be compared byte for byte.  Never linked into the
// product.
#pragma once
#include <memory>
#include <optional>
#include <stdexcept>
#include <utility>
#include <vector>

extern "C" {
#include "../../oracle/oracle.h"
}
#include "../../hyperplonk_amd/host/virtual_polynomial.hpp"

namespace oracle_cpu {

using hyperplonk::FrBytes;
using hyperplonk::G1Packet;

inline void ok(int st, const char* what) {
  if (st != ORACLE_OK)
    throw hyperplonk::BackendError(st == ORACLE_INVALID_PARAMETERS
                                       ? hyperplonk::BackendError::Kind::InvalidParameters
                                       : hyperplonk::BackendError::Kind::InvalidState,
                                   what);
}

class Mle {
 public:
  Mle() = default;
  Mle(size_t nv, std::vector<uint8_t> t)
      : s_(std::make_shared<State>(State{nv, std::move(t)})) {}
  size_t num_vars() const { return s_->nv; }
  uintptr_t id() const { return reinterpret_cast<uintptr_t>(s_.get()); }
  const uint8_t* data() const { return s_->table.data(); }
  size_t size() const { return size_t(1) << s_->nv; }

 private:
  struct State {
    size_t nv;
    std::vector<uint8_t> table;
  };
  std::shared_ptr<State> s_;
};

struct PreparedProverParam {
  size_t num_vars = 0;
  std::vector<std::vector<uint8_t>> layers;  // layer i: 2^(N-i) packets
};

struct ProverState {
  std::shared_ptr<void> raw;
  size_t max_degree = 0;
};

class OracleBackend {
 public:
  using Mle = oracle_cpu::Mle;
  using VirtualPolynomial = hyperplonk::BasicVirtualPolynomial<Mle>;
  using PreparedProverParam = oracle_cpu::PreparedProverParam;
  using ProverState = oracle_cpu::ProverState;

  explicit OracleBackend(int threads = 8) : threads_(threads) { oracle_init(); }

  Mle mle_from_evaluations(size_t nv, const std::vector<FrBytes>& e) const {
    if (e.size() != (size_t(1) << nv)) hyperplonk::invalid("wrong evaluation count");
    std::vector<uint8_t> t(e.size() * 32);
    for (size_t i = 0; i < e.size(); ++i) std::copy(e[i].begin(), e[i].end(), t.begin() + 32 * i);
    return Mle(nv, std::move(t));
  }
  Mle linear_combination(const std::vector<Mle>& ps, const std::vector<FrBytes>& cs) const {
    const size_t nv = ps.at(0).num_vars(), n = size_t(1) << nv;
    std::vector<uint8_t> tables(ps.size() * n * 32), coeffs(cs.size() * 32);
    for (size_t i = 0; i < ps.size(); ++i) std::copy(ps[i].data(), ps[i].data() + 32 * n, tables.begin() + 32 * n * i);
    for (size_t i = 0; i < cs.size(); ++i) std::copy(cs[i].begin(), cs[i].end(), coeffs.begin() + 32 * i);
    std::vector<uint8_t> out(32 * n);
    ok(oracle_mle_linear_combination(tables.data(), coeffs.data(), ps.size(), nv, out.data()), "lincomb");
    return Mle(nv, std::move(out));
  }
  FrBytes evaluate_mle(const Mle& m, const std::vector<FrBytes>& point) const {
    FrBytes out{};
    ok(oracle_mle_evaluate(m.data(), m.num_vars(), point.empty() ? nullptr : point[0].data(), out.data()),
       "evaluate");
    return out;
  }
  std::vector<FrBytes> evaluate_mles(const std::vector<Mle>& ms,
                                     const std::vector<std::vector<FrBytes>>& points) const {
    std::vector<FrBytes> out;
    for (size_t i = 0; i < ms.size(); ++i) out.push_back(evaluate_mle(ms[i], points[i]));
    return out;
  }
  Mle build_eq_x_r(const std::vector<FrBytes>& r) const {
    std::vector<uint8_t> out(32 << r.size());
    ok(oracle_eq_x_r(r[0].data(), r.size(), out.data()), "eq");
    return Mle(r.size(), std::move(out));
  }
  Mle compute_frac_poly(const std::vector<Mle>& f, const std::vector<Mle>& g) const {
    const size_t nv = f.at(0).num_vars(), n = size_t(1) << nv;
    std::vector<uint8_t> a = concat(f, n), b = concat(g, n), out(32 * n);
    ok(oracle_frac(a.data(), b.data(), f.size(), nv, out.data()), "frac");
    return Mle(nv, std::move(out));
  }
  Mle compute_product_poly(const Mle& frac) const {
    std::vector<uint8_t> out(32 * frac.size());
    ok(oracle_product(frac.data(), frac.num_vars(), out.data()), "product");
    return Mle(frac.num_vars(), std::move(out));
  }
  std::pair<std::vector<Mle>, std::vector<Mle>> compute_nums_and_denoms(
      const FrBytes& beta, const FrBytes& gamma, const std::vector<Mle>& f, const std::vector<Mle>& g,
      const std::vector<Mle>& p) const {
    const size_t nv = f.at(0).num_vars(), n = size_t(1) << nv, k = f.size();
    std::vector<uint8_t> a = concat(f, n), b = concat(g, n), c = concat(p, n);
    std::vector<uint8_t> on(32 * n * k), od(32 * n * k);
    ok(oracle_nums_denoms(beta.data(), gamma.data(), a.data(), b.data(), c.data(), k, nv, on.data(), od.data()),
       "nums_denoms");
    std::vector<Mle> nums, dens;
    for (size_t i = 0; i < k; ++i) {
      nums.emplace_back(nv, std::vector<uint8_t>(on.begin() + 32 * n * i, on.begin() + 32 * n * (i + 1)));
      dens.emplace_back(nv, std::vector<uint8_t>(od.begin() + 32 * n * i, od.begin() + 32 * n * (i + 1)));
    }
    return {nums, dens};
  }
  std::array<Mle, 2> compute_product_factors(const Mle& frac, const Mle& prod) const {
    std::vector<uint8_t> a(32 * frac.size()), b(32 * frac.size());
    ok(oracle_product_factors(frac.data(), prod.data(), frac.num_vars(), a.data(), b.data()), "factors");
    return {Mle(frac.num_vars(), std::move(a)), Mle(frac.num_vars(), std::move(b))};
  }

  PreparedProverParam generate_prover_param(size_t N, const std::vector<FrBytes>& t) const {
    PreparedProverParam pp;
    pp.num_vars = N;
    for (size_t i = 0; i <= N; ++i) {
      const size_t len = N - i;
      std::vector<uint8_t> tt;
      for (size_t j = i; j < N; ++j) tt.insert(tt.end(), t[j].begin(), t[j].end());
      std::vector<uint8_t> layer(97 << len);
      ok(oracle_srs_layer(len ? tt.data() : nullptr, len, threads_, layer.data()), "srs layer");
      pp.layers.push_back(std::move(layer));
    }
    return pp;
  }
  G1Packet commit(const PreparedProverParam& pp, const Mle& m) const {
    if (m.num_vars() > pp.num_vars) hyperplonk::invalid("nv exceeds params");
    G1Packet out{};
    ok(oracle_msm_pippenger(pp.layers[pp.num_vars - m.num_vars()].data(), m.data(), m.size(), threads_,
                            out.data()),
       "commit");
    return out;
  }
  std::vector<G1Packet> multi_commit(const PreparedProverParam& pp, const std::vector<Mle>& ms) const {
    std::vector<G1Packet> out;
    for (const auto& m : ms) out.push_back(commit(pp, m));
    return out;
  }
  std::pair<std::vector<G1Packet>, FrBytes> open(const PreparedProverParam& pp, const Mle& m,
                                                  const std::vector<FrBytes>& point) const {
    const size_t nv = m.num_vars();
    std::vector<uint8_t> layers;
    for (size_t i = pp.num_vars - nv + 1; i <= pp.num_vars; ++i)
      layers.insert(layers.end(), pp.layers[i].begin(), pp.layers[i].end());
    std::vector<uint8_t> q(97 * std::max<size_t>(nv, 1));
    FrBytes ev{};
    ok(oracle_kzg_open(layers.data(), m.data(), nv, point.empty() ? nullptr : point[0].data(), threads_,
                       q.data(), ev.data()),
       "open");
    std::vector<G1Packet> proofs(nv);
    for (size_t j = 0; j < nv; ++j) std::copy(q.begin() + 97 * j, q.begin() + 97 * (j + 1), proofs[j].begin());
    return {proofs, ev};
  }

  ProverState prover_init(const VirtualPolynomial& vp) const {
    const auto& fs = vp.flattened_ml_extensions();
    const size_t nv = vp.aux_info.num_variables;
    std::vector<uint8_t> tables = concat(fs, size_t(1) << nv), coeffs;
    std::vector<size_t> offsets{0}, idx;
    for (const auto& [c, ids] : vp.products) {
      coeffs.insert(coeffs.end(), c.begin(), c.end());
      idx.insert(idx.end(), ids.begin(), ids.end());
      offsets.push_back(idx.size());
    }
    int st = 0;
    void* raw = oracle_sumcheck_create(nv, vp.aux_info.max_degree, tables.data(), fs.size(),
                                       coeffs.empty() ? nullptr : coeffs.data(), offsets.data(),
                                       vp.products.size(), idx.empty() ? nullptr : idx.data(), idx.size(), &st);
    ok(st, "sumcheck create");
    ProverState s;
    s.raw = std::shared_ptr<void>(raw, [](void* p) { oracle_sumcheck_destroy(p); });
    s.max_degree = vp.aux_info.max_degree;
    return s;
  }
  std::vector<FrBytes> prove_round_and_update_state(ProverState& s, const std::optional<FrBytes>& c) const {
    std::vector<FrBytes> out(s.max_degree + 1);
    ok(oracle_sumcheck_round(s.raw.get(), c ? c->data() : nullptr, out[0].data(), out.size()), "round");
    return out;
  }

 private:
  static std::vector<uint8_t> concat(const std::vector<Mle>& ms, size_t n) {
    std::vector<uint8_t> out(ms.size() * n * 32);
    for (size_t i = 0; i < ms.size(); ++i) std::copy(ms[i].data(), ms[i].data() + 32 * n, out.begin() + 32 * n * i);
    return out;
  }
  int threads_;
};

}  // namespace oracle_cpu
