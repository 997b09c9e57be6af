// C++ host mirror of the reference backend traits over the C ABI.
//
// The reference's protocol layer calls three Rust traits:
//   MleBackend            backend/src/multilinear_polynomial.rs:33-121
//   MultilinearKzgBackend backend/src/multilinear_kzg.rs:20-55
//   SumCheckProver        backend/src/sum_check.rs:17-38
// and its accelerator glue implements them over a C bridge
// (backend/src/ascend910b3/{mod,multilinear_polynomial,multilinear_kzg,sum_check}.rs).
// This header is that glue for the MI355X bridge (include/hp_mi355x.h): the
// same method names and argument meaning, BackendError kinds mapped as the
// reference maps native statuses (backend/src/ascend910b3/status.rs:23-44),
// MLE handles whose identity is the device address and whose last owner
// retires them, and a VirtualPolynomial with identity-deduplicated factors
// (backend/src/virtual_polynomial.rs:45-235).
//
// Field elements cross this interface as canonical little-endian 32-byte
// values (Fr) and G1 points as 97-byte packets, exactly as the C ABI takes
// them; the VirtualPolynomial container is shared with every other backend
// the protocol layer drives (virtual_polynomial.hpp).
#pragma once
#include <array>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/hp_mi355x.h"
#include "virtual_polynomial.hpp"

namespace hyperplonk {
namespace mi355x {

using Fr = FrBytes;
using hyperplonk::G1Packet;

inline Fr fr_from_u64(uint64_t v) {
  Fr r{};
  for (int i = 0; i < 8; ++i) r[i] = static_cast<uint8_t>(v >> (8 * i));
  return r;
}

using hyperplonk::BackendError;
using hyperplonk::invalid;

inline void check(HpMi355xStatus s, const char* what) {
  if (s.kind == HP_MI355X_STATUS_SUCCESS && s.native_code == 0) return;
  const auto kind = s.kind == HP_MI355X_STATUS_INVALID_PARAMETERS
                        ? BackendError::Kind::InvalidParameters
                        : BackendError::Kind::InvalidState;
  throw BackendError(kind, std::string(what) + ": status " + std::to_string(s.kind) +
                               " native " + std::to_string(s.native_code),
                     s.native_code);
}

// Owns the native context; every handle keeps it alive.
class Context {
 public:
  explicit Context(int32_t device) { check(hp_mi355x_v1_context_create(device, &raw_), "context_create"); }
  ~Context() {
    if (raw_) {
      const HpMi355xStatus s = hp_mi355x_v1_context_destroy(raw_);
      if (s.kind != 0) std::abort();  // a failed destroy is fatal (context.rs:10-18)
    }
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  HpMi355xContext* raw() const { return raw_; }
  // The native context is single-threaded: every call into it, including the
  // retirement of handles dropped on other threads, holds this lock.
  std::unique_lock<std::recursive_mutex> guard() const {
    return std::unique_lock<std::recursive_mutex>(mu_);
  }

 private:
  HpMi355xContext* raw_ = nullptr;
  mutable std::recursive_mutex mu_;
};

// Mle trait: shared handle to an immutable resident table.
class Mle {
 public:
  Mle() = default;
  size_t num_vars() const { return state_->num_vars; }
  uintptr_t id() const { return reinterpret_cast<uintptr_t>(state_->address); }
  HpMi355xDeviceAddress address() const { return state_->address; }
  const std::shared_ptr<Context>& context() const { return state_->ctx; }
  explicit operator bool() const { return static_cast<bool>(state_); }

 private:
  friend class Mi355xBackend;
  struct State {
    std::shared_ptr<Context> ctx;
    HpMi355xDeviceAddress address;
    size_t num_vars;
    ~State() {
      if (address) {
        const auto g = ctx->guard();
        (void)hp_mi355x_v1_buffer_retire(ctx->raw(), address);
      }
    }
  };
  std::shared_ptr<State> state_;
};

using VirtualPolynomial = BasicVirtualPolynomial<Mle>;

class PreparedProverParam {
 public:
  size_t num_vars() const { return state_->num_vars; }
  const HpMi355xKzgParams* raw() const { return state_->raw; }

 private:
  friend class Mi355xBackend;
  struct State {
    std::shared_ptr<Context> ctx;
    HpMi355xKzgParams* raw;
    size_t num_vars;
    ~State() {
      if (raw) {
        const auto g = ctx->guard();
        (void)hp_mi355x_v1_multilinear_kzg_params_destroy(ctx->raw(), raw);
      }
    }
  };
  std::shared_ptr<State> state_;
};

class ProverState {
 public:
  size_t max_degree() const { return state_->max_degree; }

 private:
  friend class Mi355xBackend;
  struct State {
    std::shared_ptr<Context> ctx;
    HpMi355xSumCheckState* raw;
    size_t max_degree;
    ~State() {
      if (raw) {
        const auto g = ctx->guard();
        (void)hp_mi355x_v1_sumcheck_prover_state_destroy(ctx->raw(), raw);
      }
    }
  };
  std::shared_ptr<State> state_;
};

class Mi355xBackend {
 public:
  using Mle = mi355x::Mle;
  using VirtualPolynomial = mi355x::VirtualPolynomial;
  using PreparedProverParam = mi355x::PreparedProverParam;
  using ProverState = mi355x::ProverState;

  explicit Mi355xBackend(int32_t device_id = 0) : ctx_(std::make_shared<Context>(device_id)) {}

  // ---- MleBackend --------------------------------------------------------
  Mle mle_from_evaluations(size_t num_vars, const std::vector<Fr>& evaluations) const {
    const auto lock = ctx_->guard();
    HpMi355xDeviceAddress out = nullptr;
    check(hp_mi355x_v1_mle_upload(ctx_->raw(), num_vars,
                                  evaluations.empty() ? nullptr : evaluations[0].data(),
                                  evaluations.size(), &out),
          "mle_from_evaluations");
    return wrap(out, num_vars);
  }

  std::vector<Fr> evaluations(const Mle& mle) const {  // extension: read back
    own(mle);
    std::vector<Fr> out(size_t(1) << mle.num_vars());
    check(hp_mi355x_v1_mle_download(ctx_->raw(), mle.address(), mle.num_vars(), out[0].data(),
                                    out.size()),
          "mle_download");
    return out;
  }

  Mle linear_combination(const std::vector<Mle>& polynomials,
                         const std::vector<Fr>& coefficients) const {
    const auto lock = ctx_->guard();
    if (polynomials.empty()) invalid("linear combination requires nonempty polynomials");
    if (polynomials.size() != coefficients.size())
      invalid("linear combination requires matching counts and dimensions");
    const size_t nv = polynomials[0].num_vars();
    std::vector<HpMi355xDeviceAddress> addrs;
    for (const auto& p : polynomials) {
      own(p);
      if (p.num_vars() != nv) invalid("linear combination requires matching counts and dimensions");
      addrs.push_back(p.address());
    }
    HpMi355xDeviceAddress out = nullptr;
    check(hp_mi355x_v1_mle_linear_combination(ctx_->raw(), nv, addrs.data(), coefficients[0].data(),
                                              addrs.size(), &out),
          "linear_combination");
    return wrap(out, nv);
  }

  Fr evaluate_mle(const Mle& mle, const std::vector<Fr>& point) const {
    const auto lock = ctx_->guard();
    own(mle);
    if (point.size() != mle.num_vars()) invalid("wrong number of variables");
    Fr out{};
    check(hp_mi355x_v1_mle_evaluate(ctx_->raw(), mle.address(), mle.num_vars(),
                                    point.empty() ? nullptr : point[0].data(), out.data()),
          "evaluate_mle");
    return out;
  }

  // extension: many evaluations in one pass (mle_evaluate_batch)
  std::vector<Fr> evaluate_mles(const std::vector<Mle>& mles,
                                const std::vector<std::vector<Fr>>& points) const {
    const auto lock = ctx_->guard();
    if (mles.size() != points.size()) invalid("one point per polynomial");
    std::vector<Fr> out(mles.size());
    if (mles.empty()) return out;
    std::vector<HpMi355xDeviceAddress> addrs;
    std::vector<size_t> nvs;
    std::vector<const uint8_t*> pts;
    for (size_t i = 0; i < mles.size(); ++i) {
      own(mles[i]);
      if (points[i].size() != mles[i].num_vars()) invalid("wrong number of variables");
      addrs.push_back(mles[i].address());
      nvs.push_back(mles[i].num_vars());
      pts.push_back(points[i].empty() ? nullptr : points[i][0].data());
    }
    check(hp_mi355x_v1_mle_evaluate_batch(ctx_->raw(), addrs.data(), nvs.data(), pts.data(),
                                          out[0].data(), out.size()),
          "evaluate_mles");
    return out;
  }

  Mle build_eq_x_r(const std::vector<Fr>& r) const {
    const auto lock = ctx_->guard();
    if (r.empty()) invalid("r length is 0");
    HpMi355xDeviceAddress out = nullptr;
    check(hp_mi355x_v1_mle_build_eq_x_r(ctx_->raw(), r[0].data(), r.size(), &out), "build_eq_x_r");
    return wrap(out, r.size());
  }

  Mle compute_frac_poly(const std::vector<Mle>& fxs, const std::vector<Mle>& gxs) const {
    const auto lock = ctx_->guard();
    if (fxs.empty() || fxs.size() != gxs.size()) invalid("frac requires equal nonempty lists");
    HpMi355xDeviceAddress out = nullptr;
    const auto f = addresses(fxs), g = addresses(gxs);
    check(hp_mi355x_v1_mle_compute_frac_poly(ctx_->raw(), fxs[0].num_vars(), f.data(), g.data(),
                                             f.size(), &out),
          "compute_frac_poly");
    return wrap(out, fxs[0].num_vars());
  }

  Mle compute_product_poly(const Mle& frac) const {
    const auto lock = ctx_->guard();
    own(frac);
    HpMi355xDeviceAddress out = nullptr;
    check(hp_mi355x_v1_mle_compute_product_poly(ctx_->raw(), frac.address(), frac.num_vars(), &out),
          "compute_product_poly");
    return wrap(out, frac.num_vars());
  }

  std::pair<std::vector<Mle>, std::vector<Mle>> compute_nums_and_denoms(
      const Fr& beta, const Fr& gamma, const std::vector<Mle>& fxs, const std::vector<Mle>& gxs,
      const std::vector<Mle>& perms) const {
    const auto lock = ctx_->guard();
    const size_t k = fxs.size();
    if (k == 0 || gxs.size() != k || perms.size() != k) invalid("nums/denoms require equal lists");
    const size_t nv = fxs[0].num_vars();
    const auto f = addresses(fxs), g = addresses(gxs), p = addresses(perms);
    std::vector<HpMi355xDeviceAddress> on(k), od(k);
    check(hp_mi355x_v1_mle_compute_nums_and_denoms(ctx_->raw(), nv, beta.data(), gamma.data(),
                                                   f.data(), g.data(), p.data(), k, on.data(),
                                                   od.data()),
          "compute_nums_and_denoms");
    std::vector<Mle> nums, dens;
    for (size_t i = 0; i < k; ++i) {
      nums.push_back(wrap(on[i], nv));
      dens.push_back(wrap(od[i], nv));
    }
    return {nums, dens};
  }

  std::array<Mle, 2> compute_product_factors(const Mle& frac, const Mle& prod) const {
    const auto lock = ctx_->guard();
    own(frac);
    own(prod);
    HpMi355xDeviceAddress a = nullptr, b = nullptr;
    check(hp_mi355x_v1_mle_compute_product_factors(ctx_->raw(), frac.address(), prod.address(),
                                                   frac.num_vars(), &a, &b),
          "compute_product_factors");
    return {wrap(a, frac.num_vars()), wrap(b, frac.num_vars())};
  }

  // ---- MultilinearKzgBackend -------------------------------------------
  // Layer i holds 2^(num_vars - i) concatenated 97-byte packets.
  PreparedProverParam prepare_prover_param(size_t num_vars,
                                           const std::vector<std::vector<uint8_t>>& layers) const {
    const auto lock = ctx_->guard();
    if (layers.size() != num_vars + 1) invalid("prover parameters need num_vars + 1 layers");
    PreparedProverParam pp = new_params(num_vars);
    for (size_t i = 0; i < layers.size(); ++i)
      check(hp_mi355x_v1_multilinear_kzg_params_upload_layer(
                ctx_->raw(), pp.state_->raw, i, layers[i].data(), layers[i].size() / 97),
            "upload_layer");
    return pp;
  }

  // The same from layers in the compressed encoding of SRS files (48 bytes a
  // point, decompressed on the device); layers[i] holds 2^(num_vars - i) points.
  PreparedProverParam prepare_prover_param_compressed(
      size_t num_vars, const std::vector<const std::vector<uint8_t>*>& layers) const {
    const auto lock = ctx_->guard();
    if (layers.size() != num_vars + 1) invalid("prover parameters need num_vars + 1 layers");
    PreparedProverParam pp = new_params(num_vars);
    for (size_t i = 0; i < layers.size(); ++i)
      check(hp_mi355x_v1_multilinear_kzg_params_upload_layer_compressed(
                ctx_->raw(), pp.state_->raw, i, layers[i]->data(), layers[i]->size() / 48),
            "upload_layer_compressed");
    return pp;
  }

  // Every layer of prepared parameters, as 97-byte packets or compressed.
  std::vector<std::vector<uint8_t>> download_prover_param(const PreparedProverParam& pp,
                                                          bool compressed) const {
    const auto lock = ctx_->guard();
    std::vector<std::vector<uint8_t>> out;
    for (size_t i = 0; i <= pp.num_vars(); ++i) {
      const size_t n = size_t(1) << (pp.num_vars() - i);
      out.emplace_back(n * (compressed ? 48 : 97));
      check(compressed ? hp_mi355x_v1_multilinear_kzg_params_download_layer_compressed(
                             ctx_->raw(), pp.raw(), i, out.back().data(), n)
                       : hp_mi355x_v1_multilinear_kzg_params_download_layer(
                             ctx_->raw(), pp.raw(), i, out.back().data(), n),
            "download_layer");
    }
    return out;
  }

  PreparedProverParam generate_prover_param(size_t num_vars, const std::vector<Fr>& t) const {
    const auto lock = ctx_->guard();
    if (t.size() != num_vars) invalid("trapdoor length must equal num_vars");
    PreparedProverParam pp = new_params(num_vars);
    check(hp_mi355x_v1_multilinear_kzg_params_generate(ctx_->raw(), pp.state_->raw,
                                                       t.empty() ? nullptr : t[0].data()),
          "params_generate");
    return pp;
  }

  G1Packet commit(const PreparedProverParam& pp, const Mle& polynomial) const {
    const auto lock = ctx_->guard();
    own(polynomial);
    G1Packet out{};
    check(hp_mi355x_v1_multilinear_kzg_commit(ctx_->raw(), pp.raw(), polynomial.address(),
                                              polynomial.num_vars(), out.data()),
          "commit");
    return out;
  }

  std::vector<G1Packet> multi_commit(const PreparedProverParam& pp,
                                     const std::vector<Mle>& polynomials) const {
    const auto lock = ctx_->guard();
    std::vector<G1Packet> out(polynomials.size());
    if (polynomials.empty()) return out;
    const auto addrs = addresses(polynomials);
    std::vector<size_t> nvs;
    for (const auto& p : polynomials) nvs.push_back(p.num_vars());
    check(hp_mi355x_v1_multilinear_kzg_multi_commit(ctx_->raw(), pp.raw(), addrs.data(), nvs.data(),
                                                    addrs.size(), out[0].data()),
          "multi_commit");
    return out;
  }

  std::pair<std::vector<G1Packet>, Fr> open(const PreparedProverParam& pp, const Mle& polynomial,
                                           const std::vector<Fr>& point) const {
    const auto lock = ctx_->guard();
    own(polynomial);
    const size_t nv = polynomial.num_vars();
    if (point.size() != nv) invalid("polynomial num_vars does not match point length");
    std::vector<G1Packet> proofs(nv);
    Fr value{};
    check(hp_mi355x_v1_multilinear_kzg_open(ctx_->raw(), pp.raw(), polynomial.address(), nv,
                                            nv ? point[0].data() : nullptr,
                                            nv ? proofs[0].data() : nullptr, value.data()),
          "open");
    return {proofs, value};
  }

  // extension: commits and opens of many polynomials in one batched MSM pass
  struct KzgOp {
    bool open = false;
    Mle polynomial;
    std::vector<Fr> point;  // opens only
  };
  struct KzgResult {
    std::vector<G1Packet> points;  // commit: 1; open: num_vars quotients
    Fr value{};                    // opens only
  };
  std::vector<KzgResult> kzg_batch(const PreparedProverParam& pp, const std::vector<KzgOp>& ops) const {
    const auto lock = ctx_->guard();
    std::vector<KzgResult> res(ops.size());
    std::vector<HpMi355xKzgRequest> reqs(ops.size());
    for (size_t i = 0; i < ops.size(); ++i) {
      const KzgOp& op = ops[i];
      own(op.polynomial);
      const size_t nv = op.polynomial.num_vars();
      if (op.open && op.point.size() != nv) invalid("polynomial num_vars does not match point length");
      res[i].points.resize(op.open ? nv : 1);
      reqs[i].kind = op.open ? HP_MI355X_KZG_OPEN : HP_MI355X_KZG_COMMIT;
      reqs[i].polynomial = op.polynomial.address();
      reqs[i].num_vars = nv;
      reqs[i].point = op.open && nv ? op.point[0].data() : nullptr;
      reqs[i].out_points = res[i].points.empty() ? nullptr : res[i].points[0].data();
      reqs[i].out_evaluation = op.open ? res[i].value.data() : nullptr;
    }
    if (!reqs.empty())
      check(hp_mi355x_v1_multilinear_kzg_batch(ctx_->raw(), pp.raw(), reqs.data(), reqs.size()),
            "kzg_batch");
    return res;
  }

  // extension: the same batch, submitted to run beside later calls; the
  // results are written by wait() (call it exactly once)
  class KzgPending {
   public:
    KzgPending() = default;
    KzgPending(KzgPending&&) = default;
    KzgPending& operator=(KzgPending&&) = default;
    std::vector<KzgResult> wait() {
      if (!ticket_) invalid("batch already waited for");
      const auto lock = ctx_->guard();
      HpMi355xKzgTicket* t = ticket_;
      ticket_ = nullptr;
      check(hp_mi355x_v1_multilinear_kzg_batch_wait(ctx_->raw(), t), "kzg_batch_wait");
      return std::move(*res_);
    }
    ~KzgPending() {
      if (ticket_) {
        const auto lock = ctx_->guard();
        (void)hp_mi355x_v1_multilinear_kzg_batch_wait(ctx_->raw(), ticket_);
      }
    }

   private:
    friend class Mi355xBackend;
    std::shared_ptr<Context> ctx_;
    HpMi355xKzgTicket* ticket_ = nullptr;
    std::unique_ptr<std::vector<KzgResult>> res_;  // stable output storage
  };
  KzgPending kzg_batch_submit(const PreparedProverParam& pp, const std::vector<KzgOp>& ops) const {
    const auto lock = ctx_->guard();
    KzgPending pending;
    pending.ctx_ = ctx_;
    pending.res_ = std::make_unique<std::vector<KzgResult>>(ops.size());
    std::vector<KzgResult>& res = *pending.res_;
    std::vector<HpMi355xKzgRequest> reqs(ops.size());
    for (size_t i = 0; i < ops.size(); ++i) {
      const KzgOp& op = ops[i];
      own(op.polynomial);
      const size_t nv = op.polynomial.num_vars();
      if (op.open && op.point.size() != nv) invalid("polynomial num_vars does not match point length");
      res[i].points.resize(op.open ? nv : 1);
      reqs[i].kind = op.open ? HP_MI355X_KZG_OPEN : HP_MI355X_KZG_COMMIT;
      reqs[i].polynomial = op.polynomial.address();
      reqs[i].num_vars = nv;
      reqs[i].point = op.open && nv ? op.point[0].data() : nullptr;
      reqs[i].out_points = res[i].points.empty() ? nullptr : res[i].points[0].data();
      reqs[i].out_evaluation = op.open ? res[i].value.data() : nullptr;
    }
    check(hp_mi355x_v1_multilinear_kzg_batch_submit(ctx_->raw(), pp.raw(), reqs.data(), reqs.size(),
                                                    &pending.ticket_),
          "kzg_batch_submit");
    return pending;
  }

  // extension: device memory and this context's live / peak bytes
  struct MemoryInfo {
    size_t device_free = 0, device_total = 0, live = 0, peak = 0;
  };
  MemoryInfo memory_info(bool reset_peak = false) const {
    const auto lock = ctx_->guard();
    MemoryInfo m;
    check(hp_mi355x_v1_memory_info(ctx_->raw(), &m.device_free, &m.device_total, &m.live, &m.peak,
                                   reset_peak ? 1 : 0),
          "memory_info");
    return m;
  }

  // extension: one round of several states in shared launches
  std::vector<std::vector<Fr>> prove_rounds_batch(std::vector<ProverState*>& states,
                                                  const std::vector<std::optional<Fr>>& challenges) const {
    const auto lock = ctx_->guard();
    const size_t n = states.size();
    if (challenges.size() != n) invalid("one challenge slot per state");
    std::vector<std::vector<Fr>> out(n);
    std::vector<HpMi355xSumCheckState*> raw(n);
    std::vector<const uint8_t*> ch(n);
    std::vector<uint8_t*> outs(n);
    std::vector<size_t> counts(n);
    for (size_t i = 0; i < n; ++i) {
      out[i].resize(states[i]->max_degree() + 1);
      raw[i] = states[i]->state_->raw;
      ch[i] = challenges[i] ? challenges[i]->data() : nullptr;
      outs[i] = out[i][0].data();
      counts[i] = out[i].size();
    }
    if (n)
      check(hp_mi355x_v1_sumcheck_prover_state_round_batch(ctx_->raw(), raw.data(), ch.data(),
                                                           outs.data(), counts.data(), n),
            "prove_rounds_batch");
    return out;
  }

  // ---- SumCheckProver ----------------------------------------------------
  ProverState prover_init(const VirtualPolynomial& poly) const {
    const auto lock = ctx_->guard();
    const size_t nv = poly.aux_info.num_variables, d = poly.aux_info.max_degree;
    if (nv == 0 || d == 0) invalid("SumCheck requires positive variable count and degree");
    std::vector<uint8_t> coeffs;
    std::vector<size_t> offsets{0}, indices;
    const auto& factors = poly.flattened_ml_extensions();
    for (const auto& [c, idx] : poly.products) {
      if (idx.empty() || idx.size() > d) invalid("product arity outside 1..max_degree");
      for (size_t i : idx)
        if (i >= factors.size()) invalid("product references an out-of-range factor");
      coeffs.insert(coeffs.end(), c.begin(), c.end());
      indices.insert(indices.end(), idx.begin(), idx.end());
      offsets.push_back(indices.size());
    }
    const auto addrs = addresses(factors);
    ProverState st;
    st.state_ = std::make_shared<ProverState::State>();
    st.state_->ctx = ctx_;
    st.state_->raw = nullptr;
    st.state_->max_degree = d;
    check(hp_mi355x_v1_sumcheck_prover_state_create(
              ctx_->raw(), nv, d, addrs.empty() ? nullptr : addrs.data(), addrs.size(),
              coeffs.empty() ? nullptr : coeffs.data(), offsets.data(), poly.products.size(),
              indices.empty() ? nullptr : indices.data(), indices.size(), &st.state_->raw),
          "prover_init");
    return st;
  }

  std::vector<Fr> prove_round_and_update_state(ProverState& state,
                                               const std::optional<Fr>& challenge) const {
    const auto lock = ctx_->guard();
    std::vector<Fr> out(state.max_degree() + 1);
    check(hp_mi355x_v1_sumcheck_prover_state_round(
              ctx_->raw(), state.state_->raw, challenge ? challenge->data() : nullptr,
              out[0].data(), out.size()),
          "prove_round_and_update_state");
    return out;
  }

 private:
  Mle wrap(HpMi355xDeviceAddress address, size_t nv) const {
    Mle m;
    m.state_ = std::make_shared<Mle::State>();
    m.state_->ctx = ctx_;
    m.state_->address = address;
    m.state_->num_vars = nv;
    return m;
  }
  PreparedProverParam new_params(size_t nv) const {
    PreparedProverParam pp;
    pp.state_ = std::make_shared<PreparedProverParam::State>();
    pp.state_->ctx = ctx_;
    pp.state_->raw = nullptr;
    pp.state_->num_vars = nv;
    check(hp_mi355x_v1_multilinear_kzg_params_create(ctx_->raw(), nv, &pp.state_->raw),
          "params_create");
    return pp;
  }
  void own(const Mle& m) const {
    if (!m || m.context() != ctx_) invalid("MLE handle belongs to another backend context");
  }
  std::vector<HpMi355xDeviceAddress> addresses(const std::vector<Mle>& ms) const {
    std::vector<HpMi355xDeviceAddress> a;
    for (const auto& m : ms) {
      own(m);
      a.push_back(m.address());
    }
    return a;
  }
  std::shared_ptr<Context> ctx_;
};

}  // namespace mi355x
}  // namespace hyperplonk
