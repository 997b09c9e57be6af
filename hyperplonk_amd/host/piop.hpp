// PolyIOP protocols on the host, generic over the backend that owns the MLE
// tables and the sumcheck prover state:
//   SumCheck      subroutines/src/poly_iop/sum_check/mod.rs:130-223,
//                 verifier subroutines/src/poly_iop/sum_check/verifier.rs:17-200
//   ZeroCheck     subroutines/src/poly_iop/zero_check/mod.rs:62-123
//   ProductCheck  subroutines/src/poly_iop/prod_check/mod.rs:150-266, util.rs:24-65
//   PermutationCheck subroutines/src/poly_iop/perm_check/mod.rs:97-178
// The transcript labels and the order of appends and challenges are the
// reference's; a backend B supplies Mle, VirtualPolynomial, ProverState,
// PreparedProverParam and the trait methods of backend/src/{multilinear_polynomial,
// multilinear_kzg,sum_check}.rs with canonical-byte field elements.
#pragma once
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "ff.hpp"
#include "serialize.hpp"
#include "transcript.hpp"
#include "virtual_polynomial.hpp"

namespace hyperplonk {
namespace host {

// subroutines/src/poly_iop/errors.rs
struct PolyIOPError : std::runtime_error {
  enum class Kind { InvalidParameters, InvalidProof, InvalidVerifier, InvalidProver };
  Kind kind;
  PolyIOPError(Kind k, const std::string& what) : std::runtime_error(what), kind(k) {}
};

[[noreturn]] inline void piop_fail(PolyIOPError::Kind k, const std::string& what) {
  throw PolyIOPError(k, what);
}

struct IOPProof {
  std::vector<Fr> point;
  std::vector<std::vector<Fr>> proofs;  // IOPProverMessage evaluations per round
  bool operator==(const IOPProof& o) const { return point == o.point && proofs == o.proofs; }
};

struct SumCheckSubClaim {
  std::vector<Fr> point;
  Fr expected_evaluation;
};

struct ZeroCheckSubClaim {
  std::vector<Fr> point;
  Fr expected_evaluation;
  std::vector<Fr> init_challenge;
};

// eq(x, y) = prod_i (x_i y_i + (1 - x_i)(1 - y_i))  (backend/src/common/eq.rs:10-30)
inline Fr eq_eval(const std::vector<Fr>& x, const std::vector<Fr>& y) {
  if (x.size() != y.size()) piop_fail(PolyIOPError::Kind::InvalidParameters, "x and y have different length");
  Fr res = Fr::one();
  for (size_t i = 0; i < x.size(); ++i) {
    const Fr xy = x[i] * y[i];
    res = res * (xy + xy - x[i] - y[i] + Fr::one());
  }
  return res;
}

// out[idx] = prod_i (bit_i(idx) ? r_i : 1 - r_i), low bit first
// (backend/src/common/eq.rs:38-95); host-side, for verifier-size inputs
inline std::vector<Fr> eq_table(const std::vector<Fr>& r) {
  std::vector<Fr> out{Fr::one()};
  for (const Fr& ri : r) {
    std::vector<Fr> next(out.size() * 2);
    const Fr one_minus = Fr::one() - ri;
    for (size_t j = 0; j < out.size(); ++j) {
      next[j] = out[j] * one_minus;
      next[j + out.size()] = out[j] * ri;
    }
    out.swap(next);
  }
  return out;
}

// Value at x of the polynomial of degree < n through (i, p_i), i = 0..n-1
// (the unique interpolant; verifier.rs:202-300 computes the same value).
inline Fr interpolate_uni_poly(const std::vector<Fr>& p, const Fr& x) {
  const size_t n = p.size();
  if (n == 0) piop_fail(PolyIOPError::Kind::InvalidParameters, "empty round message");
  std::vector<Fr> diff(n);
  for (size_t i = 0; i < n; ++i) {
    diff[i] = x - Fr::from_u64(i);
    if (diff[i].is_zero()) return p[i];
  }
  // prod_j (x - j)
  Fr all = Fr::one();
  for (const Fr& d : diff) all *= d;
  // w_i = prod_{j != i} (i - j) = (-1)^(n-1-i) i! (n-1-i)!
  std::vector<Fr> fact(n);
  fact[0] = Fr::one();
  for (size_t i = 1; i < n; ++i) fact[i] = fact[i - 1] * Fr::from_u64(i);
  Fr sum = Fr::zero();
  for (size_t i = 0; i < n; ++i) {
    Fr w = fact[i] * fact[n - 1 - i];
    if ((n - 1 - i) & 1) w = -w;
    sum += p[i] * (diff[i] * w).inverse();
  }
  return sum * all;
}

// ---------------------------------------------------------------- SumCheck
template <class B>
IOPProof sumcheck_prove(const B& backend, const typename B::VirtualPolynomial& poly,
                        IOPTranscript& transcript) {
  const size_t nv = poly.aux_info.num_variables, d = poly.aux_info.max_degree;
  if (nv == 0 || d == 0)
    piop_fail(PolyIOPError::Kind::InvalidParameters,
              "SumCheck requires positive variable count and degree.");
  transcript.append_message("aux info", ser_aux(poly.aux_info));
  typename B::ProverState state = backend.prover_init(poly);
  std::optional<FrBytes> challenge;
  IOPProof proof;
  proof.point.reserve(nv);
  proof.proofs.reserve(nv);
  for (size_t round = 0; round < nv; ++round) {
    std::vector<Fr> msg = from_wire(backend.prove_round_and_update_state(state, challenge));
    transcript.append_message("prover msg", ser_frs(msg));
    proof.proofs.push_back(std::move(msg));
    const Fr c = transcript.get_and_append_challenge("Internal round");
    proof.point.push_back(c);
    challenge = c.bytes();
  }
  return proof;
}

inline SumCheckSubClaim sumcheck_verify(const Fr& claimed_sum, const IOPProof& proof,
                                        const VPAuxInfo& aux, IOPTranscript& transcript) {
  const size_t nv = aux.num_variables;
  if (nv == 0 || aux.max_degree == 0)
    piop_fail(PolyIOPError::Kind::InvalidParameters,
              "SumCheck requires positive variable count and degree.");
  if (aux.max_degree == SIZE_MAX)
    piop_fail(PolyIOPError::Kind::InvalidParameters, "SumCheck degree is too large.");
  const size_t width = aux.max_degree + 1;
  if (proof.point.size() != nv || proof.proofs.size() != nv)
    piop_fail(PolyIOPError::Kind::InvalidProof, "incorrect SumCheck dimensions");
  for (const auto& m : proof.proofs)
    if (m.size() != width) piop_fail(PolyIOPError::Kind::InvalidProof, "incorrect number of evaluations");
  transcript.append_message("aux info", ser_aux(aux));
  std::vector<Fr> challenges;
  challenges.reserve(nv);
  for (const auto& m : proof.proofs) {
    transcript.append_message("prover msg", ser_frs(m));
    challenges.push_back(transcript.get_and_append_challenge("Internal round"));
  }
  // deferred round checks: g_0(0) + g_0(1) = sum, g_k(0) + g_k(1) = g_{k-1}(r_{k-1})
  Fr expected = claimed_sum;
  for (size_t k = 0; k < nv; ++k) {
    if (proof.proofs[k][0] + proof.proofs[k][1] != expected)
      piop_fail(PolyIOPError::Kind::InvalidProof, "Prover message is not consistent with the claim.");
    expected = interpolate_uni_poly(proof.proofs[k], challenges[k]);
  }
  return {challenges, expected};
}

// ---------------------------------------------------------------- ZeroCheck
template <class B>
IOPProof zerocheck_prove(const B& backend, const typename B::VirtualPolynomial& poly,
                         IOPTranscript& transcript) {
  const size_t length = poly.aux_info.num_variables;
  const std::vector<Fr> r = transcript.get_and_append_challenge_vectors("0check r", length);
  typename B::Mle eq = backend.build_eq_x_r(to_wire(r));
  typename B::VirtualPolynomial f_hat = poly;
  f_hat.mul_by_mle(eq, Fr::one().bytes());
  return sumcheck_prove(backend, f_hat, transcript);
}

inline ZeroCheckSubClaim zerocheck_verify(const IOPProof& proof, const VPAuxInfo& fx_aux,
                                          IOPTranscript& transcript) {
  if (fx_aux.max_degree >= SIZE_MAX - 1)
    piop_fail(PolyIOPError::Kind::InvalidParameters, "ZeroCheck degree is too large.");
  VPAuxInfo hat = fx_aux;
  hat.max_degree = fx_aux.max_degree + 1;
  const std::vector<Fr> r =
      transcript.get_and_append_challenge_vectors("0check r", fx_aux.num_variables);
  SumCheckSubClaim sub = sumcheck_verify(Fr::zero(), proof, hat, transcript);
  const Fr eq = eq_eval(sub.point, r);
  return {sub.point, sub.expected_evaluation * eq.inverse(), r};
}

// ------------------------------------------------------------- ProductCheck
struct ProductCheckProof {
  IOPProof zero_check_proof;
  G1Packet prod_x_comm{};
  G1Packet frac_comm{};
};

struct ProductCheckSubClaim {
  ZeroCheckSubClaim zero_check_sub_claim;
  std::vector<Fr> final_query_point;
  Fr final_query_eval;
  Fr alpha;
};

template <class B>
std::tuple<ProductCheckProof, typename B::Mle, typename B::Mle> prodcheck_prove(
    const B& backend, const typename B::PreparedProverParam& pcs_param,
    const std::vector<typename B::Mle>& fxs, const std::vector<typename B::Mle>& gxs,
    IOPTranscript& transcript) {
  using Mle = typename B::Mle;
  if (fxs.empty()) piop_fail(PolyIOPError::Kind::InvalidParameters, "fxs is empty");
  if (fxs.size() != gxs.size())
    piop_fail(PolyIOPError::Kind::InvalidParameters,
              "fxs and gxs have different number of polynomials");
  const size_t nv = fxs[0].num_vars();
  for (const auto* list : {&fxs, &gxs})
    for (const auto& p : *list)
      if (p.num_vars() != nv)
        piop_fail(PolyIOPError::Kind::InvalidParameters,
                  "fx and gx have different number of variables");
  Mle frac = backend.compute_frac_poly(fxs, gxs);
  Mle prod = backend.compute_product_poly(frac);
  ProductCheckProof proof;
  // both commitments in one batched launch (commit order is irrelevant to the values)
  const std::vector<G1Packet> comms = backend.multi_commit(pcs_param, {frac, prod});
  proof.frac_comm = comms[0];
  proof.prod_x_comm = comms[1];
  transcript.append_message("frac(x)", ser_g1(proof.frac_comm));
  transcript.append_message("prod(x)", ser_g1(proof.prod_x_comm));
  const Fr alpha = transcript.get_and_append_challenge("alpha");
  const auto factors = backend.compute_product_factors(frac, prod);
  // Q(x) = prod(x) - p1 p2 + alpha frac g1..gk - alpha f1..fk  (prod_check/util.rs:24-65)
  typename B::VirtualPolynomial q = B::VirtualPolynomial::new_from_mle(prod, Fr::one().bytes());
  q.add_mle_list({factors[0], factors[1]}, (-Fr::one()).bytes());
  std::vector<Mle> gl = gxs;
  gl.push_back(frac);
  q.add_mle_list(gl, alpha.bytes());
  q.add_mle_list(fxs, (-alpha).bytes());
  proof.zero_check_proof = zerocheck_prove(backend, q, transcript);
  return {std::move(proof), prod, frac};
}

inline ProductCheckSubClaim prodcheck_verify(const ProductCheckProof& proof, const VPAuxInfo& aux,
                                             IOPTranscript& transcript) {
  transcript.append_message("frac(x)", ser_g1(proof.frac_comm));
  transcript.append_message("prod(x)", ser_g1(proof.prod_x_comm));
  const Fr alpha = transcript.get_and_append_challenge("alpha");
  ProductCheckSubClaim out;
  out.zero_check_sub_claim = zerocheck_verify(proof.zero_check_proof, aux, transcript);
  out.final_query_point.assign(aux.num_variables, Fr::one());
  if (!out.final_query_point.empty()) out.final_query_point[0] = Fr::zero();
  out.final_query_eval = Fr::one();
  out.alpha = alpha;
  return out;
}

// --------------------------------------------------------- PermutationCheck
struct PermutationCheckSubClaim {
  ProductCheckSubClaim product_check_sub_claim;
  Fr beta, gamma;
};

template <class B>
std::tuple<ProductCheckProof, typename B::Mle, typename B::Mle> permcheck_prove(
    const B& backend, const typename B::PreparedProverParam& pcs_param,
    const std::vector<typename B::Mle>& fxs, const std::vector<typename B::Mle>& gxs,
    const std::vector<typename B::Mle>& perms, IOPTranscript& transcript) {
  if (fxs.empty()) piop_fail(PolyIOPError::Kind::InvalidParameters, "fxs is empty");
  if (fxs.size() != gxs.size() || fxs.size() != perms.size())
    piop_fail(PolyIOPError::Kind::InvalidProof, "fxs, gxs and perms have different lengths");
  const size_t nv = fxs[0].num_vars();
  for (size_t i = 0; i < fxs.size(); ++i)
    if (fxs[i].num_vars() != nv || gxs[i].num_vars() != nv || perms[i].num_vars() != nv)
      piop_fail(PolyIOPError::Kind::InvalidParameters, "number of variables unmatched");
  const Fr beta = transcript.get_and_append_challenge("beta");
  const Fr gamma = transcript.get_and_append_challenge("gamma");
  auto [nums, dens] = backend.compute_nums_and_denoms(beta.bytes(), gamma.bytes(), fxs, gxs, perms);
  return prodcheck_prove(backend, pcs_param, nums, dens, transcript);
}

inline PermutationCheckSubClaim permcheck_verify(const ProductCheckProof& proof,
                                                 const VPAuxInfo& aux, IOPTranscript& transcript) {
  PermutationCheckSubClaim out;
  out.beta = transcript.get_and_append_challenge("beta");
  out.gamma = transcript.get_and_append_challenge("gamma");
  out.product_check_sub_claim = prodcheck_verify(proof, aux, transcript);
  return out;
}

}  // namespace host
}  // namespace hyperplonk
