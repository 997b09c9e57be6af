// Multilinear KZG batch opening on the host (prover side):
// multi_open_internal, subroutines/src/pcs/multilinear_kzg/batching.rs:49-194.
//   1. append every point ("eval_point", Vec<Fr>) and evaluation ("eval");
//   2. t <- ceil(log2 k) challenges "t"; weights eq(t, i);
//   3. group polynomials by equal point (first-seen order) and combine each
//      group with its weights (a singleton group is still a linear combination);
//   4. sumcheck of sum_groups g~_j(x) eq(x, point_j), degree 2;
//   5. g' = sum_j eq(a2, point_j) g~_j at the sumcheck point a2, opened at a2.
#pragma once
#include <vector>

#include "piop.hpp"

namespace hyperplonk {
namespace host {

struct BatchProof {
  IOPProof sum_check_proof;
  std::vector<Fr> f_i_eval_at_point_i;
  std::vector<G1Packet> g_prime_proof;  // MultilinearKzgProof.proofs
};

struct PCSError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ark_std::log2: ceiling of log2, 0 for inputs 0 and 1
inline size_t ceil_log2(size_t x) {
  size_t l = 0;
  while ((size_t(1) << l) < x) ++l;
  return l;
}

// eq(t, i) for i < 2^len(t); the zero-variable table is [1] (batching.rs:31-37)
inline std::vector<Fr> batch_weights(const std::vector<Fr>& t) { return eq_table(t); }

// first-seen-order deduplication of points; groups[j] lists the indices of point j
inline void group_points(const std::vector<std::vector<Fr>>& points,
                         std::vector<size_t>* first_of_group,
                         std::vector<std::vector<size_t>>* groups) {
  for (size_t i = 0; i < points.size(); ++i) {
    size_t g = 0;
    for (; g < first_of_group->size(); ++g)
      if (points[(*first_of_group)[g]] == points[i]) break;
    if (g == first_of_group->size()) {
      first_of_group->push_back(i);
      groups->emplace_back();
    }
    (*groups)[g].push_back(i);
  }
}

template <class B>
BatchProof multi_open(const B& backend, const typename B::PreparedProverParam& pp,
                      const std::vector<typename B::Mle>& polynomials,
                      const std::vector<std::vector<Fr>>& points, const std::vector<Fr>& evals,
                      IOPTranscript& transcript) {
  using Mle = typename B::Mle;
  const size_t k = polynomials.size();
  if (k == 0 || points.size() != k || evals.size() != k)
    throw PCSError("batch polynomials, points, and evaluations must have equal nonzero lengths");
  const size_t nv = polynomials[0].num_vars();
  for (size_t i = 0; i < k; ++i)
    if (nv == 0 || polynomials[i].num_vars() != nv || points[i].size() != nv)
      throw PCSError("batch polynomials and points must have the same positive dimension");
  for (const auto& p : points) transcript.append_message("eval_point", ser_frs(p));
  for (const auto& e : evals) transcript.append_field_element("eval", e);
  const std::vector<Fr> t = transcript.get_and_append_challenge_vectors("t", ceil_log2(k));
  const std::vector<Fr> w = batch_weights(t);

  std::vector<size_t> first;
  std::vector<std::vector<size_t>> groups;
  group_points(points, &first, &groups);
  std::vector<Mle> merged;
  merged.reserve(groups.size());
  for (const auto& g : groups) {
    std::vector<Mle> ps;
    std::vector<FrBytes> cs;
    for (size_t i : g) {
      ps.push_back(polynomials[i]);
      cs.push_back(w[i].bytes());
    }
    merged.push_back(backend.linear_combination(ps, cs));
  }
  typename B::VirtualPolynomial vp(nv);
  for (size_t j = 0; j < groups.size(); ++j) {
    Mle eq = backend.build_eq_x_r(to_wire(points[first[j]]));
    vp.add_mle_list({merged[j], eq}, Fr::one().bytes());
  }
  BatchProof out;
  try {
    out.sum_check_proof = sumcheck_prove(backend, vp, transcript);
  } catch (const PolyIOPError&) {
    throw PCSError("Sumcheck in batch proving Failed");
  }
  const std::vector<Fr>& a2 = out.sum_check_proof.point;
  std::vector<FrBytes> coeffs;
  for (size_t j = 0; j < groups.size(); ++j) coeffs.push_back(eq_eval(a2, points[first[j]]).bytes());
  Mle g_prime = backend.linear_combination(merged, coeffs);
  auto opened = backend.open(pp, g_prime, to_wire(a2));
  out.g_prime_proof = std::move(opened.first);
  out.f_i_eval_at_point_i = evals;
  return out;
}

}  // namespace host
}  // namespace hyperplonk
