// HyperPlonk SNARK on the host over any backend (hyperplonk/src/snark.rs:56-620):
// preprocess (commit selectors and permutation oracles), prove (witness
// commitments, gate zerocheck, permutation check, 3k + s + 8 evaluations and
// one batch opening) and verify (all IOP subclaims, the public-input check and
// the PCS batch verification with pairings, pairing.hpp).  Gates, parameters and
// the permutation labelling follow hyperplonk/src/{custom_gate,structs,utils}.rs.
#pragma once
#include <cstdint>
#include <initializer_list>
#include <optional>
#include <string>
#include <tuple>
#include <vector>

#include "pairing.hpp"
#include "pcs.hpp"
#include "piop.hpp"

namespace hyperplonk {
namespace host {

// (coefficient, selector index, witness indices) per monomial (custom_gate.rs:31-34)
struct Monomial {
  int64_t coeff;
  std::optional<size_t> selector;
  std::vector<size_t> witnesses;
};

struct CustomizedGates {
  std::vector<Monomial> gates;

  size_t degree() const {
    size_t r = 0;
    for (const auto& g : gates) r = std::max(r, g.witnesses.size() + (g.selector ? 1 : 0));
    return r;
  }
  size_t num_selector_columns() const {
    size_t r = 0;
    for (const auto& g : gates) r += g.selector ? 1 : 0;
    return r;
  }
  size_t num_witness_columns() const {
    size_t r = 0;
    for (const auto& g : gates)
      if (!g.witnesses.empty()) r = std::max(r, g.witnesses.back());
    return r + 1;
  }
  // q_L w1 + q_R w2 + q_O w3 + q_M w1 w2 + q_C (custom_gate.rs:86-96)
  static CustomizedGates vanilla_plonk_gate() {
    return {{{1, 0, {0}}, {1, 1, {1}}, {1, 2, {2}}, {1, 3, {0, 1}}, {1, 4, {}}}};
  }
  // jellyfish turbo gate, 5 wires and 13 selectors (custom_gate.rs:129-146)
  static CustomizedGates jellyfish_turbo_plonk_gate() {
    return {{{1, 0, {0}},
             {1, 1, {1}},
             {1, 2, {2}},
             {1, 3, {3}},
             {1, 4, {0, 1}},
             {1, 5, {2, 3}},
             {1, 6, {0, 0, 0, 0, 0}},
             {1, 7, {1, 1, 1, 1, 1}},
             {1, 8, {2, 2, 2, 2, 2}},
             {1, 9, {3, 3, 3, 3, 3}},
             {1, 10, {0, 1, 2, 3}},
             {1, 11, {4}},
             {1, 12, {}}}};
  }
};

struct HyperPlonkParams {
  size_t num_constraints = 0;
  size_t num_pub_input = 0;
  CustomizedGates gate_func;

  size_t num_variables() const { return ceil_log2(num_constraints); }
  size_t num_selector_columns() const { return gate_func.num_selector_columns(); }
  size_t num_witness_columns() const { return gate_func.num_witness_columns(); }
  // sum_i 2^i point_i over num_vars + log2(#witnesses) variables (structs.rs:73-90)
  Fr eval_id_oracle(const std::vector<Fr>& point) const {
    const size_t len = num_variables() + ceil_log2(num_witness_columns());
    if (point.size() != len) piop_fail(PolyIOPError::Kind::InvalidParameters, "ID oracle point length");
    Fr res = Fr::zero(), base = Fr::one();
    for (const Fr& v : point) {
      res += base * v;
      base = base + base;
    }
    return res;
  }
};

// Columns are held in the encoding the backend takes (canonical LE32 Fr), so
// a proof does no per-entry conversion on the host.
using Column = std::vector<FrBytes>;

struct HyperPlonkIndex {
  HyperPlonkParams params;
  Column permutation;              // #witnesses * 2^nv labels
  std::vector<Column> selectors;   // #selectors columns of 2^nv
};

template <class B>
struct HyperPlonkProvingKey {
  HyperPlonkParams params;
  std::vector<typename B::Mle> permutation_oracles;
  std::vector<typename B::Mle> selector_oracles;
  std::vector<G1Packet> selector_commitments;
  std::vector<G1Packet> permutation_commitments;
  const typename B::PreparedProverParam* pcs_param = nullptr;
};

struct HyperPlonkVerifyingKey {
  HyperPlonkParams params;
  MultilinearVerifierParam pcs_param;
  std::vector<G1Packet> selector_commitments;
  std::vector<G1Packet> perm_commitments;
};

struct HyperPlonkProof {
  std::vector<G1Packet> witness_commits;
  BatchProof batch_openings;
  IOPProof zero_check_proof;
  ProductCheckProof perm_check_proof;
};

// Byte encoding of a proof in the reference's field order with the
// ark-serialize conventions (Vec length prefixes, compressed points); used to
// compare proofs byte for byte and to hash them.
inline std::vector<uint8_t> serialize_proof(const HyperPlonkProof& p) {
  Writer w;
  auto iop = [&](const IOPProof& x) {
    w.frs(x.point);
    w.u64(x.proofs.size());
    for (const auto& m : x.proofs) w.frs(m);
  };
  w.g1s(p.witness_commits);
  iop(p.batch_openings.sum_check_proof);
  w.frs(p.batch_openings.f_i_eval_at_point_i);
  w.g1s(p.batch_openings.g_prime_proof);
  iop(p.zero_check_proof);
  iop(p.perm_check_proof.zero_check_proof);
  w.g1(p.perm_check_proof.prod_x_comm);
  w.g1(p.perm_check_proof.frac_comm);
  return w.buf;
}

struct HyperPlonkError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline Fr gate_coeff(int64_t c) { return Fr::from_i64(c); }

// ------------------------------------------------------------------ helpers
template <class B>
typename B::VirtualPolynomial build_f(const CustomizedGates& gates, size_t nv,
                                      const std::vector<typename B::Mle>& selectors,
                                      const std::vector<typename B::Mle>& witnesses) {
  for (const auto* list : {&selectors, &witnesses})
    for (const auto& m : *list)
      if (m.num_vars() != nv) throw HyperPlonkError("selector has different number of vars");
  typename B::VirtualPolynomial res(nv);
  for (const auto& g : gates.gates) {
    std::vector<typename B::Mle> list;
    if (g.selector) list.push_back(selectors.at(*g.selector));
    for (size_t wi : g.witnesses) list.push_back(witnesses.at(wi));
    res.add_mle_list(list, gate_coeff(g.coeff).bytes());
  }
  return res;
}

inline Fr eval_f(const CustomizedGates& gates, const std::vector<Fr>& sel,
                 const std::vector<Fr>& wit) {
  Fr res = Fr::zero();
  for (const auto& g : gates.gates) {
    Fr cur = gate_coeff(g.coeff);
    if (g.selector) cur *= sel.at(*g.selector);
    for (size_t wi : g.witnesses) cur *= wit.at(wi);
    res += cur;
  }
  return res;
}

// utils.rs:279-300
inline Fr eval_perm_gate(const std::vector<Fr>& prod, const std::vector<Fr>& frac,
                         const std::vector<Fr>& wit, const std::vector<Fr>& id,
                         const std::vector<Fr>& perm, const Fr& alpha, const Fr& beta,
                         const Fr& gamma, const Fr& x1) {
  const Fr p1 = frac[1] + x1 * (prod[1] - frac[1]);
  const Fr p2 = frac[2] + x1 * (prod[2] - frac[2]);
  Fr f = Fr::one(), g = Fr::one();
  for (size_t i = 0; i < wit.size(); ++i) f *= wit[i] + beta * id[i] + gamma;
  for (size_t i = 0; i < wit.size(); ++i) g *= wit[i] + beta * perm[i] + gamma;
  return prod[0] - p1 * p2 + alpha * (frac[0] * g - f);
}

// point ++ (bits of index, low first) (snark.rs:25-35)
inline std::vector<Fr> gen_eval_point(size_t index, size_t index_len, const std::vector<Fr>& point) {
  std::vector<Fr> r = point;
  for (size_t b = 0; b < index_len; ++b) r.push_back(((index >> b) & 1) ? Fr::one() : Fr::zero());
  return r;
}

// MLE evaluation on the host, point[0] = lowest table bit (public input only)
inline Fr evaluate_table(std::vector<Fr> t, const std::vector<Fr>& point) {
  for (const Fr& r : point) {
    const size_t half = t.size() / 2;
    for (size_t b = 0; b < half; ++b) t[b] = t[2 * b] + r * (t[2 * b + 1] - t[2 * b]);
    t.resize(half);
  }
  return t.at(0);
}

// ------------------------------------------------------------- preprocess
template <class B>
std::pair<HyperPlonkProvingKey<B>, HyperPlonkVerifyingKey> preprocess(
    const B& backend, const HyperPlonkIndex& index, const typename B::PreparedProverParam& pp,
    const MultilinearVerifierParam& full_vk) {
  const size_t nv = index.params.num_variables();
  HyperPlonkProvingKey<B> pk;
  pk.params = index.params;
  pk.pcs_param = &pp;
  const size_t n = size_t(1) << nv;
  for (size_t i = 0; i < index.params.num_witness_columns(); ++i) {
    Column col(index.permutation.begin() + i * n, index.permutation.begin() + (i + 1) * n);
    pk.permutation_oracles.push_back(backend.mle_from_evaluations(nv, col));
  }
  for (const auto& s : index.selectors)
    pk.selector_oracles.push_back(backend.mle_from_evaluations(ceil_log2(s.size()), s));
  pk.permutation_commitments = backend.multi_commit(pp, pk.permutation_oracles);
  pk.selector_commitments = backend.multi_commit(pp, pk.selector_oracles);
  HyperPlonkVerifyingKey vk;
  vk.params = index.params;
  vk.pcs_param = full_vk.trim(nv);
  vk.selector_commitments = pk.selector_commitments;
  vk.perm_commitments = pk.permutation_commitments;
  return {std::move(pk), std::move(vk)};
}

// ------------------------------------------------------------------ prove
template <class B>
HyperPlonkProof prove(const B& backend, const HyperPlonkProvingKey<B>& pk,
                      const std::vector<Fr>& pub_input,
                      const std::vector<Column>& witnesses) {
  using Mle = typename B::Mle;
  IOPTranscript transcript("hyperplonk");
  const HyperPlonkParams& params = pk.params;
  // prover_sanity_check (utils.rs:110-160)
  if (pub_input.size() > params.num_constraints)
    throw HyperPlonkError("Public input length is greater than num constraints");
  if (pub_input.size() != params.num_pub_input) throw HyperPlonkError("Public input length is not correct");
  if (pub_input.empty() || (pub_input.size() & (pub_input.size() - 1)))
    throw HyperPlonkError("Public input length is not power of two");
  for (const auto& w : witnesses)
    if (w.size() != params.num_constraints) throw HyperPlonkError("witness length is not correct");
  for (size_t i = 0; i < pub_input.size(); ++i)
    if (pub_input[i].bytes() != witnesses.at(0)[i])
      throw HyperPlonkError("public input does not match witness[0]");

  const size_t nv = params.num_variables();
  const size_t ell = ceil_log2(params.num_pub_input);
  const typename B::PreparedProverParam& pp = *pk.pcs_param;

  // 1. witness commitments
  std::vector<Mle> wpolys;
  for (const auto& w : witnesses) wpolys.push_back(backend.mle_from_evaluations(nv, w));
  HyperPlonkProof proof;
  proof.witness_commits = backend.multi_commit(pp, wpolys);
  for (const auto& c : proof.witness_commits) transcript.append_message("w", ser_g1(c));

  // 2. gate zerocheck
  auto fx = build_f<B>(params.gate_func, nv, pk.selector_oracles, wpolys);
  proof.zero_check_proof = zerocheck_prove(backend, fx, transcript);

  // 3. permutation check
  auto [perm_proof, prod_x, frac_poly] =
      permcheck_prove(backend, pp, wpolys, wpolys, pk.permutation_oracles, transcript);
  proof.perm_check_proof = std::move(perm_proof);
  const std::vector<Fr>& pc = proof.perm_check_proof.zero_check_proof.point;

  // 4. opening points, in the verifier's order
  std::vector<Fr> pc0{Fr::zero()}, pc1{Fr::one()}, final_q{Fr::zero()};
  pc0.insert(pc0.end(), pc.begin(), pc.end() - 1);
  pc1.insert(pc1.end(), pc.begin(), pc.end() - 1);
  for (size_t i = 1; i < nv; ++i) final_q.push_back(Fr::one());
  std::vector<Mle> polys;
  std::vector<std::vector<Fr>> points;
  auto insert = [&](const Mle& m, const std::vector<Fr>& pt) {
    polys.push_back(m);
    points.push_back(pt);
  };
  insert(prod_x, pc);
  insert(prod_x, pc0);
  insert(prod_x, pc1);
  insert(prod_x, final_q);
  insert(frac_poly, pc);
  insert(frac_poly, pc0);
  insert(frac_poly, pc1);
  for (const auto& p : pk.permutation_oracles) insert(p, pc);
  for (const auto& w : wpolys) insert(w, pc);
  for (const auto& w : wpolys) insert(w, proof.zero_check_proof.point);
  for (const auto& s : pk.selector_oracles) insert(s, proof.zero_check_proof.point);
  std::vector<Fr> r_pi = transcript.get_and_append_challenge_vectors("r_pi", ell);
  r_pi.resize(nv, Fr::zero());
  insert(wpolys[0], r_pi);
  std::vector<std::vector<FrBytes>> wire_points;
  for (const auto& pt : points) wire_points.push_back(to_wire(pt));
  const std::vector<Fr> evals = from_wire(backend.evaluate_mles(polys, wire_points));

  // 5. batch opening
  proof.batch_openings = multi_open(backend, pp, polys, points, evals, transcript);
  return proof;
}

// ----------------------------------------------------------------- verify
// Returns the pairing check's result; malformed proofs and failed IOP checks
// throw (the reference returns Err for those, snark.rs:384-620).
inline bool verify(const HyperPlonkVerifyingKey& vk, const std::vector<Fr>& pub_input,
                   const HyperPlonkProof& proof) {
  IOPTranscript transcript("hyperplonk");
  const size_t ns = vk.params.num_selector_columns();
  const size_t nw = vk.params.num_witness_columns();
  const size_t nv = vk.params.num_variables();
  const size_t ell = ceil_log2(vk.params.num_pub_input);
  if (pub_input.size() != vk.params.num_pub_input)
    throw HyperPlonkError("Public input length is not correct");
  if (proof.witness_commits.size() != nw) throw HyperPlonkError("incorrect witness commitment count");
  const size_t expected = 3 * nw + ns + 8;
  const auto& ev = proof.batch_openings.f_i_eval_at_point_i;
  if (ev.size() != expected) throw HyperPlonkError("incorrect opening evaluation count");
  const std::vector<Fr> prod_evals(ev.begin(), ev.begin() + 4);
  const std::vector<Fr> frac_evals(ev.begin() + 4, ev.begin() + 7);
  const std::vector<Fr> perm_evals(ev.begin() + 7, ev.begin() + 7 + nw);
  const std::vector<Fr> wit_perm(ev.begin() + 7 + nw, ev.begin() + 7 + 2 * nw);
  const std::vector<Fr> wit_gate(ev.begin() + 7 + 2 * nw, ev.begin() + 7 + 3 * nw);
  const std::vector<Fr> sel_evals(ev.begin() + 7 + 3 * nw, ev.begin() + 7 + 3 * nw + ns);
  const Fr pi_eval = ev[expected - 1];

  // 1. gate zerocheck
  for (const auto& c : proof.witness_commits) transcript.append_message("w", ser_g1(c));
  const ZeroCheckSubClaim zc =
      zerocheck_verify(proof.zero_check_proof, {vk.params.gate_func.degree(), nv}, transcript);
  if (eval_f(vk.params.gate_func, sel_evals, wit_gate) != zc.expected_evaluation)
    throw HyperPlonkError("zero check evaluation failed");

  // 2. permutation check
  const PermutationCheckSubClaim pcs =
      permcheck_verify(proof.perm_check_proof, {nw + 1, nv}, transcript);
  const auto& psc = pcs.product_check_sub_claim;
  if (prod_evals[3] != psc.final_query_eval)
    throw HyperPlonkError("permutation final product evaluation failed");
  const std::vector<Fr>& pc = psc.zero_check_sub_claim.point;
  if (pc.empty()) throw HyperPlonkError("empty permutation check point");
  std::vector<Fr> id_evals;
  for (size_t i = 0; i < nw; ++i)
    id_evals.push_back(vk.params.eval_id_oracle(gen_eval_point(i, ceil_log2(nw), pc)));
  if (eval_perm_gate(prod_evals, frac_evals, wit_perm, id_evals, perm_evals, psc.alpha, pcs.beta,
                     pcs.gamma, pc.back()) != psc.zero_check_sub_claim.expected_evaluation)
    throw HyperPlonkError("permutation evaluation failed");

  // 3. commitments and points of the batch opening
  std::vector<G1Packet> comms;
  std::vector<std::vector<Fr>> points;
  std::vector<Fr> pc0{Fr::zero()}, pc1{Fr::one()};
  pc0.insert(pc0.end(), pc.begin(), pc.end() - 1);
  pc1.insert(pc1.end(), pc.begin(), pc.end() - 1);
  for (const std::vector<Fr>* pt :
       std::initializer_list<const std::vector<Fr>*>{&pc, &pc0, &pc1, &psc.final_query_point}) {
    comms.push_back(proof.perm_check_proof.prod_x_comm);
    points.push_back(*pt);
  }
  for (const std::vector<Fr>* pt : std::initializer_list<const std::vector<Fr>*>{&pc, &pc0, &pc1}) {
    comms.push_back(proof.perm_check_proof.frac_comm);
    points.push_back(*pt);
  }
  for (const auto& c : vk.perm_commitments) {
    comms.push_back(c);
    points.push_back(pc);
  }
  for (const auto& c : proof.witness_commits) {
    comms.push_back(c);
    points.push_back(pc);
  }
  for (const auto& c : proof.witness_commits) {
    comms.push_back(c);
    points.push_back(zc.point);
  }
  for (const auto& c : vk.selector_commitments) {
    comms.push_back(c);
    points.push_back(zc.point);
  }
  std::vector<Fr> r_pi = transcript.get_and_append_challenge_vectors("r_pi", ell);
  if (evaluate_table(pub_input, r_pi) != pi_eval) throw HyperPlonkError("Public input eval mismatch");
  r_pi.resize(nv, Fr::zero());
  comms.push_back(proof.witness_commits[0]);
  points.push_back(r_pi);

  // 4. PCS batch verification
  return batch_verify(vk.pcs_param, comms, points, proof.batch_openings, transcript);
}

}  // namespace host
}  // namespace hyperplonk
