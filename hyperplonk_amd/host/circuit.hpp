// Seeded synthetic circuits for the prover: a satisfying witness for a
// customized gate with random (non-identity) copy constraints.
//
// Generator: SplitMix64, uniform canonical Fr by rejection of 255-bit draws
// (four words, top word masked to 63 bits), the scheme of
// backend/src/ascend910b3/native/tests/sumcheck_tensor_edge.cpp:62-80.
// Circuit (vanilla: hyperplonk/examples/full_prove_bench.rs:470-535 with its
// row-pair copies generalised to random pairs; jellyfish:
// hyperplonk/src/mock.rs:52-92):
//   * sigma pairs the #wires * 2^nv slots by a uniformly random perfect
//     matching (a Fisher-Yates shuffle, consecutive shuffled slots swapped;
//     an odd slot out stays fixed), so copies connect random rows of random
//     columns;
//   * every 2-cycle of sigma holds one random non-zero value, so the copy
//     constraints w[sigma(i)] = w[i] hold and the witness has ~half as many
//     distinct values as slots (as the reference's row pairs have);
//   * every selector except `solved` is uniform; the solved one makes the
//     row's gate equation zero (its monomial's witness product is non-zero);
//   * the permutation oracle holds the labels sigma(i), slot i = column * 2^nv + row
//     (cpu/multilinear_polynomial.rs:198-206), and the public input is
//     w0[0 .. min(4, 2^nv)] (mock.rs:93-94).
#pragma once
#include <cstdint>
#include <vector>

#include "snark.hpp"

namespace hyperplonk {
namespace host {

class SplitMix64 {
 public:
  explicit SplitMix64(uint64_t seed) : s_(seed) {}
  uint64_t next() {
    s_ += 0x9E3779B97F4A7C15ULL;
    uint64_t z = s_;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  Fr field() {
    for (;;) {
      Limbs<4> c{};
      for (int w = 0; w < 4; ++w) c[w] = next();
      c[3] &= 0x7FFFFFFFFFFFFFFFULL;
      if (limbs_less(c, FrSpec::MOD)) return Fr::from_canonical(c);
    }
  }
  Fr nonzero() {
    for (;;) {
      const Fr f = field();
      if (!f.is_zero()) return f;
    }
  }
  uint64_t below(uint64_t n) {  // uniform in [0, n) by rejection
    const uint64_t limit = UINT64_MAX - UINT64_MAX % n;
    for (;;) {
      const uint64_t x = next();
      if (x < limit) return x % n;
    }
  }

 private:
  uint64_t s_;
};

struct MockCircuit {
  std::vector<Fr> public_inputs;
  std::vector<Column> witnesses;
  HyperPlonkIndex index;
};

inline Column encode_column(const std::vector<Fr>& v) {
  Column out(v.size());
  for (size_t i = 0; i < v.size(); ++i) out[i] = v[i].bytes();
  return out;
}

inline Fr monomial_value(const Monomial& g, const std::vector<std::vector<Fr>>& sel,
                         const std::vector<std::vector<Fr>>& wit, size_t row, bool with_selector) {
  Fr v = gate_coeff(g.coeff);
  if (with_selector && g.selector) v *= sel[*g.selector][row];
  for (size_t w : g.witnesses) v *= wit[w][row];
  return v;
}

// `solved`: index of the gate monomial whose selector absorbs the residual
inline MockCircuit mock_circuit(size_t nv, const CustomizedGates& gate, size_t solved, uint64_t seed) {
  SplitMix64 rng(seed);
  const size_t n = size_t(1) << nv;
  const size_t nw = gate.num_witness_columns(), ns = gate.num_selector_columns();
  const size_t slots = nw * n;
  // random perfect matching of the slots, one value per 2-cycle
  std::vector<uint64_t> order(slots), sigma(slots);
  for (size_t i = 0; i < slots; ++i) order[i] = sigma[i] = i;
  for (size_t i = slots; i > 1; --i) std::swap(order[i - 1], order[rng.below(i)]);
  for (size_t k = 0; k + 1 < slots; k += 2) {
    sigma[order[k]] = order[k + 1];
    sigma[order[k + 1]] = order[k];
  }
  std::vector<std::vector<Fr>> wit(nw, std::vector<Fr>(n));
  std::vector<uint8_t> seen(slots, 0);
  for (size_t s = 0; s < slots; ++s) {
    if (seen[s]) continue;
    const Fr v = rng.nonzero();
    for (size_t j = s; !seen[j]; j = sigma[j]) {
      seen[j] = 1;
      wit[j / n][j % n] = v;
    }
  }
  std::vector<std::vector<Fr>> sel(ns, std::vector<Fr>(n));
  const Monomial& sg = gate.gates.at(solved);
  std::vector<Fr> residual(n), m(n);
  for (size_t row = 0; row < n; ++row) {
    for (size_t s = 0; s < ns; ++s)
      if (s != *sg.selector) sel[s][row] = rng.field();
    Fr r = Fr::zero();
    for (size_t gi = 0; gi < gate.gates.size(); ++gi)
      if (gi != solved) r += monomial_value(gate.gates[gi], sel, wit, row, true);
    residual[row] = r;
    m[row] = monomial_value(sg, sel, wit, row, false);
  }
  // batch inversion of the solved monomials (all non-zero)
  std::vector<Fr> prefix(n);
  Fr acc = Fr::one();
  for (size_t row = 0; row < n; ++row) {
    prefix[row] = acc;
    acc *= m[row];
  }
  Fr inv = acc.inverse();
  for (size_t row = n; row-- > 0;) {
    const Fr mi = inv * prefix[row];
    inv *= m[row];
    sel[*sg.selector][row] = -(residual[row] * mi);
  }
  MockCircuit c;
  for (const auto& w : wit) c.witnesses.push_back(encode_column(w));
  for (const auto& s : sel) c.index.selectors.push_back(encode_column(s));
  c.index.permutation.resize(slots);
  for (size_t i = 0; i < slots; ++i) c.index.permutation[i] = Fr::from_u64(sigma[i]).bytes();
  const size_t pub = std::min<size_t>(4, n);
  c.public_inputs.assign(wit[0].begin(), wit[0].begin() + pub);
  c.index.params.num_constraints = n;
  c.index.params.num_pub_input = pub;
  c.index.params.gate_func = gate;
  return c;
}

// vanilla: q_O (monomial 2) solved; jellyfish: q_C (monomial 12) solved
inline MockCircuit vanilla_circuit(size_t nv, uint64_t seed) {
  return mock_circuit(nv, CustomizedGates::vanilla_plonk_gate(), 2, seed);
}
inline MockCircuit jellyfish_circuit(size_t nv, uint64_t seed) {
  return mock_circuit(nv, CustomizedGates::jellyfish_turbo_plonk_gate(), 12, seed);
}

inline bool is_satisfied(const MockCircuit& c) {
  const auto& g = c.index.params.gate_func;
  auto decode = [](const std::vector<Column>& cols) {
    std::vector<std::vector<Fr>> out;
    for (const auto& col : cols) out.push_back(from_wire(col));
    return out;
  };
  const auto sel = decode(c.index.selectors), wit = decode(c.witnesses);
  for (size_t row = 0; row < c.index.params.num_constraints; ++row) {
    Fr acc = Fr::zero();
    for (const auto& m : g.gates) acc += monomial_value(m, sel, wit, row, true);
    if (!acc.is_zero()) return false;
  }
  return true;
}

}  // namespace host
}  // namespace hyperplonk
