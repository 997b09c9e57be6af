// C++ host-mirror tests: the reference CPU backend's known-answer tests
// (backend/src/cpu/multilinear_polynomial.rs:255-593, backend/src/cpu/sum_check.rs:224-253,
// subroutines/src/pcs/multilinear_kzg/mod.rs:342-426) written against the
// MI355X backend through hyperplonk_amd/host/mi355x_backend.hpp.
// Prints one "PASS <name>" line per test and exits non-zero on the first failure.
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "../../hyperplonk_amd/host/mi355x_backend.hpp"

using namespace hyperplonk::mi355x;

namespace {

// canonical r - k for small k
Fr neg(uint64_t k) {
  static const uint8_t R[32] = {0x01, 0x00, 0x00, 0x00, 0xff, 0xff, 0xff, 0xff, 0xfe, 0x5b, 0xfe,
                                0xff, 0x02, 0xa4, 0xbd, 0x53, 0x05, 0xd8, 0xa1, 0x09, 0x08, 0xd8,
                                0x39, 0x33, 0x48, 0x7d, 0x9d, 0x29, 0x53, 0xa7, 0xed, 0x73};
  Fr out{};
  int borrow = 0;
  for (int i = 0; i < 32; ++i) {
    int d = int(R[i]) - int((k >> (8 * i)) & 0xff) * (i < 8) - borrow;
    borrow = d < 0;
    out[i] = static_cast<uint8_t>(d + (borrow ? 256 : 0));
  }
  return out;
}

std::vector<Fr> v(std::initializer_list<uint64_t> xs) {
  std::vector<Fr> out;
  for (auto x : xs) out.push_back(fr_from_u64(x));
  return out;
}

void require(bool c, const std::string& what) {
  if (!c) throw std::runtime_error("check failed: " + what);
}

template <class F>
void expect_error(BackendError::Kind kind, F&& f, const std::string& what) {
  try {
    f();
  } catch (const BackendError& e) {
    require(e.kind == kind, what + " (wrong error kind)");
    return;
  }
  throw std::runtime_error("expected an error: " + what);
}

}  // namespace

int main() {
  int failures = 0;
  Mi355xBackend be(0);
  const auto P = BackendError::Kind::InvalidParameters;
  const auto S = BackendError::Kind::InvalidState;
  std::vector<std::pair<const char*, std::function<void()>>> tests = {
      {"evaluate_known_answer",
       [&] {
         auto poly = be.mle_from_evaluations(2, v({2, 3, 5, 7}));
         require(be.evaluate_mle(poly, v({2, 3})) == fr_from_u64(19), "[2,3,5,7](2,3) = 19");
         auto c = be.mle_from_evaluations(0, v({11}));
         require(be.evaluate_mle(c, {}) == fr_from_u64(11), "constant");
         auto same = be.mle_from_evaluations(2, v({2, 3, 5, 7}));
         require(same.id() != poly.id(), "identity is the allocation");
         expect_error(P, [&] { be.mle_from_evaluations(2, v({1, 2, 3})); }, "wrong length");
         expect_error(P, [&] { be.build_eq_x_r({}); }, "empty r");
       }},
      {"linear_combination_known_answer",
       [&] {
         auto p = be.mle_from_evaluations(2, v({2, 3, 5, 7}));
         auto q = be.mle_from_evaluations(2, v({11, 13, 17, 19}));
         auto comb = be.linear_combination({p, q, p}, {fr_from_u64(2), fr_from_u64(3), neg(2)});
         require(be.evaluate_mle(comb, v({2, 3})) == fr_from_u64(99), "lincomb = 99");
         expect_error(P, [&] { be.linear_combination({}, {}); }, "empty lincomb");
       }},
      {"product_tree_known_answer",
       [&] {
         auto frac = be.mle_from_evaluations(3, v({2, 3, 5, 7, 11, 13, 17, 19}));
         auto prod = be.compute_product_poly(frac);
         require(be.evaluations(prod) == v({6, 35, 143, 323, 210, 46189, 9699690, 0}), "tree");
         auto f = be.compute_product_factors(be.mle_from_evaluations(2, v({2, 3, 5, 7})),
                                             be.mle_from_evaluations(2, v({11, 13, 17, 19})));
         require(be.evaluations(f[0]) == v({2, 5, 11, 17}), "p1");
         require(be.evaluations(f[1]) == v({3, 7, 13, 19}), "p2");
       }},
      {"nums_denoms_scalar_columns",
       [&] {
         std::vector<Mle> fx, gx, pm;
         for (uint64_t x : {2, 3, 5}) fx.push_back(be.mle_from_evaluations(0, v({x})));
         for (uint64_t x : {7, 11, 13}) gx.push_back(be.mle_from_evaluations(0, v({x})));
         for (uint64_t x : {2, 0, 1}) pm.push_back(be.mle_from_evaluations(0, v({x})));
         auto [nums, dens] = be.compute_nums_and_denoms(fr_from_u64(2), fr_from_u64(3), fx, gx, pm);
         const uint64_t en[3] = {5, 8, 12}, ed[3] = {14, 14, 18};
         for (int i = 0; i < 3; ++i) {
           require(be.evaluations(nums[i])[0] == fr_from_u64(en[i]), "num");
           require(be.evaluations(dens[i])[0] == fr_from_u64(ed[i]), "den");
         }
       }},
      {"frac_rejects_zero_denominator",
       [&] {
         auto f = be.compute_frac_poly({be.mle_from_evaluations(0, v({12}))},
                                       {be.mle_from_evaluations(0, v({4}))});
         require(be.evaluations(f)[0] == fr_from_u64(3), "12/4 = 3");
         expect_error(P,
                      [&] {
                        be.compute_frac_poly({be.mle_from_evaluations(1, v({0, 2}))},
                                             {be.mle_from_evaluations(1, v({0, 3}))});
                      },
                      "zero denominator");
       }},
      {"sumcheck_round_lifecycle",
       [&] {
         auto f = be.mle_from_evaluations(2, v({2, 3, 5, 7}));
         auto vp = VirtualPolynomial::new_from_mle(f, fr_from_u64(2));
         vp.add_mle_list({f, f}, fr_from_u64(3));
         auto st = be.prover_init(vp);
         expect_error(S, [&] { be.prove_round_and_update_state(st, fr_from_u64(2)); }, "premature");
         require(be.prove_round_and_update_state(st, std::nullopt) == v({101, 194, 317}), "round 0");
         expect_error(S, [&] { be.prove_round_and_update_state(st, std::nullopt); }, "missing");
         require(be.prove_round_and_update_state(st, fr_from_u64(2)) == v({56, 261, 616}),
                 "round 1");
         expect_error(S, [&] { be.prove_round_and_update_state(st, fr_from_u64(2)); }, "exhausted");
       }},
      {"kzg_constant_commit_and_open_values",
       [&] {
         auto pp = be.generate_prover_param(3, v({5, 6, 7}));
         auto c = be.mle_from_evaluations(0, v({29}));
         // commitment of a constant 29 over the last layer [g] is 29 g
         auto g29 = be.commit(pp, c);
         auto one = be.commit(pp, be.mle_from_evaluations(0, v({1})));
         require(g29[0] == 0 && one[0] == 0 && g29 != one, "29 g is finite and differs from g");
         auto f = be.mle_from_evaluations(2, v({2, 3, 5, 7}));
         require(be.open(pp, f, v({0, 0})).second == fr_from_u64(2), "(0,0) -> 2");
         require(be.open(pp, f, v({1, 1})).second == fr_from_u64(7), "(1,1) -> 7");
         require(be.open(pp, f, v({2, 3})).second == fr_from_u64(19), "(2,3) -> 19");
         auto mc = be.multi_commit(pp, {f, c, f});
         require(mc[0] == be.commit(pp, f) && mc[1] == g29 && mc[2] == mc[0], "multi_commit");
       }},
  };
  for (auto& [name, fn] : tests) {
    try {
      fn();
      std::printf("PASS %s\n", name);
    } catch (const std::exception& e) {
      std::printf("FAIL %s: %s\n", name, e.what());
      ++failures;
    }
    std::fflush(stdout);
  }
  std::printf(failures ? "BACKEND_TEST FAIL\n" : "BACKEND_TEST PASS\n");
  return failures ? 1 : 0;
}
