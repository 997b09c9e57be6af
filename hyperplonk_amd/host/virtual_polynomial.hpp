// VirtualPolynomial / VPAuxInfo (backend/src/virtual_polynomial.rs:45-235):
// a sum of coefficient-weighted products of shared MLE handles, with factors
// deduplicated by handle identity in first-insertion order.  Generic over the
// handle type (anything with num_vars() and id()), so the same protocol code
// drives every backend.  Coefficients are canonical little-endian Fr bytes,
// the encoding the backend boundary takes.
#pragma once
#include <algorithm>
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "ff.hpp"

namespace hyperplonk {

using FrBytes = std::array<uint8_t, 32>;
using G1Packet = std::array<uint8_t, 97>;

// backend/src/errors.rs:16-21
struct BackendError : std::runtime_error {
  enum class Kind { InvalidParameters, InvalidState };
  Kind kind;
  int32_t native_code;
  BackendError(Kind k, const std::string& what, int32_t code = 0)
      : std::runtime_error(what), kind(k), native_code(code) {}
};

[[noreturn]] inline void invalid(const char* what) {
  throw BackendError(BackendError::Kind::InvalidParameters, what);
}

// VPAuxInfo serialises (CanonicalSerialize) as max_degree, num_variables: two u64 LE
struct VPAuxInfo {
  size_t max_degree = 0;
  size_t num_variables = 0;
};

template <class M>
class BasicVirtualPolynomial {
 public:
  explicit BasicVirtualPolynomial(size_t num_variables = 0) {
    aux_info.num_variables = num_variables;
  }
  static BasicVirtualPolynomial new_from_mle(const M& mle, const FrBytes& coefficient) {
    BasicVirtualPolynomial vp(mle.num_vars());
    vp.aux_info.max_degree = 1;
    vp.flattened_.push_back(mle);
    vp.lookup_[mle.id()] = 0;
    vp.products.push_back({coefficient, {0}});
    return vp;
  }
  void add_mle_list(const std::vector<M>& mles, const FrBytes& coefficient) {
    if (mles.empty()) invalid("input mle_list is empty");
    for (const auto& m : mles)
      if (m.num_vars() != aux_info.num_variables)
        invalid("product has a multiplicand with wrong number of variables");
    aux_info.max_degree = std::max(aux_info.max_degree, mles.size());
    std::vector<size_t> idx;
    idx.reserve(mles.size());
    for (const auto& m : mles) idx.push_back(index_of(m));
    products.push_back({coefficient, idx});
  }
  // multiply every product by mle and coefficient; degree grows by one
  void mul_by_mle(const M& mle, const FrBytes& coefficient) {
    if (mle.num_vars() != aux_info.num_variables)
      invalid("product has a multiplicand with wrong number of variables");
    const size_t i = index_of(mle);
    host::Fr c;
    if (!host::Fr::from_bytes(coefficient.data(), &c)) invalid("non-canonical coefficient");
    for (auto& [coeff, idx] : products) {
      idx.push_back(i);
      host::Fr p;
      if (!host::Fr::from_bytes(coeff.data(), &p)) invalid("non-canonical coefficient");
      coeff = (p * c).bytes();
    }
    aux_info.max_degree += 1;
  }
  const std::vector<M>& flattened_ml_extensions() const { return flattened_; }

  VPAuxInfo aux_info;
  std::vector<std::pair<FrBytes, std::vector<size_t>>> products;

 private:
  size_t index_of(const M& m) {
    auto it = lookup_.find(m.id());
    if (it != lookup_.end()) return it->second;
    const size_t i = flattened_.size();
    flattened_.push_back(m);
    lookup_[m.id()] = i;
    return i;
  }
  std::vector<M> flattened_;
  std::unordered_map<uintptr_t, size_t> lookup_;
};

}  // namespace hyperplonk
