// Device MLE primitives shared by the sumcheck and KZG code paths.
#pragma once
#include "field.hpp"
#include "runtime.hpp"

namespace hp {
namespace mle {

// dst[b] = src[2b] + r (src[2b+1] - src[2b]) for b < half  (src != dst)
void fold(rt::Context& c, const Fr* src, Fr* dst, const Fr* r, size_t half);
// out[0] = table evaluated at point[0..nv) (point[0] binds the lowest bit).
// Tables above 2^10 entries fold through 2^(nv-1) + 2^(nv-2) scratch entries:
// `fold_scratch` when given (nothing is allocated), else a context buffer.
void evaluate(rt::Context& c, const Fr* table, size_t nv, const Fr* point, Fr* out,
              Fr* fold_scratch = nullptr);
// Look up this device's copy of the allocation-free entry points' static
// scratch array (called once per device, with that device current, by
// context_create); failures are reported by those entry points.
void noalloc_scratch_setup(int device);
// out[x] = eq(x, r) over {0,1}^nv
void build_eq(rt::Context& c, const Fr* r, size_t nv, Fr* out);

}  // namespace mle
}  // namespace hp
