"""VirtualPolynomial: a sum of products of shared MLE handles.

Mirrors backend/src/virtual_polynomial.rs:45-235: products are
(coefficient, factor indices) over handles stored once each in first-insertion
order and deduplicated by handle identity (never by contents); `mul_by_mle`
appends the new factor to every product and raises max_degree by one.  The
ABI encoding of the products (coefficients LE32, offsets[P+1] from 0, flat
indices) follows backend/src/ascend910b3/sum_check.rs:32-86.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Iterable, List, Tuple

from .errors import InvalidParameters
from .fields import R_MODULUS, fr_to_bytes


@dataclass
class VPAuxInfo:
    max_degree: int = 0
    num_variables: int = 0


class VirtualPolynomial:
    def __init__(self, num_variables: int = 0):
        self.aux_info = VPAuxInfo(0, num_variables)
        self.products: List[Tuple[int, List[int]]] = []
        self._flattened: List[object] = []
        self._lookup: Dict[object, int] = {}

    @classmethod
    def new_from_mle(cls, mle, coefficient: int) -> "VirtualPolynomial":
        vp = cls(mle.num_vars)
        vp.aux_info.max_degree = 1
        vp._flattened = [mle]
        vp._lookup = {mle.id: 0}
        vp.products = [(coefficient % R_MODULUS, [0])]
        return vp

    def flattened_ml_extensions(self) -> List[object]:
        return list(self._flattened)

    def _index_of(self, mle) -> int:
        idx = self._lookup.get(mle.id)
        if idx is None:
            idx = len(self._flattened)
            self._flattened.append(mle)
            self._lookup[mle.id] = idx
        return idx

    def add_mle_list(self, mle_list: Iterable[object], coefficient: int) -> None:
        mles = list(mle_list)
        if not mles:
            raise InvalidParameters("input mle_list is empty")
        for m in mles:
            if m.num_vars != self.aux_info.num_variables:
                raise InvalidParameters(
                    f"product has a multiplicand with wrong number of variables {m.num_vars} vs "
                    f"{self.aux_info.num_variables}"
                )
        self.aux_info.max_degree = max(self.aux_info.max_degree, len(mles))
        indices = [self._index_of(m) for m in mles]
        self.products.append((coefficient % R_MODULUS, indices))

    def mul_by_mle(self, mle, coefficient: int) -> None:
        if mle.num_vars != self.aux_info.num_variables:
            raise InvalidParameters(
                f"product has a multiplicand with wrong number of variables {mle.num_vars} vs "
                f"{self.aux_info.num_variables}"
            )
        idx = self._index_of(mle)
        self.products = [((c * coefficient) % R_MODULUS, idxs + [idx]) for c, idxs in self.products]
        self.aux_info.max_degree += 1

    def __add__(self, other: "VirtualPolynomial") -> "VirtualPolynomial":
        res = VirtualPolynomial(self.aux_info.num_variables)
        res.aux_info = VPAuxInfo(self.aux_info.max_degree, self.aux_info.num_variables)
        res.products = [(c, list(i)) for c, i in self.products]
        res._flattened = list(self._flattened)
        res._lookup = dict(self._lookup)
        flat = other.flattened_ml_extensions()
        for c, idxs in other.products:
            res.add_mle_list([flat[i] for i in idxs], c)
        return res


def encode_products(vp: VirtualPolynomial):
    """(coefficient bytes, offsets, indices) after the checks of
    backend/src/ascend910b3/sum_check.rs:32-86."""
    nv, d = vp.aux_info.num_variables, vp.aux_info.max_degree
    if nv == 0 or d == 0:
        raise InvalidParameters("SumCheck requires positive variable count and maximum degree")
    factors = vp.flattened_ml_extensions()
    if any(f.num_vars != nv for f in factors):
        raise InvalidParameters("SumCheck factor dimensions do not match virtual polynomial metadata")
    coeffs: List[int] = []
    offsets = [0]
    indices: List[int] = []
    for c, idxs in vp.products:
        if not idxs or len(idxs) > d:
            raise InvalidParameters("SumCheck product must have between one and max_degree multiplicands")
        if any(i >= len(factors) or i < 0 for i in idxs):
            raise InvalidParameters("SumCheck product references an out-of-range factor")
        coeffs.append(c)
        indices.extend(idxs)
        offsets.append(len(indices))
    return fr_to_bytes(coeffs), offsets, indices
