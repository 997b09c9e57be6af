"""MI355X-native HyperPlonk prover backend (the `mi355x` backend).

Hand-written HIP kernels for gfx950 behind the C ABI of include/hp_mi355x.h
(built in-tree as hyperplonk_amd/libhp_mi355x.so), and a host mirror of the
reference's MleBackend / MultilinearKzgBackend / SumCheckProver traits.
"""
from .errors import BackendError, InvalidParameters, InvalidState
from .fields import Q_MODULUS, R_MODULUS
from .virtual_polynomial import VPAuxInfo, VirtualPolynomial

__all__ = [
    "BackendError",
    "InvalidParameters",
    "InvalidState",
    "Mi355xBackend",
    "Mle",
    "Q_MODULUS",
    "R_MODULUS",
    "VPAuxInfo",
    "VirtualPolynomial",
]


def __getattr__(name):
    # the backend loads the native library on first use, not at import time
    if name in ("Mi355xBackend", "Mle"):
        from . import backend

        return getattr(backend, name)
    raise AttributeError(name)
