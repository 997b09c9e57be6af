"""ctypes binding of the C ABI declared in include/hp_mi355x.h.

The library is built in-tree (hyperplonk_amd/libhp_mi355x.so).  There is no
fallback: if the library is missing, importing the backend raises.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhp_mi355x.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "hp_mi355x.h")

SUCCESS, INVALID_PARAMETERS, INVALID_STATE, NOT_IMPLEMENTED = 0, 1, 2, 3


class Status(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("native_code", ctypes.c_int32)]


_vp = ctypes.c_void_p
_vpp = ctypes.POINTER(ctypes.c_void_p)
_u8p = ctypes.c_char_p
_sz = ctypes.c_size_t
_szp = ctypes.POINTER(ctypes.c_size_t)

SIGNATURES = {
    "hp_mi355x_v1_context_create": [ctypes.c_int32, _vpp],
    "hp_mi355x_v1_context_destroy": [_vp],
    "hp_mi355x_v1_buffer_retire": [_vp, _vp],
    "hp_mi355x_v1_mle_upload": [_vp, _sz, _u8p, _sz, _vpp],
    "hp_mi355x_v1_mle_linear_combination": [_vp, _sz, _vpp, _u8p, _sz, _vpp],
    "hp_mi355x_v1_mle_evaluate": [_vp, _vp, _sz, _u8p, _u8p],
    "hp_mi355x_v1_mle_build_eq_x_r": [_vp, _u8p, _sz, _vpp],
    "hp_mi355x_v1_mle_compute_frac_poly": [_vp, _sz, _vpp, _vpp, _sz, _vpp],
    "hp_mi355x_v1_mle_compute_product_poly": [_vp, _vp, _sz, _vpp],
    "hp_mi355x_v1_mle_compute_nums_and_denoms": [_vp, _sz, _u8p, _u8p, _vpp, _vpp, _vpp, _sz, _vpp, _vpp],
    "hp_mi355x_v1_mle_compute_product_factors": [_vp, _vp, _vp, _sz, _vpp, _vpp],
    "hp_mi355x_v1_multilinear_kzg_params_create": [_vp, _sz, _vpp],
    "hp_mi355x_v1_multilinear_kzg_params_upload_layer": [_vp, _vp, _sz, _u8p, _sz],
    "hp_mi355x_v1_multilinear_kzg_params_destroy": [_vp, _vp],
    "hp_mi355x_v1_multilinear_kzg_commit": [_vp, _vp, _vp, _sz, _u8p],
    "hp_mi355x_v1_multilinear_kzg_multi_commit": [_vp, _vp, _vpp, _szp, _sz, _u8p],
    "hp_mi355x_v1_multilinear_kzg_open": [_vp, _vp, _vp, _sz, _u8p, _u8p, _u8p],
    "hp_mi355x_v1_sumcheck_prover_state_create": [_vp, _sz, _sz, _vpp, _sz, _u8p, _szp, _sz, _szp, _sz, _vpp],
    "hp_mi355x_v1_sumcheck_prover_state_round": [_vp, _vp, _u8p, _u8p, _sz],
    "hp_mi355x_v1_sumcheck_prover_state_destroy": [_vp, _vp],
    "hp_mi355x_v1_mle_download": [_vp, _vp, _sz, _u8p, _sz],
    "hp_mi355x_v1_multilinear_kzg_params_generate": [_vp, _vp, _u8p],
    "hp_mi355x_v1_multilinear_kzg_params_download_layer": [_vp, _vp, _sz, _u8p, _sz],
    "hp_mi355x_v1_multilinear_kzg_params_upload_layer_compressed": [_vp, _vp, _sz, _u8p, _sz],
    "hp_mi355x_v1_multilinear_kzg_params_download_layer_compressed": [_vp, _vp, _sz, _u8p, _sz],
    "hp_mi355x_v1_msm_profile": [
        _vp, _vp, _vp, _sz, ctypes.POINTER(ctypes.c_float), _sz, ctypes.POINTER(ctypes.c_uint32)
    ],
    "hp_mi355x_v1_multilinear_kzg_batch": [_vp, _vp, _vp, _sz],
    "hp_mi355x_v1_msm_autotune": [
        _vp, ctypes.POINTER(_sz), _sz, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
        ctypes.POINTER(ctypes.c_float),
    ],
    "hp_mi355x_v1_msm_set_tuning": [
        _vp, ctypes.POINTER(_sz), ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32), _sz
    ],
    "hp_mi355x_v1_mle_fold_profile": [_vp, _vp, _sz, _u8p, _sz, ctypes.POINTER(ctypes.c_float)],
    "hp_mi355x_v1_mle_evaluate_batch": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_sz), ctypes.POINTER(_vp), _u8p, _sz],
    "hp_mi355x_v1_debug_inject_failure": [_vp, ctypes.c_uint32, ctypes.c_int32],
    "hp_mi355x_v1_memory_info": [_vp, _szp, _szp, _szp, _szp, ctypes.c_int32],
    "hp_mi355x_v1_build_info": [ctypes.c_char_p, _sz],
    "hp_mi355x_v1_multilinear_kzg_batch_submit": [_vp, _vp, _vp, _sz, _vpp],
    "hp_mi355x_v1_multilinear_kzg_batch_wait": [_vp, _vp],
    "hp_mi355x_v1_mle_evaluate_noalloc": [_vp, _vp, _sz, _u8p, _u8p],
    "hp_mi355x_v1_device_peaks": [
        _vp, _sz, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)
    ],
    "hp_mi355x_v1_sumcheck_prover_state_round_batch": [
        _vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_sz), _sz
    ],
}

_LIB = None


class LibraryMissing(RuntimeError):
    pass


def header_symbols(path: str = HEADER_PATH):
    """Every hp_mi355x_* function declared in the public header."""
    with open(path) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(hp_mi355x_v1_[a-z0-9_]+)\s*\(", text)))


def load(path: str = LIB_PATH):
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise LibraryMissing(f"{path} is not built; run __graft_entry__.build() or make -C hyperplonk_amd/csrc")
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = Status
    _LIB = lib
    return lib


def ptr_array(values):
    arr = (ctypes.c_void_p * max(len(values), 1))()
    for i, v in enumerate(values):
        arr[i] = v
    return arr


def size_array(values):
    arr = (ctypes.c_size_t * max(len(values), 1))()
    for i, v in enumerate(values):
        arr[i] = v
    return arr


def source_digest() -> str:
    """SHA-256 of the library's sources as the build computes it: every .hip and
    .hpp file of hyperplonk_amd/csrc in name order, then include/hp_mi355x.h."""
    import hashlib

    csrc = os.path.join(HERE, "csrc")
    names = sorted(n for n in os.listdir(csrc) if n.endswith((".hip", ".hpp")))
    h = hashlib.sha256()
    for n in names:
        with open(os.path.join(csrc, n), "rb") as f:
            h.update(f.read())
    with open(HEADER_PATH, "rb") as f:
        h.update(f.read())
    return h.hexdigest()


def build_info() -> str:
    buf = ctypes.create_string_buffer(128)
    st = load().hp_mi355x_v1_build_info(buf, len(buf))
    if st.kind != SUCCESS:
        raise LibraryMissing("build_info failed")
    return buf.value.decode()
