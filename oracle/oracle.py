"""Oracle (test infrastructure only): ctypes front end of liboracle.so.

Only the test suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of ``bench.py`` import this module, and only as the checker.  The product
library never links or calls it.

Values cross this interface in the reference's canonical host encodings
(backend/src/ascend910b3/encoding.rs:12-108): Fr = 32-byte little-endian
integer, Fq = 48 bytes, G1 = 97-byte packet (tag 0 finite / 1 identity,
x LE48, y LE48).  Python ints are accepted wherever a field element is
expected, and G1 points are exchanged as ``None`` (identity) or ``(x, y)``.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence, Tuple

R_MODULUS = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
Q_MODULUS = int(
    "1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab",
    16,
)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OK, INVALID_PARAMETERS, INVALID_STATE = 0, 1, 2


class OracleError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: oracle status {status}")
        self.status = status


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise OracleError(-1, f"{path} is not built (run make -C oracle)")
        so = ctypes.CDLL(path)
        u8p = ctypes.c_char_p
        sz = ctypes.c_size_t
        i32 = ctypes.c_int
        szp = ctypes.POINTER(ctypes.c_size_t)
        sigs = {
            "oracle_init": (None, []),
            "oracle_fr_op": (i32, [i32, u8p, u8p, u8p, sz]),
            "oracle_fq_op": (i32, [i32, u8p, u8p, u8p, sz]),
            "oracle_g1_generator": (i32, [u8p]),
            "oracle_g1_add": (i32, [u8p, u8p, u8p]),
            "oracle_g1_mul": (i32, [u8p, u8p, u8p]),
            "oracle_g1_on_curve": (i32, [u8p]),
            "oracle_msm_naive": (i32, [u8p, u8p, sz, u8p]),
            "oracle_msm_pippenger": (i32, [u8p, u8p, sz, i32, u8p]),
            "oracle_mle_evaluate": (i32, [u8p, sz, u8p, u8p]),
            "oracle_mle_fold": (i32, [u8p, sz, u8p, u8p]),
            "oracle_mle_linear_combination": (i32, [u8p, u8p, sz, sz, u8p]),
            "oracle_eq_x_r": (i32, [u8p, sz, u8p]),
            "oracle_eq_eval": (i32, [u8p, u8p, sz, u8p]),
            "oracle_frac": (i32, [u8p, u8p, sz, sz, u8p]),
            "oracle_product": (i32, [u8p, sz, u8p]),
            "oracle_nums_denoms": (i32, [u8p, u8p, u8p, u8p, u8p, sz, sz, u8p, u8p]),
            "oracle_product_factors": (i32, [u8p, u8p, sz, u8p, u8p]),
            "oracle_sumcheck_create": (
                ctypes.c_void_p,
                [sz, sz, u8p, sz, u8p, szp, sz, szp, sz, ctypes.POINTER(ctypes.c_int)],
            ),
            "oracle_sumcheck_round": (i32, [ctypes.c_void_p, u8p, u8p, sz]),
            "oracle_sumcheck_destroy": (None, [ctypes.c_void_p]),
            "oracle_kzg_open": (i32, [u8p, u8p, sz, u8p, i32, u8p, u8p]),
            "oracle_srs_layer": (i32, [u8p, sz, i32, u8p]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(so, name)
            fn.restype = res
            fn.argtypes = args
        so.oracle_init()
        _LIB = so
    return _LIB


# ---------------------------------------------------------------- encodings
def fr_bytes(values: Sequence[int]) -> bytes:
    return b"".join(int(v).to_bytes(32, "little") for v in values)


def fr_ints(data: bytes) -> List[int]:
    return [int.from_bytes(data[i : i + 32], "little") for i in range(0, len(data), 32)]


def fq_bytes(values: Sequence[int]) -> bytes:
    return b"".join(int(v).to_bytes(48, "little") for v in values)


def fq_ints(data: bytes) -> List[int]:
    return [int.from_bytes(data[i : i + 48], "little") for i in range(0, len(data), 48)]


def g1_packet(point: Optional[Tuple[int, int]]) -> bytes:
    if point is None:
        return b"\x01" + bytes(96)
    x, y = point
    return b"\x00" + x.to_bytes(48, "little") + y.to_bytes(48, "little")


def g1_point(packet: bytes) -> Optional[Tuple[int, int]]:
    if packet[0] == 1:
        return None
    return (int.from_bytes(packet[1:49], "little"), int.from_bytes(packet[49:97], "little"))


def _check(status: int, what: str) -> None:
    if status != OK:
        raise OracleError(status, what)


# ------------------------------------------------------------------- fields
def field_op(field: str, op: str, a: Sequence[int], b: Optional[Sequence[int]] = None) -> List[int]:
    codes = {"add": 0, "sub": 1, "mul": 2, "neg": 3, "square": 4, "inverse": 5}
    width = 32 if field == "fr" else 48
    enc = fr_bytes if field == "fr" else fq_bytes
    dec = fr_ints if field == "fr" else fq_ints
    n = len(a)
    out = ctypes.create_string_buffer(width * max(n, 1))
    bb = enc(b) if b is not None else enc(a)
    fn = lib().oracle_fr_op if field == "fr" else lib().oracle_fq_op
    _check(fn(codes[op], enc(a), bb, out, n), f"{field} {op}")
    return dec(out.raw[: width * n])


# ----------------------------------------------------------------------- G1
def g1_generator() -> Tuple[int, int]:
    out = ctypes.create_string_buffer(97)
    _check(lib().oracle_g1_generator(out), "generator")
    return g1_point(out.raw)


def g1_add(p, q):
    out = ctypes.create_string_buffer(97)
    _check(lib().oracle_g1_add(g1_packet(p), g1_packet(q), out), "g1 add")
    return g1_point(out.raw)


def g1_mul(p, k: int):
    out = ctypes.create_string_buffer(97)
    _check(lib().oracle_g1_mul(g1_packet(p), fr_bytes([k % R_MODULUS]), out), "g1 mul")
    return g1_point(out.raw)


def msm_naive_packets(points97: bytes, scalars32: bytes, n: int) -> bytes:
    out = ctypes.create_string_buffer(97)
    _check(lib().oracle_msm_naive(points97, scalars32, n, out), "msm naive")
    return out.raw


def msm_pippenger_packets(points97: bytes, scalars32: bytes, n: int, threads: int = 8) -> bytes:
    out = ctypes.create_string_buffer(97)
    _check(lib().oracle_msm_pippenger(points97, scalars32, n, threads, out), "msm pippenger")
    return out.raw


# ---------------------------------------------------------------------- MLE
def _table_bytes(table) -> bytes:
    """A table as canonical LE32 bytes: raw bytes (or a uint8 array) pass through."""
    if isinstance(table, (bytes, bytearray)):
        return bytes(table)
    if hasattr(table, "tobytes") and getattr(table, "dtype", None) is not None and table.dtype.itemsize == 1:
        return table.tobytes()
    return fr_bytes(table)


def mle_evaluate(table, point: Sequence[int]) -> int:
    nv = len(point)
    out = ctypes.create_string_buffer(32)
    _check(lib().oracle_mle_evaluate(_table_bytes(table), nv, fr_bytes(point), out), "evaluate")
    return fr_ints(out.raw)[0]


def mle_fold(table: Sequence[int], r: int) -> List[int]:
    nv = (len(table)).bit_length() - 1
    out = ctypes.create_string_buffer(32 * (len(table) // 2))
    _check(lib().oracle_mle_fold(fr_bytes(table), nv, fr_bytes([r]), out), "fold")
    return fr_ints(out.raw)


def linear_combination(tables: Sequence[Sequence[int]], coeffs: Sequence[int]) -> List[int]:
    n = len(tables[0])
    nv = n.bit_length() - 1
    out = ctypes.create_string_buffer(32 * n)
    flat = b"".join(fr_bytes(t) for t in tables)
    _check(lib().oracle_mle_linear_combination(flat, fr_bytes(coeffs), len(tables), nv, out), "lincomb")
    return fr_ints(out.raw)


def eq_x_r(r: Sequence[int]) -> List[int]:
    out = ctypes.create_string_buffer(32 * (1 << len(r)))
    _check(lib().oracle_eq_x_r(fr_bytes(r), len(r), out), "eq")
    return fr_ints(out.raw)


def eq_eval(x: Sequence[int], y: Sequence[int]) -> int:
    out = ctypes.create_string_buffer(32)
    _check(lib().oracle_eq_eval(fr_bytes(x), fr_bytes(y), len(x), out), "eq_eval")
    return fr_ints(out.raw)[0]


def frac(nums: Sequence[Sequence[int]], dens: Sequence[Sequence[int]]) -> List[int]:
    n = len(nums[0])
    out = ctypes.create_string_buffer(32 * n)
    st = lib().oracle_frac(
        b"".join(fr_bytes(t) for t in nums), b"".join(fr_bytes(t) for t in dens), len(nums), n.bit_length() - 1, out
    )
    _check(st, "frac")
    return fr_ints(out.raw)


def product(frac_table: Sequence[int]) -> List[int]:
    n = len(frac_table)
    out = ctypes.create_string_buffer(32 * n)
    _check(lib().oracle_product(fr_bytes(frac_table), n.bit_length() - 1, out), "product")
    return fr_ints(out.raw)


def nums_denoms(beta, gamma, fxs, gxs, perms):
    n = len(fxs[0])
    k = len(fxs)
    on = ctypes.create_string_buffer(32 * n * k)
    od = ctypes.create_string_buffer(32 * n * k)
    flat = lambda ts: b"".join(fr_bytes(t) for t in ts)  # noqa: E731
    st = lib().oracle_nums_denoms(
        fr_bytes([beta]), fr_bytes([gamma]), flat(fxs), flat(gxs), flat(perms), k, n.bit_length() - 1, on, od
    )
    _check(st, "nums_denoms")
    nums = fr_ints(on.raw)
    dens = fr_ints(od.raw)
    return [nums[c * n : (c + 1) * n] for c in range(k)], [dens[c * n : (c + 1) * n] for c in range(k)]


def product_factors(frac_table, prod_table):
    n = len(frac_table)
    p1 = ctypes.create_string_buffer(32 * n)
    p2 = ctypes.create_string_buffer(32 * n)
    st = lib().oracle_product_factors(fr_bytes(frac_table), fr_bytes(prod_table), n.bit_length() - 1, p1, p2)
    _check(st, "product_factors")
    return fr_ints(p1.raw), fr_ints(p2.raw)


# ----------------------------------------------------------------- SumCheck
class SumCheckState:
    """Oracle prover state over raw tables (factor_count x 2^nv)."""

    def __init__(self, nv: int, max_degree: int, tables: Sequence[Sequence[int]], products):
        self.nv = nv
        self.max_degree = max_degree
        coeffs = [c % R_MODULUS for c, _ in products]
        offsets = [0]
        indices: List[int] = []
        for _, idx in products:
            indices.extend(idx)
            offsets.append(len(indices))
        off_arr = (ctypes.c_size_t * len(offsets))(*offsets)
        idx_arr = (ctypes.c_size_t * max(len(indices), 1))(*indices)
        status = ctypes.c_int(0)
        flat = b"".join(_table_bytes(t) for t in tables)
        self._h = lib().oracle_sumcheck_create(
            nv,
            max_degree,
            flat,
            len(tables),
            fr_bytes(coeffs),
            off_arr,
            len(products),
            idx_arr,
            len(indices),
            ctypes.byref(status),
        )
        if not self._h:
            raise OracleError(status.value, "sumcheck create")

    def round(self, challenge: Optional[int]) -> List[int]:
        out = ctypes.create_string_buffer(32 * (self.max_degree + 1))
        ch = fr_bytes([challenge]) if challenge is not None else None
        _check(lib().oracle_sumcheck_round(self._h, ch, out, self.max_degree + 1), "sumcheck round")
        return fr_ints(out.raw)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_sumcheck_destroy(self._h)
            self._h = None


# ---------------------------------------------------------------------- KZG
def kzg_open_packets(layers97: bytes, table: Sequence[int], point: Sequence[int], threads: int = 8):
    nv = len(point)
    quot = ctypes.create_string_buffer(97 * max(nv, 1))
    ev = ctypes.create_string_buffer(32)
    st = lib().oracle_kzg_open(layers97, _table_bytes(table), nv, fr_bytes(point), threads, quot, ev)
    _check(st, "kzg open")
    return [quot.raw[97 * j : 97 * (j + 1)] for j in range(nv)], fr_ints(ev.raw)[0]


def srs_layer_packets(t: Sequence[int], threads: int = 8) -> bytes:
    n = 1 << len(t)
    out = ctypes.create_string_buffer(97 * n)
    _check(lib().oracle_srs_layer(fr_bytes(t), len(t), threads, out), "srs layer")
    return out.raw
