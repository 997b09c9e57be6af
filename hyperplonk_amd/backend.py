"""Mi355xBackend: the host side of the `mi355x` backend over the C ABI.

One object implements the three traits the reference protocol layer calls:
  MleBackend            backend/src/multilinear_polynomial.rs:33-121
  MultilinearKzgBackend backend/src/multilinear_kzg.rs:20-55
  SumCheckProver        backend/src/sum_check.rs:17-38
with the same method names, argument meaning and error kinds, in the way the
reference's accelerator glue does (backend/src/ascend910b3/{mod,
multilinear_polynomial,multilinear_kzg,sum_check}.rs).  MLE handles are
device-resident tables; their identity is the device address; dropping the
last Python reference retires the handle.  Field elements are Python ints
(canonical) or, for whole tables, (n, 32) uint8 arrays; G1 results are
(x, y) tuples or None for the identity.
"""
from __future__ import annotations

import ctypes
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import abi
from .errors import InvalidParameters, InvalidState, check
from .fields import FR_BYTES, G1_BYTES, R_MODULUS, FrTable, fr_from_bytes, fr_to_bytes, g1_from_packet, table_bytes
from .virtual_polynomial import VirtualPolynomial, encode_products


class _Context:
    """Owns the native context; handles keep it alive until they are retired."""

    def __init__(self, device_id: int):
        self.lib = abi.load()
        raw = ctypes.c_void_p()
        check(self.lib.hp_mi355x_v1_context_create(device_id, ctypes.byref(raw)), "context_create")
        self.raw = raw.value

    def __del__(self):
        if getattr(self, "raw", None):
            self.lib.hp_mi355x_v1_context_destroy(self.raw)
            self.raw = None


class Mle:
    """Handle to an immutable device-resident table (`Mle` trait)."""

    __slots__ = ("_ctx", "address", "num_vars", "__weakref__")

    def __init__(self, ctx: _Context, address: int, num_vars: int):
        self._ctx = ctx
        self.address = address
        self.num_vars = num_vars

    @property
    def id(self) -> int:
        return self.address

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.raw and self.address:
            ctx.lib.hp_mi355x_v1_buffer_retire(ctx.raw, self.address)
            self.address = 0


class PreparedProverParam:
    def __init__(self, ctx: _Context, raw: int, num_vars: int):
        self._ctx = ctx
        self.raw = raw
        self.num_vars = num_vars

    def __del__(self):
        if getattr(self, "raw", None) and self._ctx.raw:
            self._ctx.lib.hp_mi355x_v1_multilinear_kzg_params_destroy(self._ctx.raw, self.raw)
            self.raw = None


class ProverState:
    def __init__(self, ctx: _Context, raw: int, num_vars: int, max_degree: int):
        self._ctx = ctx
        self.raw = raw
        self.num_vars = num_vars
        self.max_degree = max_degree

    def __del__(self):
        if getattr(self, "raw", None) and self._ctx.raw:
            self._ctx.lib.hp_mi355x_v1_sumcheck_prover_state_destroy(self._ctx.raw, self.raw)
            self.raw = None


def _fr_scalars(values: Sequence[int]) -> bytes:
    for v in values:
        if not 0 <= int(v) < R_MODULUS:
            raise InvalidParameters("field element is not canonical")
    return fr_to_bytes(values)


class _KzgRequest(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("polynomial", ctypes.c_void_p),
        ("num_vars", ctypes.c_size_t),
        ("point", ctypes.c_void_p),
        ("out_points", ctypes.c_void_p),
        ("out_evaluation", ctypes.c_void_p),
    ]


class Mi355xBackend:
    """MleBackend + MultilinearKzgBackend + SumCheckProver on one MI355X device."""

    def __init__(self, device_id: int = 0):
        self._ctx = _Context(device_id)
        self.lib = self._ctx.lib

    @property
    def context(self) -> int:
        return self._ctx.raw

    def _check_mle(self, mle: Mle) -> None:
        if not isinstance(mle, Mle) or mle._ctx is not self._ctx:
            raise InvalidParameters("MLE handle belongs to another backend context")

    def _new_mle(self, address: int, num_vars: int) -> Mle:
        return Mle(self._ctx, address, num_vars)

    # ------------------------------------------------------------ MleBackend
    def mle_from_evaluations(self, num_vars: int, evaluations: FrTable) -> Mle:
        data = table_bytes(evaluations)
        out = ctypes.c_void_p()
        count = len(data) // FR_BYTES
        if len(data) % FR_BYTES:
            raise InvalidParameters("evaluation bytes are not a whole number of Fr values")
        check(
            self.lib.hp_mi355x_v1_mle_upload(self.context, num_vars, data, count, ctypes.byref(out)),
            "mle_upload",
        )
        return self._new_mle(out.value, num_vars)

    def mle_to_evaluations(self, mle: Mle) -> np.ndarray:
        """Extension: the table back as an (n, 32) uint8 array of canonical values."""
        self._check_mle(mle)
        n = 1 << mle.num_vars
        buf = ctypes.create_string_buffer(n * FR_BYTES)
        check(self.lib.hp_mi355x_v1_mle_download(self.context, mle.address, mle.num_vars, buf, n), "mle_download")
        return np.frombuffer(buf.raw, dtype=np.uint8).reshape(n, FR_BYTES)

    def mle_values(self, mle: Mle) -> List[int]:
        return fr_from_bytes(self.mle_to_evaluations(mle).tobytes())

    def linear_combination(self, polynomials: Sequence[Mle], coefficients: Sequence[int]) -> Mle:
        if not polynomials:
            raise InvalidParameters("linear combination requires nonempty polynomials")
        if len(polynomials) != len(coefficients):
            raise InvalidParameters("linear combination requires matching counts and dimensions")
        nv = polynomials[0].num_vars
        for p in polynomials:
            self._check_mle(p)
            if p.num_vars != nv:
                raise InvalidParameters("linear combination requires matching counts and dimensions")
        out = ctypes.c_void_p()
        check(
            self.lib.hp_mi355x_v1_mle_linear_combination(
                self.context, nv, abi.ptr_array([p.address for p in polynomials]), _fr_scalars(coefficients),
                len(polynomials), ctypes.byref(out)
            ),
            "linear_combination",
        )
        return self._new_mle(out.value, nv)

    def evaluate_mle(self, mle: Mle, point: Sequence[int]) -> int:
        self._check_mle(mle)
        if len(point) != mle.num_vars:
            raise InvalidParameters(f"wrong number of variables {mle.num_vars} vs {len(point)}")
        out = ctypes.create_string_buffer(FR_BYTES)
        pt = _fr_scalars(point) if point else None
        check(self.lib.hp_mi355x_v1_mle_evaluate(self.context, mle.address, mle.num_vars, pt, out), "evaluate")
        return fr_from_bytes(out.raw)[0]

    def evaluate_mles(self, mles: Sequence[Mle], points: Sequence[Sequence[int]]) -> List[int]:
        """Many evaluations in one pass (extension: mle_evaluate_batch)."""
        n = len(mles)
        if len(points) != n:
            raise InvalidParameters("one point per polynomial")
        for m, pt in zip(mles, points):
            self._check_mle(m)
            if len(pt) != m.num_vars:
                raise InvalidParameters("wrong number of variables")
        bufs = [ctypes.create_string_buffer(_fr_scalars(pt), FR_BYTES * len(pt)) if pt else None for pt in points]
        addrs = (ctypes.c_void_p * max(n, 1))(*[m.address for m in mles])
        nvs = (ctypes.c_size_t * max(n, 1))(*[m.num_vars for m in mles])
        pts = (ctypes.c_void_p * max(n, 1))(*[ctypes.addressof(b) if b is not None else None for b in bufs])
        out = ctypes.create_string_buffer(FR_BYTES * max(n, 1))
        check(self.lib.hp_mi355x_v1_mle_evaluate_batch(self.context, addrs, nvs, pts, out, n), "evaluate batch")
        return fr_from_bytes(out.raw[: FR_BYTES * n])

    def msm_autotune(self, log_sizes: Sequence[int]) -> List[dict]:
        """Autotune window bits and task cap per layer size (extension); applies to
        parameters prepared afterwards."""
        n = len(log_sizes)
        sizes = (ctypes.c_size_t * max(n, 1))(*log_sizes)
        w = (ctypes.c_uint32 * max(n, 1))()
        cap = (ctypes.c_uint32 * max(n, 1))()
        ms = (ctypes.c_float * max(n, 1))()
        check(self.lib.hp_mi355x_v1_msm_autotune(self.context, sizes, n, w, cap, ms), "msm autotune")
        return [{"log_n": log_sizes[i], "window": w[i], "cap": cap[i], "ms": ms[i]} for i in range(n)]

    def msm_set_tuning(self, table: Sequence[dict]) -> None:
        n = len(table)
        sizes = (ctypes.c_size_t * max(n, 1))(*[t["log_n"] for t in table])
        w = (ctypes.c_uint32 * max(n, 1))(*[t["window"] for t in table])
        cap = (ctypes.c_uint32 * max(n, 1))(*[t["cap"] for t in table])
        check(self.lib.hp_mi355x_v1_msm_set_tuning(self.context, sizes, w, cap, n), "msm set tuning")

    def fold_profile(self, mle: Mle, r: int, reps: int = 20) -> float:
        """Mean ms of one fold of `mle` by its lowest variable (extension)."""
        self._check_mle(mle)
        ms = ctypes.c_float()
        check(self.lib.hp_mi355x_v1_mle_fold_profile(self.context, mle.address, mle.num_vars,
                                                     _fr_scalars([r]), reps, ctypes.byref(ms)), "fold profile")
        return ms.value

    def device_peaks(self, copy_bytes: int = 1 << 30, reps: int = 5) -> dict:
        """Extension: HBM copy rate (read + write bytes/s) and u32 x u32 -> u64
        multiply-add rate measured on this device (roofline denominators)."""
        hbm, mad = ctypes.c_double(), ctypes.c_double()
        check(self.lib.hp_mi355x_v1_device_peaks(self.context, copy_bytes, reps, ctypes.byref(hbm), ctypes.byref(mad)),
              "device_peaks")
        return {"hbm_bytes_per_s": hbm.value, "mad_per_s": mad.value, "copy_bytes": copy_bytes}

    def build_eq_x_r(self, r: Sequence[int]) -> Mle:
        if not r:
            raise InvalidParameters("r length is 0")
        out = ctypes.c_void_p()
        check(self.lib.hp_mi355x_v1_mle_build_eq_x_r(self.context, _fr_scalars(r), len(r), ctypes.byref(out)), "eq")
        return self._new_mle(out.value, len(r))

    def compute_frac_poly(self, fxs: Sequence[Mle], gxs: Sequence[Mle]) -> Mle:
        if not fxs or len(fxs) != len(gxs):
            raise InvalidParameters("frac requires nonempty, equally sized lists")
        nv = fxs[0].num_vars
        for p in list(fxs) + list(gxs):
            self._check_mle(p)
        out = ctypes.c_void_p()
        check(
            self.lib.hp_mi355x_v1_mle_compute_frac_poly(
                self.context, nv, abi.ptr_array([p.address for p in fxs]), abi.ptr_array([p.address for p in gxs]),
                len(fxs), ctypes.byref(out)
            ),
            "frac",
        )
        return self._new_mle(out.value, nv)

    def compute_product_poly(self, frac_poly: Mle) -> Mle:
        self._check_mle(frac_poly)
        out = ctypes.c_void_p()
        check(
            self.lib.hp_mi355x_v1_mle_compute_product_poly(
                self.context, frac_poly.address, frac_poly.num_vars, ctypes.byref(out)
            ),
            "product",
        )
        return self._new_mle(out.value, frac_poly.num_vars)

    def compute_nums_and_denoms(self, beta: int, gamma: int, fxs, gxs, perms):
        k = len(fxs)
        if k == 0 or len(gxs) != k or len(perms) != k:
            raise InvalidParameters("nums/denoms require nonempty, equally sized lists")
        nv = fxs[0].num_vars
        for p in list(fxs) + list(gxs) + list(perms):
            self._check_mle(p)
        on = (ctypes.c_void_p * k)()
        od = (ctypes.c_void_p * k)()
        check(
            self.lib.hp_mi355x_v1_mle_compute_nums_and_denoms(
                self.context, nv, _fr_scalars([beta]), _fr_scalars([gamma]),
                abi.ptr_array([p.address for p in fxs]), abi.ptr_array([p.address for p in gxs]),
                abi.ptr_array([p.address for p in perms]), k, on, od
            ),
            "nums_and_denoms",
        )
        return [self._new_mle(on[i], nv) for i in range(k)], [self._new_mle(od[i], nv) for i in range(k)]

    def compute_product_factors(self, frac_poly: Mle, prod_poly: Mle) -> Tuple[Mle, Mle]:
        self._check_mle(frac_poly)
        self._check_mle(prod_poly)
        if frac_poly.num_vars != prod_poly.num_vars:
            raise InvalidParameters("product factor inputs have different variable counts")
        p1, p2 = ctypes.c_void_p(), ctypes.c_void_p()
        check(
            self.lib.hp_mi355x_v1_mle_compute_product_factors(
                self.context, frac_poly.address, prod_poly.address, frac_poly.num_vars, ctypes.byref(p1),
                ctypes.byref(p2)
            ),
            "product_factors",
        )
        return self._new_mle(p1.value, frac_poly.num_vars), self._new_mle(p2.value, frac_poly.num_vars)

    # ------------------------------------------------- MultilinearKzgBackend
    def _new_params(self, num_vars: int) -> PreparedProverParam:
        raw = ctypes.c_void_p()
        check(self.lib.hp_mi355x_v1_multilinear_kzg_params_create(self.context, num_vars, ctypes.byref(raw)), "params")
        return PreparedProverParam(self._ctx, raw.value, num_vars)

    def prepare_prover_param(self, num_vars: int, powers_of_g: Iterable[bytes]) -> PreparedProverParam:
        """Consume N+1 layers (layer i: 2^(N-i) points as concatenated 97-byte packets)."""
        layers = list(powers_of_g)
        if len(layers) != num_vars + 1:
            raise InvalidParameters("prover parameters need num_vars + 1 layers")
        pp = self._new_params(num_vars)
        for i, layer in enumerate(layers):
            data = bytes(layer)
            check(
                self.lib.hp_mi355x_v1_multilinear_kzg_params_upload_layer(
                    self.context, pp.raw, i, data, len(data) // G1_BYTES
                ),
                f"upload layer {i}",
            )
        return pp

    def generate_prover_param(self, num_vars: int, t: Sequence[int]) -> PreparedProverParam:
        """Extension: build every layer on the device from a trapdoor t."""
        if len(t) != num_vars:
            raise InvalidParameters("trapdoor length must equal num_vars")
        pp = self._new_params(num_vars)
        data = _fr_scalars(t) if num_vars else None
        check(self.lib.hp_mi355x_v1_multilinear_kzg_params_generate(self.context, pp.raw, data), "params_generate")
        return pp

    def download_layer(self, pp: PreparedProverParam, layer: int) -> bytes:
        n = 1 << (pp.num_vars - layer)
        buf = ctypes.create_string_buffer(n * G1_BYTES)
        check(
            self.lib.hp_mi355x_v1_multilinear_kzg_params_download_layer(self.context, pp.raw, layer, buf, n),
            "download_layer",
        )
        return buf.raw

    def prepare_prover_param_compressed(self, num_vars: int, powers_of_g: Iterable[bytes]) -> PreparedProverParam:
        """Extension: N+1 layers in the compressed encoding of canonical SRS
        files (48 bytes a point), decompressed on the device."""
        layers = list(powers_of_g)
        if len(layers) != num_vars + 1:
            raise InvalidParameters("prover parameters need num_vars + 1 layers")
        pp = self._new_params(num_vars)
        for i, layer in enumerate(layers):
            data = bytes(layer)
            check(
                self.lib.hp_mi355x_v1_multilinear_kzg_params_upload_layer_compressed(
                    self.context, pp.raw, i, data, len(data) // 48
                ),
                f"upload compressed layer {i}",
            )
        return pp

    def download_layer_compressed(self, pp: PreparedProverParam, layer: int) -> bytes:
        n = 1 << (pp.num_vars - layer)
        buf = ctypes.create_string_buffer(n * 48)
        check(
            self.lib.hp_mi355x_v1_multilinear_kzg_params_download_layer_compressed(
                self.context, pp.raw, layer, buf, n
            ),
            "download_layer_compressed",
        )
        return buf.raw

    def commit_packet(self, pp: PreparedProverParam, polynomial: Mle) -> bytes:
        self._check_mle(polynomial)
        out = ctypes.create_string_buffer(G1_BYTES)
        check(
            self.lib.hp_mi355x_v1_multilinear_kzg_commit(
                self.context, pp.raw, polynomial.address, polynomial.num_vars, out
            ),
            "commit",
        )
        return out.raw

    def commit(self, pp: PreparedProverParam, polynomial: Mle):
        return g1_from_packet(self.commit_packet(pp, polynomial))

    def multi_commit_packets(self, pp: PreparedProverParam, polynomials: Sequence[Mle]) -> List[bytes]:
        if not polynomials:
            return []
        for p in polynomials:
            self._check_mle(p)
        out = ctypes.create_string_buffer(G1_BYTES * len(polynomials))
        check(
            self.lib.hp_mi355x_v1_multilinear_kzg_multi_commit(
                self.context, pp.raw, abi.ptr_array([p.address for p in polynomials]),
                abi.size_array([p.num_vars for p in polynomials]), len(polynomials), out
            ),
            "multi_commit",
        )
        return [out.raw[G1_BYTES * i : G1_BYTES * (i + 1)] for i in range(len(polynomials))]

    def multi_commit(self, pp: PreparedProverParam, polynomials: Sequence[Mle]):
        return [g1_from_packet(p) for p in self.multi_commit_packets(pp, polynomials)]

    def open_packets(self, pp: PreparedProverParam, polynomial: Mle, point: Sequence[int]):
        self._check_mle(polynomial)
        nv = polynomial.num_vars
        if len(point) != nv:
            raise InvalidParameters(f"Polynomial num_vars {nv} does not match point len {len(point)}")
        quot = ctypes.create_string_buffer(G1_BYTES * max(nv, 1))
        ev = ctypes.create_string_buffer(FR_BYTES)
        pt = _fr_scalars(point) if nv else None
        check(
            self.lib.hp_mi355x_v1_multilinear_kzg_open(self.context, pp.raw, polynomial.address, nv, pt, quot, ev),
            "open",
        )
        return [quot.raw[G1_BYTES * j : G1_BYTES * (j + 1)] for j in range(nv)], fr_from_bytes(ev.raw)[0]

    def open(self, pp: PreparedProverParam, polynomial: Mle, point: Sequence[int]):
        packets, value = self.open_packets(pp, polynomial, point)
        return [g1_from_packet(p) for p in packets], value

    def msm_profile(self, pp: PreparedProverParam, polynomial: Mle):
        """Extension: per-stage device times (ms) of one commit-sized MSM and its plan."""
        self._check_mle(polynomial)
        ms = (ctypes.c_float * 8)()
        info = (ctypes.c_uint32 * 4)()
        check(
            self.lib.hp_mi355x_v1_msm_profile(
                self.context, pp.raw, polynomial.address, polynomial.num_vars, ms, 8, info
            ),
            "msm_profile",
        )
        names = ("digits", "sort", "tasks", "accumulate", "bucket_sums", "reduce", "device_total")
        return {k: float(ms[i]) for i, k in enumerate(names)}, {
            "c": info[0], "windows": info[1], "buckets": info[2], "entries": info[3]
        }

    # ---------------------------------------------------------- SumCheckProver
    def prover_init(self, polynomial: VirtualPolynomial) -> ProverState:
        coeffs, offsets, indices = encode_products(polynomial)
        factors = polynomial.flattened_ml_extensions()
        for f in factors:
            self._check_mle(f)
        raw = ctypes.c_void_p()
        check(
            self.lib.hp_mi355x_v1_sumcheck_prover_state_create(
                self.context, polynomial.aux_info.num_variables, polynomial.aux_info.max_degree,
                abi.ptr_array([f.address for f in factors]), len(factors), coeffs if coeffs else None,
                abi.size_array(offsets), len(polynomial.products), abi.size_array(indices), len(indices),
                ctypes.byref(raw)
            ),
            "prover_init",
        )
        return ProverState(self._ctx, raw.value, polynomial.aux_info.num_variables, polynomial.aux_info.max_degree)

    def prove_round_and_update_state(self, state: ProverState, challenge: Optional[int]) -> List[int]:
        count = state.max_degree + 1
        out = ctypes.create_string_buffer(FR_BYTES * count)
        ch = _fr_scalars([challenge]) if challenge is not None else None
        check(
            self.lib.hp_mi355x_v1_sumcheck_prover_state_round(self.context, state.raw, ch, out, count),
            "sumcheck round",
        )
        return fr_from_bytes(out.raw)

    _KzgRequestType = _KzgRequest

    def _kzg_requests(self, ops: Sequence[tuple]):
        reqs = (_KzgRequest * max(len(ops), 1))()
        keep = []
        for i, op in enumerate(ops):
            kind, mle = op[0], op[1]
            self._check_mle(mle)
            nv = mle.num_vars
            r = reqs[i]
            r.polynomial = mle.address
            r.num_vars = nv
            if kind == "commit":
                r.kind = 0
                out = ctypes.create_string_buffer(97)
                keep.append((out, None))
            else:
                r.kind = 1
                point = op[2]
                if len(point) != nv:
                    raise InvalidParameters("polynomial num_vars does not match point length")
                pbuf = ctypes.create_string_buffer(_fr_scalars(point), FR_BYTES * nv) if nv else None
                out = ctypes.create_string_buffer(97 * max(nv, 1))
                ev = ctypes.create_string_buffer(FR_BYTES)
                keep.append((out, ev, pbuf))
                r.point = ctypes.addressof(pbuf) if pbuf is not None else None
                r.out_evaluation = ctypes.addressof(ev)
            r.out_points = ctypes.addressof(out)
        return reqs, keep

    @staticmethod
    def _kzg_results(ops: Sequence[tuple], keep) -> List[tuple]:
        res = []
        for op, k in zip(ops, keep):
            if op[0] == "commit":
                res.append((k[0].raw,))
            else:
                nv = op[1].num_vars
                res.append(([k[0].raw[97 * j : 97 * (j + 1)] for j in range(nv)], fr_from_bytes(k[1].raw)[0]))
        return res

    def kzg_batch(self, pp: PreparedProverParam, ops: Sequence[tuple]) -> List[tuple]:
        """Commits ("commit", mle) and opens ("open", mle, point) with all their
        MSMs in one batch (extension: multilinear_kzg_batch).  Returns packets
        per op: (commitment,) or (quotient packets, evaluation)."""
        reqs, keep = self._kzg_requests(ops)
        check(self.lib.hp_mi355x_v1_multilinear_kzg_batch(self.context, pp.raw, reqs, len(ops)), "kzg batch")
        return self._kzg_results(ops, keep)

    def kzg_batch_submit(self, pp: PreparedProverParam, ops: Sequence[tuple]) -> "KzgPending":
        """kzg_batch queued beside later calls (extension: batch_submit); the
        returned object's wait() gives the same results."""
        reqs, keep = self._kzg_requests(ops)
        raw = ctypes.c_void_p()
        check(self.lib.hp_mi355x_v1_multilinear_kzg_batch_submit(self.context, pp.raw, reqs, len(ops),
                                                                  ctypes.byref(raw)), "kzg batch submit")
        return KzgPending(self, pp, list(ops), reqs, keep, raw.value)

    def prove_rounds_batch(self, states: Sequence[ProverState], challenges: Sequence[Optional[int]]) -> List[List[int]]:
        """One round of each distinct state in shared launches (extension)."""
        n = len(states)
        if len(challenges) != n:
            raise InvalidParameters("one challenge slot per state")
        outs = [ctypes.create_string_buffer(FR_BYTES * (s.max_degree + 1)) for s in states]
        chs = [ctypes.create_string_buffer(_fr_scalars([c]), FR_BYTES) if c is not None else None for c in challenges]
        st_arr = (ctypes.c_void_p * max(n, 1))(*[s.raw for s in states])
        ch_arr = (ctypes.c_void_p * max(n, 1))(*[ctypes.addressof(c) if c is not None else None for c in chs])
        out_arr = (ctypes.c_void_p * max(n, 1))(*[ctypes.addressof(o) for o in outs])
        cnt_arr = (ctypes.c_size_t * max(n, 1))(*[s.max_degree + 1 for s in states])
        check(
            self.lib.hp_mi355x_v1_sumcheck_prover_state_round_batch(self.context, st_arr, ch_arr, out_arr, cnt_arr, n),
            "sumcheck round batch",
        )
        return [fr_from_bytes(o.raw) for o in outs]


class KzgPending:
    """A submitted KZG batch: its request and output buffers stay alive here
    until wait() (called once) writes and returns the results."""

    def __init__(self, be: Mi355xBackend, pp: PreparedProverParam, ops, reqs, keep, raw: int):
        self._be, self._pp, self._ops, self._reqs, self._keep, self._raw = be, pp, ops, reqs, keep, raw

    def wait(self) -> List[tuple]:
        if not self._raw:
            raise InvalidParameters("batch already waited for")
        raw, self._raw = self._raw, None
        check(self._be.lib.hp_mi355x_v1_multilinear_kzg_batch_wait(self._be.context, raw), "kzg batch wait")
        return Mi355xBackend._kzg_results(self._ops, self._keep)

    def __del__(self):
        if getattr(self, "_raw", None) and self._be._ctx.raw:
            self._be.lib.hp_mi355x_v1_multilinear_kzg_batch_wait(self._be.context, self._raw)
            self._raw = None
