"""Lifecycle and failure contracts of the C ABI, after the reference
accelerator backend's integration tests (backend/src/ascend910b3/integration_tests.rs)
and its bridge's validation and poisoning rules:

* zero denominators found across the whole table at num_vars 12, at the first
  and the last entry, x/0 and 0/0, as validation failures that leave the
  context usable (integration_tests.rs:721-813);
* two contexts reject each other's tables, parameters and prover states and
  stay usable (:1261-1315);
* parameters, input tables and the virtual polynomial may be dropped while a
  prover state still runs; the context and a fresh one on the same device
  keep working (:1317-1386);
* handles are looked up before anything is read through them: double
  destroys, use after destroy and a handle of the wrong kind are
  INVALID_PARAMETERS;
* a device failure (injected from the host: the HIP call completes, its
  status is reported as failed) poisons the prover state and the context,
  later calls report the first error, releases still succeed and
  context_destroy reports the first error (bridge_sumcheck.cpp:528-550,
  bridge_kzg.cpp:160-172, bridge_runtime.cpp:1208-1319)."""
import ctypes

import pytest

from hyperplonk_amd import InvalidParameters, VirtualPolynomial, abi
from hyperplonk_amd.backend import Mi355xBackend
from hyperplonk_amd.fields import R_MODULUS as R
from hyperplonk_amd.fields import fr_from_bytes
from tests import pyfield as pf

pytestmark = pytest.mark.gpu


def fr(v):
    return int(v).to_bytes(32, "little")


def frs(vals):
    return b"".join(fr(v) for v in vals)


def nonzero_table(seed, n):
    rng = pf.SplitMix64(seed)
    out = []
    while len(out) < n:
        v = rng.field()
        if v:
            out.append(v)
    return out


class RawCtx:
    """A context driven through the raw C ABI (no Python-side ownership checks)."""

    def __init__(self):
        self.lib = abi.load()
        raw = ctypes.c_void_p()
        st = self.lib.hp_mi355x_v1_context_create(0, ctypes.byref(raw))
        assert st.kind == 0
        self.raw = raw.value

    def upload(self, values):
        out = ctypes.c_void_p()
        nv = len(values).bit_length() - 1
        st = self.lib.hp_mi355x_v1_mle_upload(self.raw, nv, frs(values), len(values), ctypes.byref(out))
        return st, out.value

    def evaluate(self, handle, nv, point):
        out = ctypes.create_string_buffer(32)
        st = self.lib.hp_mi355x_v1_mle_evaluate(self.raw, handle, nv, frs(point) if nv else None, out)
        return st, int.from_bytes(out.raw, "little")

    def params(self, nv, t):
        raw = ctypes.c_void_p()
        st = self.lib.hp_mi355x_v1_multilinear_kzg_params_create(self.raw, nv, ctypes.byref(raw))
        assert st.kind == 0
        st = self.lib.hp_mi355x_v1_multilinear_kzg_params_generate(self.raw, raw.value, frs(t))
        assert st.kind == 0
        return raw.value

    def commit(self, params, handle, nv, ctx=None):
        out = ctypes.create_string_buffer(97)
        st = self.lib.hp_mi355x_v1_multilinear_kzg_commit(ctx or self.raw, params, handle, nv, out)
        return st, out.raw

    def state(self, nv, degree, handles, products):
        coeffs = frs([c for c, _ in products])
        offsets, idx = [0], []
        for _, ids in products:
            idx.extend(ids)
            offsets.append(len(idx))
        out = ctypes.c_void_p()
        st = self.lib.hp_mi355x_v1_sumcheck_prover_state_create(
            self.raw, nv, degree, abi.ptr_array(handles), len(handles), coeffs or None,
            abi.size_array(offsets), len(products), abi.size_array(idx), len(idx), ctypes.byref(out))
        return st, out.value

    def round(self, state, challenge, count, ctx=None):
        buf = ctypes.create_string_buffer(32 * count)
        ch = fr(challenge) if challenge is not None else None
        st = self.lib.hp_mi355x_v1_sumcheck_prover_state_round(ctx or self.raw, state, ch, buf, count)
        return st, fr_from_bytes(buf.raw)

    def destroy(self):
        st = self.lib.hp_mi355x_v1_context_destroy(self.raw)
        self.raw = None
        return st


def test_frac_zero_denominators_across_tiles_without_poisoning(backend):
    nv = 12
    n = 1 << nv
    nums = nonzero_table(0xF0, n)
    nums[0] = 1
    nums[n - 1] = 9
    dens = nonzero_table(0xF1, n)
    num_h = backend.mle_from_evaluations(nv, nums)
    den_h = backend.mle_from_evaluations(nv, dens)
    for pos in (0, n - 1):
        zeroed = list(dens)
        zeroed[pos] = 0
        zero_h = backend.mle_from_evaluations(nv, zeroed)
        with pytest.raises(InvalidParameters):  # x / 0
            backend.compute_frac_poly([num_h], [zero_h])
        zn = list(nums)
        zn[pos] = 0
        with pytest.raises(InvalidParameters):  # 0 / 0
            backend.compute_frac_poly([backend.mle_from_evaluations(nv, zn)], [zero_h])
        with pytest.raises(InvalidParameters):  # one factor of several carries the zero
            backend.compute_frac_poly([num_h], [den_h, zero_h])  # (counts differ: refused as well)
        with pytest.raises(InvalidParameters):  # the zero in the second of two factors of both lists
            backend.compute_frac_poly([num_h, num_h], [den_h, zero_h])
    # validation failures, not poisoning: the same context computes exactly
    expect = [a * pow(b, R - 2, R) % R for a, b in zip(nums, dens)]
    res = backend.compute_frac_poly([num_h], [den_h])
    assert backend.mle_values(res) == expect
    comb = backend.linear_combination([res], [1])
    assert backend.mle_values(comb) == expect


def test_independent_contexts_isolate_resources():
    a, b = RawCtx(), RawCtx()
    try:
        vals = [2, 3, 5, 7]
        _, fa = a.upload(vals)
        _, fb = b.upload(vals)
        assert a.evaluate(fa, 2, [2, 3])[1] == 19
        assert b.evaluate(fb, 2, [2, 3])[1] == 19
        assert a.evaluate(fb, 2, [2, 3])[0].kind == abi.INVALID_PARAMETERS
        assert b.evaluate(fa, 2, [2, 3])[0].kind == abi.INVALID_PARAMETERS
        out = ctypes.c_void_p()
        st = a.lib.hp_mi355x_v1_mle_linear_combination(a.raw, 2, abi.ptr_array([fa, fb]), frs([1, 1]), 2,
                                                       ctypes.byref(out))
        assert st.kind == abi.INVALID_PARAMETERS and out.value is None
        t = [11, 13]
        pa, pb = a.params(2, t), b.params(2, t)
        want = pf.aff_mul(pf.G, sum(v * e for v, e in zip(vals, [(1 - t[0]) * (1 - t[1]), t[0] * (1 - t[1]),
                                                                  (1 - t[0]) * t[1], t[0] * t[1]])) % R)
        st, pkt = a.commit(pa, fa, 2)
        assert st.kind == 0 and pf.unpacket(pkt) == want
        st, pkt = b.commit(pb, fb, 2)
        assert st.kind == 0 and pf.unpacket(pkt) == want
        assert a.commit(pa, fb, 2)[0].kind == abi.INVALID_PARAMETERS
        assert a.commit(pb, fa, 2)[0].kind == abi.INVALID_PARAMETERS
        assert b.commit(pa, fb, 2)[0].kind == abi.INVALID_PARAMETERS
        quot = ctypes.create_string_buffer(97 * 2)
        ev = ctypes.create_string_buffer(32)
        st = a.lib.hp_mi355x_v1_multilinear_kzg_open(a.raw, pb, fa, 2, frs([2, 3]), quot, ev)
        assert st.kind == abi.INVALID_PARAMETERS
        # both contexts stay usable
        assert a.evaluate(fa, 2, [2, 3])[1] == 19 and b.evaluate(fb, 2, [2, 3])[1] == 19
        # a prover state is bound to the context that created it
        _, g = a.upload([4, 9])
        st, s = a.state(1, 1, [g], [(1, [0])])
        assert st.kind == 0
        assert a.round(s, None, 2, ctx=b.raw)[0].kind == abi.INVALID_PARAMETERS
        assert a.lib.hp_mi355x_v1_sumcheck_prover_state_destroy(b.raw, s).kind == abi.INVALID_PARAMETERS
        assert a.lib.hp_mi355x_v1_multilinear_kzg_params_destroy(b.raw, pa).kind == abi.INVALID_PARAMETERS
        assert a.lib.hp_mi355x_v1_buffer_retire(b.raw, fa).kind == abi.INVALID_PARAMETERS
        st, msg = a.round(s, None, 2)
        assert st.kind == 0 and msg == [4, 9]
        for ctx, h in ((a, s),):
            assert ctx.lib.hp_mi355x_v1_sumcheck_prover_state_destroy(ctx.raw, h).kind == 0
        assert a.lib.hp_mi355x_v1_multilinear_kzg_params_destroy(a.raw, pa).kind == 0
        assert b.lib.hp_mi355x_v1_multilinear_kzg_params_destroy(b.raw, pb).kind == 0
    finally:
        assert a.destroy().kind == 0
        assert b.destroy().kind == 0


def test_resources_survive_ordered_drops_and_same_device_reuse(oracle):
    be = Mi355xBackend(0)
    t = [5, 6]
    pp = be.generate_prover_param(2, t)
    tables = [nonzero_table(0x3C + i, 4) for i in range(3)]
    handles = [be.mle_from_evaluations(2, tb) for tb in tables]
    prods = [(2, [0, 1]), (3, [2])]
    vp = VirtualPolynomial(2)
    for c, idx in prods:
        vp.add_mle_list([handles[i] for i in idx], c)
    ost = oracle.SumCheckState(2, 2, tables, prods)
    st = be.prover_init(vp)
    del pp  # the prepared SRS retires while the state still owns its inputs
    msgs = [be.prove_round_and_update_state(st, None)]
    del vp, handles  # so may the inputs and the virtual polynomial
    msgs.append(be.prove_round_and_update_state(st, 5))
    assert msgs == [ost.round(None), ost.round(5)]
    del st
    alive = be.mle_from_evaluations(2, [2, 3, 5, 7])
    assert be.evaluate_mle(alive, [2, 3]) == 19
    extra = be.mle_from_evaluations(1, [1, 2])
    del be  # handles keep the context until they are retired
    assert alive._ctx.raw
    del alive, extra
    nxt = Mi355xBackend(0)
    h = nxt.mle_from_evaluations(2, [2, 3, 5, 7])
    assert nxt.evaluate_mle(h, [2, 3]) == 19


def test_stale_and_wrong_kind_handles_are_rejected():
    c = RawCtx()
    try:
        _, f = c.upload([2, 3, 5, 7])
        st, s = c.state(2, 1, [f], [(1, [0])])
        assert st.kind == 0
        assert c.lib.hp_mi355x_v1_sumcheck_prover_state_destroy(c.raw, s).kind == 0
        assert c.lib.hp_mi355x_v1_sumcheck_prover_state_destroy(c.raw, s).kind == abi.INVALID_PARAMETERS
        assert c.round(s, None, 2)[0].kind == abi.INVALID_PARAMETERS
        p = c.params(2, [3, 4])
        assert c.lib.hp_mi355x_v1_multilinear_kzg_params_destroy(c.raw, p).kind == 0
        assert c.lib.hp_mi355x_v1_multilinear_kzg_params_destroy(c.raw, p).kind == abi.INVALID_PARAMETERS
        assert c.commit(p, f, 2)[0].kind == abi.INVALID_PARAMETERS
        # live handles of the other kind are refused too
        p2 = c.params(2, [3, 4])
        st, s2 = c.state(2, 1, [f], [(1, [0])])
        assert c.commit(s2, f, 2)[0].kind == abi.INVALID_PARAMETERS
        assert c.round(p2, None, 2)[0].kind == abi.INVALID_PARAMETERS
        assert c.lib.hp_mi355x_v1_multilinear_kzg_params_destroy(c.raw, s2).kind == abi.INVALID_PARAMETERS
        assert c.lib.hp_mi355x_v1_sumcheck_prover_state_destroy(c.raw, p2).kind == abi.INVALID_PARAMETERS
        assert c.lib.hp_mi355x_v1_buffer_retire(c.raw, f).kind == 0
        assert c.lib.hp_mi355x_v1_buffer_retire(c.raw, f).kind == abi.INVALID_PARAMETERS
        assert c.evaluate(f, 2, [1, 1])[0].kind == abi.INVALID_PARAMETERS
        # the state still owns the retired table
        st, msg = c.round(s2, None, 2)
        assert st.kind == 0 and msg == [2 + 5, 3 + 7]
        # sizes beyond 2^31 entries are refused before any work
        out = ctypes.c_void_p()
        st = c.lib.hp_mi355x_v1_mle_upload(c.raw, 32, b"\0" * 32, 1 << 32, ctypes.byref(out))
        assert st.kind == abi.INVALID_PARAMETERS and out.value is None
        _, g = c.upload([1, 2])
        st, _ = c.state(32, 1, [g], [(1, [0])])
        assert st.kind == abi.INVALID_PARAMETERS
        assert c.evaluate(g, 32, [1] * 32)[0].kind == abi.INVALID_PARAMETERS
        # the context is intact after all of these
        assert c.lib.hp_mi355x_v1_sumcheck_prover_state_destroy(c.raw, s2).kind == 0
        assert c.lib.hp_mi355x_v1_multilinear_kzg_params_destroy(c.raw, p2).kind == 0
        _, h = c.upload([2, 3, 5, 7])
        assert c.evaluate(h, 2, [2, 3])[1] == 19
    finally:
        assert c.destroy().kind == 0


@pytest.mark.parametrize("after", [1, 2])  # the round's launch, then its synchronisation
def test_device_failure_poisons_state_and_context(after):
    c = RawCtx()
    _, f = c.upload(list(range(3, 3 + 16)))
    _, g = c.upload(list(range(40, 40 + 16)))
    st, s = c.state(4, 2, [f, g], [(3, [0, 1]), (5, [1])])
    assert st.kind == 0
    st, first = c.round(s, None, 3)
    assert st.kind == 0
    assert c.lib.hp_mi355x_v1_debug_inject_failure(c.raw, 0, 0).kind == 0  # disarm is a no-op
    assert c.lib.hp_mi355x_v1_debug_inject_failure(c.raw, 1, 0).kind == abi.INVALID_PARAMETERS
    # validation failures check no HIP call and do not consume the injection
    assert c.lib.hp_mi355x_v1_debug_inject_failure(c.raw, after, 777).kind == 0
    assert c.round(s, None, 3)[0].kind == abi.INVALID_STATE  # ordering error, state untouched
    st, _ = c.round(s, 9, 3)
    assert (st.kind, st.native_code) == (abi.INVALID_STATE, 777)
    # the state and the context report the first error from now on
    st, _ = c.round(s, 9, 3)
    assert (st.kind, st.native_code) == (abi.INVALID_STATE, 777)
    st, _ = c.upload([1, 2])
    assert (st.kind, st.native_code) == (abi.INVALID_STATE, 777)
    assert c.evaluate(f, 4, [1, 2, 3, 4])[0].native_code == 777
    # releases still succeed; the destroy reports the first error
    assert c.lib.hp_mi355x_v1_sumcheck_prover_state_destroy(c.raw, s).kind == 0
    assert c.lib.hp_mi355x_v1_buffer_retire(c.raw, f).kind == 0
    assert c.lib.hp_mi355x_v1_buffer_retire(c.raw, g).kind == 0
    st = c.destroy()
    assert (st.kind, st.native_code) == (abi.INVALID_STATE, 777)
    # a new context on the same device is unaffected
    d = RawCtx()
    _, h = d.upload([2, 3, 5, 7])
    assert d.evaluate(h, 2, [2, 3])[1] == 19
    assert d.destroy().kind == 0


def test_commit_failure_poisons_context_and_params_still_release():
    c = RawCtx()
    p = c.params(3, [3, 4, 5])
    _, f = c.upload(list(range(1, 9)))
    st, good = c.commit(p, f, 3)
    assert st.kind == 0
    assert c.lib.hp_mi355x_v1_debug_inject_failure(c.raw, 4, 31337).kind == 0
    st, _ = c.commit(p, f, 3)
    assert (st.kind, st.native_code) == (abi.INVALID_STATE, 31337)
    st, _ = c.commit(p, f, 3)
    assert (st.kind, st.native_code) == (abi.INVALID_STATE, 31337)
    assert c.lib.hp_mi355x_v1_multilinear_kzg_params_destroy(c.raw, p).kind == 0
    assert c.lib.hp_mi355x_v1_buffer_retire(c.raw, f).kind == 0
    st = c.destroy()
    assert (st.kind, st.native_code) == (abi.INVALID_STATE, 31337)
