"""hp_mi355x_v1_mle_evaluate_noalloc (evaluate_mle over a static per-GPU
scratch array, no allocation) against hp_mi355x_v1_mle_evaluate through the C
ABI: byte-identical values on calls in a row with shrinking tables (the
second call reuses what the first left in the shared array), the size limit
of the scratch refused with its own native code, and the argument checks of
the allocating entry point."""
import ctypes

import pytest

from hyperplonk_amd import abi
from hyperplonk_amd.fields import fr_from_bytes

pytestmark = pytest.mark.gpu


def evaluate(backend, fn, handle, nv, point_bytes):
    out = ctypes.create_string_buffer(b"\xa5" * 32, 32)
    st = fn(backend.context, handle, nv, point_bytes, out)
    return st, out.raw


def test_noalloc_evaluate_matches_allocating_entry_point(backend, rng_factory):
    lib = backend.lib
    rng = rng_factory(0xA110C)
    limit = 20  # HP_MI355X_NOALLOC_MAX_NUM_VARS
    for nv in (limit, 13, 11, 10, 3, 0):
        tab = rng.fr_bytes(1 << nv)
        m = backend.mle_from_evaluations(nv, tab)
        point = rng.fr_bytes(nv).tobytes() if nv else None
        st_a, want = evaluate(backend, lib.hp_mi355x_v1_mle_evaluate, m.address, nv, point)
        assert st_a.kind == 0, nv
        for _ in range(2):
            st_b, got = evaluate(backend, lib.hp_mi355x_v1_mle_evaluate_noalloc, m.address, nv, point)
            assert (st_b.kind, st_b.native_code) == (0, 0), nv
            assert got == want, nv
    # the known answer: [2, 3, 5, 7] at (2, 3) is 19
    m = backend.mle_from_evaluations(2, [2, 3, 5, 7])
    pt = (2).to_bytes(32, "little") + (3).to_bytes(32, "little")
    st, got = evaluate(backend, lib.hp_mi355x_v1_mle_evaluate_noalloc, m.address, 2, pt)
    assert st.kind == 0 and fr_from_bytes(got) == [19]


def test_noalloc_evaluate_limits_and_checks(backend, rng_factory):
    lib = backend.lib
    rng = rng_factory(0xA110D)
    nv = 21  # one above the scratch limit
    m = backend.mle_from_evaluations(nv, rng.fr_bytes(1 << nv))
    point = rng.fr_bytes(nv).tobytes()
    st, out = evaluate(backend, lib.hp_mi355x_v1_mle_evaluate_noalloc, m.address, nv, point)
    assert (st.kind, st.native_code) == (abi.INVALID_PARAMETERS, 1001)
    assert out == b"\xa5" * 32
    # the allocating entry point still evaluates it
    st, _ = evaluate(backend, lib.hp_mi355x_v1_mle_evaluate, m.address, nv, point)
    assert st.kind == 0
    small = backend.mle_from_evaluations(2, [2, 3, 5, 7])
    pt = (2).to_bytes(32, "little") + (3).to_bytes(32, "little")
    fn = lib.hp_mi355x_v1_mle_evaluate_noalloc
    assert fn(backend.context, small.address, 2, None, ctypes.create_string_buffer(32)).kind == abi.INVALID_PARAMETERS
    assert fn(backend.context, small.address, 2, pt, None).kind == abi.INVALID_PARAMETERS
    assert fn(backend.context, small.address, 3, pt + pt[:32], ctypes.create_string_buffer(32)).kind \
        == abi.INVALID_PARAMETERS
    bad = b"\xff" * 32 + pt[32:]
    assert fn(backend.context, small.address, 2, bad, ctypes.create_string_buffer(32)).kind == abi.INVALID_PARAMETERS
    assert fn(None, small.address, 2, pt, ctypes.create_string_buffer(32)).kind == abi.INVALID_PARAMETERS
    # the context is unaffected by the refusals
    st, got = evaluate(backend, fn, small.address, 2, pt)
    assert st.kind == 0 and fr_from_bytes(got) == [19]
