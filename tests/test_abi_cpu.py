"""Host-side checks that need no GPU: the C ABI library loads and exports every
declared entry point, and the host mirror of the reference's VirtualPolynomial
keeps its semantics (backend/src/virtual_polynomial.rs:273-343)."""
import ctypes

import pytest

from hyperplonk_amd import InvalidParameters, VirtualPolynomial, abi
from hyperplonk_amd.fields import R_MODULUS
from hyperplonk_amd.virtual_polynomial import encode_products


def test_library_exports_every_header_symbol():
    lib = abi.load()
    declared = abi.header_symbols()
    assert len(declared) == 39
    for name in declared:
        assert hasattr(lib, name), name
    # the 20 entry points of the reference bridge (bridge.h:43-62) are all present
    ref_names = [
        "context_create", "context_destroy", "buffer_retire", "mle_upload", "mle_linear_combination",
        "mle_evaluate", "mle_build_eq_x_r", "mle_compute_frac_poly", "mle_compute_product_poly",
        "mle_compute_nums_and_denoms", "mle_compute_product_factors", "multilinear_kzg_params_create",
        "multilinear_kzg_params_upload_layer", "multilinear_kzg_params_destroy", "multilinear_kzg_commit",
        "multilinear_kzg_multi_commit", "multilinear_kzg_open", "sumcheck_prover_state_create",
        "sumcheck_prover_state_round", "sumcheck_prover_state_destroy",
    ]
    for n in ref_names:
        assert "hp_mi355x_v1_" + n in declared
    assert set(declared) == set(abi.SIGNATURES)


def test_library_was_built_from_the_sources_in_the_tree():
    """The library reports the SHA-256 of the sources it was compiled from;
    it must equal the digest of the files in this tree (a stale or foreign
    build fails here)."""
    info = abi.build_info()
    assert info == f"sources-sha256={abi.source_digest()} arch=gfx950", info


def test_null_and_unknown_context_are_rejected_without_a_device():
    lib = abi.load()
    st = lib.hp_mi355x_v1_context_destroy(None)
    assert st.kind == abi.INVALID_PARAMETERS
    bogus = ctypes.c_void_p(0x1234)
    st = lib.hp_mi355x_v1_buffer_retire(bogus, None)
    assert st.kind == abi.INVALID_PARAMETERS
    out = ctypes.c_void_p(5)
    st = lib.hp_mi355x_v1_mle_upload(bogus, 1, b"\0" * 64, 2, ctypes.byref(out))
    assert st.kind == abi.INVALID_PARAMETERS and out.value is None
    st = lib.hp_mi355x_v1_context_create(0, None)
    assert st.kind == abi.INVALID_PARAMETERS


def test_static_archive_links_into_a_c_program():
    """libhp_mi355x.a links into a plain C program with the HIP runtime and the
    C++ runtime only (the link the reference's `kind = "static"` binding makes,
    backend/src/ascend910b3/ffi.rs:34); the program's calls are argument checks
    that run before any device use."""
    import os
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    prog = os.path.join(root, "_build", "static_link_test")
    if not os.path.exists(prog):
        subprocess.run(["make", "-C", os.path.join(root, "tests", "cpp"), prog], check=True, capture_output=True)
    libs = subprocess.run(["ldd", prog], capture_output=True, text=True).stdout
    assert "libhp_mi355x" not in libs
    res = subprocess.run([prog], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0 and "STATIC_LINK PASS" in res.stdout, (res.returncode, res.stdout, res.stderr)


class FakeMle:
    """Metadata-only handle (the reference's test fake, virtual_polynomial.rs:243-271)."""

    _next = 1

    def __init__(self, num_vars):
        self.num_vars = num_vars
        self.id = FakeMle._next
        FakeMle._next += 1


def test_virtual_polynomial_dedups_by_identity_and_keeps_repeats():
    a, b, c = FakeMle(3), FakeMle(3), FakeMle(3)
    vp = VirtualPolynomial.new_from_mle(a, 2)
    vp.add_mle_list([a, b, a], 3)
    vp.add_mle_list([c], 5)
    assert vp.aux_info.max_degree == 3
    assert [m.id for m in vp.flattened_ml_extensions()] == [a.id, b.id, c.id]
    assert vp.products == [(2, [0]), (3, [0, 1, 0]), (5, [2])]
    vp.mul_by_mle(b, 7)
    assert vp.aux_info.max_degree == 4
    assert vp.products == [(14, [0, 1]), (21, [0, 1, 0, 1]), (35, [2, 1])]
    coeffs, offsets, indices = encode_products(vp)
    assert offsets == [0, 2, 6, 8] and indices == [0, 1, 0, 1, 0, 1, 2, 1]
    assert coeffs[:32] == (14).to_bytes(32, "little")
    with pytest.raises(InvalidParameters):
        vp.add_mle_list([], 1)
    with pytest.raises(InvalidParameters):
        vp.add_mle_list([FakeMle(2)], 1)
    with pytest.raises(InvalidParameters):
        vp.mul_by_mle(FakeMle(4), 1)


def test_equal_contents_distinct_handles_are_distinct_factors():
    a, b = FakeMle(2), FakeMle(2)
    vp = VirtualPolynomial(2)
    vp.add_mle_list([a, b], 1)
    assert len(vp.flattened_ml_extensions()) == 2


def test_product_encoding_rejects_bad_metadata():
    # backend/src/ascend910b3/sum_check.rs:262-280 style checks
    with pytest.raises(InvalidParameters):
        encode_products(VirtualPolynomial(0))
    a = FakeMle(2)
    vp = VirtualPolynomial.new_from_mle(a, 1)
    vp.products.append((1, []))
    with pytest.raises(InvalidParameters):
        encode_products(vp)
    vp = VirtualPolynomial.new_from_mle(a, 1)
    vp.products.append((1, [3]))
    with pytest.raises(InvalidParameters):
        encode_products(vp)
    vp = VirtualPolynomial.new_from_mle(a, 1)
    vp.products.append((1, [0, 0, 0]))  # arity above max_degree
    with pytest.raises(InvalidParameters):
        encode_products(vp)
    # an empty whole sum with positive dimensions is valid
    vp = VirtualPolynomial(2)
    vp.aux_info.max_degree = 2
    assert encode_products(vp) == (b"", [0], [])


def test_vp_add_merges_products():
    a, b = FakeMle(1), FakeMle(1)
    p = VirtualPolynomial.new_from_mle(a, 3)
    q = VirtualPolynomial.new_from_mle(b, 4)
    s = p + q
    assert s.products == [(3, [0]), (4, [1])]
    assert s.aux_info.max_degree == 1
    assert (R_MODULUS - 1) % R_MODULUS == R_MODULUS - 1


def test_synthetic_stream_matches_sequential_sampler():
    from hyperplonk_amd.synth import SplitMix64Stream, limbs_to_ints
    from tests.pyfield import SplitMix64

    s = SplitMix64Stream(99)
    got = limbs_to_ints(s.fr_limbs(33)) + limbs_to_ints(s.fr_limbs(5))
    ref = SplitMix64(99)
    assert got == [ref.field() for _ in range(38)]
    assert all(v < R_MODULUS for v in got)
