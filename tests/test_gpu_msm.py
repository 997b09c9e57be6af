"""G1 MSM / multilinear KZG on the GPU.

Pinned three ways: the Python-integer closed-form fixtures
(tests/golden/golden.json: expected (sum a_i s_i) G over bases a_i G, including
zero / ones / identity / cancel / alternating / boundary / exception patterns),
the device SRS against the oracle's SRS layers, and commitments / openings
against the oracle's Pippenger and KZG open on the same bytes."""
import pytest

from hyperplonk_amd import InvalidParameters
from hyperplonk_amd.fields import R_MODULUS as R
from hyperplonk_amd.fields import fr_from_bytes
from tests import pyfield as pf

pytestmark = pytest.mark.gpu


def H(v):
    return int(v, 16)


@pytest.mark.parametrize("window", [None, 4, 7, 15, 17])
def test_golden_msm_cases(backend, golden, window, monkeypatch):
    if window:
        monkeypatch.setenv("HP_MI355X_MSM_WINDOW", str(window))
    small = [pf.point_from_hex(p) for p in golden["g1_small_multiples"]]

    def base(a):
        return small[a] if a <= 300 else pf.aff_neg(small[R - a])

    for case in golden["msm"]:
        coeffs = [H(c) for c in case["coeffs"]]
        scalars = [H(s) for s in case["scalars"]]
        n = len(coeffs)
        N = max(1, (max(n, 1) - 1).bit_length())
        size = 1 << N
        pts = [base(a) for a in coeffs] + [None] * (size - n)
        layers = [b"".join(pf.packet(p) for p in pts)]
        for i in range(1, N + 1):
            layers.append(b"".join(pf.packet(small[1]) for _ in range(1 << (N - i))))
        pp = backend.prepare_prover_param(N, layers)
        m = backend.mle_from_evaluations(N, scalars + [0] * (size - n))
        got = backend.commit(pp, m)
        assert got == pf.point_from_hex(case["expected"]), (case["name"], window)


@pytest.mark.parametrize("window", [None, 5, 15, 16, 17])
def test_msm_scalars_around_half_modulus(backend, window, monkeypatch):
    """Scalars above (r - 1) / 2 are folded to r - s against the negated point:
    the values on both sides of the fold, the extremes and 2^254 +- 1, against
    the closed form (sum a_i s_i) G over bases a_i G."""
    if window:
        monkeypatch.setenv("HP_MI355X_MSM_WINDOW", str(window))
    half = (R - 1) // 2
    scalars = [half - 1, half, half + 1, half + 2, R - 1, R - 2, 1, 2,
               (1 << 254) - 1, 1 << 254, (1 << 254) + 1, (1 << 253) + 7,
               half, half + 1, 3, R - 3]
    coeffs = [5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 5, 5, 7, 7]
    N = 4
    pts = [pf.aff_mul(pf.G, a) for a in coeffs]
    layers = [b"".join(pf.packet(p) for p in pts)]
    for i in range(1, N + 1):
        layers.append(b"".join(pf.packet(pf.G) for _ in range(1 << (N - i))))
    pp = backend.prepare_prover_param(N, layers)
    m = backend.mle_from_evaluations(N, scalars)
    want = pf.aff_mul(pf.G, sum(a * s for a, s in zip(coeffs, scalars)) % R)
    assert backend.commit(pp, m) == want
    # a pair that cancels across the fold: s P + (r - s) P = identity
    m2 = backend.mle_from_evaluations(N, [half, R - half] + [0] * 14)
    pts2 = [pf.G, pf.G] + pts[2:]
    pp2 = backend.prepare_prover_param(N, [b"".join(pf.packet(p) for p in pts2)] + layers[1:])
    assert backend.commit(pp2, m2) is None


def test_srs_generation_matches_oracle(backend, oracle, rng_factory):
    rng = rng_factory(42)
    N = 6
    t = fr_from_bytes(rng.fr_bytes(N).tobytes())
    pp = backend.generate_prover_param(N, t)
    for i in range(N + 1):
        assert backend.download_layer(pp, i) == oracle.srs_layer_packets(t[i:], threads=4), i


@pytest.fixture(scope="module")
def srs20(backend, rng_factory):
    """A device-generated SRS with N = 20 (layer k has 2^(20 - k) points)."""
    rng = rng_factory(2020)
    N = 20
    t = fr_from_bytes(rng.fr_bytes(N).tobytes())
    return N, t, backend.generate_prover_param(N, t)


@pytest.mark.parametrize("nv", [0, 1, 2, 5, 10, 12, 16])
def test_commit_matches_oracle(backend, oracle, rng_factory, srs20, nv):
    N, t, pp = srs20
    rng = rng_factory(600 + nv)
    tab = rng.fr_bytes(1 << nv)
    m = backend.mle_from_evaluations(nv, tab)
    layer = backend.download_layer(pp, N - nv)
    want = oracle.msm_pippenger_packets(layer, tab.tobytes(), 1 << nv, 8)
    assert backend.commit_packet(pp, m) == want


def test_commit_2_20_closed_form(backend, rng_factory, srs20):
    """At 2^20 the commitment of f against layer 0 equals g * f(t) (the SRS
    trapdoor makes layer 0 the eq(t, .) multiples of g): a size-independent check."""
    N, t, pp = srs20
    rng = rng_factory(620)
    tab = rng.fr_bytes(1 << N)
    m = backend.mle_from_evaluations(N, tab)
    value = backend.evaluate_mle(m, t)
    assert backend.commit(pp, m) == pf.aff_mul(pf.G, value)


def test_constant_and_special_commits(backend, srs20):
    N, t, pp = srs20
    # constant polynomial commits to g * 29 (pcs/multilinear_kzg/mod.rs:342-363 pattern)
    c = backend.mle_from_evaluations(0, [29])
    assert backend.commit(pp, c) == pf.aff_mul(pf.G, 29)
    z = backend.mle_from_evaluations(8, [0] * 256)
    assert backend.commit(pp, z) is None
    ones = backend.mle_from_evaluations(4, [R - 1] * 16)
    value = backend.evaluate_mle(ones, t[N - 4 :])
    assert backend.commit(pp, ones) == pf.aff_mul(pf.G, value)
    with pytest.raises(InvalidParameters):
        backend.commit(pp, backend.mle_from_evaluations(N + 1, bytes(32 << (N + 1))))


def test_multi_commit_uneven_matches_single(backend, rng_factory, srs20):
    N, t, pp = srs20
    rng = rng_factory(700)
    nvs = [3, 10, 0, 7, 12, 3, 1, 9, 11, 5, 2, 8, 6]
    mles = [backend.mle_from_evaluations(nv, rng.fr_bytes(1 << nv)) for nv in nvs]
    mles.append(mles[1])  # duplicates are kept
    batch = backend.multi_commit_packets(pp, mles)
    assert batch == [backend.commit_packet(pp, m) for m in mles]
    assert backend.multi_commit(pp, []) == []


@pytest.mark.parametrize("nv", [0, 1, 4, 9, 10, 11, 14])
def test_open_matches_oracle(backend, oracle, rng_factory, srs20, nv):
    N, t, pp = srs20
    rng = rng_factory(800 + nv)
    tab = rng.fr_bytes(1 << nv)
    m = backend.mle_from_evaluations(nv, tab)
    point = fr_from_bytes(rng.fr_bytes(nv).tobytes()) if nv else []
    proofs, value = backend.open_packets(pp, m, point)
    ignored = N - nv + 1
    layers = b"".join(backend.download_layer(pp, ignored + j) for j in range(nv))
    want_proofs, want_value = oracle.kzg_open_packets(layers, fr_from_bytes(tab.tobytes()), point, threads=8)
    assert value == want_value
    assert proofs == want_proofs
    with pytest.raises(InvalidParameters):
        backend.open(pp, m, point + [1])


def test_open_known_values(backend, srs20):
    # pcs/multilinear_kzg/mod.rs:391-426: [2,3,5,7] opened at (0,0), (1,1), (2,3)
    N, t, pp = srs20
    f = backend.mle_from_evaluations(2, [2, 3, 5, 7])
    for point, value in (([0, 0], 2), ([1, 1], 7), ([2, 3], 19)):
        assert backend.open(pp, f, point)[1] == value


def test_upload_layer_rejections(backend):
    g = pf.G
    good = pf.packet(g)
    for bad in (
        b"\x02" + bytes(96),  # bad tag
        b"\x01" + b"\x01" + bytes(95),  # identity with coordinates
        b"\x00" + bytes(96),  # finite-tag (0, 0)
        pf.packet((1, 1)),  # off-curve
        b"\x00" + (pf.Q).to_bytes(48, "little") + g[1].to_bytes(48, "little"),  # non-canonical x
    ):
        with pytest.raises(InvalidParameters):
            backend.prepare_prover_param(1, [good + bad, good])
    with pytest.raises(InvalidParameters):
        backend.prepare_prover_param(1, [good, good])  # wrong layer size
    # an on-curve point outside the subgroup is accepted (no subgroup re-check)
    pp = backend.prepare_prover_param(1, [pf.packet((0, 2)) + good, good])
    m = backend.mle_from_evaluations(1, [0, 3])
    assert backend.commit(pp, m) == pf.aff_mul(g, 3)


def test_commit_before_layer_upload_is_invalid_state(backend):
    import ctypes

    raw = ctypes.c_void_p()
    st = backend.lib.hp_mi355x_v1_multilinear_kzg_params_create(backend.context, 2, ctypes.byref(raw))
    assert st.kind == 0
    m = backend.mle_from_evaluations(2, [1, 2, 3, 4])
    out = ctypes.create_string_buffer(97)
    st = backend.lib.hp_mi355x_v1_multilinear_kzg_commit(backend.context, raw.value, m.address, 2, out)
    assert st.kind == 2
    assert backend.lib.hp_mi355x_v1_multilinear_kzg_params_destroy(backend.context, raw.value).kind == 0


def test_heavy_buckets(backend, oracle, rng_factory, srs20):
    """Tables of repeated small values put most digits into a few buckets
    (selector columns full of 0/1 do this).  sum_x eq(t, x) = 1, so committing
    the constant table k gives k * g against any layer."""
    N, t, pp = srs20
    for nv in (16, 20):
        ones = backend.mle_from_evaluations(nv, [1] * (1 << nv))
        assert backend.commit(pp, ones) == pf.G
        minus = backend.mle_from_evaluations(nv, [R - 1] * (1 << nv))
        assert backend.commit(pp, minus) == pf.aff_neg(pf.G)
    nv = 16
    rng = rng_factory(900)
    tab = rng.fr_bytes(1 << nv)
    tab[::2] = 0
    tab[::2, 0] = 1  # every other entry is 1
    m = backend.mle_from_evaluations(nv, tab)
    layer = backend.download_layer(pp, N - nv)
    assert backend.commit_packet(pp, m) == oracle.msm_pippenger_packets(layer, tab.tobytes(), 1 << nv, 8)


def test_kzg_batch_matches_single_calls(backend, rng_factory, srs20):
    """Commits and opens of mixed sizes in one batch give exactly the packets
    and values of the one-at-a-time calls (which the tests above pin to the
    oracle); an invalid request anywhere leaves every output untouched."""
    N, t, pp = srs20
    rng = rng_factory(901)
    nvs = [12, 0, 5, 12, 9, 1, 14]
    mles = [backend.mle_from_evaluations(nv, rng.fr_bytes(1 << nv)) for nv in nvs]
    ops = []
    for i, (nv, m) in enumerate(zip(nvs, mles)):
        if i % 2:
            ops.append(("open", m, fr_from_bytes(rng.fr_bytes(nv).tobytes()) if nv else []))
        else:
            ops.append(("commit", m))
    ops.append(("commit", mles[0]))  # a repeat
    got = backend.kzg_batch(pp, ops)
    for op, g in zip(ops, got):
        if op[0] == "commit":
            assert g[0] == backend.commit_packet(pp, op[1])
        else:
            assert (g[0], g[1]) == backend.open_packets(pp, op[1], op[2])
    with pytest.raises(InvalidParameters):
        backend.kzg_batch(pp, [("commit", mles[0]), ("open", mles[2], [1, 2])])
    assert backend.kzg_batch(pp, []) == []


def test_autotune_changes_plans_not_results(backend, oracle, rng_factory):
    """The autotuner only picks window bits and task caps: commits through
    parameters prepared after tuning still equal the oracle's MSM."""
    table = backend.msm_autotune([6, 9])
    assert [t["log_n"] for t in table] == [6, 9]
    for t in table:
        assert 4 <= t["window"] <= 20 and t["cap"] in (2, 4, 8, 16, 32, 64) and t["ms"] > 0
    try:
        rng = rng_factory(930)
        N = 10
        t = fr_from_bytes(rng.fr_bytes(N).tobytes())
        pp = backend.generate_prover_param(N, t)
        for nv in (6, 9):
            tab = rng.fr_bytes(1 << nv)
            m = backend.mle_from_evaluations(nv, tab)
            layer = backend.download_layer(pp, N - nv)
            want = oracle.msm_pippenger_packets(layer, tab.tobytes(), 1 << nv, 8)
            assert backend.commit_packet(pp, m) == want
        # an explicit table: windows outside 4..20 are rejected
        with pytest.raises(InvalidParameters):
            backend.msm_set_tuning([{"log_n": 6, "window": 3, "cap": 4}])
    finally:
        backend.msm_set_tuning([{"log_n": 6, "window": 0, "cap": 0}, {"log_n": 9, "window": 0, "cap": 0}])


def test_kzg_batch_submit_overlaps_and_matches(backend, oracle, rng_factory, srs20):
    """An asynchronous KZG batch (side stream) gives the packets and values of
    the synchronous batch and of the oracle while sumcheck rounds, table
    uploads and a retire of one of its tables run on the context meanwhile;
    invalid requests are refused before any work and a ticket is waited for
    once."""
    import ctypes

    from hyperplonk_amd import VirtualPolynomial, abi

    N, t, pp = srs20
    rng = rng_factory(0xA5C)
    nvs = [16, 3, 12, 0, 9]
    tabs = [rng.fr_bytes(1 << nv) for nv in nvs]
    mles = [backend.mle_from_evaluations(nv, tb) for nv, tb in zip(nvs, tabs)]
    ops = [("commit", mles[0]), ("open", mles[1], fr_from_bytes(rng.fr_bytes(3).tobytes())),
           ("commit", mles[2]), ("open", mles[3], []), ("commit", mles[4])]
    want = backend.kzg_batch(pp, ops)
    ticket = backend.kzg_batch_submit(pp, ops)
    # other work on the context while the batch runs
    nv = 10
    st_tabs = [rng.fr_bytes(1 << nv) for _ in range(3)]
    st_mles = [backend.mle_from_evaluations(nv, tb) for tb in st_tabs]
    vp = VirtualPolynomial(nv)
    vp.add_mle_list([st_mles[0], st_mles[1]], 3)
    vp.add_mle_list([st_mles[2]], 5)
    st = backend.prover_init(vp)
    ost = oracle.SumCheckState(nv, 2, st_tabs, [(3, [0, 1]), (5, [2])])
    ch = None
    srng = pf.SplitMix64(0xA5D)
    for _ in range(nv):
        assert backend.prove_round_and_update_state(st, ch) == ost.round(ch)
        ch = srng.nonboolean()
    # a table of the pending batch retired: the ticket keeps its memory alive
    assert backend.lib.hp_mi355x_v1_buffer_retire(backend.context, mles[2].address).kind == 0
    mles[2].address = 0
    got = ticket.wait()
    assert got == want
    assert got[0][0] == oracle.msm_pippenger_packets(backend.download_layer(pp, N - 16), tabs[0].tobytes(),
                                                     1 << 16, 16)
    with pytest.raises(InvalidParameters):
        ticket.wait()
    # a bad request: nothing is submitted
    out = ctypes.c_void_p()
    req = (backend._KzgRequestType * 1)()
    req[0].kind = 7
    st_ = backend.lib.hp_mi355x_v1_multilinear_kzg_batch_submit(backend.context, pp.raw, req, 1, ctypes.byref(out))
    assert st_.kind == abi.INVALID_PARAMETERS and out.value is None
