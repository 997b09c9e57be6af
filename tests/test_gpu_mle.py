"""MLE operations on the GPU: the reference CPU backend's own known-answer
tests (backend/src/cpu/multilinear_polynomial.rs:255-593) run against the
device backend, then seeded random tables checked byte-exactly against the
oracle."""
import pytest

from hyperplonk_amd import InvalidParameters
from hyperplonk_amd.fields import R_MODULUS as R
from hyperplonk_amd.fields import fr_from_bytes

pytestmark = pytest.mark.gpu


def mle(be, nv, values):
    return be.mle_from_evaluations(nv, [v % R for v in values])


def test_upload_roundtrip_edges(backend):
    vals = [0, 1, R - 1, R - 2, 2**64 - 1, 2**128, 2**254, 7, 11, 13, 2**32, R // 2, 3, 5, 8, 9]
    m = mle(backend, 4, vals)
    assert backend.mle_values(m) == vals
    with pytest.raises(InvalidParameters):
        backend.mle_from_evaluations(2, [0, 1, 2])  # wrong length
    with pytest.raises(InvalidParameters):
        backend.mle_from_evaluations(1, (R).to_bytes(32, "little") + bytes(32))  # non-canonical


def test_owned_evaluations_preserve_layout_and_validate_dimensions(backend):
    # cpu/multilinear_polynomial.rs:278-312
    poly = mle(backend, 2, [2, 3, 5, 7])
    assert backend.evaluate_mle(poly, [2, 3]) == 19
    same = mle(backend, 2, [2, 3, 5, 7])
    assert poly.id != same.id
    const = mle(backend, 0, [11])
    assert backend.evaluate_mle(const, []) == 11
    for n in (0, 2, 4):
        with pytest.raises(InvalidParameters):
            backend.evaluate_mle(mle(backend, 3, [2, 3, 5, 7, 11, 13, 17, 19]), [0] * n)
    with pytest.raises(InvalidParameters):
        backend.build_eq_x_r([])


def test_linear_combination_known_answers(backend):
    # :314-373
    p = mle(backend, 2, [2, 3, 5, 7])
    q = mle(backend, 2, [11, 13, 17, 19])
    comb = backend.linear_combination([p, q, p], [2, 3, R - 2])
    assert backend.evaluate_mle(comb, [2, 3]) == 99
    assert backend.evaluate_mle(p, [2, 3]) == 19
    assert backend.evaluate_mle(q, [2, 3]) == 33
    for coeffs in ([1, R - 1], [0, 0]):
        zero = backend.linear_combination([p, p], coeffs)
        assert zero.num_vars == 2
        assert backend.evaluate_mle(zero, [2, 3]) == 0
    c = backend.linear_combination([mle(backend, 0, [11]), mle(backend, 0, [3])], [2, R - 3])
    assert backend.evaluate_mle(c, []) == 13
    with pytest.raises(InvalidParameters):
        backend.linear_combination([], [])
    with pytest.raises(InvalidParameters):
        backend.linear_combination([mle(backend, 1, [2, 3])], [])
    with pytest.raises(InvalidParameters):
        backend.linear_combination([mle(backend, 1, [2, 3]), mle(backend, 2, [2, 3, 5, 7])], [1, 1])


def test_fraction_known_answers(backend):
    # :376-420
    fxs = [mle(backend, 2, [0, 2, 3, 4]), mle(backend, 2, [2, 3, 5, 7])]
    gxs = [mle(backend, 2, [1, 2, 3, 4]), mle(backend, 2, [7, 5, 6, 7])]
    frac = backend.compute_frac_poly(fxs, gxs)
    expect = [n * pow(d, R - 2, R) % R for n, d in zip([0, 6, 15, 28], [7, 10, 18, 28])]
    assert backend.mle_values(frac) == expect
    assert backend.mle_values(fxs[0]) == [0, 2, 3, 4]
    f = backend.compute_frac_poly([mle(backend, 0, [12])], [mle(backend, 0, [4])])
    assert backend.evaluate_mle(f, []) == 3
    for nv, fv, gv in ((0, [0], [0]), (1, [0, 2], [0, 3])):
        with pytest.raises(InvalidParameters):
            backend.compute_frac_poly([mle(backend, nv, fv)], [mle(backend, nv, gv)])


def test_product_tree_known_answers(backend):
    # :422-460
    for nv, values, expected in (
        (1, [2, 3], [6, 0]),
        (2, [2, 3, 5, 7], [6, 35, 210, 0]),
        (3, [2, 3, 5, 7, 11, 13, 17, 19], [6, 35, 143, 323, 210, 46189, 9699690, 0]),
    ):
        frac = mle(backend, nv, values)
        prod = backend.compute_product_poly(frac)
        assert backend.mle_values(prod) == expected
        p1, p2 = backend.compute_product_factors(frac, prod)
        a, b = backend.mle_values(p1), backend.mle_values(p2)
        for row in range(len(expected) - 1):
            assert expected[row] == a[row] * b[row] % R
    with pytest.raises(InvalidParameters):
        backend.compute_product_poly(mle(backend, 0, [2]))


def test_nums_denoms_known_answers(backend):
    # :462-535
    fxs = [mle(backend, 2, [2 + i, 3 + i, 5 + i, 7 + i]) for i in range(3)]
    gxs = [mle(backend, 2, [0, 11 + i, 13 + i, 17 + i]) for i in range(3)]
    perms = [mle(backend, 2, [11 - 4 * i, 10 - 4 * i, 9 - 4 * i, 8 - 4 * i]) for i in range(3)]
    fv = [backend.mle_values(x) for x in fxs]
    gv = [backend.mle_values(x) for x in gxs]
    pv = [backend.mle_values(x) for x in perms]
    for beta, gamma in ((2, 5), (0, 5), (R - 2, 0), (0, 0)):
        nums, dens = backend.compute_nums_and_denoms(beta, gamma, fxs, gxs, perms)
        for c in range(3):
            n, d = backend.mle_values(nums[c]), backend.mle_values(dens[c])
            for row in range(4):
                assert n[row] == (fv[c][row] + beta * (4 * c + row) + gamma) % R
                assert d[row] == (gv[c][row] + beta * pv[c][row] + gamma) % R
    nums, dens = backend.compute_nums_and_denoms(
        2, 3, [mle(backend, 0, [v]) for v in (2, 3, 5)], [mle(backend, 0, [v]) for v in (7, 11, 13)],
        [mle(backend, 0, [v]) for v in (2, 0, 1)],
    )
    assert [backend.mle_values(x)[0] for x in nums] == [5, 8, 12]
    assert [backend.mle_values(x)[0] for x in dens] == [14, 14, 18]


def test_product_factors_known_answers(backend):
    # :537-593
    frac = mle(backend, 2, [2, 3, 5, 7])
    prod = mle(backend, 2, [11, 13, 17, 19])
    p1, p2 = backend.compute_product_factors(frac, prod)
    assert backend.mle_values(p1) == [2, 5, 11, 17]
    assert backend.mle_values(p2) == [3, 7, 13, 19]
    p1, p2 = backend.compute_product_factors(mle(backend, 1, [2, 3]), mle(backend, 1, [5, 7]))
    assert backend.mle_values(p1) == [2, 5] and backend.mle_values(p2) == [3, 7]
    with pytest.raises(InvalidParameters):
        backend.compute_product_factors(mle(backend, 3, list(range(8))), mle(backend, 1, [2, 3]))
    with pytest.raises(InvalidParameters):
        backend.compute_product_factors(mle(backend, 0, [2]), mle(backend, 0, [2]))


@pytest.mark.parametrize("nv", [1, 3, 10, 11, 14, 16])
def test_random_tables_match_oracle(backend, oracle, rng_factory, nv):
    rng = rng_factory(1000 + nv)
    n = 1 << nv
    tabs = [rng.fr_bytes(n) for _ in range(6)]
    ints = [fr_from_bytes(t.tobytes()) for t in tabs]
    m = [backend.mle_from_evaluations(nv, t) for t in tabs]
    point = fr_from_bytes(rng.fr_bytes(nv).tobytes())
    assert backend.evaluate_mle(m[0], point) == oracle.mle_evaluate(ints[0], point)
    coeffs = fr_from_bytes(rng.fr_bytes(3).tobytes())
    lc = backend.linear_combination(m[:3], coeffs)
    assert backend.mle_values(lc) == oracle.linear_combination(ints[:3], coeffs)
    eq = backend.build_eq_x_r(point)
    assert backend.mle_values(eq) == oracle.eq_x_r(point)
    fr = backend.compute_frac_poly(m[:2], m[2:4])
    assert backend.mle_values(fr) == oracle.frac(ints[:2], ints[2:4])
    frv = backend.mle_values(fr)
    prod = backend.compute_product_poly(fr)
    assert backend.mle_values(prod) == oracle.product(frv)
    p1, p2 = backend.compute_product_factors(fr, prod)
    assert (backend.mle_values(p1), backend.mle_values(p2)) == oracle.product_factors(frv, backend.mle_values(prod))
    beta, gamma = fr_from_bytes(rng.fr_bytes(2).tobytes())
    nums, dens = backend.compute_nums_and_denoms(beta, gamma, m[:2], m[2:4], m[4:6])
    on, od = oracle.nums_denoms(beta, gamma, ints[:2], ints[2:4], ints[4:6])
    assert [backend.mle_values(x) for x in nums] == on
    assert [backend.mle_values(x) for x in dens] == od


@pytest.mark.parametrize("nv", [12, 18, 22])
def test_eq_large_tables_match_closed_form(backend, rng_factory, nv):
    # eq(x, r) at sampled indices against the product formula (size-independent check)
    rng = rng_factory(2000 + nv)
    r = fr_from_bytes(rng.fr_bytes(nv).tobytes())
    eq = backend.build_eq_x_r(r)
    vals = backend.mle_to_evaluations(eq)
    for x in (0, 1, (1 << nv) - 1, 0x5A5A5 & ((1 << nv) - 1), 12345 % (1 << nv)):
        expect = 1
        for i in range(nv):
            expect = expect * (r[i] if (x >> i) & 1 else 1 - r[i]) % R
        assert int.from_bytes(vals[x].tobytes(), "little") == expect
    # sum over the hypercube of eq(x, r) is 1
    total = sum(int.from_bytes(v.tobytes(), "little") for v in vals[: min(len(vals), 1 << 12)])
    if nv <= 12:
        assert total % R == 1


def test_evaluate_large_table_matches_eq_weighted_sum(backend, rng_factory):
    nv = 20
    rng = rng_factory(77)
    tab = rng.fr_bytes(1 << nv)
    m = backend.mle_from_evaluations(nv, tab)
    point = fr_from_bytes(rng.fr_bytes(nv).tobytes())
    got = backend.evaluate_mle(m, point)
    # f(point) = <f, eq(., point)>, computed with the device eq table and numpy-free ints
    eq = backend.mle_to_evaluations(backend.build_eq_x_r(point))
    vals = tab
    acc = 0
    for i in range(1 << nv):
        acc += int.from_bytes(vals[i].tobytes(), "little") * int.from_bytes(eq[i].tobytes(), "little")
    assert got == acc % R


def test_evaluate_batch_matches_oracle(backend, oracle, rng_factory):
    """Batched evaluations: mixed sizes (including a constant and a table above
    one dot block), repeated and shared points, eq tables of one to three
    chunks built in one batch, byte-exact against the oracle."""
    rng = rng_factory(77)
    nvs = [0, 1, 3, 12, 12, 12, 15, 17, 5, 21]
    tabs = [rng.fr_bytes(1 << nv) for nv in nvs]
    mles = [backend.mle_from_evaluations(nv, t) for nv, t in zip(nvs, tabs)]
    p12 = fr_from_bytes(rng.fr_bytes(12).tobytes())
    points = []
    for nv in nvs:
        points.append(p12 if nv == 12 else fr_from_bytes(rng.fr_bytes(nv).tobytes()) if nv else [])
    points[5] = [0, 1] * 6  # a boolean point picks one entry
    got = backend.evaluate_mles(mles, points)
    for nv, t, pt, g in zip(nvs, tabs, points, got):
        assert g == oracle.mle_evaluate(fr_from_bytes(t.tobytes()), pt)
    assert got[5] == fr_from_bytes(tabs[5].tobytes())[int("10" * 6, 2)]
    assert backend.evaluate_mles([], []) == []
