"""Pins the CPU oracle before it is trusted as the parity checker.

Every check compares the oracle with something it does not share code with:
Python-integer fixtures (tests/golden/golden.json, written from the
mathematics by tests/golden/gen_golden.py), the reference's own known-answer
tests (cited per test), closed forms and brute force on tiny inputs.
"""
import random

import pytest

from tests import pyfield as pf

R, Q = pf.R, pf.Q


def H(v):
    return int(v, 16)


@pytest.mark.parametrize("field,modulus", [("fr", R), ("fq", Q)])
def test_field_ops_match_python_integers(oracle, golden, field, modulus):
    case = golden[field]
    lhs = [H(v) for v in case["lhs"]]
    rhs = [H(v) for v in case["rhs"]]
    for op in ("add", "sub", "mul"):
        assert oracle.field_op(field, op, lhs, rhs) == [H(v) for v in case[op]], op
    assert oracle.field_op(field, "neg", lhs) == [H(v) for v in case["neg"]]
    assert oracle.field_op(field, "square", lhs) == [H(v) for v in case["square"]]
    inv_in = [H(v) for v in case["inverse_input"]]
    assert oracle.field_op(field, "inverse", inv_in) == [H(v) for v in case["inverse"]]


@pytest.mark.parametrize("field,modulus", [("fr", R), ("fq", Q)])
def test_field_rejects_noncanonical(oracle, field, modulus):
    with pytest.raises(oracle.OracleError):
        oracle.field_op(field, "add", [modulus], [0])
    with pytest.raises(oracle.OracleError):
        oracle.field_op(field, "inverse", [0])


def test_generator_and_multiples(oracle, golden):
    g = oracle.g1_generator()
    assert g == pf.point_from_hex(golden["generator"])
    small = [pf.point_from_hex(p) for p in golden["g1_small_multiples"]]
    for k in (1, 2, 3, 7, 64, 255, 300):
        assert oracle.g1_mul(g, k) == small[k]
    for entry in golden["g1_big_multiples"]:
        assert oracle.g1_mul(g, H(entry["k"])) == pf.point_from_hex(entry["point"])
    assert oracle.g1_mul(g, 0) is None
    assert oracle.g1_add(small[5], small[7]) == small[12]
    assert oracle.g1_add(small[9], small[9]) == small[18]
    assert oracle.g1_add(small[9], pf.aff_neg(small[9])) is None
    assert oracle.g1_add(None, small[4]) == small[4]


def test_on_curve_non_subgroup_point_is_accepted_by_curve_check_only(oracle):
    # (0, 2) is on y^2 = x^3 + 4 but outside the prime-order subgroup
    # (backend/src/ascend910b3/encoding.rs:224-231); the oracle only checks the curve.
    assert oracle.lib().oracle_g1_on_curve(pf.packet((0, 2))) == 1
    assert oracle.lib().oracle_g1_on_curve(pf.packet((1, 1))) == 0


def _msm_case(golden, case):
    small = [pf.point_from_hex(p) for p in golden["g1_small_multiples"]]

    def base(a):
        if a <= 300:
            return small[a]
        return pf.aff_neg(small[R - a])

    coeffs = [H(c) for c in case["coeffs"]]
    scalars = [H(s) for s in case["scalars"]]
    points = b"".join(pf.packet(base(a)) for a in coeffs)
    scal = b"".join(s.to_bytes(32, "little") for s in scalars)
    return points, scal, len(coeffs), pf.point_from_hex(case["expected"])


def test_msm_closed_form_cases(oracle, golden):
    for case in golden["msm"]:
        points, scal, n, expected = _msm_case(golden, case)
        got = oracle.msm_pippenger_packets(points, scal, n, threads=3)
        assert pf.unpacket(got) == expected, case["name"]
        if n <= 65:
            assert pf.unpacket(oracle.msm_naive_packets(points, scal, n)) == expected, case["name"]


def test_mle_known_answers(oracle):
    # backend/src/cpu/multilinear_polynomial.rs:281-284: [2,3,5,7] at (2,3) = 19
    assert oracle.mle_evaluate([2, 3, 5, 7], [2, 3]) == 19
    assert oracle.mle_evaluate([11], []) == 11
    # :310-322: 2p + 3q - 2p at (2,3) = 99 with q = [11,13,17,19]
    comb = oracle.linear_combination([[2, 3, 5, 7], [11, 13, 17, 19], [2, 3, 5, 7]], [2, 3, R - 2])
    assert oracle.mle_evaluate(comb, [2, 3]) == 99
    # :364-376: frac = [0,6,15,28] / [7,10,18,28]
    fr = oracle.frac([[0, 2, 3, 4], [2, 3, 5, 7]], [[1, 2, 3, 4], [7, 5, 6, 7]])
    assert fr == [n * pow(d, R - 2, R) % R for n, d in zip([0, 6, 15, 28], [7, 10, 18, 28])]
    assert oracle.frac([[12]], [[4]]) == [3]
    for f, g in (([0], [0]), ([0, 2], [0, 3])):
        with pytest.raises(oracle.OracleError):
            oracle.frac([f], [g])
    # :401-435: product tree
    for values, expected in (
        ([2, 3], [6, 0]),
        ([2, 3, 5, 7], [6, 35, 210, 0]),
        ([2, 3, 5, 7, 11, 13, 17, 19], [6, 35, 143, 323, 210, 46189, 9699690, 0]),
    ):
        prod = oracle.product(values)
        assert prod == expected
        p1, p2 = oracle.product_factors(values, prod)
        for row in range(len(prod) - 1):
            assert prod[row] == p1[row] * p2[row] % R
    # :516-557: factors
    assert oracle.product_factors([2, 3, 5, 7], [11, 13, 17, 19]) == ([2, 5, 11, 17], [3, 7, 13, 19])
    assert oracle.product_factors([2, 3], [5, 7]) == ([2, 5], [3, 7])
    # :497-513: scalar nums/denoms [5,8,12] / [14,14,18]
    nums, dens = oracle.nums_denoms(2, 3, [[2], [3], [5]], [[7], [11], [13]], [[2], [0], [1]])
    assert [n[0] for n in nums] == [5, 8, 12]
    assert [d[0] for d in dens] == [14, 14, 18]


def test_nums_denoms_formula(oracle):
    # backend/src/cpu/multilinear_polynomial.rs:438-495
    fxs = [[2 + i, 3 + i, 5 + i, 7 + i] for i in range(3)]
    gxs = [[0, 11 + i, 13 + i, 17 + i] for i in range(3)]
    perms = [[11 - 4 * i, 10 - 4 * i, 9 - 4 * i, 8 - 4 * i] for i in range(3)]
    for beta, gamma in ((2, 5), (0, 5), (R - 2, 0), (0, 0)):
        nums, dens = oracle.nums_denoms(beta, gamma, fxs, gxs, perms)
        for c in range(3):
            for row in range(4):
                assert nums[c][row] == (fxs[c][row] + beta * (4 * c + row) + gamma) % R
                assert dens[c][row] == (gxs[c][row] + beta * perms[c][row] + gamma) % R


def test_eq_table_matches_naive_product(oracle):
    # backend/src/common/eq.rs:106-152 (build_eq_x_r_vec vs naive)
    rng = pf.SplitMix64(11)
    for nv in range(1, 9):
        r = [rng.field() for _ in range(nv)]
        assert oracle.eq_x_r(r) == pf.eq_naive(r)
        x = [rng.field() for _ in range(nv)]
        expect = 1
        for xi, ri in zip(x, r):
            expect = expect * (xi * ri + (1 - xi) * (1 - ri)) % R
        assert oracle.eq_eval(x, r) == expect
    with pytest.raises(oracle.OracleError):
        oracle.eq_x_r([])


def test_evaluate_and_fold_match_multilinear_closed_form(oracle):
    # f(x) = sum_b f[b] eq(b, x): evaluation equals the eq-weighted sum.
    rng = pf.SplitMix64(12)
    for nv in range(1, 8):
        table = [rng.field() for _ in range(1 << nv)]
        point = [rng.field() for _ in range(nv)]
        weights = pf.eq_naive(point)
        expect = sum(a * w for a, w in zip(table, weights)) % R
        assert oracle.mle_evaluate(table, point) == expect
        folded = oracle.mle_fold(table, point[0])
        assert oracle.mle_evaluate(folded, point[1:]) == expect if nv > 1 else folded[0] == expect


def test_sumcheck_reference_known_answer(oracle):
    # backend/src/cpu/sum_check.rs:225-253: 2 f + 3 f^2 over f = [2,3,5,7], challenge 2
    st = oracle.SumCheckState(2, 2, [[2, 3, 5, 7]], [(2, [0]), (3, [0, 0])])
    with pytest.raises(oracle.OracleError) as e:
        st.round(2)
    assert e.value.status == oracle.INVALID_STATE
    assert st.round(None) == [101, 194, 317]
    with pytest.raises(oracle.OracleError):
        st.round(None)
    assert st.round(2) == [56, 261, 616]
    with pytest.raises(oracle.OracleError) as e:
        st.round(2)
    assert e.value.status == oracle.INVALID_STATE


# The nine cases of backend/src/ascend910b3/native/tests/sumcheck_tensor_edge.cpp:164-170
EDGE_CASES = [
    ("n1_d2", 1, 2, 0), ("n2_d2", 2, 2, 0), ("n3_multi", 3, 2, 1), ("n9_tail", 9, 2, 0),
    ("n12_d2", 12, 2, 0), ("n12_multi", 12, 2, 1), ("n9_mixed", 9, 2, 2), ("n9_d3", 9, 3, 3),
    ("n2_empty", 2, 2, 4),
]


def edge_case_inputs(n, kind, seed):
    """Inputs shaped like sumcheck_tensor_edge.cpp:189-213 (factor 2 repeats factor 0)."""
    rng = pf.SplitMix64(seed)
    tables = [[rng.field() for _ in range(1 << n)] for _ in range(2)]
    if n > 1:
        tables[0][0], tables[0][1] = 0, R - 1
        tables[1][0], tables[1][1] = R - 1, 1
    tables.append(tables[0])
    products = {
        0: [(rng.nonboolean(), [2, 1])],
        1: [(rng.nonboolean(), [2, 1]), (R - 1, [0, 2]), (rng.nonboolean(), [1, 1])],
        2: [(3, [0]), (7, [2, 1])],
        3: [(rng.nonboolean(), [0, 1, 2]), (R - 1, [1])],
        4: [],
    }[kind]
    return tables, products, rng


@pytest.mark.parametrize("name,n,degree,kind", [c for c in EDGE_CASES if c[1] <= 9])
def test_sumcheck_rounds_match_textbook(oracle, name, n, degree, kind):
    tables, products, rng = edge_case_inputs(n, kind, 0x534354454E534F52 + 17 * n + kind)
    st = oracle.SumCheckState(n, degree, tables, products)
    cur = [list(t) for t in tables]
    prev = None
    for rnd in range(n):
        ch = None
        if rnd:
            ch = rng.nonboolean()
            cur = pf.fold_tables(cur, ch)
        got = st.round(ch)
        assert got == pf.textbook_message(cur, products, degree), (name, rnd)
        if prev is not None:
            # sumcheck invariant: g_r(0) + g_r(1) = g_{r-1}(challenge)
            assert (got[0] + got[1]) % R == interpolate_at(prev, ch)
        prev = got


def interpolate_at(values, x):
    n = len(values)
    total = 0
    for j in range(n):
        num, den = 1, 1
        for m in range(n):
            if m != j:
                num = num * (x - m) % R
                den = den * (j - m) % R
        total += values[j] * num * pow(den, R - 2, R)
    return total % R


def test_kzg_open_closed_form(oracle):
    # layer k of an SRS with trapdoor t holds g * eq(t[k..], x); an open quotient
    # commitment is therefore g * q_j(t suffix), checked against Python ints.
    rng = pf.SplitMix64(21)
    N, nv = 6, 4
    t = [rng.field() for _ in range(N)]
    layers = {k: oracle.srs_layer_packets(t[k:], threads=2) for k in range(N + 1)}
    # spot-check layer points directly
    for k in (0, 3, N):
        pts = layers[k]
        eq = pf.eq_naive(t[k:]) if k < N else [1]
        for x in (0, len(eq) - 1):
            assert pf.unpacket(pts[97 * x : 97 * (x + 1)]) == pf.aff_mul(pf.G, eq[x])
    table = [rng.field() for _ in range(1 << nv)]
    point = [rng.field() for _ in range(nv)]
    ignored = N - nv + 1
    concat = b"".join(layers[ignored + j] for j in range(nv))
    quotients, value = oracle.kzg_open_packets(concat, table, point, threads=2)
    assert value == pf.evaluate(table, point)
    f = list(table)
    for j in range(nv):
        q = [(f[2 * b + 1] - f[2 * b]) % R for b in range(len(f) // 2)]
        f = [(f[2 * b] + point[j] * q[b]) % R for b in range(len(f) // 2)]
        k = ignored + j
        expect = pf.aff_mul(pf.G, pf.evaluate(q, t[k:]) if k < N else q[0])
        assert pf.unpacket(quotients[j]) == expect, j


def test_pippenger_matches_naive_random(oracle):
    rng = random.Random(5)
    small = [None]
    for _ in range(40):
        small.append(pf.aff_add(small[-1], pf.G))
    for n in (1, 5, 40):
        pts = b"".join(pf.packet(small[1 + (i * 7) % 40]) for i in range(n))
        sc = b"".join(rng.randrange(R).to_bytes(32, "little") for _ in range(n))
        assert oracle.msm_pippenger_packets(pts, sc, n, 4) == oracle.msm_naive_packets(pts, sc, n)
