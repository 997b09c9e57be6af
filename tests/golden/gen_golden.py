#!/usr/bin/env python3
"""Write tests/golden/golden.json: exact BLS12-381 fixtures from Python integers.

Everything here follows from the published curve constants and textbook
arithmetic (modular +,-,*, Fermat inverse, affine chord-and-tangent addition on
y^2 = x^3 + 4); nothing is computed by the oracle or by the device code, so
the fixtures pin both.  Run:  python3 tests/golden/gen_golden.py

Case shapes follow the reference's own fixture plan
(backend/src/ascend910b3/native/tests/generate_golden.py:26-69 for field edge
values, :322-415 for the MSM closed form (sum a_i s_i) * G over bases a_i * G
and its zero / ones / identity / cancel / alternating / boundary / exception
patterns); the generator itself is written from the mathematics.
"""
import json
import os
import random

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
Q = int(
    "1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab", 16
)
GX = int(
    "17f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb", 16
)
GY = int(
    "08b3f481e3aaa0f1a09e30ed741d8ae4fcf5e095d5d00af600db18cb2c04b3edd03cc744a2888ae40caa232946c5e7e1", 16
)
G = (GX, GY)


def on_curve(p):
    return p is None or (p[1] * p[1] - p[0] ** 3 - 4) % Q == 0


def aff_add(p, q):
    if p is None:
        return q
    if q is None:
        return p
    (x1, y1), (x2, y2) = p, q
    if x1 == x2:
        if (y1 + y2) % Q == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, Q - 2, Q) % Q
    else:
        lam = (y2 - y1) * pow(x2 - x1, Q - 2, Q) % Q
    x3 = (lam * lam - x1 - x2) % Q
    return (x3, (lam * (x1 - x3) - y1) % Q)


def aff_neg(p):
    return None if p is None else (p[0], (-p[1]) % Q)


def aff_mul(p, k):
    k %= R
    acc = None
    for bit in bin(k)[2:] if k else "":
        acc = aff_add(acc, acc)
        if bit == "1":
            acc = aff_add(acc, p)
    return acc


def hexs(v):
    return format(v, "x")


def pt(p):
    return None if p is None else [hexs(p[0]), hexs(p[1])]


def field_edges(p):
    values = [0, 1, p - 1, p - 2, p // 2, (p + 1) // 2]
    for bits in range(16, p.bit_length(), 16):
        for v in ((1 << bits) - 1, 1 << bits, (1 << bits) + 1):
            if v < p:
                values.append(v)
    return values


def field_case(p, seed, count):
    rng = random.Random(seed)
    edges = field_edges(p)
    lhs = [edges[i] if i < len(edges) else rng.randrange(p) for i in range(count)]
    rhs = [edges[(i + 3) % len(edges)] if i < len(edges) else rng.randrange(p) for i in range(count)]
    inv_in = [a if a else 1 for a in lhs]
    return {
        "lhs": [hexs(v) for v in lhs],
        "rhs": [hexs(v) for v in rhs],
        "add": [hexs((a + b) % p) for a, b in zip(lhs, rhs)],
        "sub": [hexs((a - b) % p) for a, b in zip(lhs, rhs)],
        "mul": [hexs(a * b % p) for a, b in zip(lhs, rhs)],
        "neg": [hexs(-a % p) for a in lhs],
        "square": [hexs(a * a % p) for a in lhs],
        "inverse_input": [hexs(v) for v in inv_in],
        "inverse": [hexs(pow(a, p - 2, p)) for a in inv_in],
    }


def main():
    out = {"r": hexs(R), "q": hexs(Q), "generator": pt(G)}
    out["fr"] = field_case(R, 20261203, 96)
    out["fq"] = field_case(Q, 20261204, 96)

    # k * G for k = 1..300 by repeated addition, plus random large multiples.
    small = [None]
    for _ in range(300):
        small.append(aff_add(small[-1], G))
    assert all(on_curve(p) for p in small)
    rng = random.Random(20261205)
    big = []
    for k in [R - 1, R - 2, (R - 1) // 2, 1 << 128, (1 << 254) + 1] + [rng.randrange(1, R) for _ in range(11)]:
        big.append({"k": hexs(k), "point": pt(aff_mul(G, k))})
    assert aff_mul(G, R) is None
    out["g1_small_multiples"] = [pt(p) for p in small]  # index k -> k*G
    out["g1_big_multiples"] = big

    def base(a):
        a %= R
        if a <= 300:
            return small[a]
        if R - a <= 300:
            return aff_neg(small[R - a])
        raise ValueError("base coefficient outside the small table")

    cases = []

    def add_case(name, coeffs, scalars, check_direct):
        coeffs = [c % R for c in coeffs]
        expected = aff_mul(G, sum(a * s for a, s in zip(coeffs, scalars)) % R)
        if check_direct:
            direct = None
            for a, s in zip(coeffs, scalars):
                direct = aff_add(direct, aff_mul(base(a), s))
            assert direct == expected, name
        cases.append(
            {"name": name, "coeffs": [hexs(c) for c in coeffs], "scalars": [hexs(s) for s in scalars],
             "expected": pt(expected)}
        )

    for count in (0, 1, 2, 3, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257):
        rng = random.Random(20261206 + count)
        coeffs = list(range(1, count + 1))
        scalars = [rng.randrange(R) for _ in range(count)]
        add_case(f"uniform-{count}", coeffs, scalars, count <= 17)
    add_case("zero-257", list(range(1, 258)), [0] * 257, False)
    add_case("ones-257", list(range(1, 258)), [1] * 257, False)
    add_case("identity-65", [0] * 65, [random.Random(7).randrange(R) for _ in range(65)], True)
    rng = random.Random(20261207)
    cancel_c, cancel_s = [], []
    for i in range(64):
        s = rng.randrange(R)
        cancel_c += [i + 1, -(i + 1)]
        cancel_s += [s, s]
    add_case("cancel-128", cancel_c, cancel_s, False)
    cyc = (R - 1, 0, 1 << 254, (1 << 254) + 1)
    add_case("alternating-129", list(range(1, 130)), [cyc[i % 4] for i in range(129)], False)
    paired = list(range(1, 258))
    rng.shuffle(paired)
    add_case("boundary-257", paired, list(range(256)) + [R - 1], False)
    cyc = (0, 1, 2, 3, -6, 0, 5, -5, 0)
    add_case("mixed-exceptions-65", [cyc[i % len(cyc)] for i in range(65)], [1] * 65, True)
    coeffs, scalars = [], []
    for bucket in range(1, 49):
        pattern = (1, 2, 3, -6) if bucket % 3 == 0 else (1, 2, 4, 8)
        coeffs += list(pattern)
        scalars += [bucket] * 4
    add_case("bucket-exceptions-192", coeffs, scalars, True)
    # every point in one bucket (hot), with equal points repeated: doubling paths
    add_case("hot-equal-200", [7] * 200, [5] * 200, False)
    add_case("negated-pairs-100", [3, -3] * 50, [11, 11] * 50, False)
    out["msm"] = cases

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
