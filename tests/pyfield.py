"""Python-integer arithmetic used by the tests as an independent check.

Textbook modular arithmetic and affine chord-and-tangent addition on the
BLS12-381 G1 curve y^2 = x^3 + 4; it shares no code with the oracle (C) or the
device library (HIP), so agreement pins both.
"""
import random

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
Q = int(
    "1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab", 16
)
G = (
    int("17f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb", 16),
    int("08b3f481e3aaa0f1a09e30ed741d8ae4fcf5e095d5d00af600db18cb2c04b3edd03cc744a2888ae40caa232946c5e7e1", 16),
)


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


def point_from_hex(h):
    return None if h is None else (int(h[0], 16), int(h[1], 16))


def packet(p):
    if p is None:
        return b"\x01" + bytes(96)
    return b"\x00" + p[0].to_bytes(48, "little") + p[1].to_bytes(48, "little")


def unpacket(b):
    if b[0] == 1:
        return None
    return (int.from_bytes(b[1:49], "little"), int.from_bytes(b[49:97], "little"))


def evaluate(table, point):
    """ark-poly fix_variables chain: point[0] binds the lowest table bit."""
    t = [v % R for v in table]
    for r in point:
        t = [(t[2 * b] + r * (t[2 * b + 1] - t[2 * b])) % R for b in range(len(t) // 2)]
    return t[0]


def eq_naive(r):
    out = []
    for x in range(1 << len(r)):
        v = 1
        for i, ri in enumerate(r):
            v = v * (ri if (x >> i) & 1 else (1 - ri)) % R
        out.append(v)
    return out


def textbook_message(tables, products, degree):
    """Round message by direct evaluation of every product at t = 0..degree."""
    pairs = len(tables[0]) // 2
    out = []
    for t in range(degree + 1):
        acc = 0
        for b in range(pairs):
            for coeff, idx in products:
                term = coeff
                for f in idx:
                    lo, hi = tables[f][2 * b], tables[f][2 * b + 1]
                    term = term * (lo + t * (hi - lo)) % R
                acc += term
        out.append(acc % R)
    return out


def fold_tables(tables, c):
    return [[(t[2 * b] + c * (t[2 * b + 1] - t[2 * b])) % R for b in range(len(t) // 2)] for t in tables]


class SplitMix64:
    """SplitMix64 with 255-bit rejection sampling of canonical Fr values
    (the sampling scheme of backend/src/ascend910b3/native/tests/sumcheck_tensor_edge.cpp:62-80)."""

    def __init__(self, seed):
        self.state = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def field(self):
        while True:
            v = 0
            for w in range(4):
                bits = self.next()
                if w == 3:
                    bits &= 0x7FFFFFFFFFFFFFFF
                v |= bits << (64 * w)
            if v < R:
                return v

    def nonboolean(self):
        while True:
            v = self.field()
            if v > 1:
                return v
