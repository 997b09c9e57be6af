"""Radix-2^29 Fq arithmetic and G1 group law of the MSM bucket stages
(hyperplonk_amd/csrc/field29.hpp, curve29.hpp), run on the host from the same
source the kernels compile (tests/cpp/field29_test.cpp).

Pinned against the 32-bit-limb arithmetic (field.hpp / curve.hpp, itself
pinned through the oracle's Python-integer fixtures): products and sums of
lazy operands up to 31 p, subtractions with every k p, the exact zero test on
k p and on values that share k p's low limb, and mixed / general additions
and doublings with every exceptional case compared as affine points."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "_build", "field29_test")


def test_radix29_field_and_group_law():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), BIN], check=True)
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "FIELD29 PASS" in res.stdout
