"""Prover service (hyperplonk_amd/host/serve.hpp): a small mixed trace of
vanilla and jellyfish requests served with continuous batching and again
serially; every request's proof must be byte-identical in the two runs (alone
versus in the mix), and sampled proofs must verify."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "_build", "prove_bench")

pytestmark = pytest.mark.gpu


def test_served_proofs_identical_to_serial_proofs():
    assert os.path.exists(BIN), "build the host programs first (make -C hyperplonk_amd/host)"
    res = subprocess.run([BIN, "serve", "40", "auto", "5", "10", "0.3", "6", "11"],
                         capture_output=True, text=True, timeout=600)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
    out = json.loads(res.stdout.strip().splitlines()[-1])
    assert out["identical_proofs"] == out["requests"] == 40
    assert out["batched"]["errors"] == 0 and out["serial"]["errors"] == 0
    assert out["verified"] == out["verify_checked"] > 0
    # the mix really overlapped: more than one request on the device at once
    assert out["batched"]["max_concurrent"] > 1
    assert out["batched"]["ops"] > out["batched"]["ticks"]


def test_served_proofs_under_a_binding_budget_match_the_oracle():
    """A budget of 2.5 of the largest measured request footprints holds
    requests back (admission waits with slots free); served proofs still
    equal the serial ones, and sampled served proofs equal the CPU oracle
    backend's proofs of the same circuits (tests/cpp/serve_oracle_test.cpp)."""
    prog = os.path.join(ROOT, "_build", "serve_oracle_test")
    assert os.path.exists(prog), "build the test programs first (make -C tests/cpp)"
    res = subprocess.run([prog, "24", "5", "9", "8", "2.5", "21"], capture_output=True, text=True, timeout=600)
    print(res.stdout, res.stderr)
    assert res.returncode == 0 and "SERVE_ORACLE PASS" in res.stdout, res.stdout + res.stderr
