"""HyperPlonk proofs from the MI355X backend, byte for byte against the same
protocol over the CPU oracle backend (identical seeded circuits and SRS), each
GPU proof verified with pairings (tests/cpp/protocol_gpu_test.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "_build", "protocol_gpu_test")

pytestmark = pytest.mark.gpu


def test_gpu_proofs_match_oracle_proofs():
    assert os.path.exists(BIN), "build the test programs first (make -C tests/cpp)"
    cases = ["3", "vanilla", "11", "5", "jellyfish", "12", "8", "vanilla", "13", "12", "vanilla", "14"]
    res = subprocess.run([BIN] + cases, capture_output=True, text=True, timeout=600)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "PROTOCOL_GPU PASS" in res.stdout
    assert res.stdout.count("MATCH") == 4 and "MISMATCH" not in res.stdout


def test_srs_file_round_trip_on_device(tmp_path):
    assert os.path.exists(BIN), "build the test programs first (make -C tests/cpp)"
    res = subprocess.run([BIN, "srs", "10", str(tmp_path)], capture_output=True, text=True, timeout=600)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "SRS_GPU PASS" in res.stdout
