"""The C++ host mirror of the reference traits (hyperplonk_amd/host/mi355x_backend.hpp)
runs the reference CPU backend's known-answer tests against the GPU
(tests/cpp/backend_test.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "_build", "backend_test")

pytestmark = pytest.mark.gpu


def test_cpp_host_mirror_known_answers():
    assert os.path.exists(BIN), "build the test programs first (make -C tests/cpp)"
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "BACKEND_TEST PASS" in res.stdout
