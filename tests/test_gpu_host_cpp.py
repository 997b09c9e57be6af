"""The C++ host mirror of the reference traits (hyperplonk_amd/host/mi355x_backend.hpp)
runs the reference CPU backend's known-answer tests against the GPU
(tests/cpp/backend_test.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "_build", "backend_test")

pytestmark = pytest.mark.gpu


def build_backend_test():
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    subprocess.run(
        ["g++", "-std=c++17", "-O2", "-Wall", os.path.join(ROOT, "tests", "cpp", "backend_test.cpp"), "-o", BIN,
         "-L" + os.path.join(ROOT, "hyperplonk_amd"), "-lhp_mi355x", "-Wl,-rpath," + os.path.join(ROOT, "hyperplonk_amd")],
        check=True,
    )


def test_cpp_host_mirror_known_answers():
    if not os.path.exists(BIN):
        build_backend_test()
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "BACKEND_TEST PASS" in res.stdout
