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


def test_static_archive_program_known_answers():
    """A C program linked against libhp_mi355x.a (no shared library), the way
    the reference's Rust glue links its bridge statically, runs the evaluate
    and sumcheck known answers on the GPU (tests/cpp/static_gpu_test.c)."""
    prog = os.path.join(ROOT, "_build", "static_gpu_test")
    assert os.path.exists(prog), "build the test programs first (make -C tests/cpp)"
    res = subprocess.run([prog], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and "STATIC_GPU PASS" in res.stdout, (res.returncode, res.stdout, res.stderr)


def test_loaded_library_matches_the_sources_on_this_machine():
    """On the GPU machine too, the library the tests load was compiled from the
    sources in this tree (its embedded source digest equals theirs)."""
    from hyperplonk_amd import abi

    assert abi.build_info() == f"sources-sha256={abi.source_digest()} arch=gfx950"


def test_runtime_audit_of_online_calls():
    """Every HIP allocation, copy, memset, synchronisation and launch the
    archive makes is counted per entry-point call (--wrap shims,
    tests/cpp/runtime_audit_test.c, as the reference audits its bridge:
    backend/src/ascend910b3/native/tests/runtime_audit.h:10-38).  After
    warm-up, evaluate, commit, open and every sumcheck round allocate nothing,
    keep their copies within byte budgets and synchronise at most once;
    buffer_retire neither copies, synchronises nor launches."""
    prog = os.path.join(ROOT, "_build", "runtime_audit_test")
    assert os.path.exists(prog), "build the test programs first (make -C tests/cpp)"
    res = subprocess.run([prog], capture_output=True, text=True, timeout=180)
    print(res.stdout)
    assert res.returncode == 0 and "RUNTIME_AUDIT PASS" in res.stdout, (res.returncode, res.stdout, res.stderr)
