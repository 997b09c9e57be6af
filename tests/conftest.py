import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (liboracle.so), built from oracle/ when it is missing."""
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    from oracle import oracle as o

    o.lib()
    return o


@pytest.fixture(scope="session")
def backend():
    """One Mi355xBackend (device 0) shared by the GPU tests."""
    from hyperplonk_amd.backend import Mi355xBackend

    return Mi355xBackend(0)


@pytest.fixture(scope="session")
def rng_factory():
    from hyperplonk_amd.synth import SplitMix64Stream

    return SplitMix64Stream
