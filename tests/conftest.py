"""pytest configuration: the `gpu` marker and import paths.

GPU tests call libmpm_gpu.so through its C-ABI (mpm_3d_amd/gpu.py) and compare against
the CPU oracle (oracle/pyoracle.py) on the same seeded inputs.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def examples():
    return os.path.join(ROOT, "tests", "golden", "examples")
