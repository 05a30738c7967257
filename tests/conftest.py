import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers",
                            "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def qlib():
    """(hip, qcs) ctypes libraries; GPU tests require a device."""
    import qcs_amd
    hip, qcs = qcs_amd.load()
    if not hip.qcs_hip_device_available():
        pytest.fail("no HIP device: GPU tests must run on an MI355X")
    return hip, qcs


@pytest.fixture(scope="session")
def cl():
    import qcs_amd
    return qcs_amd.libc()


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return np.load(os.path.join(TESTS, "golden", "ref_vectors.npz"))
