import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the HIP library")


def _has_gpu():
    try:
        from pathtracer_amd import _lib as L
        return L.lib.pt_device_synchronize() == 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _has_gpu():
        pytest.fail("GPU test run without a visible HIP device")
    from pathtracer_amd import _lib as L
    return L


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
