import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def hpg():
    """The HIP package, initialized on device 0 (GPU tests only)."""
    import hpg_amd
    hpg_amd.initialize(0)
    return hpg_amd


@pytest.fixture(scope="session")
def gridone_golden():
    import numpy as np
    a = np.fromfile(os.path.join(ROOT, "tests", "golden", "gridone.bin"), dtype="<f8")
    a = a.reshape(91, 91, 2)
    return a[..., 0] + 1j * a[..., 1]  # [x][y]
