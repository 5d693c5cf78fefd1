import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "libaccint-public_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

try:  # one HIP runtime per process: torch's, loaded before the package's library
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def golden(name):
    return os.path.join(GOLDEN, name)


@pytest.fixture(scope="session")
def h2o_atoms():
    from libaccint import systems
    return systems.h2o_sto3g_geometry()


@pytest.fixture(scope="session")
def density7():
    rng = np.random.default_rng(1)
    C = rng.standard_normal((7, 5)) / np.sqrt(7)
    return C @ C.T
