import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lowlevelfem.jl_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def hooke_symgrad_D(E=2e5, nu=0.3):
    """6x6 Hooke matrix in SymGrad row order, as test/test_dsl.jl:16-26 builds it."""
    mu = E / (2 * (1 + nu))
    lam = 2 * mu * nu / (1 - 2 * nu)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = lam + 2 * mu
    D[np.arange(3, 6), np.arange(3, 6)] = mu
    return D


@pytest.fixture(scope="session")
def gpu_ctx():
    from lowlevelfem import _lib
    return _lib.default_context()
