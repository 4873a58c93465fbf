import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")
    config.addinivalue_line("markers", "slow: longer oracle runs")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def hip_lib():
    """The HIP library; on a GPU run a missing library or device is an error, not a skip."""
    from tpsc_conductivity_amd import capi

    L = capi.lib()
    return L


def relerr(a, b):
    """normwise relative error max|a-b| / max|b|"""
    import numpy as np

    a = np.asarray(a)
    b = np.asarray(b)
    den = float(np.abs(b).max()) if b.size else 0.0
    return float(np.abs(a - b).max()) / (den if den > 0 else 1.0)
