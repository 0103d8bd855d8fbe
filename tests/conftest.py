import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")


def rel_err(a, b, floor=1e-300):
    """Normwise relative error ||a - b||_inf / max(||b||_inf, floor)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), floor))


@pytest.fixture(scope="session")
def hip():
    """The HIP runtime; GPU tests fail (not skip) when it cannot be created."""
    from pylbm_amd import runtime
    return runtime.get()
