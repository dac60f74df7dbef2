import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _has_gpu():
    """Count HIP devices through the HIP runtime (no torch import)."""
    import ctypes
    if not os.path.exists("/dev/kfd"):
        return False
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


HAS_GPU = _has_gpu()


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx():
    if not HAS_GPU:
        pytest.skip("no HIP device")
    from bsr_amd import Context
    c = Context(0)
    yield c
    c.close()


def rel_err(a, b, floor_frac=1e-3):
    """max_i |a_i-b_i| / max(|b_i|, floor_frac*max|b|): per-element relative
    error with a floor for elements that cancel to near zero."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if b.size == 0:
        return 0.0
    scale = np.abs(b).max()
    if scale == 0.0:
        return float(np.abs(a).max())
    den = np.maximum(np.abs(b), floor_frac * scale)
    return float(np.max(np.abs(a - b) / den))
