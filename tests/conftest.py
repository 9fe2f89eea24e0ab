import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("ho-cxx_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    config.addinivalue_line("markers", "slow: long CPU-only oracle runs")


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a); b = np.asarray(b)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


@pytest.fixture(scope="session")
def gpu_available():
    import hoc
    try:
        h = hoc.HO2D(K=2)
        h.close()
    except Exception as e:  # the product fails loudly; GPU tests must not silently pass
        pytest.fail(f"HIP path unavailable: {e}")
    return True
