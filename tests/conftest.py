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


def fixture_check(g, name, tol):
    """Compare a handle's fields with the oracle fingerprints tests/golden/flow_<name>.npz
    (tests/golden/make_flow_fixtures.py): 4096 sampled values, the norm and 32 seeded random
    projections per field, each within tol (relative)."""
    import numpy as np
    z = np.load(os.path.join(ROOT, "tests", "golden", f"flow_{name}.npz"))
    assert abs(g.time - float(z["time"])) < 1e-12
    idx = z["idx"]
    out = {}
    for fld, x in (("omega", g.get_vorticity()), ("psi", g.get_streamfunction())):
        rng = np.random.default_rng(1234)
        ridx = np.sort(rng.choice(len(x), size=min(4096, len(x)), replace=False))
        assert np.array_equal(ridx, idx)
        proj = np.array([rng.standard_normal(len(x)) @ x for _ in range(len(z[fld + "_proj"]))])
        e_sample = rel_l2(x[idx], z[fld + "_sample"])
        e_proj = float(np.sqrt(np.mean((proj - z[fld + "_proj"]) ** 2)) / z[fld + "_norm"])
        e_norm = abs(np.linalg.norm(x) - float(z[fld + "_norm"])) / float(z[fld + "_norm"])
        out[fld] = (e_sample, e_proj, e_norm)
        print(name, fld, "sampled rel L2", e_sample, "projected estimate of rel L2", e_proj, "norm", e_norm)
    for fld, (es, ep, en) in out.items():
        assert es < tol and ep < tol and en < tol, (fld, es, ep, en)
