"""Writes oracle fingerprints of multi-step runs too large to repeat inside a GPU test.

The FP64 oracle (oracle/liboracle.so, all host threads) advances each case and the script
stores, per field (vorticity, stream function): the values at 4096 seeded sample dofs, the
field's L2 norm, and its products with 32 seeded Gaussian vectors (drawn again by the test
from the same seed).  A GPU run of the same case is compared on the samples and, through the
random products, on the whole field: for a difference d, each product is N(0, |d|^2), so the
root mean square of the 32 differences estimates |d| (within a factor of about 1.3 at 99%).

Cases (data only, tests/golden/flow_<case>.npz):
  cavity32_K4 lid-driven cavity 32 x 32, K=4, Re=100, dt=5e-4 (the configured 1e-3 exceeds the
              RK4 limit of this mesh), 2000 RK4 steps from rest to t = 1, tol 1e-13
  bfs384_K5   backward-facing step 384 x 64 (jitter 0.1, seed 2), K=5, Re=400, dt=5e-4,
              10 RK4 steps from t = 20/3 (inflow ramp 2|sin(pi t/40)| = 1), tol 1e-14

Run: python tests/golden/make_flow_fixtures.py [case ...]   (about 10 minutes on 8 threads)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("ho-cxx_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))
from hoc import meshes  # noqa: E402
import pyoracle  # noqa: E402

N_SAMPLE, N_PROJ, SEED = 4096, 32, 1234


def cases():
    return {
        "cavity32_K4": dict(mesh=lambda: meshes.cavity(32, 2), K=4,
                            kw=dict(problem_type=1, Re=100.0, dt=5e-4, time_integration_type=4, advection_type=1,
                                    tol=1e-13, maxiter=500), t0=0.0, steps=2000),
        "bfs384_K5": dict(mesh=lambda: meshes.backward_facing_step(384, 64, 2, jitter=0.1, seed=2), K=5,
                          kw=dict(problem_type=2, Re=400.0, dt=5e-4, time_integration_type=4, advection_type=1,
                                  tol=1e-14, maxiter=500), t0=20.0 / 3.0, steps=10),
    }


def fingerprint(x, seed=SEED):
    """Sample indices, sampled values, norm and random products of x (same seed -> same draws)."""
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(len(x), size=min(N_SAMPLE, len(x)), replace=False)).astype(np.int64)
    proj = np.array([rng.standard_normal(len(x)) @ x for _ in range(N_PROJ)])
    return idx, x[idx], float(np.linalg.norm(x)), proj


def main(names):
    pyoracle.set_threads(max(1, min(16, len(os.sched_getaffinity(0)))))
    for name, c in cases().items():
        if names and name not in names:
            continue
        t0 = time.time()
        m = c["mesh"]()
        o = pyoracle.Oracle(m, c["K"], **c["kw"])
        o.set_time(c["t0"])
        iters = []
        for s in range(c["steps"]):
            o.step(1)
            iters.append(o.last_solve()[0])
            if s % max(1, c["steps"] // 20) == 0 or s == c["steps"] - 1:
                print(name, "step", s, "last solve", o.last_solve(), f"{time.time() - t0:.0f} s", flush=True)
        out = {"steps": c["steps"], "t0": c["t0"], "time": o.time(), "oracle_iters": np.array(iters)}
        for fld, x in (("omega", o.vorticity()), ("psi", o.streamfunction())):
            idx, vals, nrm, proj = fingerprint(x)
            out["idx"] = idx
            out[fld + "_sample"], out[fld + "_norm"], out[fld + "_proj"] = vals, nrm, proj
        np.savez_compressed(os.path.join(HERE, f"flow_{name}.npz"), **out)
        print(name, "written", f"{time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
