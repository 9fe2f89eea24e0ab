"""The named workloads at their own sizes.

- Backward-facing step, 384 x 64 jittered grid (22 528 elements), K=5, Re=400, dt=5e-4:
  Poisson solve and stage residual against the oracle on the full mesh; FP64 vs mixed
  after 20 RK4 steps on the full mesh (size-independent agreement property).
- Lid-driven cavity 32 x 32, K=4, Re=100: 100 RK4 steps against the oracle run in the test, and
  2000 steps (to t = 1, the configured end time) against the oracle's fingerprints.  The
  configured dt=1e-3 exceeds the explicit RK4 diffusion limit of this mesh (7.3e-4, from
  rho(J^-1 A) = 3.83e5; the oracle itself diverges within 15 steps), so the runs use 5e-4.
"""
import numpy as np
import pytest

import hoc
import pyoracle
from hoc import meshes
from pyoracle import Oracle
from conftest import fixture_check, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bfs_full():
    return meshes.backward_facing_step(384, 64, 2, jitter=0.1, seed=2)


def oracle_threads():
    import os
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 1


def test_bfs_full_poisson_and_stage(gpu_available, bfs_full):
    m = bfs_full
    assert m.n_elem == 22528
    K = 5
    prev = pyoracle.set_threads(oracle_threads())
    try:
        o = Oracle(m, K, problem_type=2, Re=400.0, dt=5e-4, time_integration_type=4, advection_type=1, tol=1e-12,
                   maxiter=300)
        g = hoc.HO2D(K=K, problem_type=2, Re=400.0, dt=5e-4, time_integration_type=4, advection_type=1, tol=1e-12,
                     maxiter=300)
        g.load_mesh(m)
        g.setup()
        xy = o.table("sps_xy").reshape(-1, 2)
        w = np.sin(0.7 * xy[:, 0]) * np.cos(2.0 * xy[:, 1]) * 3.0
        t = 7.0   # a point of the inflow ramp 2|sin(pi t / 40)|
        p_ref = o.poisson(w, t)
        p_got = g.poisson(w, t)
        ep = rel_l2(p_got, p_ref)
        k_ref = o.stage(w, t)
        k_got = g.stage(w, t)
        ek = rel_l2(k_got, k_ref)
    finally:
        pyoracle.set_threads(prev)
    print("bfs 384x64 K=5: psi", ep, "stage", ek, "gpu its", g.stats()["last_iters"], "krylov", g.stats()["krylov"])
    assert ep < 1e-10 and ek < 1e-9


def test_bfs_full_mixed_vs_fp64(gpu_available, bfs_full):
    # The wall-vorticity flux divides the wall gradient of psi by dt/Re (8e5 here), so two runs
    # that each stop at a relative residual of 1e-10 differ by ~1e-8 in omega whatever their
    # precision (where a Krylov run stops depends on its initial guess); at 1e-12 the comparison
    # measures the mixed-precision error.
    kw = dict(K=5, problem_type=2, Re=400.0, dt=5e-4, time_integration_type=4, advection_type=1, tol=1e-12,
              maxiter=500)
    out = {}
    for prec in (hoc.HOC_FP64, hoc.HOC_MIXED):
        g = hoc.HO2D(precision=prec, **kw)
        g.load_mesh(bfs_full)
        g.setup()
        g.set_time(20.0 / 3.0)         # inflow ramp 2|sin(pi t/40)| = 1: the full parabolic inflow
        g.step(20)
        out[prec] = (g.get_vorticity(), g.get_streamfunction(), g.solution_areas())
    (w64, p64, a), (w32, p32, _) = out[hoc.HOC_FP64], out[hoc.HOC_MIXED]
    ew, ep = rel_l2(w32, w64), rel_l2(p32, p64)
    print("bfs full, 20 steps: mixed vs fp64 omega", ew, "psi", ep)
    assert np.all(np.isfinite(w64)) and ew < 1e-8 and ep < 1e-8


def test_cavity32_K4_100_steps(gpu_available):
    m = meshes.cavity(32, 2)
    kw = dict(problem_type=1, Re=100.0, dt=5e-4, time_integration_type=4, advection_type=1, tol=1e-13, maxiter=500)
    o = Oracle(m, 4, **kw)
    g = hoc.HO2D(K=4, **kw)
    g.load_mesh(m)
    g.setup()
    prev = pyoracle.set_threads(oracle_threads())
    try:
        o.step(100)
    finally:
        pyoracle.set_threads(prev)
    g.step(100)
    ew = rel_l2(g.get_vorticity(), o.vorticity())
    ep = rel_l2(g.get_streamfunction(), o.streamfunction())
    print("cavity 32x32 K=4, 100 steps: omega", ew, "psi", ep, "t", g.time)
    assert abs(g.time - 0.05) < 1e-12
    assert ew < 1e-11 and ep < 1e-11


def test_cavity32_K4_2000_steps_vs_oracle(gpu_available):
    m = meshes.cavity(32, 2)
    g = hoc.HO2D(K=4, problem_type=1, Re=100.0, dt=5e-4, time_integration_type=4, advection_type=1, tol=1e-13,
                 maxiter=500)
    g.load_mesh(m)
    g.setup()
    g.step(2000)
    assert abs(g.time - 1.0) < 1e-12
    fixture_check(g, "cavity32_K4", 1e-11)
