"""Ensembles: independent meshes with their own Re and dt advanced by the same launches
(one handle), checked instance by instance against separate FP64 oracle runs."""
import numpy as np
import pytest

import hoc
from hoc import meshes
from pyoracle import Oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def run_ensemble(parts, K, Re, dt, steps, problem_type=1, precision=hoc.HOC_FP64, tol=1e-13, t0=0.0):
    m, off = meshes.concat(parts)
    g = hoc.HO2D(K=K, problem_type=problem_type, Re=Re[0], dt=dt[0], time_integration_type=4, advection_type=1,
                 tol=tol, maxiter=800, precision=precision)
    g.load_mesh(m)
    g.set_instances(off, Re, dt)
    g.setup()
    if t0:
        g.set_time(t0)
    g.step(steps)
    return g, off


def test_uneven_cavity_ensemble_vs_oracle(gpu_available):
    parts = [meshes.cavity(4, 2), meshes.cavity(6, 2, seed=3, rotate=True), meshes.cavity(5, 2, seed=5, jitter=0.1)]
    K = 4
    Re = np.array([100.0, 400.0, 1600.0])
    dt = np.array([2e-4, 1e-4, 1.5e-4])
    steps = 5
    g, off = run_ensemble(parts, K, Re, dt, steps)
    w, p = g.get_vorticity(), g.get_streamfunction()
    assert np.allclose(g.instance_times(), steps * dt, rtol=1e-14, atol=0)
    for i, part in enumerate(parts):
        o = Oracle(part, K, problem_type=1, Re=Re[i], dt=dt[i], time_integration_type=4, advection_type=1,
                   tol=1e-13, maxiter=800)
        o.step(steps)
        sl = slice(off[i] * K * K, off[i + 1] * K * K)
        ew, ep = rel_l2(w[sl], o.vorticity()), rel_l2(p[sl], o.streamfunction())
        print("instance", i, "omega", ew, "psi", ep)
        assert ew < 1e-10 and ep < 1e-10


def test_corner_coupled_ensemble_vs_oracle(gpu_available):
    """Two cavities that share one corner vertex (the second translated by (1, 1), its corner
    node merged with the first's): the Q1 coarse space couples them, so the instances run one
    joint PCG with the per-instance stopping test instead of the batched per-instance PCG.
    Each instance against its own FP64 oracle run."""
    a = meshes.cavity(6, 2)
    b0 = meshes.cavity(8, 2)
    b = meshes.QuadMesh(b0.lnod, b0.xy + 1.0, b0.elem, b0.bedges, b0.bedge_bdry, list(b0.bnames))
    m, off = meshes.concat([a, b])
    ia = int(np.flatnonzero(np.all(np.abs(a.xy - 1.0) < 1e-14, axis=1))[0])
    ib = int(np.flatnonzero(np.all(np.abs(b0.xy) < 1e-14, axis=1))[0]) + a.n_nodes
    m.elem[m.elem == ib] = ia
    m.bedges[m.bedges == ib] = ia
    assert np.array_equal(m.xy[ia], m.xy[ib])
    K = 4
    Re = np.array([100.0, 400.0])
    dt = np.array([2e-4, 1e-4])
    steps = 3
    g = hoc.HO2D(K=K, problem_type=1, Re=Re[0], dt=dt[0], time_integration_type=4, advection_type=1,
                 tol=1e-13, maxiter=800)
    g.load_mesh(m)
    g.set_instances(off, Re, dt)
    g.setup()
    g.step(steps)
    w, p = g.get_vorticity(), g.get_streamfunction()
    for i, part in enumerate([a, b]):
        o = Oracle(part, K, problem_type=1, Re=Re[i], dt=dt[i], time_integration_type=4, advection_type=1,
                   tol=1e-13, maxiter=800)
        o.step(steps)
        sl = slice(off[i] * K * K, off[i + 1] * K * K)
        ew, ep = rel_l2(w[sl], o.vorticity()), rel_l2(p[sl], o.streamfunction())
        print("corner-coupled instance", i, "omega", ew, "psi", ep)
        assert ew < 1e-10 and ep < 1e-10


def test_ensemble_with_ramped_channels(gpu_available):
    """Backward-facing steps with time-dependent inflow: each instance ramps on its own clock,
    from t = 20/3 where the ramp 2|sin(pi t/40)| (src/HO_2D.cpp:2014) is 1."""
    parts = [meshes.backward_facing_step(16, 6, 2, jitter=0.05, seed=1),
             meshes.backward_facing_step(20, 8, 2, jitter=0.05, seed=2)]
    K = 3
    Re = np.array([200.0, 400.0])
    dt = np.array([5e-3, 2e-3])
    steps = 3
    t0 = 20.0 / 3.0
    g, off = run_ensemble(parts, K, Re, dt, steps, problem_type=2, t0=t0)
    assert np.allclose(g.instance_times(), t0 + steps * dt, rtol=1e-14, atol=0)
    w = g.get_vorticity()
    for i, part in enumerate(parts):
        o = Oracle(part, K, problem_type=2, Re=Re[i], dt=dt[i], time_integration_type=4, advection_type=1,
                   tol=1e-13, maxiter=800)
        o.set_time(t0)
        o.step(steps)
        sl = slice(off[i] * K * K, off[i + 1] * K * K)
        ew = rel_l2(w[sl], o.vorticity())
        print("channel", i, "omega", ew)
        assert ew < 1e-10


def test_mixed_ensemble_matches_fp64(gpu_available):
    parts = [meshes.cavity(8, 2), meshes.cavity(12, 2), meshes.cavity(10, 2, seed=1, rotate=True)]
    Re = np.array([100.0, 800.0, 3200.0])
    dt = np.array([5e-5, 2e-5, 3e-5])
    g64, _ = run_ensemble(parts, 5, Re, dt, 3, tol=1e-10)
    g32, _ = run_ensemble(parts, 5, Re, dt, 3, tol=1e-10, precision=hoc.HOC_MIXED)
    ew = rel_l2(g32.get_vorticity(), g64.get_vorticity())
    ep = rel_l2(g32.get_streamfunction(), g64.get_streamfunction())
    print("mixed ensemble omega", ew, "psi", ep)
    assert ew < 1e-8 and ep < 1e-8


def test_ensemble_errors(gpu_available):
    m, off = meshes.concat([meshes.cavity(3, 2), meshes.cavity(3, 2)])
    g = hoc.HO2D(K=3, dt=1e-4)
    with pytest.raises(hoc.HocError) as e:
        g.set_instances(off, [100.0, 100.0], [1e-4, 1e-4])      # before a mesh
    assert e.value.code == hoc.HOC_ESTATE
    g.load_mesh(m)
    with pytest.raises(hoc.HocError) as e:
        g.set_instances(np.array([0, 5, 17]), [100.0, 100.0], [1e-4, 1e-4])
    assert e.value.code == hoc.HOC_EINVAL
    one = meshes.cavity(4, 2)                                    # one connected mesh split in two
    g2 = hoc.HO2D(K=3, dt=1e-4)
    g2.load_mesh(one)
    g2.set_instances(np.array([0, 8, 16]), [100.0, 100.0], [1e-4, 1e-4])
    with pytest.raises(hoc.HocError) as e:
        g2.setup()
    assert e.value.code == hoc.HOC_EINVAL
    g.set_instances(off, [100.0, 200.0], [1e-4, 2e-4])
    g.setup()
    with pytest.raises(hoc.HocError) as e:
        g.solveto(1e-3, 2, 4, 100.0)
    assert e.value.code == hoc.HOC_EINVAL


def test_64_instance_ensemble_full_size(gpu_available):
    """The batched-ensemble workload at its size: 64 cavity instances with seeded sizes between
    16^2 and 256^2 elements (uneven, 1.5 M elements in all), K=5, Re log-uniform in [100, 3200],
    lid speed 1, per-instance dt (half the RK4 diffusion limit, capped by an advective CFL of
    0.05 per element width).  A few RK4 steps in FP64 and in mixed precision: every instance's
    mixed fields within 1e-8 of its FP64 fields, and the two smallest instances against
    separate FP64 oracle runs."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from bench import cavity_dt
    rng = np.random.default_rng(2024)
    ns = rng.integers(16, 257, size=64)
    Re = np.exp(rng.uniform(np.log(100.0), np.log(3200.0), size=64))
    K = 5
    dt = np.array([min(cavity_dt(int(n), K, r), 0.05 / n) for n, r in zip(ns, Re)])
    parts = [meshes.cavity(int(n), 2) for n in ns]
    steps = 3
    g64, off = run_ensemble(parts, K, Re, dt, steps, tol=1e-10)
    g32, _ = run_ensemble(parts, K, Re, dt, steps, tol=1e-10, precision=hoc.HOC_MIXED)
    assert off[-1] == int(np.sum(ns * ns))
    w64, p64 = g64.get_vorticity(), g64.get_streamfunction()
    w32, p32 = g32.get_vorticity(), g32.get_streamfunction()
    worst = 0.0
    for i in range(64):
        sl = slice(off[i] * K * K, off[i + 1] * K * K)
        worst = max(worst, rel_l2(w32[sl], w64[sl]), rel_l2(p32[sl], p64[sl]))
    print("64 instances,", off[-1], "elements: worst per-instance mixed-vs-FP64 error", worst)
    assert worst < 1e-8
    for i in np.argsort(ns)[:2]:
        o = Oracle(parts[i], K, problem_type=1, Re=Re[i], dt=dt[i], time_integration_type=4, advection_type=1,
                   tol=1e-12, maxiter=800)
        o.step(steps)
        sl = slice(off[i] * K * K, off[i + 1] * K * K)
        ew = rel_l2(w64[sl], o.vorticity())
        print("instance", i, "n", ns[i], "omega vs oracle", ew)
        assert ew < 1e-8
