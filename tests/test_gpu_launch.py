"""Krylov iterations replayed from a captured hipGraph of iteration chunks (launch_mode 2)
against the same iterations queued kernel by kernel (launch_mode 1): the kernels and their
order are the same, so the fields must agree bit for bit after several RK4 steps, with the
same iteration counts, for PCG (affine cavity), BiCGSTAB (jittered step, non-symmetric
operator) and the mixed-precision path (FP32 V-cycle inside the FP64 Krylov method); each
replay runs between the program's own eager kernels and synchronisations (the stage
residuals, the guesses, the host waits).  launch_mode 0 times both at setup and reports the
ratio; its choice changes no result."""
import numpy as np
import pytest

import hoc
from hoc import meshes

pytestmark = pytest.mark.gpu

CASES = [
    ("cavity12_K5_pcg", lambda: meshes.cavity(12, 2, seed=3, rotate=True), 5, dict(problem_type=1, Re=300.0, dt=2e-4),
     hoc.HOC_FP64, hoc.KRYLOV_PCG),
    ("cavity12_K5_mixed", lambda: meshes.cavity(12, 2, seed=3, rotate=True), 5,
     dict(problem_type=1, Re=300.0, dt=2e-4), hoc.HOC_MIXED, hoc.KRYLOV_PCG),
    ("bfs_jit_K4_bicgstab", lambda: meshes.backward_facing_step(24, 8, 2, jitter=0.08, seed=2), 4,
     dict(problem_type=2, Re=400.0, dt=5e-4), hoc.HOC_FP64, hoc.KRYLOV_BICGSTAB),
]


def run(m, K, kw, precision, mode, steps=4):
    g = hoc.HO2D(K=K, time_integration_type=4, advection_type=1, tol=1e-12, maxiter=400, precision=precision,
                 launch_mode=mode, **kw)
    g.load_mesh(m)
    g.setup()
    if kw["problem_type"] == 2:
        g.set_time(20.0 / 3.0)
    g.step(steps)
    out = (g.get_vorticity(), g.get_streamfunction(), g.stats())
    g.close()
    return out


@pytest.mark.parametrize("name,mk,K,kw,precision,krylov", CASES)
def test_graph_replay_equals_direct_launches(gpu_available, name, mk, K, kw, precision, krylov):
    m = mk()
    w1, p1, s1 = run(m, K, kw, precision, 1)
    w2, p2, s2 = run(m, K, kw, precision, 2)
    print(name, "iterations direct", s1["total_iters"], "graph", s2["total_iters"], "graphs", s1["launch_graphs"],
          s2["launch_graphs"])
    assert s1["krylov"] == krylov
    assert s1["launch_graphs"] == 0 and s2["launch_graphs"] == 1
    assert s1["total_iters"] == s2["total_iters"] and s1["n_solves"] == s2["n_solves"]
    assert np.array_equal(w1, w2) and np.array_equal(p1, p2)


def test_launch_mode_autotune_reports(gpu_available):
    m = meshes.cavity(12, 2)
    w0, p0, s0 = run(m, 5, dict(problem_type=1, Re=100.0, dt=2e-4), hoc.HOC_FP64, 0, steps=2)
    w1, p1, s1 = run(m, 5, dict(problem_type=1, Re=100.0, dt=2e-4), hoc.HOC_FP64, 1, steps=2)
    print("12x12 K=5: graph/direct time at setup", s0["launch_tune_ratio"], "graphs chosen", s0["launch_graphs"])
    assert s0["launch_tune_ratio"] > 0 and s1["launch_tune_ratio"] == 0
    assert s0["launch_graphs"] == (1 if s0["launch_tune_ratio"] < 0.85 else 0)
    assert np.array_equal(w0, w1) and np.array_equal(p0, p1)
    with pytest.raises(hoc.HocError) as e:
        hoc.HO2D(K=3, launch_mode=3)
    assert e.value.code == hoc.HOC_EINVAL
