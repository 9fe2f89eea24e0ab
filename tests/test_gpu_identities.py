"""The oracle's identity pins replayed on the HIP path, and the named workloads at their sizes.

Circulation identity (the wall-vorticity flux of src/HO_2D.cpp:2647-2664 with the
velocity_jump of :3249-3253, the advective residual's conservation and the Poisson solve
together): inside no-slip walls every stage k = dt (adv + diff) carries circulation
sum w J k = Gamma_wall - Gamma(w), whatever dt, Re, K or mesh, where Gamma_wall is the wall
velocity's circulation (-1 for the unit lid at speed 1).  So one Euler step gives Gamma_wall,
and RK4 shrinks the excess Gamma - Gamma_wall by p(-1) = 3/8 per step.  These hold at any size,
the 512 x 512, K = 5 bench workload included.

Named sizes against the oracle: the backward-facing step 384 x 64 (K=5, Re=400), 10 RK4 steps
with the inflow at full strength, against the oracle's fingerprints
(tests/golden/make_flow_fixtures.py), and the 96 x 96, K=5 cavity of the bench family (the
largest the oracle's solver converges on), 2 RK4 steps, against the oracle run in the test."""
import os

import numpy as np
import pytest

import hoc
from hoc import meshes
from conftest import fixture_check, rel_l2

pytestmark = pytest.mark.gpu



def handle(m, K, **kw):
    g = hoc.HO2D(K=K, **kw)
    g.load_mesh(m)
    g.setup()
    return g


def circ(g, w):
    return float(np.sum(g.solution_areas() * w))


IDENT_MESHES = [
    ("affine_rot", lambda: meshes.cavity(8, 2, seed=1, rotate=True, flip_bedges=True), 4),
    ("jittered", lambda: meshes.cavity(7, 2, seed=2, rotate=True, jitter=0.2), 5),
    ("curved_q", lambda: meshes.cavity(6, 3, seed=3, rotate=True, curve=0.05), 4),
]


@pytest.mark.parametrize("adv", [1, 2])
@pytest.mark.parametrize("name,mk,K", IDENT_MESHES)
def test_stage_circulation_identity(gpu_available, name, mk, K, adv):
    g = handle(mk(), K, problem_type=1, Re=700.0, dt=3e-4, time_integration_type=4, advection_type=adv,
               tol=1e-14, maxiter=600)
    xy = g.solution_points()
    rng = np.random.default_rng(7)
    w = np.cos(3 * xy[:, 0]) * np.sin(5 * xy[:, 1] + 0.5) * 4 + 0.5 * rng.standard_normal(len(xy))
    k = g.stage(w, 0.0)
    got, want = circ(g, k), -1.0 - circ(g, w)
    print(name, "adv", adv, "circulation of k", got, "expected", want)
    assert abs(got - want) < 1e-11 * max(1.0, abs(want))


@pytest.mark.parametrize("integ", [1, 4])
def test_step_circulation_identity(gpu_available, integ):
    """Euler sets the wall circulation in one step from any field; RK4 relaxes the excess by 3/8
    per step.  The lid carries a non-uniform profile u(x) = 16 x^2 (1-x)^2 (wall circulation
    -8/15), set through the wall velocity (BC_parl_vel = -u along the lid)."""
    m = meshes.cavity(10, 2, seed=4, rotate=True, jitter=0.1)
    g = handle(m, 4, problem_type=1, Re=300.0, dt=2e-4, time_integration_type=integ, advection_type=1,
               tol=1e-14, maxiter=600)
    bp = g.boundary_points()
    v = g.wall_velocity()
    lid = v == -1.0
    assert lid.sum() == 10 * 4
    v[lid] = -16 * bp[lid, 0] ** 2 * (1 - bp[lid, 0]) ** 2
    g.set_wall_velocity(v)
    xy = g.solution_points()
    g.set_vorticity(np.sin(4 * xy[:, 0]) * np.cos(3 * xy[:, 1]) * 2.0)
    target = -8.0 / 15.0
    e = circ(g, g.get_vorticity()) - target
    for s in range(3):
        g.step(1)
        e *= 0.0 if integ == 1 else 0.375
        got = circ(g, g.get_vorticity())
        print("integ", integ, "step", s, "circulation", got, "expected", target + e)
        assert abs(got - (target + e)) < 1e-11


@pytest.mark.parametrize("precision", [hoc.HOC_FP64, hoc.HOC_MIXED])
def test_bench_workload_circulation(gpu_available, precision):
    """The 512 x 512, K = 5 cavity of the bench (dt = half the RK4 diffusion limit): from rest,
    Gamma_n = -1 + (3/8)^n after n RK4 steps, to the Poisson tolerance."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from bench import cavity_dt
    n, K = 512, 5
    g = handle(meshes.cavity(n, 2), K, problem_type=1, Re=100.0, dt=cavity_dt(n, K, 100.0),
               time_integration_type=4, advection_type=1, tol=1e-12, maxiter=600, precision=precision)
    e = 1.0
    for s in range(3):
        g.step(1)
        e *= 0.375
        got = circ(g, g.get_vorticity())
        print("512x512 K=5 precision", precision, "step", s, "circulation", got, "expected", -1.0 + e)
        assert abs(got - (-1.0 + e)) < 1e-9


def test_bfs_full_10_steps_vs_oracle(gpu_available):
    """Backward-facing step 384 x 64 (22 528 elements, jitter 0.1, seed 2), K=5, Re=400,
    dt=5e-4: 10 RK4 steps from t = 20/3 (full inflow) against the FP64 oracle's run."""
    m = meshes.backward_facing_step(384, 64, 2, jitter=0.1, seed=2)
    g = handle(m, 5, problem_type=2, Re=400.0, dt=5e-4, time_integration_type=4, advection_type=1, tol=1e-14,
               maxiter=600)
    g.set_time(20.0 / 3.0)
    g.step(10)
    fixture_check(g, "bfs384_K5", 1e-11)


def test_cavity96_K5_vs_oracle(gpu_available):
    """The bench family at the largest size the oracle's solver reaches its tolerance within 300
    iterations (its AMGCL restatement takes 26 per solve at 64^2 and 96^2; at 128^2, K=5 it
    stalls): 96 x 96, K=5, Re=100, dt = half the RK4 limit, 2 RK4 steps from rest, tol 1e-12.
    Every stage's Poisson solve, velocity, wall-vorticity flux, advection and diffusion at the
    bench's K, Re and dt rule, against the oracle run in the test on the host's cores."""
    import sys
    import pyoracle
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from bench import cavity_dt
    n, K = 96, 5
    kw = dict(problem_type=1, Re=100.0, dt=cavity_dt(n, K, 100.0), time_integration_type=4, advection_type=1,
              tol=1e-12, maxiter=300)
    m = meshes.cavity(n, 2)
    prev = pyoracle.set_threads(16)
    try:
        o = pyoracle.Oracle(m, K, **kw)
        o.step(2)
        its = o.solve_counts()
    finally:
        pyoracle.set_threads(prev)
    for prec in (hoc.HOC_FP64, hoc.HOC_MIXED):
        g = handle(m, K, precision=prec, **kw)
        g.step(2)
        ew = rel_l2(g.get_vorticity(), o.vorticity())
        ep = rel_l2(g.get_streamfunction(), o.streamfunction())
        print("cavity 96x96 K=5 precision", prec, "omega", ew, "psi", ep, "gpu its", g.stats()["total_iters"],
              "oracle solves/its", its)
        assert ew < 1e-10 and ep < 1e-10
        g.close()
