"""Discrete adjoint of the RK4 trajectory (hoc_adjoint_forward/backward, hoc_stage_residual_vjp)
against central finite differences: of the stage map omega -> k (quadratic in omega once the
upwind switches are fixed, so a central difference has no truncation error) and of the final
enstrophy after a run of steps, with respect to the initial vorticity and the wall velocity;
the directional derivative also against central differences of the FP64 CPU oracle."""
import numpy as np
import pytest

import hoc
from hoc import meshes
from pyoracle import Oracle

pytestmark = pytest.mark.gpu


def handle(n=6, K=4, Re=1000.0, dt=2e-3, tol=1e-14):
    m = meshes.cavity(n, 2)
    g = hoc.HO2D(K=K, problem_type=1, Re=Re, dt=dt, time_integration_type=4, advection_type=1, tol=tol,
                 maxiter=800)
    g.load_mesh(m)
    g.setup()
    return m, g


def rough_field(g, seed, amp=0.1):
    """Vorticity with O(amp) jumps across every face (so no flux point sits near the 1e-4
    upwind threshold)."""
    rng = np.random.default_rng(seed)
    xy = g.solution_points()
    return amp * (np.sin(3 * xy[:, 0] + 1) * np.cos(2 * xy[:, 1]) + rng.uniform(-1, 1, g.D))


def test_stage_vjp_dot_product(gpu_available):
    m, g = handle()
    rng = np.random.default_rng(1)
    w = rough_field(g, 2)
    lam = rng.standard_normal(g.D)
    gw, gp = g.stage_vjp(w, lam)
    eps = 1e-5
    for trial in range(3):
        v = rng.standard_normal(g.D)
        fd = (g.stage(w + eps * v) - g.stage(w - eps * v)) / (2 * eps)
        lhs, rhs = gw @ v, lam @ fd
        print("omega direction", trial, lhs, rhs, abs(lhs - rhs) / abs(rhs))
        assert abs(lhs - rhs) <= 1e-7 * abs(rhs)
    # the wall velocity enters linearly
    pv0 = g.wall_velocity()
    assert np.count_nonzero(pv0) > 0
    vp = rng.standard_normal(pv0.size)
    g.set_wall_velocity(pv0 + vp)
    kp = g.stage(w)
    g.set_wall_velocity(pv0 - vp)
    km = g.stage(w)
    g.set_wall_velocity(pv0)
    lhs, rhs = gp @ vp, lam @ (kp - km) / 2
    print("wall velocity direction", lhs, rhs)
    assert abs(lhs - rhs) <= 1e-9 * abs(rhs)


def enstrophy_run(g, w0, pvel, nsteps, every=None):
    a = g.solution_areas()
    g.set_wall_velocity(pvel)
    g.set_vorticity(w0)
    lib = hoc.lib()
    g._chk(lib.hoc_set_time(g.h, 0.0))
    if every is not None:
        g.adjoint_forward(nsteps, every)
    else:
        g.step(nsteps)
    wN = g.get_vorticity()
    return 0.5 * np.sum(a * wN * wN), a * wN


@pytest.mark.parametrize("nsteps,every", [(7, 3), (4, 10)])
def test_trajectory_gradient_vs_finite_differences(gpu_available, nsteps, every):
    m, g = handle(n=6, K=4, dt=5e-3)
    rng = np.random.default_rng(7)
    w0 = rough_field(g, 3)
    pv0 = g.wall_velocity()
    J, wbar = enstrophy_run(g, w0, pv0, nsteps, every)
    gw, gp = g.adjoint_backward(wbar)
    # the handle is left at the end of the forward run
    assert np.allclose(g.get_vorticity(), wbar / g.solution_areas(), rtol=1e-13, atol=1e-15)
    eps = 1e-6
    for trial in range(2):
        d = rng.standard_normal(g.D)
        Jp, _ = enstrophy_run(g, w0 + eps * d, pv0, nsteps)
        Jm, _ = enstrophy_run(g, w0 - eps * d, pv0, nsteps)
        fd, ad = (Jp - Jm) / (2 * eps), gw @ d
        print("omega0 direction", trial, ad, fd, abs(ad - fd) / abs(fd))
        assert abs(ad - fd) <= 1e-6 * abs(fd)
    lid = pv0 != 0
    dp = np.where(lid, rng.standard_normal(pv0.size), 0.)
    Jp, _ = enstrophy_run(g, w0, pv0 + eps * dp, nsteps)
    Jm, _ = enstrophy_run(g, w0, pv0 - eps * dp, nsteps)
    fd, ad = (Jp - Jm) / (2 * eps), gp @ dp
    print("wall velocity direction", ad, fd, abs(ad - fd) / abs(fd))
    assert abs(ad - fd) <= 1e-6 * abs(fd)


def test_gradient_vs_cpu_oracle_tangent(gpu_available):
    """The directional derivative of the GPU adjoint against central differences of the FP64
    CPU oracle's own trajectory (uniform lid)."""
    n, K, Re, dt, nsteps = 5, 4, 400.0, 4e-3, 5
    m, g = handle(n=n, K=K, Re=Re, dt=dt)
    # jumps of O(1) across the faces keep every flux point far from the 1e-4 upwind threshold
    # for these steps (a crossing between the two runs of a difference quotient adds a jump of
    # J over 2 eps to it; the oracle's quotients at 1e-6 and 1e-7 agree to 3e-8 for this field)
    w0 = rough_field(g, 9, amp=1.0)
    J, wbar = enstrophy_run(g, w0, g.wall_velocity(), nsteps, 2)
    gw, _ = g.adjoint_backward(wbar)
    a = g.solution_areas()
    d = np.random.default_rng(4).standard_normal(g.D)

    def J_cpu(w):
        o = Oracle(m, K, problem_type=1, Re=Re, dt=dt, time_integration_type=4, advection_type=1, tol=1e-14,
                   maxiter=800)
        o.set_vorticity(w)
        o.step(nsteps)
        wN = o.vorticity()
        return 0.5 * np.sum(a * wN * wN)

    eps = 1e-6
    fd = (J_cpu(w0 + eps * d) - J_cpu(w0 - eps * d)) / (2 * eps)
    ad = gw @ d
    pv0 = g.wall_velocity()
    fdg = (enstrophy_run(g, w0 + eps * d, pv0, nsteps)[0] - enstrophy_run(g, w0 - eps * d, pv0, nsteps)[0]) / (2 * eps)
    print("GPU adjoint", ad, "CPU tangent", fd, abs(ad - fd) / abs(fd), "GPU central difference", fdg,
          "J gpu", J, "cpu", J_cpu(w0))
    assert abs(ad - fd) <= 1e-6 * abs(fd)
    assert abs(fdg - fd) <= 1e-7 * abs(fd)


def test_boundary_points_and_wall_velocity_match_oracle(gpu_available):
    m = meshes.cavity(4, 2, seed=2, rotate=True)
    g = hoc.HO2D(K=4, problem_type=1, Re=100.0, dt=1e-3)
    g.load_mesh(m)
    g.setup()
    o = Oracle(m, 4, problem_type=1, Re=100.0, dt=1e-3)
    assert np.allclose(g.boundary_points().ravel(), o.table("edge_xy"), rtol=0, atol=1e-14)
    assert np.array_equal(g.wall_velocity(), o.table("BC_parl_vel"))


def test_adjoint_errors(gpu_available):
    m, g = handle(n=3, K=3)
    with pytest.raises(hoc.HocError) as e:
        g.adjoint_backward(np.zeros(g.D))
    assert e.value.code == hoc.HOC_ESTATE
    with pytest.raises(hoc.HocError) as e:
        g.adjoint_forward(-1, 1)
    assert e.value.code == hoc.HOC_EINVAL
    with pytest.raises(hoc.HocError) as e:
        g.set_wall_velocity(np.zeros(3))
    assert e.value.code == hoc.HOC_EINVAL
    g.adjoint_forward(0, 1)
    gw, gp = g.adjoint_backward(np.arange(g.D, dtype=float))
    assert np.array_equal(gw, np.arange(g.D)) and not gp.any()
    gm = hoc.HO2D(K=3, problem_type=1, Re=100.0, dt=1e-3, precision=hoc.HOC_MIXED)
    gm.load_mesh(m)
    gm.setup()
    with pytest.raises(hoc.HocError) as e:
        gm.adjoint_forward(1, 1)
    assert e.value.code == hoc.HOC_EINVAL
    c = meshes.backward_facing_step(8, 4, 2)
    gb = hoc.HO2D(K=3, problem_type=2, Re=100.0, dt=1e-3)
    gb.load_mesh(c)
    gb.setup()
    with pytest.raises(hoc.HocError) as e:
        gb.stage_vjp(np.zeros(gb.D), np.zeros(gb.D))
    assert e.value.code == hoc.HOC_EINVAL


def test_torch_harness(gpu_available):
    torch = pytest.importorskip("torch")
    from hoc.gradient import CavityEnstrophy, enstrophy, gaussian_field
    prob = CavityEnstrophy(n=6, K=4, Re=400.0, dt=4e-3, nsteps=4, every=3, tol=1e-14, n_modes=3)
    w0 = rough_field(prob.g, 9, amp=1.0) + gaussian_field(prob.g.solution_points(), seed=1)
    a0 = np.array([0.1, -0.05, 0.02])
    wt = torch.tensor(w0, requires_grad=True)
    at = torch.tensor(a0, requires_grad=True)
    J = enstrophy(wt, at, prob)
    (2.0 * J).backward()
    gw, ga = wt.grad.numpy() / 2.0, at.grad.numpy() / 2.0
    assert abs(float(J) - prob.value(w0, a0)) <= 1e-13 * abs(float(J))
    eps = 1e-6
    for mo in range(3):
        e = np.zeros(3)
        e[mo] = eps
        fd = (prob.value(w0, a0 + e) - prob.value(w0, a0 - e)) / (2 * eps)
        print("lid mode", mo + 1, ga[mo], fd)
        assert abs(ga[mo] - fd) <= 1e-6 * abs(fd)
    d = np.random.default_rng(3).standard_normal(prob.g.D)
    fd = (prob.value(w0 + eps * d, a0) - prob.value(w0 - eps * d, a0)) / (2 * eps)
    print("omega0 direction", gw @ d, fd)
    assert abs(gw @ d - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("every", [0, -1, -2, -3])
def test_checkpoint_schedules_agree(gpu_available, every):
    """The same trajectory gradient from uniform segments (every 3), the recorded trajectory
    (every = 0: every step's vorticity and stage solutions kept by the forward run, no
    recompute) and binomial checkpointing with 1, 2 or 3 snapshots (every < 0).  The schedules
    differ only in which Poisson solves are repeated (each to the tolerance, from its own initial
    guess), so the gradients agree to the solve tolerance."""
    m, g = handle(n=6, K=4, dt=5e-3)
    w0 = rough_field(g, 3)
    pv0 = g.wall_velocity()
    nsteps = 9
    J, wbar = enstrophy_run(g, w0, pv0, nsteps, 3)
    gw_ref, gp_ref = g.adjoint_backward(wbar)
    J2, wbar2 = enstrophy_run(g, w0, pv0, nsteps, every)
    assert abs(J2 - J) <= 1e-13 * abs(J)
    gw, gp = g.adjoint_backward(wbar2)
    ew = np.linalg.norm(gw - gw_ref) / np.linalg.norm(gw_ref)
    ep = np.linalg.norm(gp - gp_ref) / np.linalg.norm(gp_ref)
    print("schedule", every, "gradient vs uniform segments: omega", ew, "wall", ep)
    assert ew < 1e-10 and ep < 1e-10
    # the handle is left at the end of the forward run
    assert np.allclose(g.get_vorticity(), wbar2 / g.solution_areas(), rtol=1e-13, atol=1e-15)
