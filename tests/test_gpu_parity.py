"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle.

Tolerances: operator-level results (Laplacian apply, stage residual with the same
stream function) agree to rounding (<=1e-12 relative L2); fields after Poisson
solves / time steps agree to <=1e-11 relative L2 with both solvers driven to a
relative residual of 1e-14."""
import numpy as np
import pytest

import hoc
from hoc import meshes
from pyoracle import Oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def make_pair(mesh, K, problem_type=1, Re=100.0, dt=1e-4, integ=4, adv=1, tol=1e-14, maxiter=500, krylov=0):
    o = Oracle(mesh, K, problem_type=problem_type, Re=Re, dt=dt, time_integration_type=integ,
               advection_type=adv, tol=tol, maxiter=maxiter)
    g = hoc.HO2D(K=K, problem_type=problem_type, Re=Re, dt=dt, time_integration_type=integ,
                 advection_type=adv, tol=tol, maxiter=maxiter, krylov=krylov)
    g.load_mesh(mesh)
    g.setup()
    return o, g


CASES = [
    ("cavity6_K3", lambda: meshes.cavity(6, 2), 3),
    ("cavity5_rot_K4", lambda: meshes.cavity(5, 2, seed=3, rotate=True, flip_bedges=True), 4),
    ("cavity4_jit_K5", lambda: meshes.cavity(4, 2, seed=4, rotate=True, jitter=0.15), 5),
    ("cavity4_curved_L3_K4", lambda: meshes.cavity(4, 3, seed=5, rotate=True, curve=0.04), 4),
]


@pytest.mark.parametrize("name,mk,K", CASES)
def test_laplacian_apply(gpu_available, name, mk, K):
    m = mk()
    o, g = make_pair(m, K)
    x = np.random.default_rng(1).standard_normal(o.D)
    ref = o.laplacian_mf(x)
    got = g.apply_laplacian(x)
    print(name, "rel", rel_l2(got, ref))
    assert rel_l2(got, ref) < 1e-12


@pytest.mark.parametrize("name,mk,K", CASES)
def test_poisson(gpu_available, name, mk, K):
    m = mk()
    o, g = make_pair(m, K)
    xy = o.table("sps_xy").reshape(-1, 2)
    w = np.sin(3 * xy[:, 0] + 0.3) * np.cos(2 * xy[:, 1]) * 5.0
    p_ref = o.poisson(w)
    p_got = g.poisson(w)
    print(name, "rel", rel_l2(p_got, p_ref), "oracle its", o.last_solve(), "gpu", g.stats()["last_iters"], g.stats()["last_rel_residual"])
    assert rel_l2(p_got, p_ref) < 1e-11


@pytest.mark.parametrize("name,mk,K", CASES)
def test_stage_residual(gpu_available, name, mk, K):
    m = mk()
    o, g = make_pair(m, K, dt=1e-4)
    xy = o.table("sps_xy").reshape(-1, 2)
    w = np.sin(3 * xy[:, 0] + 0.3) * np.cos(2 * xy[:, 1]) * 5.0
    k_ref = o.stage(w, 0.0)
    k_got = g.stage(w, 0.0)
    print(name, "rel", rel_l2(k_got, k_ref))
    assert rel_l2(k_got, k_ref) < 1e-10


def bfs_small():
    return meshes.backward_facing_step(24, 8, 2, jitter=0.08, seed=2)


@pytest.mark.parametrize("K,adv", [(3, 1), (4, 2)])
def test_bfs_poisson_stage(gpu_available, K, adv):
    """Backward-facing step: Neumann outlet (psi and omega), ramped parabolic inflow."""
    m = bfs_small()
    o, g = make_pair(m, K, problem_type=2, Re=400.0, dt=5e-4, adv=adv)
    x = np.random.default_rng(2).standard_normal(o.D)
    assert rel_l2(g.apply_laplacian(x), o.laplacian_mf(x)) < 1e-12
    xy = o.table("sps_xy").reshape(-1, 2)
    w = np.sin(0.5 * xy[:, 0]) * np.cos(2 * xy[:, 1])
    t = 3.0
    p_ref = o.poisson(w, t)
    p_got = g.poisson(w, t)
    print("bfs psi", rel_l2(p_got, p_ref), o.last_solve(), g.stats()["last_iters"], g.stats()["krylov"])
    assert rel_l2(p_got, p_ref) < 1e-11
    k_ref = o.stage(w, t)
    k_got = g.stage(w, t)
    print("bfs k", rel_l2(k_got, k_ref))
    assert rel_l2(k_got, k_ref) < 1e-10


# Time at which the inflow ramp 2|sin(pi t/40)| of problem types 2 and 3 (src/HO_2D.cpp:2014,
# 2062) is 1: multi-step runs of those problems start here, with the inflow at full strength.
T_FULL_INFLOW = 20.0 / 3.0


@pytest.mark.parametrize("adv", [1, 2])
def test_bfs_steps(gpu_available, adv):
    m = bfs_small()
    o, g = make_pair(m, 3, problem_type=2, Re=400.0, dt=5e-4, adv=adv)
    o.set_time(T_FULL_INFLOW)
    g.set_time(T_FULL_INFLOW)
    o.step(6)
    g.step(6)
    ew = rel_l2(g.get_vorticity(), o.vorticity())
    ep = rel_l2(g.get_streamfunction(), o.streamfunction())
    print("bfs steps omega", ew, "psi", ep)
    assert ew < 1e-11 and ep < 1e-11


@pytest.mark.parametrize("integ", [4, 2, 1])
def test_cavity_steps(gpu_available, integ):
    m = meshes.cavity(8, 2, seed=7, rotate=True)
    o, g = make_pair(m, 4, dt=2e-4, integ=integ)
    o.step(10)
    g.step(10)
    ew = rel_l2(g.get_vorticity(), o.vorticity())
    ep = rel_l2(g.get_streamfunction(), o.streamfunction())
    print("integ", integ, "omega", ew, "psi", ep)
    assert ew < 1e-11 and ep < 1e-11


# ------------------------------------------------------------------ shipped reference meshes
import os  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def shipped(case):
    z = np.load(os.path.join(GOLD, f"mesh_{case}.npz"))
    return meshes.QuadMesh(int(z["lnod"]), z["xy"], z["elem"], z["bedges"], z["bedge_bdry"], [str(s) for s in z["bnames"]])


# input/input_*.dat: K=3, RK4, advection 1; cavity Re 1e4 dt 1e-3; BFS Re 50 dt 5e-3; cylinder Re 20 dt 5e-4
@pytest.mark.parametrize("case,ptype,Re,dt", [("cavity", 1, 10000.0, 1e-3), ("bfs", 2, 50.0, 5e-3),
                                             ("cylinder", 3, 20.0, 5e-4)])
def test_shipped_mesh_steps(gpu_available, case, ptype, Re, dt):
    m = shipped(case)
    o, g = make_pair(m, 3, problem_type=ptype, Re=Re, dt=dt, tol=1e-14, maxiter=800)
    if ptype in (2, 3):
        o.set_time(T_FULL_INFLOW)
        g.set_time(T_FULL_INFLOW)
    o.step(4)
    g.step(4)
    ew = rel_l2(g.get_vorticity(), o.vorticity())
    ep = rel_l2(g.get_streamfunction(), o.streamfunction())
    ev = rel_l2(g.get_velocity(), o.velocity())
    print(case, "omega", ew, "psi", ep, "vel", ev, "gpu its", g.stats()["last_iters"], "oracle its", o.last_solve())
    assert ew < 1e-11 and ep < 1e-11 and ev < 1e-11


def test_solveto_matches_steps(gpu_available):
    """solveto_d (src/HO_2D.cpp:4194-4233): dt = outer_dt/nsub, then nsub steps."""
    m = meshes.cavity(6, 2, seed=2, rotate=True)
    o, g = make_pair(m, 3, dt=1.0)
    g.solveto(1e-3, 5, 4, 100.0)
    o2 = Oracle(m, 3, problem_type=1, Re=100.0, dt=2e-4, time_integration_type=4, advection_type=1, tol=1e-14,
                maxiter=500)
    o2.step(5)
    assert abs(g.time - 1e-3) < 1e-15
    assert rel_l2(g.get_vorticity(), o2.vorticity()) < 1e-11


def test_api_errors(gpu_available):
    g = hoc.HO2D(K=3)
    with pytest.raises(hoc.HocError) as e:
        g.step(1)                                   # before setup
    assert e.value.code == hoc.HOC_ESTATE
    m = meshes.cavity(3, 2)
    g.load_mesh(m)
    g.setup()
    with pytest.raises(hoc.HocError) as e:
        g.set_vorticity(np.zeros(5))                # wrong length
    assert e.value.code == hoc.HOC_EINVAL
    bad = meshes.cavity(3, 2)
    bad.bedges = bad.bedges[:-1]                    # a boundary side that is not listed
    bad.bedge_bdry = bad.bedge_bdry[:-1]
    g2 = hoc.HO2D(K=3)
    g2.load_mesh(bad)
    with pytest.raises(hoc.HocError) as e:
        g2.setup()
    assert e.value.code == hoc.HOC_EMESH
    g3 = hoc.HO2D(K=3, tol=1e-30, maxiter=2, fail_on_noconv=True)
    g3.load_mesh(m)
    g3.setup()
    xy = g3.solution_points()
    with pytest.raises(hoc.HocError) as e:
        g3.poisson(np.sin(xy[:, 0]))
    assert e.value.code == hoc.HOC_ENOCONV


def test_nan_detected(gpu_available):
    m = meshes.cavity(3, 2)
    g = hoc.HO2D(K=3, dt=1e-4)
    g.load_mesh(m)
    g.setup()
    w = np.zeros(g.D)
    w[5] = np.nan
    with pytest.raises(hoc.HocError) as e:
        g.poisson(w)
    assert e.value.code == hoc.HOC_ENAN
