"""Mixed precision (HOC_MIXED): the FP64 Krylov method (PCG or BiCGSTAB) with the FP32
V-cycle as its preconditioner.  The Krylov vectors, the operator application, the residual
and the stopping test are FP64, so fields must match the FP64 oracle to the same tolerance
as the plain path; the
acceptance bar for mixed mode is a relative L2 field error <= 1e-8 (plain and
quadrature-weighted norms), asserted here at 1e-10 on the small cases."""
import numpy as np
import pytest

import hoc
from hoc import meshes
from pyoracle import Oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def qrel_l2(a, b, areas):
    """Relative L2 error in the quadrature norm sqrt(sum w_i w_j J e^2)."""
    d = np.sqrt(np.sum(areas * b * b))
    return float(np.sqrt(np.sum(areas * (a - b) ** 2)) / (d if d > 0 else 1.0))


def make_mixed(mesh, K, problem_type=1, Re=100.0, dt=1e-4, integ=4, adv=1, tol=1e-14, maxiter=500, **kw):
    g = hoc.HO2D(K=K, problem_type=problem_type, Re=Re, dt=dt, time_integration_type=integ, advection_type=adv,
                 tol=tol, maxiter=maxiter, precision=hoc.HOC_MIXED, **kw)
    g.load_mesh(mesh)
    g.setup()
    return g


CASES = [
    ("cavity6_K3", lambda: meshes.cavity(6, 2), 3),
    ("cavity5_rot_K4", lambda: meshes.cavity(5, 2, seed=3, rotate=True, flip_bedges=True), 4),
    ("cavity4_jit_K5", lambda: meshes.cavity(4, 2, seed=4, rotate=True, jitter=0.15), 5),
    ("cavity4_curved_L3_K4", lambda: meshes.cavity(4, 3, seed=5, rotate=True, curve=0.04), 4),
]


@pytest.mark.parametrize("name,mk,K", CASES)
def test_mixed_poisson(gpu_available, name, mk, K):
    m = mk()
    o = Oracle(m, K, problem_type=1, Re=100.0, dt=1e-4, time_integration_type=4, advection_type=1, tol=1e-14,
               maxiter=500)
    g = make_mixed(m, K)
    xy = o.table("sps_xy").reshape(-1, 2)
    w = np.sin(3 * xy[:, 0] + 0.3) * np.cos(2 * xy[:, 1]) * 5.0
    p_ref = o.poisson(w)
    p_got = g.poisson(w)
    st = g.stats()
    print(name, "rel", rel_l2(p_got, p_ref), "its", st["last_iters"], "refinements", st["last_refinements"],
          "res", st["last_rel_residual"], "krylov", st["krylov"])
    assert st["precision"] == hoc.HOC_MIXED
    assert st["last_refinements"] == 0          # no refinement passes: one FP64 Krylov solve
    assert st["last_rel_residual"] <= 1e-14
    assert rel_l2(p_got, p_ref) < 1e-11


def test_mixed_zero_rhs(gpu_available):
    """Zero vorticity and zero wall values: psi = 0 without iterating (as the FP64 path)."""
    m = meshes.cavity(4, 2)
    g = make_mixed(m, 3)
    p = g.poisson(np.zeros(g.D))
    assert np.all(p == 0.0)
    assert g.stats()["last_iters"] == 0


@pytest.mark.parametrize("integ", [4, 2])
def test_mixed_cavity_steps(gpu_available, integ):
    m = meshes.cavity(8, 2, seed=7, rotate=True)
    o = Oracle(m, 4, problem_type=1, Re=100.0, dt=2e-4, time_integration_type=integ, advection_type=1, tol=1e-12,
               maxiter=500)
    g = make_mixed(m, 4, dt=2e-4, integ=integ, tol=1e-12)
    o.step(10)
    g.step(10)
    a = g.solution_areas()
    ew = rel_l2(g.get_vorticity(), o.vorticity())
    ep = rel_l2(g.get_streamfunction(), o.streamfunction())
    qw = qrel_l2(g.get_vorticity(), o.vorticity(), a)
    qp = qrel_l2(g.get_streamfunction(), o.streamfunction(), a)
    print("integ", integ, "omega", ew, qw, "psi", ep, qp, "its", g.stats()["total_iters"])
    assert max(ew, ep, qw, qp) < 1e-10


def test_mixed_bfs_steps(gpu_available):
    """Jittered backward-facing step: non-symmetric W A -> FP64 BiCGSTAB with the FP32 V-cycle."""
    m = meshes.backward_facing_step(24, 8, 2, jitter=0.08, seed=2)
    o = Oracle(m, 3, problem_type=2, Re=400.0, dt=5e-4, time_integration_type=4, advection_type=1, tol=1e-12,
               maxiter=500)
    g = make_mixed(m, 3, problem_type=2, Re=400.0, dt=5e-4, tol=1e-12)
    o.step(4)
    g.step(4)
    ew = rel_l2(g.get_vorticity(), o.vorticity())
    ep = rel_l2(g.get_streamfunction(), o.streamfunction())
    print("bfs mixed omega", ew, "psi", ep, "krylov", g.stats()["krylov"], "its", g.stats()["total_iters"])
    assert g.stats()["krylov"] == hoc.KRYLOV_BICGSTAB
    assert ew < 1e-10 and ep < 1e-10


def test_mixed_matches_plain_gpu(gpu_available):
    """Mixed vs plain FP64 GPU on the same inputs, through the public step API."""
    m = meshes.cavity(16, 2)
    kw = dict(K=5, problem_type=1, Re=100.0, dt=1e-5, time_integration_type=4, advection_type=1, tol=1e-10,
              maxiter=500)
    g64 = hoc.HO2D(**kw)
    g64.load_mesh(m)
    g64.setup()
    g32 = make_mixed(m, 5, dt=1e-5, tol=1e-10)
    g64.step(5)
    g32.step(5)
    ew = rel_l2(g32.get_vorticity(), g64.get_vorticity())
    ep = rel_l2(g32.get_streamfunction(), g64.get_streamfunction())
    print("mixed vs plain omega", ew, "psi", ep, "its", g64.stats()["total_iters"], g32.stats()["total_iters"])
    assert ew < 1e-8 and ep < 1e-8


def test_mixed_inner_tol_validation(gpu_available):
    with pytest.raises(hoc.HocError) as e:
        hoc.HO2D(K=3, precision=hoc.HOC_MIXED, mixed_inner_tol=1.5)
    assert e.value.code == hoc.HOC_EINVAL
