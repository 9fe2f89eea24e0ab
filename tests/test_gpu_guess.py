"""Initial guesses of the stage Poisson solves (hoc_config.psi_guess): the reference's warm start
from the previous solution (1), linear extrapolation in the stage time (2), the RK4 stage
combination (3), that combination with the previous step's stage differences (4), and the
projection onto an As-orthonormal basis of earlier solutions (5, the default with the symmetric
operator of affine meshes).  Only the starting point of each solve differs, so the fields
agree to the solver tolerance, match the oracle, and the better guesses need fewer iterations."""
import numpy as np
import pytest

import hoc
from hoc import meshes
from pyoracle import Oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def run(m, K, guess, steps, tol, integ=4, **kw):
    g = hoc.HO2D(K=K, problem_type=1, Re=400.0, dt=2e-4, time_integration_type=integ, advection_type=1, tol=tol,
                 maxiter=500, psi_guess=guess, **kw)
    g.load_mesh(m)
    g.setup()
    g.step(steps)
    st = g.stats()
    return g.get_vorticity(), g.get_streamfunction(), st["total_iters"] / st["n_solves"]


def test_guess_modes_agree_and_save_iterations(gpu_available):
    m = meshes.cavity(16, 2, seed=1, rotate=True)  # affine elements: symmetric operator, PCG
    out = {gm: run(m, 4, gm, 12, 1e-13) for gm in (1, 2, 3, 4, 5)}
    o = Oracle(m, 4, problem_type=1, Re=400.0, dt=2e-4, time_integration_type=4, advection_type=1, tol=1e-13,
               maxiter=500)
    o.step(12)
    assert out[5][2] < out[4][2]  # the projection beats the fixed combinations
    for gm, (w, p, it) in out.items():
        ew, ep = rel_l2(w, o.vorticity()), rel_l2(p, o.streamfunction())
        print("psi_guess", gm, "iterations per solve", it, "omega vs oracle", ew, "psi", ep)
        assert ew < 1e-11 and ep < 1e-11
    assert all(out[gm][2] < out[1][2] for gm in (2, 3, 4, 5))


@pytest.mark.parametrize("integ", [1, 2])
def test_guess_with_other_integrators(gpu_available, integ):
    """With Euler and RK2 the RK4 modes 3 and 4 fall back to the extrapolation (mode 2), and the
    default projection (symmetric operator here) needs no more iterations than the reference's
    warm start; all give the same fields."""
    m = meshes.cavity(8, 2)
    w1, p1, it1 = run(m, 3, 1, 6, 1e-13, integ=integ)
    w0, p0, it0 = run(m, 3, 0, 6, 1e-13, integ=integ)
    w2, p2, it2 = run(m, 3, 2, 6, 1e-13, integ=integ)
    w3, p3, it3 = run(m, 3, 3, 6, 1e-13, integ=integ)
    w4, p4, it4 = run(m, 3, 4, 6, 1e-13, integ=integ)
    print("integrator", integ, "iterations (modes 1, default, 2, 3, 4)", it1, it0, it2, it3, it4)
    for w, p in ((w0, p0), (w2, p2), (w3, p3), (w4, p4)):
        assert rel_l2(w, w1) < 1e-11 and rel_l2(p, p1) < 1e-11
    assert it3 == it2 and it4 == it2
    assert it0 <= it1


def test_guess_nonsymmetric_projection(gpu_available):
    """On a jittered step (non-affine elements: non-symmetric operator, BiCGSTAB) the default is the
    minimal-residual form of the projection: the same fields as the reference's warm start and mode
    4, in fewer iterations than either."""
    m = meshes.backward_facing_step(12, 4, 2, jitter=0.1, seed=3)
    kw = dict(K=3, problem_type=2, Re=200.0, dt=1e-3, time_integration_type=4, advection_type=1, tol=1e-13,
              maxiter=500)
    res = {}
    for gm in (0, 1, 4, 5):
        g = hoc.HO2D(psi_guess=gm, **kw)
        g.load_mesh(m)
        g.setup()
        g.step(8)
        st = g.stats()
        assert st["krylov"] == hoc.KRYLOV_BICGSTAB
        res[gm] = (g.get_vorticity(), g.get_streamfunction(), st["total_iters"] / st["n_solves"])
    print("iterations per solve (default, 1, 4, 5)", [res[gm][2] for gm in (0, 1, 4, 5)])
    assert np.array_equal(res[0][0], res[5][0]) and res[0][2] == res[5][2]
    for gm in (0, 4):
        assert rel_l2(res[gm][0], res[1][0]) < 1e-11 and rel_l2(res[gm][1], res[1][1]) < 1e-11
    assert res[0][2] < res[4][2] < res[1][2]


def test_guess_mixed_precision(gpu_available):
    m = meshes.cavity(12, 2)
    w64, p64, _ = run(m, 5, 3, 6, 1e-11)
    w32, p32, _ = run(m, 5, 3, 6, 1e-11, precision=hoc.HOC_MIXED)
    assert rel_l2(w32, w64) < 1e-8 and rel_l2(p32, p64) < 1e-8


def test_guess_rejects_bad_mode():
    with pytest.raises(hoc.HocError):
        hoc.HO2D(K=3, psi_guess=6)
