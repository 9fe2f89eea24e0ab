"""Cylinder in a channel on an O-grid of curved quadratic (Lnod = 3) quads at K = 6 (P = 5):
the non-affine metric path, BiCGSTAB (W A is not symmetric on curved elements), problem
type 3 boundaries, and mixed precision.  Small grids against the oracle; the full
256 x 512 grid (131 072 elements) mixed against FP64 on the GPU."""
import numpy as np
import pytest

import hoc
from hoc import meshes
from pyoracle import Oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


# Start where the inflow ramp 2|sin(pi t/40)| of problem_type 3 (src/HO_2D.cpp:2062) is 1: the
# uniform inflow u = 1 at full strength.
T0 = 20.0 / 3.0


def omega0(xy, seed=3):
    """Seeded vorticity perturbation of amplitude 1e-3 at the solution points."""
    return 1e-3 * np.random.default_rng(seed).uniform(-1.0, 1.0, len(xy))


@pytest.mark.parametrize("precision", [hoc.HOC_FP64, hoc.HOC_MIXED])
def test_ogrid_K6_vs_oracle(gpu_available, precision):
    m = meshes.cylinder_ogrid(8, 24, 3)
    # With the inflow on, the wall vorticity flux slip * Re / dt (src/HO_2D.cpp:2656) turns the
    # O(1) velocity's solve-tolerance differences into vorticity changes that the 1e-3 field
    # shows relative to itself: tol 1e-12 leaves 3e-10 between two converged solvers, so both
    # sides solve to 1e-14.
    kw = dict(problem_type=3, Re=100.0, dt=2e-5, time_integration_type=4, advection_type=1, tol=1e-14, maxiter=600)
    o = Oracle(m, 6, **kw)
    g = hoc.HO2D(K=6, precision=precision, **kw)
    g.load_mesh(m)
    g.setup()
    w0 = omega0(g.solution_points())
    o.set_vorticity(w0)
    g.set_vorticity(w0)
    o.set_time(T0)
    g.set_time(T0)
    o.step(2)
    g.step(2)
    assert abs(g.time - o.time()) < 1e-14
    ew = rel_l2(g.get_vorticity(), o.vorticity())
    ep = rel_l2(g.get_streamfunction(), o.streamfunction())
    st = g.stats()
    print("ogrid K=6 precision", precision, "omega", ew, "psi", ep, "krylov", st["krylov"], "its", st["total_iters"])
    assert st["krylov"] == hoc.KRYLOV_BICGSTAB
    assert ew < 1e-10 and ep < 1e-10


def test_ogrid_full_mixed_vs_fp64(gpu_available):
    # Both solves stop at tol 1e-13: two solves converged to 1e-11 stop at stream functions about
    # 5e-9 apart, which the no-slip wall flux (slip * Re / dt, Re / dt = 1e8) turns into about
    # 2e-7 of vorticity whatever the preconditioner's precision.
    m = meshes.cylinder_ogrid(256, 512, 3)
    assert m.n_elem == 131072
    kw = dict(K=6, problem_type=3, Re=100.0, dt=1e-6, time_integration_type=4, advection_type=1, tol=1e-13,
              maxiter=800)
    out = {}
    for prec in (hoc.HOC_FP64, hoc.HOC_MIXED):
        g = hoc.HO2D(precision=prec, **kw)
        g.load_mesh(m)
        g.setup()
        g.set_vorticity(omega0(g.solution_points()))
        g.set_time(T0)
        g.step(3)
        out[prec] = (g.get_vorticity(), g.get_streamfunction(), g.stats())
        g.close()
    (w64, p64, s64), (w32, p32, s32) = out[hoc.HOC_FP64], out[hoc.HOC_MIXED]
    ew, ep = rel_l2(w32, w64), rel_l2(p32, p64)
    print("ogrid 256x512 K=6, 3 steps: mixed vs fp64 omega", ew, "psi", ep,
          "its", s64["total_iters"], s32["total_iters"])
    assert np.all(np.isfinite(w64)) and ew < 1e-8 and ep < 1e-8
