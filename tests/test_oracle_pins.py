"""Pins of the FP64 CPU oracle against what the mathematics fixes.

The reference has no tests or golden outputs for this path (its Poisson backend,
AMGCL, is absent here), so each oracle function is checked against closed forms,
library routines (numpy's Gauss-Legendre and Legendre series), polynomial
exactness and invariants, chosen so that a dropped term, a wrong sign or index,
or a transposed operand fails.  CPU only."""
import numpy as np
import numpy.polynomial.legendre as npleg
import pytest
import scipy.sparse as sp

from hoc import meshes
from pyoracle import Oracle
from conftest import rel_l2


def oracle(mesh, K, **kw):
    kw.setdefault("Re", 100.0)
    kw.setdefault("dt", 1e-4)
    kw.setdefault("tol", 1e-14)
    kw.setdefault("maxiter", 400)
    return Oracle(mesh, K, **kw)


# ---------------------------------------------------------------- 1-D tables (src/HO_2D.cpp:486-673)
@pytest.mark.parametrize("K", [2, 3, 4, 5, 6, 7])
def test_gauss_legendre(K):
    o = oracle(meshes.cavity(2, 2), K)
    x, w = npleg.leggauss(K)
    assert np.allclose(o.table("sps"), x, atol=2e-15)
    assert np.allclose(o.table("wgt"), w, atol=2e-15)


@pytest.mark.parametrize("K", [2, 3, 4, 5, 6])
def test_lagrange_tables(K):
    o = oracle(meshes.cavity(2, 2), K)
    x = o.table("sps")
    D = o.table("sps_sps_grad_basis").reshape(K, K)        # D[m][k] = l_m'(x_k)
    bb = o.table("sps_boundary_basis").reshape(K, 2)       # l_m(-1), l_m(+1)
    bgb = o.table("sps_boundary_grad_basis").reshape(K, 2)
    for deg in range(K):                                    # exact for degree <= K-1
        c = np.zeros(deg + 1); c[-1] = 1.0
        f = np.polynomial.polynomial.polyval(x, c)
        df = np.polynomial.polynomial.polyval(x, np.polynomial.polynomial.polyder(c))
        assert np.allclose(f @ D, df, atol=1e-11)
        assert np.allclose(f @ bb[:, 0], (-1.0) ** deg, atol=1e-12)
        assert np.allclose(f @ bb[:, 1], 1.0, atol=1e-12)
        dpol = np.polynomial.polynomial.polyder(c)
        assert np.allclose(f @ bgb[:, 0], np.polynomial.polynomial.polyval(-1.0, dpol), atol=1e-10)
        assert np.allclose(f @ bgb[:, 1], np.polynomial.polynomial.polyval(1.0, dpol), atol=1e-10)


@pytest.mark.parametrize("K", [2, 3, 4, 5, 6])
def test_radau_correction(K):
    """g_DG: left Radau R_L = (-1)^K/2 (P_K - P_{K-1}), R_L(-1)=1, R_L(1)=0, R_R(x)=R_L(-x)
    (src/HO_2D.cpp:648-670).  The interface coefficient g' the reference uses,
    -+K(K+1)/2 (:547-550), is not R_L'(-1) = -K^2/2 (the value its comment at :2853 quotes);
    the oracle and the GPU path follow the code."""
    o = oracle(meshes.cavity(2, 2), K)
    x = o.table("sps")
    rad = o.table("sps_radau").reshape(K, 2)
    grad = o.table("sps_grad_radau").reshape(K, 2)
    cL = np.zeros(K + 1); cL[K] = 1.0; cL[K - 1] = -1.0
    cL *= 0.5 * (-1.0) ** K
    RL = lambda t: npleg.legval(t, cL)
    dRL = lambda t: npleg.legval(t, npleg.legder(cL))
    assert abs(RL(-1.0) - 1.0) < 1e-14 and abs(RL(1.0)) < 1e-14
    assert np.allclose(rad[:, 0], RL(x), atol=1e-13)
    assert np.allclose(rad[:, 1], RL(-x), atol=1e-13)
    assert np.allclose(grad[:, 0], dRL(x), atol=1e-11)
    assert np.allclose(grad[:, 1], -dRL(-x), atol=1e-11)
    assert abs(dRL(-1.0) - (-0.5 * K * K)) < 1e-10


# ---------------------------------------------------------------- metrics (src/HO_2D.cpp:1240-1388)
def test_affine_metrics():
    n, K = 4, 3
    o = oracle(meshes.cavity(n, 2, seed=3, rotate=True), K)
    J = o.table("vol_jac")
    assert np.allclose(J, (1.0 / n) ** 2 / 4.0, rtol=1e-13)
    G = o.table("G").reshape(-1, 2, 2)
    assert np.allclose(G[:, 0, 1], 0.0, atol=1e-13)      # orthogonal grid lines
    assert np.allclose(G[:, 0, 0], 1.0, rtol=1e-13) and np.allclose(G[:, 1, 1], 1.0, rtol=1e-13)
    assert np.allclose(o.table("face_Anorm"), 0.5 / n, rtol=1e-13)


@pytest.mark.parametrize("lnod,jit,curve", [(2, 0.15, 0.0), (3, 0.0, 0.05), (4, 0.0, 0.05)])
def test_area_quadrature(lnod, jit, curve):
    """sum w_i w_j J over all sps = domain area (exact for these maps at K >= lnod)."""
    K = 4
    o = oracle(meshes.cavity(5, lnod, seed=9, rotate=True, jitter=jit, curve=curve), K)
    assert abs(o.table("areas").sum() - 1.0) < 1e-12


# ---------------------------------------------------------------- Laplacian (src/HO_2D.cpp:1533-1804)
LAP_MESHES = [
    ("affine_rot", lambda: meshes.cavity(5, 2, seed=1, rotate=True, flip_bedges=True), 4),
    ("jittered", lambda: meshes.cavity(4, 2, seed=2, rotate=True, jitter=0.2), 5),
    ("curved_q", lambda: meshes.cavity(4, 3, seed=3, rotate=True, curve=0.05), 3),
    ("curved_c", lambda: meshes.cavity(3, 4, seed=4, rotate=True, curve=0.05), 5),
]


@pytest.mark.parametrize("name,mk,K", LAP_MESHES)
def test_assembled_equals_matrix_free(name, mk, K):
    """The assembled matrix (blocks of :1568-1792 + CSR filter :3029-3043) equals the FR
    diffusion operator of :2666-2831 with zero psi boundary values."""
    o = oracle(mk(), K)
    ptr, col, val = o.csr()
    A = sp.csr_matrix((val, col, ptr), shape=(o.D, o.D))
    for seed in range(3):
        x = np.random.default_rng(seed).standard_normal(o.D)
        assert rel_l2(A @ x, o.laplacian_mf(x)) < 1e-14


def test_assembled_equals_matrix_free_neumann():
    m = meshes.backward_facing_step(20, 8, 2, jitter=0.1, seed=5)
    o = oracle(m, 4, problem_type=2, Re=400.0)
    assert set(np.unique(o.table("BC_switch_Poisson"))) == {0.0, 1.0}
    ptr, col, val = o.csr()
    A = sp.csr_matrix((val, col, ptr), shape=(o.D, o.D))
    x = np.random.default_rng(0).standard_normal(o.D)
    assert rel_l2(A @ x, o.laplacian_mf(x)) < 1e-13


@pytest.mark.parametrize("name,mk,K", LAP_MESHES[:1] + [("bfs", None, 4), ("bfs_rotated", None, 3)])
def test_boundary_source_is_operator_bc_part(name, mk, K):
    """boundary_source . BC (:1631-1698) is the FR operator applied to a zero field with
    those boundary values, Neumann values taken with the owner face's outward sign
    2(ijp%2)-1 (the sign of :1685): A x + S g = L(x; s g).  The rotated step puts Neumann
    faces of both signs on the outlet."""
    if mk is None:
        o = oracle(meshes.backward_facing_step(20, 8, 2, jitter=0.1, seed=5, rotate=name.endswith("rotated")), K,
                   problem_type=2, Re=400.0)
    else:
        o = oracle(mk(), K)
    rng = np.random.default_rng(7)
    g = rng.standard_normal((o.n_bedges, K))
    neu = o.table("BC_switch_Poisson") > 0.5
    s_out = np.array([2 * ([2, 1, 3, 0][int(sd)] % 2) - 1 for sd in o.table("be_side")])
    if name.endswith("rotated"):
        assert set(s_out[neu]) == {-1, 1}
    x = rng.standard_normal(o.D)
    S = o.boundary_source()
    be = o.table("be_elem").astype(int)
    Sg = np.zeros((o.D // (K * K), K * K))
    np.add.at(Sg, be, np.einsum("eik,ek->ei", S, g))
    ptr, col, val = o.csr()
    A = sp.csr_matrix((val, col, ptr), shape=(o.D, o.D))
    gs = g.copy()
    gs[neu] *= s_out[neu][:, None]
    assert rel_l2(A @ x + Sg.ravel(), o.laplacian_bc(x, gs)) < 1e-13


@pytest.mark.parametrize("K", [3, 4, 5])
def test_laplacian_polynomial_exact(K):
    """psi = x(1-x)y(1-y) is in the element space (K>=3) and vanishes on the boundary,
    so on affine meshes L(psi) = J * lap(psi) exactly at the sps."""
    o = oracle(meshes.cavity(4, 2, seed=11, rotate=True, flip_bedges=True), K)
    xy = o.table("sps_xy").reshape(-1, 2)
    x, y = xy[:, 0], xy[:, 1]
    psi = x * (1 - x) * y * (1 - y)
    lap = -2 * y * (1 - y) - 2 * x * (1 - x)
    assert rel_l2(o.laplacian_mf(psi), o.table("vol_jac") * lap) < 1e-12


@pytest.mark.parametrize("K", [3, 4])
def test_velocity_polynomial_exact(K):
    """u = dpsi/dy, v = -dpsi/dx from the FR gradient (:3174-3366) is exact for psi in the
    element space with exact (zero) Dirichlet values on affine meshes."""
    o = oracle(meshes.cavity(4, 2, seed=12, rotate=True), K)
    xy = o.table("sps_xy").reshape(-1, 2)
    x, y = xy[:, 0], xy[:, 1]
    psi = x * (1 - x) * y * (1 - y)
    uv = o.velocity_from_psi(psi)
    assert np.allclose(uv[:, 0], x * (1 - x) * (1 - 2 * y), atol=1e-13)
    assert np.allclose(uv[:, 1], -(1 - 2 * x) * y * (1 - y), atol=1e-13)


def _manufactured_error(n, K):
    """Reference debug case (src/HO_2D.cpp:1440, 1862-1869): psi = exp(x^2+y^2),
    omega = -4(1+x^2+y^2) exp(x^2+y^2), Dirichlet psi on the boundary."""
    o = oracle(meshes.cavity(n, 2), K)
    exy = o.table("edge_xy").reshape(-1, 2)
    o.set_bc("BC_Poisson", np.exp(exy[:, 0] ** 2 + exy[:, 1] ** 2))
    xy = o.table("sps_xy").reshape(-1, 2)
    r2 = xy[:, 0] ** 2 + xy[:, 1] ** 2
    psi = o.solve_poisson_bc(-4.0 * (1 + r2) * np.exp(r2))
    return rel_l2(psi, np.exp(r2))


@pytest.mark.parametrize("K", [3, 4])
def test_manufactured_poisson_converges(K):
    e1, e2 = _manufactured_error(4, K), _manufactured_error(8, K)
    assert e2 < 1e-4
    assert e1 / e2 > 2 ** (K - 0.5)      # spectral/h convergence of order ~K


def test_poisson_solve_residual():
    o = oracle(meshes.cavity(6, 2, seed=1, rotate=True, jitter=0.1), 5)
    xy = o.table("sps_xy").reshape(-1, 2)
    w = np.sin(3 * xy[:, 0]) * np.cos(2 * xy[:, 1])
    psi = o.poisson(w)
    it, rel = o.last_solve()
    ptr, col, val = o.csr()
    A = sp.csr_matrix((val, col, ptr), shape=(o.D, o.D))
    b = -w * o.table("vol_jac")
    assert rel_l2(A @ psi, b) < 1e-12 and rel < 1e-13


# ---------------------------------------------------------------- RK / invariants
def test_zero_state_is_steady_without_lid():
    """No lid, zero vorticity: every stage residual is zero (all BCs homogeneous)."""
    m = meshes.cavity(4, 2)
    m.bnames = ["bottom_wall", "right_wall", "lid_off", "left_wall"]
    o = oracle(m, 3)
    o.step(2)
    assert np.abs(o.vorticity()).max() == 0.0


def test_rk4_linear_amplification():
    """For a linear problem (lid off, advection negligible at tiny amplitude), one RK4 step
    is p(dt L) w with p(z) = 1+z+z^2/2+z^3/6+z^4/24; compare with the stage operator."""
    m = meshes.cavity(3, 2)
    m.bnames = ["bottom_wall", "right_wall", "lid_off", "left_wall"]
    dt = 1e-3
    o = oracle(m, 3, Re=100.0, dt=dt, tol=1e-15)
    rng = np.random.default_rng(3)
    w0 = 1e-9 * rng.standard_normal(o.D)
    Lop = lambda w: o.stage(w, 0.0) / dt           # k/dt, linear in w up to O(|w|^2)
    k1 = dt * Lop(w0); k2 = dt * Lop(w0 + 0.5 * k1); k3 = dt * Lop(w0 + 0.5 * k2); k4 = dt * Lop(w0 + k3)
    ref = w0 + (k1 + 2 * k2 + 2 * k3 + k4) / 6
    o2 = oracle(m, 3, Re=100.0, dt=dt, tol=1e-15)
    o2.set_vorticity(w0)
    o2.step(1)
    assert rel_l2(o2.vorticity(), ref) < 1e-9


def test_amgcl_default_spai0_stalls_at_K5(monkeypatch):
    """Documents why the oracle relaxes with Chebyshev: SA + spai0 (AMGCL defaults) does not
    reach 1e-10 in 150 BiCGSTAB iterations for the K = 5 FR Laplacian."""
    monkeypatch.setenv("HO_ORACLE_RELAX", "0")
    monkeypatch.setenv("HO_ORACLE_UNSCALED", "1")
    o = Oracle(meshes.cavity(16, 2), 5, Re=100.0, dt=1e-5, tol=1e-10, maxiter=150)
    xy = o.table("sps_xy").reshape(-1, 2)
    o.poisson(np.sin(3 * xy[:, 0]) * np.cos(2 * xy[:, 1]))
    it, rel = o.last_solve()
    assert it == 150 and rel > 1e-6
