"""Pins of the oracle's flow terms against what the physics and the mathematics fix.

The advective residual (src/HO_2D.cpp:2245-2645), the wall-vorticity flux
(:2647-2664 with the velocity_jump of :3249-3253), the Neumann common values with data
(:2905-2920), the inflow programmes of problem types 2 and 3 (:1979-2217) and the
hybrid boundary programme (:962-1120, :2104-2191, :4050-4117) have no closed form in
the reference and no reference build here.  Each is pinned by an identity that holds
for the continuous problem and, exactly, for the FR discretisation, computed here from
the mesh geometry and numpy alone:

- discrete conservation of the advective residual (its integral is the net boundary
  flux) and element-by-element upwind fluxes of piecewise-constant vorticity;
- exactness of the advective residual for polynomial velocity and vorticity;
- the circulation identity of the wall-vorticity flux: one Euler step makes the
  circulation equal the wall velocity's circulation, whatever the state, Re, dt or K;
- exactness of the FR Laplacian for polynomials with Dirichlet and Neumann data;
- mass-flux identities of the inflow and hybrid boundary values.

A dropped term, a wrong sign or a wrong face pairing breaks one of these.  CPU only."""
import numpy as np
import pytest

from hoc import meshes
from pyoracle import Oracle

F2I = [2, 1, 3, 0]      # element side (S, E, N, W) -> face index ijp (W, E, S, N)


def oracle(mesh, K, **kw):
    kw.setdefault("Re", 100.0)
    kw.setdefault("dt", 1e-4)
    kw.setdefault("tol", 1e-14)
    kw.setdefault("maxiter", 400)
    return Oracle(mesh, K, **kw)


def side_sign(o):
    """+1 where a boundary edge is the element's xi+ or eta+ face, -1 otherwise: the factor
    that turns an outward normal component into the component along the element's own axis."""
    return np.array([2 * (F2I[int(sd)] % 2) - 1 for sd in o.table("be_side")], dtype=float)


def element_faces(m):
    """Straight faces of every element from its corner nodes (Gmsh order: CCW corners first):
    face f of element e runs from corner f to corner f+1.  Returns the outward normal times
    the length [N,4,2], and the neighbour across each face [N,4] (-1 on the boundary) with the
    boundary edge index [N,4] (-1 inside), found from shared corner pairs."""
    c = m.elem[:, :4]
    xy = m.xy
    N = len(c)
    nl = np.zeros((N, 4, 2))
    owner = {}
    for e in range(N):
        for f in range(4):
            a, b = c[e, f], c[e, (f + 1) % 4]
            d = xy[b] - xy[a]
            nl[e, f] = (d[1], -d[0])
            owner.setdefault(frozenset((int(a), int(b))), []).append((e, f))
    nbr = -np.ones((N, 4), dtype=int)
    bedge = -np.ones((N, 4), dtype=int)
    for lst in owner.values():
        if len(lst) == 2:
            (e1, f1), (e2, f2) = lst
            nbr[e1, f1], nbr[e2, f2] = e2, e1
    for i, nodes in enumerate(m.bedges):
        (e, f), = owner[frozenset((int(nodes[0]), int(nodes[1])))]
        bedge[e, f] = i
    return nl, nbr, bedge


def element_integrals(o, r):
    """sum_ij w_i w_j J r over each element."""
    K = o.K
    return (o.table("areas") * r).reshape(-1, K * K).sum(1)


# ---------------------------------------------------------------- advection (src/HO_2D.cpp:2245-2645)
CONS_MESHES = [
    ("affine_rot", lambda: meshes.cavity(6, 2, seed=1, rotate=True, flip_bedges=True), 4),
    ("jittered", lambda: meshes.cavity(5, 2, seed=2, rotate=True, jitter=0.2), 5),
    ("curved_q", lambda: meshes.cavity(5, 3, seed=3, rotate=True, curve=0.05), 4),
    ("curved_c", lambda: meshes.cavity(4, 4, seed=4, rotate=True, curve=0.05), 5),
]


@pytest.mark.parametrize("adv", [1, 2])
@pytest.mark.parametrize("name,mk,K", CONS_MESHES)
def test_advection_conserves_on_walls(name, mk, K, adv):
    """Inside walls the advective residual of a full stage (velocity from the Poisson solve)
    integrates to zero: interior common fluxes cancel in pairs and walls carry no normal
    velocity.  The field is rough, so most interior faces take the upwind branch."""
    o = oracle(mk(), K, advection_type=adv)
    xy = o.table("sps_xy").reshape(-1, 2)
    rng = np.random.default_rng(5)
    w = np.sin(4 * xy[:, 0] + 1) * np.cos(3 * xy[:, 1]) + 0.3 * rng.standard_normal(o.D)
    o.stage(w, 0.0)
    r = o.table("RHS_advective")
    a = o.table("areas")
    scale = np.sum(a * np.abs(r))
    assert scale > 0.1
    assert abs(np.sum(a * r)) < 1e-14 * scale


@pytest.mark.parametrize("adv", [1, 2])
@pytest.mark.parametrize("name,mk,K", [CONS_MESHES[0], CONS_MESHES[1],
                                       ("bfs_rot", lambda: meshes.backward_facing_step(
                                           12, 4, 2, length=6.0, step_len=2.0, seed=6, rotate=True), 3)])
def test_advection_element_fluxes_are_upwind(name, mk, K, adv):
    """Uniform velocity u, vorticity constant in each element: the residual of each element
    integrates to minus the net outflow through its faces, sum_f (u.n_f) |f| w*_f, where w*
    is the upwind element's value.  Advection type 1 takes the upwind value only where the
    jump exceeds 1e-4 and the mean below that (src/HO_2D.cpp:2370); type 2 always upwinds
    (:2594-2597).  Boundary faces carry the given boundary vorticity and the normal velocity
    u.n.  The faces, normals and neighbours are computed here from the corner nodes."""
    m = mk()
    o = oracle(m, K, advection_type=adv, problem_type=1)
    N, K2 = m.n_elem, K * K
    rng = np.random.default_rng(11)
    u = np.array([np.cos(0.3), np.sin(0.3)]) * 1.7
    wel = rng.standard_normal(N)
    wel[rng.random(N) < 0.35] = 0.25          # runs of equal values: zero jumps
    small = rng.random(N) < 0.2
    wel[small] = 0.25 + rng.choice([-1, 1], small.sum()) * rng.uniform(1e-6, 8e-5, small.sum())
    w = np.repeat(wel, K2)
    uv = np.tile(u, (o.D, 1))
    nl, nbr, bedge = element_faces(m)
    be_w = rng.standard_normal(m.n_bedges)
    exy = o.table("edge_xy").reshape(m.n_bedges, K, 2)
    un_out = np.zeros(m.n_bedges)
    for e in range(N):
        for f in range(4):
            if bedge[e, f] >= 0:
                un_out[bedge[e, f]] = nl[e, f] @ u / np.linalg.norm(nl[e, f])
    bc_w = np.repeat(be_w, K).reshape(m.n_bedges, K)
    bc_un = (un_out * side_sign(o))[:, None] * np.ones((1, K))
    r = o.advection(w, uv, bc_w, bc_un)
    got = element_integrals(o, r)
    want = np.zeros(N)
    n_up = n_mean = 0
    for e in range(N):
        for f in range(4):
            q = nl[e, f] @ u                       # (u.n) |f|, outward from e
            if nbr[e, f] < 0:
                ws = be_w[bedge[e, f]]
            else:
                a, b = wel[e], wel[nbr[e, f]]
                if adv == 1 and abs(a - b) <= 1e-4:
                    ws = 0.5 * (a + b)
                    n_mean += abs(a - b) > 0
                else:
                    ws = a if q > 0 else b
                    n_up += 1
            want[e] -= q * ws
    assert n_up > N and (adv == 2 or n_mean > 4)
    assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("adv", [1, 2])
@pytest.mark.parametrize("K,rot", [(4, True), (5, True), (5, False)])
def test_advection_polynomial_exact(K, rot, adv):
    """Polynomial velocity (degree 2) and vorticity (degree K-3), continuous across faces, with
    exact boundary values: the flux u*w has degree K-1, so the FR residual equals -div(u w) at
    every solution point.  The divergence comes from numpy's polynomial algebra.  There are no
    jumps, so the switch of :2370 stays off."""
    P = np.polynomial.polynomial
    m = meshes.cavity(4, 2, seed=7, rotate=rot, flip_bedges=rot)
    o = oracle(m, K, advection_type=adv)
    cu = _add2(_add2(_mono(0, 0, 0.3), _mono(1, 1, 1.0)), _mono(0, 2, -0.5))       # u = 0.3 + xy - y^2/2
    cv = _add2(_add2(_mono(0, 0, -0.7), _mono(2, 0, 0.4)), _mono(1, 1, -1.0))      # v = -0.7 + 0.4x^2 - xy
    lin = _add2(_add2(_mono(0, 0, 1.2), _mono(1, 0, -0.8)), _mono(0, 1, 0.5))
    cw = _mono(0, 0, 1.0)
    for _ in range(K - 3):
        cw = _mul2(cw, lin)
    cw = _add2(cw, _mono(K - 3, 0, 0.3))                                            # w = lin^(K-3) + 0.3 x^(K-3)
    div = _add2(P.polyder(_mul2(cu, cw), axis=0), P.polyder(_mul2(cv, cw), axis=1))
    xy = o.table("sps_xy").reshape(-1, 2)
    x, y = xy[:, 0], xy[:, 1]
    exact = P.polyval2d(x, y, div)
    uv = np.stack([P.polyval2d(x, y, cu), P.polyval2d(x, y, cv)], 1)
    w = P.polyval2d(x, y, cw)
    exy = o.table("edge_xy").reshape(-1, 2)
    ex, ey = exy[:, 0], exy[:, 1]
    nout = np.zeros_like(exy)
    nout[ey < 1e-12] = (0, -1); nout[ex > 1 - 1e-12] = (1, 0)
    nout[ey > 1 - 1e-12] = (0, 1); nout[ex < 1e-12] = (-1, 0)
    assert np.all(np.abs(nout).sum(1) == 1)
    un = P.polyval2d(ex, ey, cu) * nout[:, 0] + P.polyval2d(ex, ey, cv) * nout[:, 1]
    un = un.reshape(-1, K) * side_sign(o)[:, None]
    r = o.advection(w, uv, P.polyval2d(ex, ey, cw).reshape(-1, K), un)
    assert np.max(np.abs(r + exact)) < 1e-11 * np.max(np.abs(exact))


def _mono(i, j, c):
    a = np.zeros((i + 1, j + 1)); a[i, j] = c
    return a


def _add2(a, b):
    s = np.zeros((max(a.shape[0], b.shape[0]), max(a.shape[1], b.shape[1])))
    s[:a.shape[0], :a.shape[1]] += a
    s[:b.shape[0], :b.shape[1]] += b
    return s


def _mul2(a, b):
    s = np.zeros((a.shape[0] + b.shape[0] - 1, a.shape[1] + b.shape[1] - 1))
    for i in range(a.shape[0]):
        for j in range(a.shape[1]):
            s[i:i + b.shape[0], j:j + b.shape[1]] += a[i, j] * b
    return s


# ---------------------------------------------------------------- wall-vorticity flux (src/HO_2D.cpp:2647-2664)
def circulation(o, w):
    return float(np.sum(o.table("areas") * w))


@pytest.mark.parametrize("adv", [1, 2])
@pytest.mark.parametrize("name,mk,K", CONS_MESHES + [("affine_k3", lambda: meshes.cavity(7, 2, seed=8), 3)])
def test_euler_step_sets_wall_circulation(name, mk, K, adv):
    """The wall-vorticity flux velocity_jump / (dt Re^-1) (:2654-2656) is the Neumann datum
    of the diffusion term on no-slip walls.  Integrated over the domain, the diffusion term
    is the boundary sum of that flux, i.e. (1/dt) times the slip between the fluid's wall
    velocity (d psi/dn of the Poisson solution, whose boundary sum is minus the circulation
    by Green's identity) and the wall's; advection adds nothing inside walls.  So one Euler
    step from any field leaves the circulation of the wall velocity, -1 for the unit lid
    moving at speed 1 (BC_parl_vel = -1 on the lid, :717-725), whatever Re, dt, K or mesh."""
    rng = np.random.default_rng(abs(hash(name)) % 1000 + adv)
    Re, dt = float(rng.uniform(50, 5000)), float(rng.uniform(1e-6, 1e-3))
    o = oracle(mk(), K, Re=Re, dt=dt, time_integration_type=1, advection_type=adv, tol=1e-15)
    xy = o.table("sps_xy").reshape(-1, 2)
    o.set_vorticity(np.cos(3 * xy[:, 0]) * np.sin(5 * xy[:, 1] + 0.5) + 0.2 * rng.standard_normal(o.D))
    o.step(1)
    assert abs(circulation(o, o.vorticity()) + 1.0) < 1e-12
    o.step(1)
    assert abs(circulation(o, o.vorticity()) + 1.0) < 1e-12


def test_rk4_relaxes_circulation_by_stability_polynomial():
    """Each stage k_s of the same flux carries circulation -1 - Gamma(w_s); the RK4 combination
    turns an excess e = Gamma + 1 into e * p(-1) = 3/8 e per step (p = 1+z+z^2/2+z^3/6+z^4/24)."""
    m = meshes.cavity(6, 2, seed=9, rotate=True, jitter=0.15)
    o = oracle(m, 4, Re=400.0, dt=2e-4, time_integration_type=4, tol=1e-15)
    assert abs(circulation(o, o.vorticity())) == 0.0
    e = 1.0
    for _ in range(3):
        o.step(1)
        e *= 0.375
        assert abs(circulation(o, o.vorticity()) - (-1.0 + e)) < 1e-12


@pytest.mark.parametrize("K", [3, 5])
def test_nonuniform_lid_circulation(K):
    """A lid profile u(x) = 16 x^2 (1-x)^2 set through BC_parl_vel (= -u along the lid) gives
    circulation -int_0^1 u dx = -8/15 after one Euler step (the profile has degree 4 <= 2K-1
    on each lid edge, so the Gauss quadrature of the flux is exact)."""
    m = meshes.cavity(5, 2, seed=10, rotate=True)
    o = oracle(m, K, Re=250.0, dt=3e-4, time_integration_type=1, tol=1e-15)
    lid = (m.bedge_bdry == m.bnames.index("top"))
    ex = o.table("edge_xy").reshape(-1, K, 2)[:, :, 0]
    pv = o.table("BC_parl_vel").reshape(-1, K)
    assert np.all(pv[lid] == -1.0) and np.all(pv[~lid] == 0.0)
    pv[lid] = -16 * ex[lid] ** 2 * (1 - ex[lid]) ** 2
    o.set_bc("BC_parl_vel", pv)
    o.step(1)
    assert abs(circulation(o, o.vorticity()) + 8.0 / 15.0) < 1e-12


# ---------------------------------------------------------------- Laplacian with boundary data (:2872-2927)
@pytest.mark.parametrize("K,rot", [(3, True), (4, True), (5, False)])
def test_laplacian_exact_with_dirichlet_and_neumann_data(K, rot):
    """A polynomial psi of total degree K-1 with its exact boundary values on the Dirichlet edges
    and its exact outward normal derivative on the Neumann outlet (taken along the element's
    own axis, as the flux is) gives L(psi) = J lap(psi) at every solution point of an affine
    mesh: pins the Dirichlet and the Neumann common values (the Neumann K x K solve) with
    non-zero data.  The rotated step has outlet faces of both orientations."""
    P = np.polynomial.polynomial
    m = meshes.backward_facing_step(12, 4, 2, length=6.0, step_len=2.0, seed=6, rotate=rot)
    o = oracle(m, K, problem_type=2, Re=400.0)
    neu = o.table("BC_switch_Poisson") > 0.5
    assert neu.any() and not neu.all()
    rng = np.random.default_rng(K)
    c = np.zeros((K, K))
    for i in range(K):
        for j in range(K - i):
            c[i, j] = rng.uniform(-1, 1) / (1 + i + j)
    xy = o.table("sps_xy").reshape(-1, 2)
    psi = P.polyval2d(xy[:, 0], xy[:, 1], c)
    lap = P.polyval2d(xy[:, 0], xy[:, 1], _add2(P.polyder(c, 2, axis=0), P.polyder(c, 2, axis=1)))
    exy = o.table("edge_xy").reshape(-1, 2)
    g = P.polyval2d(exy[:, 0], exy[:, 1], c).reshape(-1, K)
    dpdx = P.polyval2d(exy[:, 0], exy[:, 1], P.polyder(c, 1, axis=0)).reshape(-1, K)
    assert np.allclose(exy.reshape(-1, K, 2)[neu][:, :, 0], 6.0)          # outlet x = 6, normal +x
    g[neu] = dpdx[neu] * side_sign(o)[neu][:, None]
    if rot:
        assert set(side_sign(o)[neu]) == {-1.0, 1.0}
    got = o.laplacian_bc(psi, g)
    want = o.table("vol_jac") * lap
    assert np.max(np.abs(got - want)) < 1e-11 * np.max(np.abs(want))


# ---------------------------------------------------------------- inflow programmes (src/HO_2D.cpp:1979-2217)
def _boundary_tables(o, m, name):
    b = m.bedge_bdry == m.bnames.index(name)
    K = o.K
    exy = o.table("edge_xy").reshape(-1, K, 2)[b]
    s = side_sign(o)[b][:, None]
    tab = {k: o.table(k).reshape(-1, K)[b] for k in ("BC_Poisson", "BC_normal_vel", "BC_diffusion", "BC_parl_vel")}
    return exy, s, tab


def _edge_lengths(m, name):
    b = m.bedge_bdry == m.bnames.index(name)
    e = m.bedges[b]
    return np.linalg.norm(m.xy[e[:, 1]] - m.xy[e[:, 0]], axis=1)


@pytest.mark.parametrize("t,ramp", [(20.0 / 3.0, 1.0), (20.0, 2.0), (5.0, None), (47.0, None)])
def test_step_inflow_is_a_consistent_flow(t, ramp):
    """Backward-facing step (problem_type 2): at any time the inlet values describe one
    parabolic flow.  u = -(outward normal velocity), psi_inlet is its integral in y and meets
    the walls' psi at both ends (so psi_top - psi_bottom is the volume flux), the inlet
    vorticity is -du/dy; at t = 20/3 the ramp 2|sin(pi t/40)| (:2014) is 1 and the profile is
    u = 1.5 (1 - (2y - 3)^2) on y in [1, 2]; at t = 20 it is 2."""
    m = meshes.backward_facing_step(12, 4, 2, length=6.0, step_len=2.0, seed=6, rotate=True)
    K = 4
    o = oracle(m, K, problem_type=2, Re=400.0)
    for _ in range(2):                                 # values are reset each call, not accumulated
        o.update_bcs(t)
    exy, s, tab = _boundary_tables(o, m, "inlet")
    assert np.allclose(exy[:, :, 0], 0.0)
    y = exy[:, :, 1].ravel()
    u = -(tab["BC_normal_vel"] * s).ravel()
    psi = tab["BC_Poisson"].ravel()
    wb = tab["BC_diffusion"].ravel()
    cp = np.polynomial.polynomial.polyfit(y, psi, 3)
    assert np.allclose(np.polynomial.polynomial.polyval(y, cp), psi, atol=1e-13)
    assert np.allclose(np.polynomial.polynomial.polyval(y, np.polynomial.polynomial.polyder(cp)), u, atol=1e-12)
    cu = np.polynomial.polynomial.polyfit(y, u, 2)
    assert np.allclose(-np.polynomial.polynomial.polyval(y, np.polynomial.polynomial.polyder(cu)), wb, atol=1e-12)
    psi_top = _boundary_tables(o, m, "top")[2]["BC_Poisson"]
    psi_bot = _boundary_tables(o, m, "bottom")[2]["BC_Poisson"]
    assert np.ptp(psi_top) == 0.0 and np.all(psi_bot == 0.0)
    assert abs(np.polynomial.polynomial.polyval(1.0, cp) - 0.0) < 1e-13
    assert abs(np.polynomial.polynomial.polyval(2.0, cp) - psi_top[0, 0]) < 1e-13
    wq = o.table("wgt")
    flux = np.sum((u.reshape(-1, K) * wq[None, :]) * (_edge_lengths(m, "inlet") / 2)[:, None])
    assert abs(flux - psi_top[0, 0]) < 1e-13
    if ramp is not None:
        assert abs(psi_top[0, 0] - ramp) < 1e-14
        assert np.allclose(u, ramp * 1.5 * (1 - (2 * y - 3) ** 2), atol=1e-13)


def _channel(nx=10, ny=4, lnod=2, names=None, seed=1, rotate=True):
    m = meshes.backward_facing_step(nx, ny, lnod, length=5.0, height=2.0, step_len=0.0, step_h=0.0,
                                    seed=seed, rotate=rotate)
    if names:
        uniq = []
        for nm in m.bnames:
            if names[nm] not in uniq:
                uniq.append(names[nm])
        ids = np.array([uniq.index(names[nm]) for nm in m.bnames])[m.bedge_bdry].astype(np.int32)
        m = meshes.QuadMesh(m.lnod, m.xy, m.elem, m.bedges, ids, uniq)
    return m


@pytest.mark.parametrize("t,ramp", [(20.0 / 3.0, 1.0), (13.0, None)])
def test_cylinder_channel_inflow_is_a_uniform_stream(t, ramp):
    """problem_type 3 (:2062-2072) on a straight channel: the inlet, top and bottom values are
    those of a uniform stream u = ramp along x: psi = ramp (y - y_min) on the inlet, its volume
    flux on the top, the outward normal velocity -ramp on the inlet, and the stream's tangential
    velocity along the counter-clockwise boundary on top (-ramp) and bottom (+ramp)."""
    m = _channel()
    K = 3
    o = oracle(m, K, problem_type=3, Re=100.0)
    o.update_bcs(t)
    o.update_bcs(t)
    rmp = 2 * abs(np.sin(np.pi * t / 40))
    if ramp is not None:
        assert abs(rmp - ramp) < 1e-15
    exy, s, tab = _boundary_tables(o, m, "inlet")
    y = exy[:, :, 1]
    assert np.allclose(exy[:, :, 0], 0.0)
    assert np.allclose(tab["BC_Poisson"], rmp * (y - 0.0), atol=1e-14)
    assert np.allclose(tab["BC_normal_vel"] * s, -rmp, atol=1e-15)
    flux = np.sum(rmp * o.table("wgt")[None, :] * (_edge_lengths(m, "inlet") / 2)[:, None])
    exy, s, tab = _boundary_tables(o, m, "top")
    assert np.allclose(tab["BC_Poisson"], flux, atol=1e-14)
    assert np.allclose(tab["BC_parl_vel"], -rmp, atol=1e-15)
    exy, s, tab = _boundary_tables(o, m, "bottom")
    assert np.allclose(tab["BC_Poisson"], 0.0) and np.allclose(tab["BC_parl_vel"], rmp, atol=1e-15)


def test_cylinder_object_lies_on_the_dividing_streamline():
    """On the O-grid the cylinder's psi (:2062-2072) is the uniform stream's value at the height
    of its centre: half the inlet's volume flux."""
    m = meshes.cylinder_ogrid(3, 16, 3, r_in=0.5, r_out=4.0)
    o = oracle(m, 3, problem_type=3, Re=100.0)
    t = 20.0 / 3.0
    o.update_bcs(t)
    exy, s, tab = _boundary_tables(o, m, "inlet")
    y = exy[:, :, 1]
    ymin = m.xy[m.bedges[m.bedge_bdry == m.bnames.index("inlet")]][:, :, 1].min()
    assert abs(ymin + 4.0 * np.sin(np.pi / 4)) < 1e-12
    assert np.allclose(tab["BC_Poisson"], y - ymin, atol=1e-13)
    obj = _boundary_tables(o, m, "object")[2]["BC_Poisson"]
    assert np.allclose(obj, 0.0 - ymin, atol=1e-13)


# ---------------------------------------------------------------- hybrid programme (:962-1120, :2104-2191)
@pytest.mark.parametrize("b_slope", [0.0, 0.6])
def test_hybrid_inlet_psi_is_the_marched_volume_flux(b_slope):
    """Hybrid inlets and outlets carry normal velocities vn (positive into the domain); psi is
    marched along the boundary (:962-1120).  For a channel with inflow u = a + b y at x = 0 and
    a uniform outflow that balances it: psi_top - psi_bottom is the inflow's volume flux, the
    inlet's psi rises along y with slope vn edge by edge from psi_bottom to psi_top, the outward
    normal velocity there is -vn, and the inlet vorticity of every inlet edge not next to a wall
    is the central difference of vn along the boundary, -du/dy = -b."""
    m = _channel(8, 6, names={"inlet": "inlet", "outlet": "outlet", "top": "wall", "bottom": "wall"})
    K = 3
    nval = np.zeros(m.n_bedges, dtype=np.int32)
    vals = np.zeros((m.n_bedges, 2))
    e_in = np.where(m.bedge_bdry == m.bnames.index("inlet"))[0]
    e_out = np.where(m.bedge_bdry == m.bnames.index("outlet"))[0]
    ymid = 0.5 * (m.xy[m.bedges[:, 0], 1] + m.xy[m.bedges[:, 1], 1])
    leng = np.linalg.norm(m.xy[m.bedges[:, 1]] - m.xy[m.bedges[:, 0]], axis=1)
    a = 0.7
    nval[e_in] = 1
    vals[e_in, 0] = a + b_slope * ymid[e_in]
    Q = float(np.sum(leng[e_in] * vals[e_in, 0]))
    nval[e_out] = 1
    vals[e_out, 0] = -Q / np.sum(leng[e_out])
    o = Oracle(m, K, problem_type=11, Re=100.0, dt=1e-3, tol=1e-14, maxiter=400, edge_nval=nval, edge_vals=vals)
    o.update_bcs(0.0)
    psi = o.table("BC_Poisson").reshape(-1, K)
    wall = m.bedge_bdry == m.bnames.index("wall")
    top = wall & (ymid > 1.0)
    bot = wall & (ymid < 1.0)
    assert np.ptp(psi[top]) == 0.0 and np.ptp(psi[bot]) == 0.0
    assert abs((psi[top][0, 0] - psi[bot][0, 0]) - Q) < 1e-13
    exy = o.table("edge_xy").reshape(-1, K, 2)
    order = np.argsort(ymid[e_in])
    ys = m.xy[m.bedges[e_in][order]][:, :, 1]
    lo, hi = ys.min(1), ys.max(1)
    base = psi[bot][0, 0] + np.concatenate([[0.0], np.cumsum((hi - lo) * vals[e_in][order, 0])])[:-1]
    for q, e in enumerate(e_in[order]):
        want = base[q] + (exy[e, :, 1] - lo[q]) * vals[e, 0]
        assert np.allclose(psi[e], want, atol=1e-13)
    s = side_sign(o)
    nv = o.table("BC_normal_vel").reshape(-1, K)
    assert np.allclose(nv[e_in] * s[e_in][:, None], -vals[e_in, 0][:, None], atol=1e-15)
    wb = o.table("BC_diffusion").reshape(-1, K)
    inner = e_in[order][1:-1]
    assert len(inner) >= 2
    assert np.allclose(wb[inner], -b_slope, atol=1e-12)


def test_hybrid_open_boundary_values_blend_in_time():
    """Open boundaries (:4050-4117, :2104-2191): a uniform velocity (u, v) given at the open points
    becomes the Neumann datum of psi, d psi/dn = u n_y - v n_x (the tangential velocity along the
    clockwise boundary), and the open vorticity becomes the Dirichlet datum of the vorticity.
    Between two settings the values move affinely across the solveto window, taking one setting
    at each end of it."""
    m = meshes.cavity(4, 2, seed=3, rotate=True)
    names = {"bottom_wall": "wall", "right_wall": "wall", "left_wall": "wall", "top": "open"}
    uniq = ["wall", "open"]
    ids = np.array([uniq.index(names[nm]) for nm in m.bnames])[m.bedge_bdry].astype(np.int32)
    m = meshes.QuadMesh(m.lnod, m.xy, m.elem, m.bedges, ids, uniq)
    K = 3
    o = Oracle(m, K, problem_type=11, Re=100.0, dt=1.0, tol=1e-14, maxiter=400)
    npt = len(o.open_points())
    A, B = np.array([0.8, -0.3]), np.array([-0.2, 0.5])
    o.set_open_vels(np.tile(A, (npt, 1))); o.set_open_vort(np.full(npt, 2.0))
    o.set_open_vels(np.tile(B, (npt, 1))); o.set_open_vort(np.full(npt, -1.0))
    o.solveto(0.5, 1, 1, 100.0)                     # window [0, 0.5]
    openb = m.bedge_bdry == 1
    got = {}
    for t in (0.0, 0.25, 0.5):
        o.update_bcs(t)
        got[t] = (o.table("BC_Poisson").reshape(-1, K)[openb], o.table("BC_diffusion").reshape(-1, K)[openb])
    dpsi = lambda V: V[0] * 1.0 - V[1] * 0.0        # the open lid's outward normal is (0, 1)
    ends = {round(dpsi(A), 12), round(dpsi(B), 12)}
    assert {round(float(got[0.0][0].mean()), 12), round(float(got[0.5][0].mean()), 12)} == ends
    assert np.ptp(got[0.0][0]) < 1e-14 and np.ptp(got[0.5][0]) < 1e-14
    assert np.allclose(got[0.25][0], 0.5 * (dpsi(A) + dpsi(B)), atol=1e-14)
    assert {float(got[0.0][1].mean()), float(got[0.5][1].mean())} == {2.0, -1.0}
    assert np.allclose(got[0.25][1], 0.5, atol=1e-14)
