"""Pins of the oracle's node-averaged output (save_smooth_vtk, src/HO_2D.cpp:3369-3531) and
sample points (read_process_sample_points, src/HO_2D.cpp:319-474; values :3593-3612) by what
the mathematics fixes:

- a field linear in x and y is a polynomial of degree <= 2 per direction on an element with
  quadratic geometry, so K >= 3 solution points extrapolate it exactly to the geometry nodes,
  and every node away from the boundary (whose values come from the boundary arrays) averages
  to the field's value there;
- the isoparametric shape functions reproduce a linear field at the point they are evaluated
  at, so on affine elements (where Newton's method lands on the exact local coordinates in one
  step) the sampled value is exact, and on curved elements it is within the gradient times the
  reference's Newton tolerance (its shape functions are those of the iterate before the last
  update);
- the located element's map sends the local coordinates back to the point (to the reference's
  Newton tolerance), with local coordinates inside [-1, 1];
- points outside the mesh fall back to the element of the nearest mesh node.
A planted mistake (a swapped tensor index, a wrong extrapolation row, a dropped incidence)
breaks the exactness at interior nodes or sample points."""
import numpy as np
import pytest

from hoc import meshes
from pyoracle import Oracle


def lagrange(nodes, x):
    out = np.ones(len(nodes))
    for k in range(len(nodes)):
        for j in range(len(nodes)):
            if j != k:
                out[k] *= (x - nodes[j]) / (nodes[k] - nodes[j])
    return out


GPS = {2: [-1.0, 1.0], 3: [-1.0, 0.0, 1.0], 4: [-1.0, -1 / 3, 1 / 3, 1.0]}
SLOT = {2: [[0, 1], [3, 2]], 3: [[0, 4, 1], [7, 8, 5], [3, 6, 2]],
        4: [[0, 4, 5, 1], [11, 12, 13, 6], [10, 15, 14, 7], [3, 9, 8, 2]]}


def iso_map(m, el, xi, eta):
    L = m.lnod
    bx, by = lagrange(GPS[L], xi), lagrange(GPS[L], eta)
    xy = np.asarray(m.xy).reshape(-1, 2)
    en = np.asarray(m.elem).reshape(m.n_elem, -1)[el]
    p = np.zeros(2)
    for j in range(L):
        for i in range(L):
            p += bx[i] * by[j] * xy[en[SLOT[L][j][i]]]
    return p


def boundary_nodes(m):
    b = set()
    for e in np.asarray(m.bedges).reshape(m.n_bedges, -1):
        b.update(int(v) for v in e)
    return b


def linear(xy):
    return 0.3 + 1.7 * xy[..., 0] - 0.9 * xy[..., 1]


@pytest.mark.parametrize("case", ["affine", "jittered", "curved"])
def test_node_average_reproduces_linear_fields(case):
    if case == "affine":
        m, K, pt = meshes.cavity(6, 2), 3, 1
    elif case == "jittered":
        m, K, pt = meshes.cavity(6, 2, seed=4, jitter=0.15), 4, 1
    else:
        m, K, pt = meshes.cylinder_ogrid(4, 12, 3), 4, 3
    o = Oracle(m, K, problem_type=pt, Re=100.0, dt=1e-4, tol=1e-12)
    sxy = o.table("sps_xy").reshape(-1, 2)
    o.set_vorticity(linear(sxy))
    o.set_streamfunction(2.0 * linear(sxy) - 1.0)
    nf = o.smooth_fields()
    xy = np.asarray(m.xy).reshape(-1, 2)
    inner = np.array(sorted(set(range(m.n_nodes)) - boundary_nodes(m)))
    ew = np.abs(nf[inner, 0] - linear(xy[inner])).max()
    ep = np.abs(nf[inner, 1] - (2.0 * linear(xy[inner]) - 1.0)).max()
    print(case, "interior nodes", len(inner), "err w", ew, "psi", ep)
    assert len(inner) > 0 and ew < 1e-12 and ep < 1e-12


@pytest.mark.parametrize("case", ["affine", "curved"])
def test_sample_points_locate_and_interpolate(case):
    rng = np.random.default_rng(11)
    if case == "affine":
        m, K, pt = meshes.cavity(8, 2), 3, 1
        pts = rng.uniform(0.13, 0.87, size=(40, 2))
    else:
        m, K, pt = meshes.cylinder_ogrid(6, 16, 3), 4, 3
        r = rng.uniform(1.2, 3.0, 40)
        th = rng.uniform(0, 2 * np.pi, 40)
        pts = np.stack([r * np.cos(th), r * np.sin(th)], axis=1)
    o = Oracle(m, K, problem_type=pt, Re=100.0, dt=1e-4, tol=1e-12)
    sxy = o.table("sps_xy").reshape(-1, 2)
    o.set_vorticity(linear(sxy))
    o.set_sample_points(pts)
    el, loc = o.sample_location()
    vals = o.sample_values()
    bnodes = boundary_nodes(m)
    en = np.asarray(m.elem).reshape(m.n_elem, -1)
    worst_map, worst_val = 0.0, 0.0
    for p in range(len(pts)):
        assert np.all(np.abs(loc[p]) <= 1.0 + 1e-9)
        mapped = iso_map(m, el[p], *loc[p])
        worst_map = max(worst_map, np.abs(mapped - pts[p]).max())
        if not bnodes.intersection(int(v) for v in en[el[p]]):  # values of interior elements only
            worst_val = max(worst_val, abs(vals[p, 0] - linear(pts[p])))
    print(case, "map error", worst_map, "value error", worst_val)
    # Newton stops when both steps are <= 1e-3 in local coordinates; the map error follows.  The
    # values use the shape functions of the iterate before the last update (as the reference
    # keeps them): exact on affine elements, within |grad| x (1e-3 of an element) on curved ones.
    assert worst_map < (1e-12 if case == "affine" else 1e-3)
    assert worst_val < (1e-12 if case == "affine" else 5e-3)


def test_sample_point_outside_takes_nearest_node_element():
    m = meshes.cavity(4, 2)
    o = Oracle(m, 3, problem_type=1, Re=100.0, dt=1e-4, tol=1e-12)
    o.set_sample_points(np.array([[1.3, 1.2], [-0.4, 0.1]]))
    el, _ = o.sample_location()
    xy = np.asarray(m.xy).reshape(-1, 2)
    en = np.asarray(m.elem).reshape(m.n_elem, -1)
    for p, q in zip(el, [[1.3, 1.2], [-0.4, 0.1]]):
        d = np.hypot(*(xy - q).T)
        assert np.isclose(d[en[p]].min(), d.min())
