"""Node-averaged output (save_smooth_vtk, src/HO_2D.cpp:3369-3531) and sample points
(read_process_sample_points, src/HO_2D.cpp:319-474; values :3593-3612) on the GPU against the
oracle after the same steps: the state of the last stage (trigger_write) and the refreshed
state of a main-loop dump (src/HO_2D.cpp:1411-1414), on the lid-driven cavity (moving-wall
boundary velocity), the jittered backward-facing step (inflow normal velocity, Neumann outlet)
and the quadratic O-grid around a cylinder (curved faces).  Then the file writers and the
error paths."""
import numpy as np
import pytest

import hoc
from hoc import meshes
from pyoracle import Oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = {
    "cavity": (lambda: meshes.cavity(8, 2, seed=3, rotate=True), 4, dict(problem_type=1, Re=100.0, dt=2e-4)),
    "bfs": (lambda: meshes.backward_facing_step(24, 8, 2, jitter=0.1, seed=5), 4,
            dict(problem_type=2, Re=400.0, dt=1e-3)),
    "cylinder": (lambda: meshes.cylinder_ogrid(6, 16, 3), 4, dict(problem_type=3, Re=100.0, dt=2e-4)),
}


def pair(name, steps=3):
    mk, K, kw = CASES[name]
    m = mk()
    common = dict(time_integration_type=4, advection_type=1, tol=1e-14, maxiter=400, **kw)
    o = Oracle(m, K, **common)
    g = hoc.HO2D(K=K, **common)
    g.load_mesh(m)
    g.setup()
    o.step(steps)
    g.step(steps)
    return m, o, g


def sample_points(m, n=60, seed=1):
    rng = np.random.default_rng(seed)
    xy = np.asarray(m.xy).reshape(-1, 2)
    lo, hi = xy.min(0), xy.max(0)
    pts = rng.uniform(lo, hi, size=(n, 2))
    if n >= 10:  # mesh nodes, edge midpoints and points outside the mesh as well
        en = np.asarray(m.elem).reshape(m.n_elem, -1)
        pts[:5] = xy[en[:5, 0]]
        pts[5:8] = 0.5 * (xy[en[5:8, 0]] + xy[en[5:8, 1]])
        pts[8:10] = hi + [0.3, 0.1], lo - [0.2, 0.4]
    return pts


def scaled_errors(got, ref, scale_from):
    """Max differences of vorticity, psi and the velocity vector, each relative to the largest
    magnitude of that quantity in scale_from (the node fields): a component that is nearly zero
    somewhere does not inflate the error."""
    sw = np.abs(scale_from[:, 0]).max()
    sp = np.abs(scale_from[:, 1]).max()
    su = np.hypot(scale_from[:, 2], scale_from[:, 3]).max()
    if got.shape[1] == 4:
        return [np.abs(got[:, 0] - ref[:, 0]).max() / sw, np.abs(got[:, 1] - ref[:, 1]).max() / sp,
                np.hypot(*(got[:, 2:] - ref[:, 2:]).T).max() / su]
    return [np.abs(got[:, 0] - ref[:, 0]).max() / sw, np.hypot(*(got[:, 1:] - ref[:, 1:]).T).max() / su]


@pytest.mark.parametrize("name", ["cavity", "bfs", "cylinder"])
def test_node_fields_and_samples_match_oracle(gpu_available, name):
    m, o, g = pair(name)
    ref = o.smooth_fields()
    got = g.node_fields()
    errs = scaled_errors(got, ref, ref)
    pts = sample_points(m)
    o.set_sample_points(pts)
    g.set_sample_points(pts)
    el_o, loc_o = o.sample_location()
    el_g, loc_g = g.sample_location()
    sv_o, sv_g = o.sample_values(), g.sample_values()
    es = scaled_errors(sv_g, sv_o, ref)
    print(name, "node fields w psi |u|", errs, "samples w |u|", es, "loc", np.abs(loc_g - loc_o).max())
    assert np.array_equal(el_g, el_o)
    assert np.abs(loc_g - loc_o).max() < 1e-12
    # the fields themselves agree with the oracle's to ~1e-11 after these steps
    assert max(errs) < 1e-10 and max(es) < 1e-10
    # the boundary replacement is exercised: boundary-node velocities are the boundary values
    assert np.abs(ref[:, 2:]).max() > 0


@pytest.mark.parametrize("name", ["cavity", "bfs"])
def test_refreshed_dump_matches_oracle(gpu_available, name):
    m, o, g = pair(name, steps=2)
    o.refresh()
    ref = o.smooth_fields()
    got = g.node_fields(refresh=True)
    errs = scaled_errors(got, ref, ref)
    print(name, "refreshed node fields", errs)
    assert max(errs) < 1e-10
    # the refresh recomputed psi and the velocity for the current vorticity
    assert rel_l2(g.get_streamfunction(), o.streamfunction()) < 1e-10


def test_output_files(gpu_available, tmp_path):
    m, o, g = pair("cavity", steps=1)
    pts = sample_points(m, n=7)
    g.set_sample_points(pts)
    vtk = tmp_path / "smooth.vtk"
    dat = tmp_path / "samples.dat"
    g.write_smooth_vtk(str(vtk))
    g.write_samples(str(dat))
    txt = vtk.read_text().splitlines()
    assert txt[0].startswith("# vtk DataFile") and f"POINTS {m.n_nodes} double" in txt
    assert txt.count("SCALARS vorticity double 1 ") == 1 and "VECTORS velocity double " in txt
    rows = np.loadtxt(dat)
    assert rows.shape == (7, 5)
    assert np.allclose(rows[:, :2], pts, rtol=1e-5, atol=1e-6)
    assert np.allclose(rows[:, 2:], g.sample_values(), rtol=1e-5, atol=1e-6)
    # the file reader drops a repeated last pair, as the reference does
    spf = tmp_path / "pts.txt"
    spf.write_text("\n".join(f"{x} {y}" for x, y in pts) + f"\n{pts[-1][0]} {pts[-1][1]}\n")
    xy = g.load_sample_points(str(spf))
    assert len(xy) == 7


def test_output_errors(gpu_available):
    m, o, g = pair("cavity", steps=0)
    with pytest.raises(hoc.HocError) as e:
        g.node_fields(n_nodes=m.n_nodes + 1)
    assert e.value.code == hoc.HOC_EINVAL
    g.set_sample_points(np.array([[0.5, 0.5]]))
    buf = np.zeros(6)
    assert hoc.lib().hoc_get_sample_values(g.h, 0, 2, buf.ctypes.data) == hoc.HOC_EINVAL
    with pytest.raises(hoc.HocError) as e:
        g.set_sample_points(np.array([[np.nan, 0.5]]))
    assert e.value.code == hoc.HOC_EINVAL
    g.set_sample_points(np.zeros((0, 2)))
    assert g.sample_values().shape == (0, 3)
