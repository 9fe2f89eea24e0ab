"""The C++ facade HO_2D_GPU (ho-cxx_amd/csrc/HO_2D_GPU.hpp, the reference's HO_2D method names,
src/HO_2D.hpp:250-290) driven by a compiled C++ caller, ho-cxx_amd/examples/cavity_facade.cpp:
load_mesh_arrays_d + add_boundary, process_mesh_input, read_process_sample_points, solveto_d,
the getters (getallvorts_d, getsolnpts_d, getsolnareas_d, get_hoquad_weights_d) and
trigger_write.  Its fields are compared with the oracle on the same mesh."""
import os
import subprocess

import numpy as np
import pytest

from hoc import meshes
from pyoracle import Oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ho-cxx_amd", "examples", "cavity_facade")


def facade_mesh(n):
    """The example's mesh: row-major nodes, elements CCW from the SW corner, boundary "wall"
    (bottom, right, left) then "top" (the lid), edges in the example's order."""
    xy = np.array([[i / n, j / n] for j in range(n + 1) for i in range(n + 1)])
    nid = lambda i, j: j * (n + 1) + i
    elem = np.array([[nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)] for j in range(n) for i in range(n)],
                    dtype=np.int32)
    wall = [(nid(i, 0), nid(i + 1, 0)) for i in range(n)] + [(nid(n, j), nid(n, j + 1)) for j in range(n)]
    top = [(nid(i, n), nid(i - 1, n)) for i in range(n, 0, -1)]
    wall += [(nid(0, j), nid(0, j - 1)) for j in range(n, 0, -1)]
    bed = np.array(wall + top, dtype=np.int32)
    bid = np.array([0] * len(wall) + [1] * len(top), dtype=np.int32)
    return meshes.QuadMesh(2, xy, elem, bed, bid, ["wall", "top"])


def test_facade_example_vs_oracle(gpu_available, tmp_path):
    assert os.path.exists(EXE), "build ho-cxx_amd (make) first: the example is built with libhoc.so"
    n, K, steps, dt = 8, 4, 5, 1e-4
    pts = np.array([[0.5, 0.5], [0.25, 0.8], [0.9, 0.1]])
    sp = tmp_path / "sample_points.dat"
    sp.write_text("".join(f"{x:.17g} {y:.17g}\n" for x, y in pts))
    out = tmp_path / "fields.bin"
    r = subprocess.run([EXE, str(n), str(K), str(steps), repr(dt), str(out), str(sp)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    D = n * n * K * K
    v = np.fromfile(out, dtype=np.float64)
    assert v.size == 7 * D + K * K
    w, psi, uv, xy, area, hq = (v[:D], v[D:2 * D], v[2 * D:4 * D].reshape(-1, 2), v[4 * D:6 * D].reshape(-1, 2),
                                v[6 * D:7 * D], v[7 * D:])
    o = Oracle(facade_mesh(n), K, problem_type=1, Re=100.0, dt=dt, time_integration_type=4, advection_type=1,
               tol=1e-14, maxiter=500)
    assert np.allclose(xy, o.table("sps_xy").reshape(-1, 2), rtol=0, atol=1e-15)
    assert np.allclose(area, o.table("areas"), rtol=1e-14, atol=0)
    wq = o.table("wgt")
    assert np.allclose(hq, 0.25 * np.outer(wq, wq).ravel(), rtol=1e-15, atol=0)
    o.step(steps)
    ew, ep, ev = rel_l2(w, o.vorticity()), rel_l2(psi, o.streamfunction()), rel_l2(uv, o.velocity())
    print("facade vs oracle: omega", ew, "psi", ep, "velocity", ev)
    assert ew < 1e-11 and ep < 1e-11 and ev < 1e-11
    files = sorted(os.listdir(tmp_path))
    assert "vorticity1_K4_00000.vtk" in files and "problem1_K4_00000.vtk" in files, files
    assert "sampling_points1_K4_L2_00000.dat" in files, files
    o.set_sample_points(pts)
    want = o.sample_values()
    rows = np.loadtxt(tmp_path / "sampling_points1_K4_L2_00000.dat").reshape(-1, 5)
    assert np.allclose(rows[:, :2], pts, rtol=1e-5, atol=1e-6)        # the file's text precision
    assert np.allclose(rows[:, 2:], want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
