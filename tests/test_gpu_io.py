"""Mesh file input (hoc_load_gmsh = Mesh::read_msh_file) and VTK output (hoc_write_vtk)
through the C ABI: a mesh read from a Gmsh 4.1 file gives the same solution as the same
mesh passed as arrays, and the VTK file carries the solution-point fields."""
import numpy as np
import pytest

import hoc
from hoc import meshes
from conftest import rel_l2
from test_gmsh_reader import write_gmsh41

pytestmark = pytest.mark.gpu


def test_gmsh_file_equals_arrays(gpu_available, tmp_path):
    m = meshes.cavity(6, 3, seed=4, rotate=True, curve=0.03)
    order = np.argsort(m.bedge_bdry, kind="stable")      # the file groups edges by boundary
    m2 = meshes.QuadMesh(m.lnod, m.xy, m.elem, m.bedges[order], m.bedge_bdry[order], m.bnames)
    path = str(tmp_path / "cav.msh")
    write_gmsh41(m, path)
    kw = dict(K=4, problem_type=1, Re=100.0, dt=1e-4, time_integration_type=4, advection_type=1, tol=1e-13)
    ga = hoc.HO2D(**kw)
    ga.load_mesh(m2)
    ga.setup()
    gf = hoc.HO2D(**kw)
    gf.load_gmsh(path)
    gf.setup()
    assert gf.D == ga.D
    ga.step(3)
    gf.step(3)
    assert rel_l2(gf.get_vorticity(), ga.get_vorticity()) == 0.0
    assert rel_l2(gf.get_streamfunction(), ga.get_streamfunction()) == 0.0


def test_gmsh_errors(gpu_available, tmp_path):
    g = hoc.HO2D(K=3)
    with pytest.raises(hoc.HocError) as e:
        g.load_gmsh(str(tmp_path / "none.msh"))
    assert e.value.code == hoc.HOC_EMESH


def test_write_vtk(gpu_available, tmp_path):
    m = meshes.cavity(3, 2)
    g = hoc.HO2D(K=3, dt=1e-4, Re=100.0, tol=1e-12)
    g.load_mesh(m)
    g.setup()
    g.step(1)
    path = tmp_path / "out.vtk"
    g.write_vtk(str(path))
    txt = path.read_text().splitlines()
    assert txt[0].startswith("# vtk DataFile")
    npts = int([ln for ln in txt if ln.startswith("POINTS")][0].split()[1])
    ncell = int([ln for ln in txt if ln.startswith("CELLS")][0].split()[1])
    assert npts == g.D and ncell == m.n_elem * 4
    i = txt.index("SCALARS vorticity double 1")
    w = np.array([float(v) for v in txt[i + 2:i + 2 + npts]])
    assert rel_l2(w, g.get_vorticity()) < 1e-15
