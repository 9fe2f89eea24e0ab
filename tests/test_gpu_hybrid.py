"""Hybrid operation (problem_type 11; src/HO_2D.cpp:893-1134, 2104-2191, 3694-4278): a
driver feeds open-boundary velocities and vorticity, a vorticity target with weights, and
marches with solveto_d; inlets and outlets carry normal velocities and psi is marched
along the boundary.  GPU against the oracle after every outer step."""
import numpy as np
import pytest

import hoc
from hoc import meshes
from pyoracle import Oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def rename(m, mapping):
    names = []
    for nm in m.bnames:
        if mapping[nm] not in names:
            names.append(mapping[nm])
    ids = np.array([names.index(mapping[nm]) for nm in m.bnames])[m.bedge_bdry]
    return meshes.QuadMesh(m.lnod, m.xy, m.elem, m.bedges, ids.astype(np.int32), names)


def pair(m, K, Re, nval=None, vals=None, tol=1e-13):
    kw = dict(problem_type=11, Re=Re, dt=1.0, time_integration_type=4, advection_type=2, tol=tol, maxiter=500)
    o = Oracle(m, K, edge_nval=nval, edge_vals=vals, **kw)
    g = hoc.HO2D(K=K, **kw)
    g.load_mesh(m)
    if nval is not None:
        g.set_boundary_values(nval, vals)
    g.setup()
    return o, g


@pytest.mark.parametrize("lnod,K", [(2, 4), (3, 3)])
def test_open_lid_driver_sequence(gpu_available, lnod, K):
    m = rename(meshes.cavity(8, lnod, seed=2, rotate=True),
               {"bottom_wall": "wall", "right_wall": "wall", "left_wall": "wall", "top": "open"})
    o, g = pair(m, K, 100.0)
    op = g.open_points()
    assert np.allclose(op, o.open_points(), rtol=0, atol=1e-14)
    xy = g.solution_points()
    wgt = np.where(xy[:, 1] > 0.8, 0.3, 0.0)
    o.set_p2g_weights(wgt); g.set_p2g_weights(wgt)
    for s in range(3):
        uv = np.stack([np.sin(np.pi * op[:, 0]) * (1 + 0.1 * s), 0.05 * op[:, 0]], 1)
        wo = np.cos(np.pi * op[:, 0]) * 0.5 * (1 + s)
        wt = np.sin(2 * np.pi * xy[:, 0]) * xy[:, 1] * (s + 1)
        for side in (o, g):
            if side is o:
                o.set_open_vels(uv); o.set_open_vort(wo); o.set_soln_vort(wt)
            else:
                g.set_open_velocities(uv); g.set_open_vorticity(wo); g.set_vorticity_target(wt)
        o.solveto(2e-3, 2, 4, 100.0)
        g.solveto(2e-3, 2, 4, 100.0)
        ew = rel_l2(g.get_vorticity(), o.vorticity())
        ep = rel_l2(g.get_streamfunction(), o.streamfunction())
        print("outer step", s, "omega", ew, "psi", ep, "t", g.time)
        assert abs(g.time - o.time()) < 1e-15
        assert ew < 1e-10 and ep < 1e-10


def test_inlet_outlet_channel(gpu_available):
    m = meshes.backward_facing_step(16, 4, 2, length=4.0, height=1.0, step_len=0.0, step_h=0.0)
    m = rename(m, {"inlet": "inlet", "outlet": "outlet", "bottom": "wall", "top": "wall"})
    nval = np.zeros(m.n_bedges, dtype=np.int32)
    vals = np.zeros((m.n_bedges, 2))
    for e in range(m.n_bedges):
        nm = m.bnames[m.bedge_bdry[e]]
        if nm in ("inlet", "outlet"):
            nval[e] = 1
            vals[e, 0] = -1.0 if nm == "inlet" else 1.0
    o, g = pair(m, 3, 50.0, nval, vals)
    for s in range(2):
        o.solveto(5e-3, 2, 4, 50.0)
        g.solveto(5e-3, 2, 4, 50.0)
        ew = rel_l2(g.get_vorticity(), o.vorticity())
        ep = rel_l2(g.get_streamfunction(), o.streamfunction())
        print("channel outer step", s, "omega", ew, "psi", ep)
        assert ew < 1e-10 and ep < 1e-10


def test_hybrid_errors(gpu_available):
    g = hoc.HO2D(K=3, problem_type=1, dt=1e-4)
    g.load_mesh(meshes.cavity(3, 2))
    g.setup()
    with pytest.raises(hoc.HocError) as e:
        g.set_vorticity_target(np.zeros(g.D))
    assert e.value.code == hoc.HOC_ESTATE
    m = meshes.backward_facing_step(8, 4, 2, length=4.0, height=1.0, step_len=0.0, step_h=0.0)
    m = rename(m, {"inlet": "inlet", "outlet": "outlet", "bottom": "wall", "top": "wall"})
    g2 = hoc.HO2D(K=3, problem_type=11, dt=1.0)
    g2.load_mesh(m)
    with pytest.raises(hoc.HocError) as e:       # inlets need their normal velocities
        g2.setup()
    assert e.value.code == hoc.HOC_EMESH
