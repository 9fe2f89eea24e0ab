"""Host-side premises of the batched ensemble solve (CPU only).  The Q1 coarse space is built on
the element corners (Gmsh slots 0-3 of every element), so instances that share no corner keep
every level of the preconditioner block-diagonal and iterate each on its own; instances that do
share one fall back to one joint iteration.  meshes.concat must give the first case, and the
corner-merged pair of tests/test_gpu_ensemble.py the second."""
import numpy as np

from hoc import meshes


def corner_sets(m, off):
    c = m.elem[:, :4]
    return [set(np.unique(c[off[i]:off[i + 1]]).tolist()) for i in range(len(off) - 1)]


def test_concat_instances_share_no_corner():
    parts = [meshes.cavity(4, 2), meshes.cavity(6, 2, seed=1, rotate=True),
             meshes.backward_facing_step(16, 6, 2, jitter=0.05, seed=2)]
    m, off = meshes.concat(parts)
    m.validate()
    assert list(off) == [0, 16, 52, 52 + parts[2].n_elem]
    sets = corner_sets(m, off)
    for i in range(len(sets)):
        for j in range(i + 1, len(sets)):
            assert not (sets[i] & sets[j])
    n0 = 0
    for i, p in enumerate(parts):  # every instance's elements reference only its own nodes
        e = m.elem[off[i]:off[i + 1]]
        assert e.min() >= n0 and e.max() < n0 + p.n_nodes
        n0 += p.n_nodes


def test_corner_merged_pair_shares_one_corner():
    a = meshes.cavity(6, 2)
    b0 = meshes.cavity(8, 2)
    b = meshes.QuadMesh(b0.lnod, b0.xy + 1.0, b0.elem, b0.bedges, b0.bedge_bdry, list(b0.bnames))
    m, off = meshes.concat([a, b])
    ia = int(np.flatnonzero(np.all(np.abs(a.xy - 1.0) < 1e-14, axis=1))[0])
    ib = int(np.flatnonzero(np.all(np.abs(b0.xy) < 1e-14, axis=1))[0]) + a.n_nodes
    m.elem[m.elem == ib] = ia
    m.bedges[m.bedges == ib] = ia
    m.validate()
    sets = corner_sets(m, off)
    assert sets[0] & sets[1] == {ia}
    # no face is shared: no element edge (a pair of corners) appears in both instances
    def edges(lo, hi):
        c = m.elem[lo:hi, :4]
        return {tuple(sorted((int(c[k, q]), int(c[k, (q + 1) % 4])))) for k in range(len(c)) for q in range(4)}
    assert not (edges(off[0], off[1]) & edges(off[1], off[2]))
