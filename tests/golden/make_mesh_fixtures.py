"""Writes the shipped reference meshes and the reference's own connectivity as fixtures.

Run in a container where the reference tree exists (/root/reference) after building
oracle/_ref (make -C oracle ref).  Outputs (data only):
  tests/golden/mesh_<case>.npz  - nodes, elements, boundary edges of input/mesh_*.msh,
                                  read with hoc.meshes.read_gmsh41 (OCC reorientation applied)
  tests/golden/refconn_<case>.npz - elem_neighbor / boundary_edges produced by the
                                  reference's src/Mesh.cpp (read_msh_file + process_mesh)
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "ho-cxx_amd"))
from hoc import meshes  # noqa: E402

REF = "/root/reference/input"
CASES = {"cavity": ("mesh_cavity.msh", 0), "bfs": ("mesh_BFS_linux.msh", 0), "cylinder": ("mesh_cylinder.msh", 1)}


def ref_connectivity(path, occ):
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_mesh.so"))
    L.ref_mesh_load.restype = C.c_void_p
    L.ref_mesh_load.argtypes = [C.c_char_p, C.c_int]
    for fn in ("ref_mesh_counts", "ref_mesh_edge_nodes", "ref_mesh_elements"):
        getattr(L, fn).argtypes = [C.c_void_p, C.c_void_p]
    L.ref_mesh_nodes.argtypes = [C.c_void_p, C.c_void_p]
    L.ref_mesh_connectivity.argtypes = [C.c_void_p] + [C.c_void_p] * 6
    L.ref_mesh_bname.argtypes = [C.c_void_p, C.c_int]
    L.ref_mesh_bname.restype = C.c_char_p
    L.ref_mesh_bedges.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.ref_mesh_free.argtypes = [C.c_void_p]
    m = L.ref_mesh_load(path.encode(), occ)
    cnt = np.zeros(8, dtype=np.int32)
    L.ref_mesh_counts(m, cnt.ctypes.data)
    N, E, Ln, NB, NN = (int(v) for v in cnt[:5])
    arrs = {k: np.zeros(N * 4, dtype=np.int32) for k in ("nbr", "nbr_side", "is_bnd", "bnd_index")}
    be_elem = np.zeros(E, dtype=np.int32)
    be_side = np.zeros(E, dtype=np.int32)
    L.ref_mesh_connectivity(m, *(arrs[k].ctypes.data for k in ("nbr", "nbr_side", "is_bnd", "bnd_index")),
                            be_elem.ctypes.data, be_side.ctypes.data)
    edge_nodes = np.zeros(E * Ln, dtype=np.int32)
    L.ref_mesh_edge_nodes(m, edge_nodes.ctypes.data)
    elems = np.zeros(N * Ln * Ln, dtype=np.int32)
    L.ref_mesh_elements(m, elems.ctypes.data)
    xy = np.zeros(2 * NN)
    L.ref_mesh_nodes(m, xy.ctypes.data)
    names, bed = [], []
    for b in range(NB):
        names.append(L.ref_mesh_bname(m, b).decode())
        tmp = np.zeros(E, dtype=np.int32)
        n = L.ref_mesh_bedges(m, b, tmp.ctypes.data)
        bed.append(tmp[:n])
    L.ref_mesh_free(m)
    # element corner coordinates identify elements independently of node renumbering
    el_xy = xy.reshape(-1, 2)[elems.reshape(N, Ln * Ln)[:, :4]]
    edge_xy = xy.reshape(-1, 2)[edge_nodes.reshape(E, Ln)]
    out = dict(arrs, be_elem=be_elem, be_side=be_side, elem_corner_xy=el_xy, edge_node_xy=edge_xy,
               bnames=np.array(names), bedge_lists=np.array([",".join(map(str, b)) for b in bed]))
    return out


def main():
    for case, (fn, occ) in CASES.items():
        path = os.path.join(REF, fn)
        m = meshes.read_gmsh41(path, occ=occ)
        np.savez_compressed(os.path.join(HERE, f"mesh_{case}.npz"), lnod=m.lnod, xy=m.xy, elem=m.elem,
                            bedges=m.bedges, bedge_bdry=m.bedge_bdry, bnames=np.array(m.bnames), occ=occ)
        if os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref_mesh.so")):
            np.savez_compressed(os.path.join(HERE, f"refconn_{case}.npz"), **ref_connectivity(path, occ))
        print(case, m.n_elem, "elements", m.n_bedges, "boundary edges", m.bnames)


if __name__ == "__main__":
    main()
