"""The library's C++ Gmsh 4.1 reader (hoc_load_gmsh; Mesh::read_msh_file,
src/Mesh.cpp:93-293) on the CPU: compiled with g++ from the product's host source and
checked (1) on the reference's shipped meshes against the committed fixtures (made by
tests/golden/make_mesh_fixtures.py, whose connectivity matches the reference's own
src/Mesh.cpp), when the reference tree is present, and (2) on generated meshes of every
geometry order written as Gmsh 4.1 files, against the arrays they were written from."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from hoc import meshes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ho-cxx_amd", "csrc")
GOLD = os.path.join(ROOT, "tests", "golden")
REF_INPUT = "/root/reference/input"


@pytest.fixture(scope="module")
def dump_exe(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("gmsh") / "gmsh_dump"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "cpp", "gmsh_dump.cpp"),
                    os.path.join(CSRC, "hoc_host.cpp"), "-o", str(exe)], check=True)
    return str(exe)


def read_cpp(exe, path, occ):
    out = subprocess.run([exe, path, str(occ)], capture_output=True, text=True)
    lines = out.stdout.splitlines()
    if lines and lines[0].startswith("ERROR"):
        raise ValueError(lines[0])
    i = 0
    lnod = int(lines[i].split()[1]); i += 1
    nn = int(lines[i].split()[1]); i += 1
    xy = np.array([[float(v) for v in lines[i + k].split()] for k in range(nn)]); i += nn
    ne = int(lines[i].split()[1]); i += 1
    elem = np.array([[int(v) for v in lines[i + k].split()] for k in range(ne)], dtype=np.int32); i += ne
    nb = int(lines[i].split()[1]); i += 1
    rows = [[int(v) for v in lines[i + k].split()] for k in range(nb)]; i += nb
    bid = np.array([r[0] for r in rows], dtype=np.int32)
    bed = np.array([r[1:] for r in rows], dtype=np.int32).reshape(nb, lnod)
    nn2 = int(lines[i].split()[1]); i += 1
    names = lines[i:i + nn2]
    return meshes.QuadMesh(lnod, xy, elem, bed, bid, names)


def write_gmsh41(m, path):
    """Minimal Gmsh 4.1 ASCII writer: one curve entity per boundary (physical tag b+1),
    one surface; node tags are 10 + index (a gap, as Gmsh files have)."""
    L = m.lnod
    ltype = {2: 1, 3: 8, 4: 26}[L]
    qtype = {2: 3, 3: 10, 4: 36}[L]
    nb = len(m.bnames)
    with open(path, "w") as f:
        f.write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$PhysicalNames\n")
        f.write(f"{nb + 1}\n")
        for b, nm in enumerate(m.bnames):
            f.write(f'1 {b + 1} "{nm}"\n')
        f.write('2 99 "fluid"\n$EndPhysicalNames\n$Entities\n')
        f.write(f"0 {nb + 1} 1 0\n")
        for b in range(nb):
            f.write(f"{b + 1} 0 0 0 1 1 0 1 {b + 1} 0\n")
        f.write(f"{nb + 1} 0 0 0 1 1 0 0 0\n")   # an untagged curve: contributes no edges
        f.write("1 0 0 0 1 1 0 1 99 0\n$EndEntities\n$Nodes\n")
        n = m.n_nodes
        f.write(f"1 {n} 10 {n + 9}\n2 1 0 {n}\n")
        for k in range(n):
            f.write(f"{10 + k}\n")
        for k in range(n):
            f.write(f"{float(m.xy[k, 0])!r} {float(m.xy[k, 1])!r} 0\n")
        f.write("$EndNodes\n$Elements\n")
        blocks = [(1, b + 1, ltype, m.bedges[m.bedge_bdry == b]) for b in range(nb)]
        blocks.append((1, nb + 1, ltype, m.bedges[:1]))
        blocks.append((2, 1, qtype, m.elem))
        tot = sum(len(x[3]) for x in blocks)
        f.write(f"{len(blocks)} {tot} 1 {tot}\n")
        t = 1
        for dim, tag, typ, arr in blocks:
            f.write(f"{dim} {tag} {typ} {len(arr)}\n")
            for row in arr:
                f.write(f"{t} " + " ".join(str(10 + int(v)) for v in row) + "\n")
                t += 1
        f.write("$EndElements\n")


@pytest.mark.parametrize("lnod,rot", [(2, False), (3, True), (4, True)])
def test_roundtrip_generated(dump_exe, tmp_path, lnod, rot):
    m = meshes.cavity(5, lnod, seed=2, rotate=rot, curve=0.03 if lnod > 2 else 0.0)
    # boundary edges grouped per boundary in the file: reorder the source the same way
    order = np.argsort(m.bedge_bdry, kind="stable")
    path = str(tmp_path / "m.msh")
    write_gmsh41(m, path)
    got = read_cpp(dump_exe, path, 0)
    assert got.lnod == lnod and got.bnames == list(m.bnames)
    assert np.array_equal(got.xy, m.xy)
    assert np.array_equal(got.elem, m.elem)
    assert np.array_equal(got.bedges, m.bedges[order]) and np.array_equal(got.bedge_bdry, m.bedge_bdry[order])
    py = meshes.read_gmsh41(path)
    assert np.array_equal(py.elem, got.elem) and np.array_equal(py.bedges, got.bedges)


def test_occ_reflection_matches_python_reader(dump_exe, tmp_path):
    m = meshes.cavity(4, 3, seed=1)
    path = str(tmp_path / "o.msh")
    write_gmsh41(m, path)
    got = read_cpp(dump_exe, path, 1)
    py = meshes.read_gmsh41(path, occ=1)
    assert np.array_equal(got.elem, py.elem)
    assert not np.array_equal(got.elem, m.elem)


def test_errors(dump_exe, tmp_path):
    with pytest.raises(ValueError):
        read_cpp(dump_exe, str(tmp_path / "missing.msh"), 0)
    p = tmp_path / "bad.msh"
    p.write_text("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
    with pytest.raises(ValueError):
        read_cpp(dump_exe, str(p), 0)


@pytest.mark.skipif(not os.path.isdir(REF_INPUT), reason="reference meshes not present")
@pytest.mark.parametrize("case,fname,occ", [("cavity", "mesh_cavity.msh", 0), ("bfs", "mesh_BFS_linux.msh", 0),
                                           ("cylinder", "mesh_cylinder.msh", 1)])
def test_reference_meshes_match_fixtures(dump_exe, case, fname, occ):
    got = read_cpp(dump_exe, os.path.join(REF_INPUT, fname), occ)
    z = np.load(os.path.join(GOLD, f"mesh_{case}.npz"))
    assert got.lnod == int(z["lnod"])
    assert np.array_equal(got.xy, z["xy"])
    assert np.array_equal(got.elem, z["elem"])
    assert np.array_equal(got.bedges, z["bedges"]) and np.array_equal(got.bedge_bdry, z["bedge_bdry"])
    assert got.bnames == [str(s) for s in z["bnames"]]


@pytest.mark.parametrize("block_count", [7, -1])
def test_malformed_node_blocks_rejected(dump_exe, tmp_path, block_count):
    """A node block larger than the $Nodes header's count (or negative) is an error, not a write
    past the coordinate array.  The file is a valid one with its first node block's count edited."""
    m = meshes.cavity(2, 2)
    good = tmp_path / "g.msh"
    write_gmsh41(m, str(good))
    lines = good.read_text().splitlines()
    k = lines.index("$Nodes")
    n = m.n_nodes
    assert lines[k + 1] == f"1 {n} 10 {n + 9}" and lines[k + 2] == f"2 1 0 {n}"
    lines[k + 1] = f"1 {n - 2} 10 {n + 9}" if block_count > 0 else lines[k + 1]
    lines[k + 2] = f"2 1 0 {n if block_count > 0 else block_count}"
    bad = tmp_path / "b.msh"
    bad.write_text("\n".join(lines) + "\n")
    with pytest.raises(ValueError, match="node block exceeds"):   # checked before any write
        read_cpp(dump_exe, str(bad), 0)
