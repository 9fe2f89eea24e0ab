"""Unstructured quad meshes for the HO-CXX hot path, in the form the C ABI takes.

A mesh is a :class:`QuadMesh`: node coordinates, element node lists in Gmsh
order (CCW corners first, then edge nodes, then interior nodes), and the
boundary edges as Gmsh lines (two end nodes, then interior nodes) each tagged
with a named boundary.  This is the information the reference's
``HO_2D::load_mesh_arrays_d`` (src/HO_2D.cpp:3823-3881) and
``Mesh::read_msh_file`` (src/Mesh.cpp:93-340) hand to ``Mesh::process_mesh``.

Generators here build the structured-as-unstructured workloads (lid-driven
cavity, backward-facing step) and read Gmsh 4.1 ASCII files.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import numpy as np


@dataclass
class QuadMesh:
    lnod: int                       # geometry nodes per edge (2, 3 or 4)
    xy: np.ndarray                  # (n_nodes, 2) float64
    elem: np.ndarray                # (n_elem, lnod*lnod) int32, Gmsh order
    bedges: np.ndarray              # (n_bedges, lnod) int32, Gmsh line order
    bedge_bdry: np.ndarray          # (n_bedges,) int32 boundary id
    bnames: list = field(default_factory=list)

    @property
    def n_elem(self) -> int:
        return int(self.elem.shape[0])

    @property
    def n_nodes(self) -> int:
        return int(self.xy.shape[0])

    @property
    def n_bedges(self) -> int:
        return int(self.bedges.shape[0])

    def validate(self) -> None:
        L2 = self.lnod * self.lnod
        if self.elem.ndim != 2 or self.elem.shape[1] != L2:
            raise ValueError("elem must be (n_elem, lnod^2)")
        if self.bedges.ndim != 2 or self.bedges.shape[1] != self.lnod:
            raise ValueError("bedges must be (n_bedges, lnod)")
        if self.elem.min() < 0 or self.elem.max() >= self.n_nodes:
            raise ValueError("element node id out of range")
        if self.n_bedges and (self.bedges.min() < 0 or self.bedges.max() >= self.n_nodes):
            raise ValueError("boundary edge node id out of range")
        if self.n_bedges and (self.bedge_bdry.min() < 0 or self.bedge_bdry.max() >= len(self.bnames)):
            raise ValueError("boundary id out of range")


def tensor_to_gmsh(lnod: int) -> np.ndarray:
    """Gmsh quad node index at tensor position (i along xi, j along eta): t[j, i].

    Same table as src/Mesh.cpp:34-91 (tensor2FEM)."""
    if lnod == 2:
        return np.array([[0, 1], [3, 2]])
    if lnod == 3:
        return np.array([[0, 4, 1], [7, 8, 5], [3, 6, 2]])
    if lnod == 4:
        return np.array([[0, 4, 5, 1], [11, 12, 13, 6], [10, 15, 14, 7], [3, 9, 8, 2]])
    raise ValueError("lnod must be 2, 3 or 4")


def line_order(lnod: int) -> list:
    """Tensor position -> Gmsh line node slot (src/Mesh.cpp:27-32, tensor2FEM(i))."""
    return [0 if i == 0 else (1 if i == lnod - 1 else i + 1) for i in range(lnod)]


def _structured(nx: int, ny: int, lnod: int, x_of, keep=None, rng=None, rotate=False,
                flip_bedges=False):
    """Structured grid of nx*ny cells (tensor product) emitted as an unstructured mesh.

    x_of(s, t) maps lattice coordinates in [0,nx]x[0,ny] (fractional for edge nodes)
    to physical coordinates.  keep(ci, cj) -> bool drops cells.  Boundary sides are
    classified by a callback-free rule: a side with no kept neighbour is a boundary."""
    p = lnod - 1
    NX, NY = nx * p + 1, ny * p + 1
    gid = -np.ones((NY, NX), dtype=np.int64)
    cells = [(ci, cj) for cj in range(ny) for ci in range(nx) if keep is None or keep(ci, cj)]
    kept = np.zeros((nx, ny), dtype=bool)
    for ci, cj in cells:
        kept[ci, cj] = True
    # number only nodes used by kept cells
    n = 0
    for ci, cj in cells:
        for b in range(lnod):
            for a in range(lnod):
                J, I = cj * p + b, ci * p + a
                if gid[J, I] < 0:
                    gid[J, I] = n
                    n += 1
    xy = np.zeros((n, 2))
    for J in range(NY):
        for I in range(NX):
            if gid[J, I] >= 0:
                xy[gid[J, I]] = x_of(I / p, J / p)
    t2g = tensor_to_gmsh(lnod)
    elem = np.zeros((len(cells), lnod * lnod), dtype=np.int32)
    rot = np.zeros(len(cells), dtype=np.int64)
    if rotate:
        rot = rng.integers(0, 4, size=len(cells))
    for e, (ci, cj) in enumerate(cells):
        T = gid[cj * p:cj * p + lnod, ci * p:ci * p + lnod]  # T[j, i]
        T = np.rot90(T, int(rot[e]))                         # rotation keeps CCW orientation
        for j in range(lnod):
            for i in range(lnod):
                elem[e, t2g[j, i]] = T[j, i]
    return xy, elem, gid, kept, cells


def _side_nodes(gid, ci, cj, side, lnod):
    """Lattice node ids along cell side (CCW w.r.t. the cell), tensor order."""
    p = lnod - 1
    I0, J0 = ci * p, cj * p
    if side == 0:    # south, west->east
        return [gid[J0, I0 + a] for a in range(lnod)]
    if side == 1:    # east, south->north
        return [gid[J0 + a, I0 + p] for a in range(lnod)]
    if side == 2:    # north, east->west
        return [gid[J0 + p, I0 + p - a] for a in range(lnod)]
    return [gid[J0 + p - a, I0] for a in range(lnod)]  # west, north->south


def _emit_bedges(tensor_lists, lnod, rng=None, flip=False):
    lo = line_order(lnod)
    out = np.zeros((len(tensor_lists), lnod), dtype=np.int32)
    for k, nodes in enumerate(tensor_lists):
        nodes = list(nodes)
        if flip and rng is not None and rng.integers(0, 2):
            nodes = nodes[::-1]
        for i in range(lnod):
            out[k, lo[i]] = nodes[i]
    return out


def cavity(n: int, lnod: int = 2, *, seed: int | None = None, rotate: bool = False,
           jitter: float = 0.0, curve: float = 0.0, flip_bedges: bool = False) -> QuadMesh:
    """Unit-square lid-driven cavity, n x n quads stored as unstructured lists.

    Boundaries (src/HO_2D.cpp:700-726, problem_type 1): "bottom_wall", "right_wall",
    "top" (the lid), "left_wall", named as in the reference's input/mesh_cavity.msh.
    rotate: random quarter-turn of each element's local frame (exercises every
    neighbour-side/orientation pairing); jitter: interior-node displacement as a
    fraction of h; curve: amplitude of a smooth interior warp (curved elements
    when lnod >= 3)."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    p = lnod - 1
    NX = n * p + 1
    jit = np.zeros((NX, NX, 2))
    if jitter > 0:
        jit = rng.uniform(-jitter * h, jitter * h, size=(NX, NX, 2))
        jit[0, :, :] = jit[-1, :, :] = jit[:, 0, :] = jit[:, -1, :] = 0.0

    def x_of(s, t):
        x, y = s * h, t * h
        if curve:
            dx = curve * np.sin(np.pi * x) * np.sin(2 * np.pi * y)
            dy = curve * np.sin(2 * np.pi * x) * np.sin(np.pi * y)
            x, y = x + dx, y + dy
        I, J = int(round(s * p)), int(round(t * p))
        return np.array([x + jit[J, I, 0], y + jit[J, I, 1]])

    xy, elem, gid, kept, cells = _structured(n, n, lnod, x_of, rng=rng, rotate=rotate)
    names = ["bottom_wall", "right_wall", "top", "left_wall"]
    lists, ids = [], []
    for b, side in enumerate([0, 1, 2, 3]):
        for k in range(n):
            ci, cj = {0: (k, 0), 1: (n - 1, k), 2: (n - 1 - k, n - 1), 3: (0, n - 1 - k)}[side]
            lists.append(_side_nodes(gid, ci, cj, side, lnod))
            ids.append(b)
    bedges = _emit_bedges(lists, lnod, rng, flip_bedges)
    m = QuadMesh(lnod, xy, elem, bedges, np.asarray(ids, dtype=np.int32), names)
    m.validate()
    return m


def backward_facing_step(nx: int = 384, ny: int = 64, lnod: int = 2, *, length: float = 30.0,
                         height: float = 2.0, step_len: float = 5.0, step_h: float = 1.0,
                         jitter: float = 0.0, seed: int | None = None,
                         rotate: bool = False) -> QuadMesh:
    """Channel [0,length]x[0,height] minus the step [0,step_len]x[0,step_h].

    Boundaries as problem_type 2 expects (src/HO_2D.cpp:728-807): "inlet" (x=0 above
    the step), "outlet" (x=length), "bottom" (step faces and floor), "top"."""
    rng = np.random.default_rng(seed)
    hx, hy = length / nx, height / ny
    p = lnod - 1
    sx, sy = int(round(step_len / hx)), int(round(step_h / hy))

    def keep(ci, cj):
        return not (ci < sx and cj < sy)

    NX, NY = nx * p + 1, ny * p + 1
    jit = np.zeros((NY, NX, 2))
    if jitter > 0:
        jit = rng.uniform(-1, 1, size=(NY, NX, 2)) * jitter * np.array([hx, hy])

    def on_boundary(I, J):
        x, y = I / p, J / p
        if I == 0 or I == NX - 1 or J == 0 or J == NY - 1:
            return True
        if x <= sx and y <= sy and (abs(x - sx) < 1e-12 or abs(y - sy) < 1e-12):
            return True
        return False

    def x_of(s, t):
        I, J = int(round(s * p)), int(round(t * p))
        d = np.zeros(2) if on_boundary(I, J) else jit[J, I]
        return np.array([s * hx + d[0], t * hy + d[1]])

    xy, elem, gid, kept, cells = _structured(nx, ny, lnod, x_of, keep=keep, rng=rng, rotate=rotate)
    names = ["inlet", "bottom", "top", "outlet"]
    lists, ids = [], []
    for ci, cj in cells:
        for side, (di, dj) in enumerate([(0, -1), (1, 0), (0, 1), (-1, 0)]):
            ni, nj = ci + di, cj + dj
            if 0 <= ni < nx and 0 <= nj < ny and kept[ni, nj]:
                continue
            if side == 3 and ci == 0:
                b = 0
            elif side == 1 and ci == nx - 1:
                b = 3
            elif side == 2 and cj == ny - 1:
                b = 2
            else:
                b = 1
            lists.append(_side_nodes(gid, ci, cj, side, lnod))
            ids.append(b)
    order = np.argsort(np.asarray(ids), kind="stable")
    lists = [lists[i] for i in order]
    ids = [ids[i] for i in order]
    bedges = _emit_bedges(lists, lnod)
    m = QuadMesh(lnod, xy, elem, bedges, np.asarray(ids, dtype=np.int32), names)
    m.validate()
    return m


def cylinder_ogrid(n_r: int = 256, n_theta: int = 512, lnod: int = 3, *, r_in: float = 0.5, r_out: float = 10.0,
                   seed: int | None = None, rotate: bool = False) -> QuadMesh:
    """O-grid of n_r (radial) x n_theta (azimuthal) isoparametric quads around a cylinder of
    radius r_in centred at the origin, out to radius r_out, with geometric radial stretching.
    Geometry nodes of every order lie on the exact circles and rays.  Element axes: xi along
    the radius, eta along the angle, which makes every element counter-clockwise.

    Boundaries as problem_type 3 expects (src/HO_2D.cpp:809-891): "object" (the cylinder),
    and the outer circle split by angle into "outlet" (-45..45 deg), "top" (45..135),
    "inlet" (135..225) and "bottom" (225..315)."""
    rng = np.random.default_rng(seed)
    p = lnod - 1
    NT, NR = n_theta * p, n_r * p + 1
    gid = np.arange(NT * NR, dtype=np.int64).reshape(NT, NR)      # [theta][r], theta periodic
    th = 2.0 * np.pi * np.arange(NT) / NT
    rr = r_in * (r_out / r_in) ** (np.arange(NR) / (NR - 1))
    xy = np.zeros((NT * NR, 2))
    xy[:, 0] = (np.cos(th)[:, None] * rr[None, :]).ravel()
    xy[:, 1] = (np.sin(th)[:, None] * rr[None, :]).ravel()
    t2g = tensor_to_gmsh(lnod)
    cells = [(ci, cj) for cj in range(n_theta) for ci in range(n_r)]
    elem = np.zeros((len(cells), lnod * lnod), dtype=np.int32)
    rot = rng.integers(0, 4, size=len(cells)) if rotate else np.zeros(len(cells), dtype=np.int64)
    sides = {"object": [], "outer": []}
    for e, (ci, cj) in enumerate(cells):
        T = np.array([[gid[(cj * p + j) % NT, ci * p + i] for i in range(lnod)] for j in range(lnod)])  # T[j, i]
        if ci == 0:           # xi = -1 side: the cylinder, CCW order (eta decreasing)
            sides["object"].append([T[p - a, 0] for a in range(lnod)])
        if ci == n_r - 1:     # xi = +1 side: the outer circle, CCW order (eta increasing)
            sides["outer"].append([T[a, p] for a in range(lnod)])
        T = np.rot90(T, int(rot[e]))
        for j in range(lnod):
            for i in range(lnod):
                elem[e, t2g[j, i]] = T[j, i]
    names = ["inlet", "outlet", "top", "bottom", "object"]
    bed, bid = [], []
    for nodes in sides["outer"]:
        mid = xy[nodes[0]] + xy[nodes[-1]]
        ang = np.degrees(np.arctan2(mid[1], mid[0])) % 360.0
        nm = "outlet" if (ang < 45 or ang >= 315) else "top" if ang < 135 else "inlet" if ang < 225 else "bottom"
        bed.append(nodes)
        bid.append(names.index(nm))
    for nodes in sides["object"]:
        bed.append(nodes)
        bid.append(names.index("object"))
    m = QuadMesh(lnod, xy, elem, _emit_bedges(bed, lnod), np.array(bid, dtype=np.int32), names)
    m.validate()
    return m


def read_gmsh41(path: str, occ: int = 0) -> QuadMesh:
    """Read a Gmsh 4.1 ASCII mesh as src/Mesh.cpp:93-293 does.

    Only curves carrying a physical tag contribute boundary edges, in file order;
    2-D elements keep file order; OCC-format files have their element nodes
    reflected into CCW order (src/Mesh.cpp:283-293)."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    pos = {ln.strip(): i for i, ln in enumerate(lines) if ln.startswith("$")}
    # physical names of dimension 1
    i = pos["$PhysicalNames"] + 1
    nphys = int(lines[i]); i += 1
    phys_tags, names = [], []
    for _ in range(nphys):
        parts = lines[i].split(None, 2); i += 1
        if int(parts[0]) != 1:
            continue
        phys_tags.append(int(parts[1]))
        names.append(parts[2].strip().strip('"'))
    # entities: curve tag -> first physical tag (0 if none)
    i = pos["$Entities"] + 1
    npt, ncu, nsu, nvo = map(int, lines[i].split()[:4]); i += 1
    i += npt
    curve_phys = {}
    for _ in range(ncu):
        parts = lines[i].split(); i += 1
        tag = int(parts[0])
        nph = int(parts[7])
        curve_phys[tag] = int(parts[8]) if nph > 0 else 0
    # nodes
    i = pos["$Nodes"] + 1
    nblocks, nnodes, _, _ = map(int, lines[i].split()); i += 1
    tag2idx = {}
    xy = np.zeros((nnodes, 2))
    c = 0
    for _ in range(nblocks):
        _, _, _, nb = map(int, lines[i].split()); i += 1
        tags = [int(lines[i + k]) for k in range(nb)]
        i += nb
        for k in range(nb):
            xs = lines[i + k].split()
            tag2idx[tags[k]] = c
            xy[c] = float(xs[0]), float(xs[1])
            c += 1
        i += nb
    # elements
    i = pos["$Elements"] + 1
    nblocks = int(lines[i].split()[0]); i += 1
    edge_n = {1: 2, 8: 3, 26: 4, 27: 5, 28: 6}
    face_n = {3: 4, 10: 9, 16: 8, 36: 16, 37: 25}
    bed, bid, els = [], [], []
    lnod = None
    for _ in range(nblocks):
        dim, etag, etype, nb = map(int, lines[i].split()); i += 1
        block = lines[i:i + nb]; i += nb
        if dim == 1:
            if etype not in edge_n:
                raise ValueError(f"unsupported line element type {etype}")
            lnod = edge_n[etype]
            ph = curve_phys.get(etag, 0)
            if ph == 0:
                continue
            b = phys_tags.index(ph)
            for ln in block:
                v = [tag2idx[int(t)] for t in ln.split()[1:]]
                bed.append(v)
                bid.append(b)
        elif dim == 2:
            if etype not in face_n:
                raise ValueError(f"unsupported face element type {etype}")
            for ln in block:
                els.append([tag2idx[int(t)] for t in ln.split()[1:]])
    elem = np.asarray(els, dtype=np.int32)
    if lnod is None:
        lnod = int(round(np.sqrt(elem.shape[1])))
    if occ:
        t2g = tensor_to_gmsh(lnod)
        new = elem.copy()
        L1 = lnod - 1
        for jo in range(lnod):
            for io in range(lnod):
                jn, inew = L1 - io, L1 - jo
                new[:, t2g[jn, inew]] = elem[:, t2g[jo, io]]
        elem = new
    m = QuadMesh(lnod, xy, elem, np.asarray(bed, dtype=np.int32),
                 np.asarray(bid, dtype=np.int32), names)
    m.validate()
    return m


def concat(parts):
    """Union of independent meshes (an ensemble): nodes, elements and boundary edges are
    renumbered instance by instance; boundaries are merged by name.  Returns the union and
    the element offsets of the instances (length len(parts) + 1)."""
    lnod = parts[0].lnod
    if any(p.lnod != lnod for p in parts):
        raise ValueError("all instances must have the same geometry order")
    names = []
    for p in parts:
        for nm in p.bnames:
            if nm not in names:
                names.append(nm)
    xy, elem, bed, bid, off = [], [], [], [], [0]
    n0 = 0
    for p in parts:
        xy.append(p.xy)
        elem.append(p.elem + n0)
        bed.append(p.bedges + n0)
        remap = np.array([names.index(nm) for nm in p.bnames], dtype=np.int32)
        bid.append(remap[p.bedge_bdry])
        n0 += p.n_nodes
        off.append(off[-1] + p.n_elem)
    m = QuadMesh(lnod, np.concatenate(xy), np.concatenate(elem).astype(np.int32),
                 np.concatenate(bed).astype(np.int32), np.concatenate(bid).astype(np.int32), names)
    return m, np.array(off, dtype=np.int32)
