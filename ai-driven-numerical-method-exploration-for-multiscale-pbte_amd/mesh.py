"""Triangle meshes: gmsh 2.2 ASCII reading, face numbering, connectivity,
geometry, and writers for the structured / jittered unit-square meshes.

Face numbering follows ``solver-python/pybte/mesh/topology.py:125-274``:
elements are visited in file order; a line element appends a boundary face
(tagged with the *last* matching boundary index); a triangle appends each of
its edges not yet present as an undirected node pair.  Triangle ids are the
tie-breaker of the sweep order, so the numbering is load-bearing.

Geometry follows ``mesh/geometry.py:35-79``: signed area
``0.5(x3y1 - x2y1 + x1y2 - x3y2 - x1y3 + x2y3)``, outward unit normal of edge
``j`` (node j -> node j+1) ``(-dy, dx)/L`` with ``d = p_j - p_{j+1}``, and
heights ``2A/L``.

The writers follow the reference's mesh family convention
(``solver-python/tools/make_meshes.py:60-140``): nodes numbered x-major
(``id = i*(n+1) + j + 1``), boundary lines south (+x), north (-x), east (+y),
west (-y) with physical ids 11..14, and per-quad diagonals chosen by
``(2i >= n) == (2j >= n)`` (on ``j`` only for the periodic channel).  The
square with a circular hole (physical id 16) is not in the reference: the
reference has no mesher for it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .constants import BC_SYMMETRY, BC_THERMALISING

__all__ = ["Mesh", "GmshMesh", "read_gmsh22", "build_mesh", "triangle_geometry",
           "write_square_mesh", "square_mesh_text"]

LINE_ELEMENT = 1
TRIANGLE_ELEMENT = 2


@dataclass
class GmshMesh:
    nodes: np.ndarray
    elem_type: np.ndarray
    elem_phyid: np.ndarray
    elem_nodes: np.ndarray      # [n_elem, 3], -1 after a line's two nodes
    physical_names: dict


def read_gmsh22(path) -> GmshMesh:
    """A gmsh 2.2 ASCII file (``mesh/gmsh_reader.py:45-122``), parsed by
    ``pbte_gmsh_read``: nodes (ids must run 1..N in order: the node index is
    the position in the file), line and triangle elements (at least two
    tags; the first is the physical id), physical names.  Any other element
    type, a format other than 2.2 ASCII or a short section is a
    ``ValueError``.  ``elem_nodes`` is ``[n_elem, 3]`` (0-based; -1 after a
    line's two nodes)."""
    from . import _native
    lib = _native.lib()
    err = C.create_string_buffer(1024)
    h = lib.pbte_gmsh_read(str(path).encode(), err, 1024)
    if not h:
        raise ValueError(err.value.decode())
    try:
        n = (C.c_int64 * 3)()
        lib.pbte_gmsh_sizes(h, n)
        nn, ne, nb = int(n[0]), int(n[1]), int(n[2])
        xy = np.empty((nn, 2))
        kind = np.empty(ne, dtype=np.int32)
        phy = np.empty(ne, dtype=np.int32)
        conn = np.empty((ne, 3), dtype=np.int32)
        names = C.create_string_buffer(max(nb, 1))
        lib.pbte_gmsh_copy(h, xy.ctypes.data, kind.ctypes.data, phy.ctypes.data, conn.ctypes.data,
                           C.cast(names, C.c_void_p))
    finally:
        lib.pbte_gmsh_free(h)
    table = {}
    for line in names.raw[:nb].decode().splitlines():
        dim, tag, label = line.split(None, 2)
        table[(int(dim), int(tag))] = label.strip().strip('"')
    return GmshMesh(nodes=xy, elem_type=kind.astype(np.int64), elem_phyid=phy.astype(np.int64),
                    elem_nodes=conn, physical_names=table)


def triangle_geometry(nodes: np.ndarray, tri_nodes: np.ndarray):
    """``(area, normal[n,3,2], hmin_per_tri, hmin_global)``."""
    x = nodes[tri_nodes, 0]
    y = nodes[tri_nodes, 1]
    x1, x2, x3 = x[:, 0], x[:, 1], x[:, 2]
    y1, y2, y3 = y[:, 0], y[:, 1], y[:, 2]
    area = 0.5 * (x3 * y1 - x2 * y1 + x1 * y2 - x3 * y2 - x1 * y3 + x2 * y3)
    if np.any(area < 0):
        bad = np.flatnonzero(area < 0)
        raise ValueError(f"{bad.size} triangles have negative area; first: {bad[:5].tolist()}")
    xn = np.concatenate([x, x[:, :1]], axis=1)
    yn = np.concatenate([y, y[:, :1]], axis=1)
    dx = xn[:, :3] - xn[:, 1:]
    dy = yn[:, :3] - yn[:, 1:]
    length = np.sqrt(dx * dx + dy * dy)
    normal = np.empty((tri_nodes.shape[0], 3, 2))
    normal[:, :, 0] = -dy / length
    normal[:, :, 1] = dx / length
    heights = 2.0 * area[:, None] / np.sqrt(dx ** 2 + dy ** 2)
    return area, normal, heights.min(axis=1), min(1.0, float(heights.min()))


@dataclass
class Mesh:
    nodes: np.ndarray
    tri_nodes: np.ndarray
    tri_faces: np.ndarray
    tri_area: np.ndarray
    tri_normal: np.ndarray
    tri_hmin: np.ndarray
    hmin: float
    face_nodes: np.ndarray
    face_bc: np.ndarray
    face_pair: np.ndarray
    face_tri: np.ndarray
    face_lfc: np.ndarray
    face_len: np.ndarray
    tri_neighbour: np.ndarray
    bc_codes: np.ndarray

    @property
    def n_nodes(self) -> int:
        return self.nodes.shape[0]

    @property
    def n_tris(self) -> int:
        return self.tri_nodes.shape[0]

    @property
    def n_faces(self) -> int:
        return self.face_nodes.shape[0]

    @property
    def n_faces_b(self) -> int:
        return int(np.count_nonzero(self.face_bc >= 0))

    @property
    def n_faces_i(self) -> int:
        return self.n_faces - self.n_faces_b

    def adjacent(self, face: int):
        """``(triangle, local edge)`` touching a boundary face; the (+) side
        is preferred, as in ``topology.py:115-122``."""
        if self.face_tri[face, 0] < 0:
            return int(self.face_tri[face, 1]), int(self.face_lfc[face, 1])
        return int(self.face_tri[face, 0]), int(self.face_lfc[face, 0])

    def adjacent_arrays(self):
        """Vectorised :meth:`adjacent` over all faces."""
        minus = self.face_tri[:, 0] < 0
        tri = np.where(minus, self.face_tri[:, 1], self.face_tri[:, 0])
        lfc = np.where(minus, self.face_lfc[:, 1], self.face_lfc[:, 0])
        return tri.astype(np.int64), lfc.astype(np.int64)


def build_mesh(path, bc_phyid, bc_codes, bc_names=None, bc_xoff=None, bc_yoff=None) -> Mesh:
    g = read_gmsh22(path)
    nbc = len(bc_phyid)
    bc_phyid = np.asarray(bc_phyid, dtype=np.int64)
    bc_codes = np.asarray(bc_codes, dtype=np.int64)
    bc_names = list(bc_names) if bc_names is not None else [""] * nbc
    bc_xoff = np.asarray(bc_xoff if bc_xoff is not None else np.zeros(nbc), dtype=float)
    bc_yoff = np.asarray(bc_yoff if bc_yoff is not None else np.zeros(nbc), dtype=float)

    # faces in order of first appearance: boundary lines as read, then the
    # edges (j, j+1) of every triangle; a face keeps the orientation of its
    # first appearance
    et = np.asarray(g.elem_type)
    is_line = et == LINE_ELEMENT
    if np.count_nonzero(~is_line) == 0:
        raise ValueError(f"{path}: no triangles")
    line_idx = np.flatnonzero(is_line)
    tri_idx = np.flatnonzero(~is_line)
    tri_nodes_a = g.elem_nodes[tri_idx].astype(np.int64).reshape(-1, 3)
    lines_a = g.elem_nodes[line_idx, :2].astype(np.int64).reshape(-1, 2)
    # occurrence stream in element order
    pos_line = line_idx
    pos_tri = tri_idx
    ea = np.concatenate([lines_a[:, 0], tri_nodes_a[:, [0, 1, 2]].reshape(-1)])
    eb = np.concatenate([lines_a[:, 1], tri_nodes_a[:, [1, 2, 0]].reshape(-1)])
    order_key = np.concatenate([pos_line * 4, (pos_tri[:, None] * 4 + np.arange(3)[None, :]).reshape(-1)])
    srt = np.argsort(order_key, kind="stable")
    ea, eb = ea[srt], eb[srt]
    is_l = np.concatenate([np.ones(len(lines_a), bool), np.zeros(3 * len(tri_nodes_a), bool)])[srt]
    lo = np.minimum(ea, eb)
    hi = np.maximum(ea, eb)
    nmax = int(max(int(hi.max()) + 1, 1))
    key = lo * nmax + hi
    ukey, first, inv = np.unique(key, return_index=True, return_inverse=True)
    # a boundary line must be the first appearance of its edge
    seen = first[inv] != np.arange(key.size)
    if np.any(seen & is_l):
        o = int(np.flatnonzero(seen & is_l)[0])
        raise ValueError(f"{path}: duplicate boundary edge {(int(ea[o]) + 1, int(eb[o]) + 1)}")
    rank = np.empty(ukey.size, dtype=np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(ukey.size)
    face_of = rank[inv]                                  # face id of every occurrence
    n_faces = int(ukey.size)
    fo = np.sort(first)
    face_nodes_a = np.stack([ea[fo], eb[fo]], axis=1).astype(np.int32)
    face_bc_a = np.full(n_faces, -1, dtype=np.int32)
    if line_idx.size:
        phy = np.asarray(g.elem_phyid)[line_idx]
        bcl = np.full(phy.size, -1, dtype=np.int32)
        for j in range(nbc):
            bcl = np.where(phy == bc_phyid[j], j, bcl)
        face_bc_a[face_of[is_l]] = bcl                   # lines keep their file order
    n_tris = len(tri_nodes_a)
    tri_nodes_a = tri_nodes_a.astype(np.int32)
    face_pair = np.full(n_faces, -1, dtype=np.int32)

    # periodic pairing between the boundaries named 'Master' and 'Slave'
    id_master = next((k for k, s in enumerate(bc_names) if str(s).strip() == "Master"), -1)
    id_slave = next((k for k, s in enumerate(bc_names) if str(s).strip() == "Slave"), -1)
    if id_master >= 0:
        if id_slave < 0:
            raise ValueError("a 'Master' boundary needs a matching 'Slave'")
        masters = np.flatnonzero(face_bc_a == id_master)
        slaves = np.flatnonzero(face_bc_a == id_slave)
        xo, yo = bc_xoff[id_master], bc_yoff[id_master]
        P = g.nodes
        for m in masters:
            a, b = face_nodes_a[m]
            pa = (P[a, 0] + xo, P[a, 1] + yo)
            pb = (P[b, 0] + xo, P[b, 1] + yo)
            for s in slaves:
                c, d = face_nodes_a[s]
                pc = (P[c, 0], P[c, 1])
                pd = (P[d, 0], P[d, 1])
                if (pa == pc and pb == pd) or (pa == pd and pb == pc):
                    face_pair[m] = s
                    face_pair[s] = m
                    break
        unpaired = [int(k) for k in np.concatenate([masters, slaves]) if face_pair[k] < 0]
        if unpaired:
            raise ValueError(f"unpaired periodic faces: {unpaired[:8]}")

    area, normal, hmin_tri, hmin = triangle_geometry(g.nodes, tri_nodes_a)
    fn = g.nodes[face_nodes_a]
    d = fn[:, 0, :] - fn[:, 1, :]
    face_len = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1])

    tri_faces = face_of[~is_l].reshape(n_tris, 3).astype(np.int32)
    face_tri = np.full((n_faces, 2), -1, dtype=np.int32)
    face_lfc = np.full((n_faces, 2), -1, dtype=np.int32)
    ta = tri_nodes_a[:, [0, 1, 2]]
    same = face_nodes_a[tri_faces, 0] == ta                  # first-appearance orientation
    tri_id = np.repeat(np.arange(n_tris, dtype=np.int32)[:, None], 3, axis=1)
    lfc_id = np.repeat(np.arange(3, dtype=np.int32)[None, :], n_tris, axis=0)
    for slot, m in ((0, same), (1, ~same)):
        face_tri[tri_faces[m], slot] = tri_id[m]
        face_lfc[tri_faces[m], slot] = lfc_id[m]

    # wall / symmetry queue index overwrites face_pair (topology.py:238-243)
    counter = 0
    for k in range(n_faces):
        bc = face_bc_a[k]
        if bc >= 0 and bc_codes[bc] in (BC_THERMALISING, BC_SYMMETRY):
            face_pair[k] = counter
            counter += 1

    t1 = face_tri[tri_faces, 0]
    t2 = face_tri[tri_faces, 1]
    own = np.arange(n_tris)[:, None]
    tri_neighbour = np.where(t1 == own, t2, np.where(t2 == own, t1, -1)).astype(np.int32)

    mesh = Mesh(nodes=g.nodes, tri_nodes=tri_nodes_a, tri_faces=tri_faces, tri_area=area,
                tri_normal=normal, tri_hmin=hmin_tri, hmin=hmin, face_nodes=face_nodes_a,
                face_bc=face_bc_a, face_pair=face_pair, face_tri=face_tri,
                face_lfc=face_lfc, face_len=face_len, tri_neighbour=tri_neighbour,
                bc_codes=bc_codes)
    interior = mesh.face_bc < 0
    if np.any((mesh.face_tri[interior] < 0).any(axis=1)):
        bad = np.flatnonzero(interior & (mesh.face_tri < 0).any(axis=1))
        raise ValueError(f"interior faces with a missing neighbour: {bad[:8].tolist()}")
    boundary = ~interior
    if np.any((mesh.face_tri[boundary] >= 0).all(axis=1)):
        bad = np.flatnonzero(boundary & (mesh.face_tri >= 0).all(axis=1))
        raise ValueError(f"boundary faces with two neighbours: {bad[:8].tolist()}")
    return mesh


# ---------------------------------------------------------------------------
# writers
# ---------------------------------------------------------------------------
def _smootherstep(t):
    return t ** 3 * (10.0 - 15.0 * t + 6.0 * t ** 2)


_HEADER = ("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$PhysicalNames\n5\n"
           '1 11 "{s} (physical id 11)"\n1 12 "{n} (physical id 12)"\n'
           '1 13 "{e} (physical id 13)"\n1 14 "{w} (physical id 14)"\n'
           '2 15 "Fluid (physical id 15)"\n$EndPhysicalNames\n')


def square_mesh_text(n: int, layout: str = "cavity", distribution: str = "uniform",
                     jitter: float = 0.0, seed: int = 0, hole=None) -> str:
    """gmsh 2.2 text of an ``n x n`` triangulated unit square.

    ``jitter`` moves every interior node by ``U(-jitter*h, +jitter*h)`` in x and
    y, drawn from ``numpy.random.default_rng(seed)`` in node-id order (x draw
    then y draw per node); ``h = 1/n``.  Boundary nodes stay on the boundary
    lines so that wall physical ids are unchanged.

    ``hole = (x0, y0, r)`` removes every triangle whose centroid lies within
    ``r`` of ``(x0, y0)`` (after the jitter); the edges left with one triangle
    inside the square become the boundary "Hole" (physical id 16), each line
    oriented as its triangle's edge (the domain on its left, as the outer
    walls), and the nodes no triangle uses are dropped (ids renumbered in
    order).  The hole is a staircase of mesh edges around the circle.
    """
    if distribution not in ("uniform", "smootherstep"):
        raise ValueError(f"unknown distribution {distribution!r}")
    t = np.arange(n + 1) / n
    cy = _smootherstep(t) if distribution == "smootherstep" else t.astype(float)
    cx = cy.copy() if layout != "channel" else t.astype(float)
    nn = n + 1
    X = np.repeat(cx[:, None], nn, axis=1)          # X[i, j] = x_i
    Y = np.repeat(cy[None, :], nn, axis=0)          # Y[i, j] = y_j
    if jitter:
        rng = np.random.default_rng(seed)
        h = 1.0 / n
        d = rng.uniform(-jitter * h, jitter * h, size=(nn, nn, 2))
        inner = np.zeros((nn, nn), dtype=bool)
        inner[1:n, 1:n] = True
        X = np.where(inner, X + d[:, :, 0], X)
        Y = np.where(inner, Y + d[:, :, 1], Y)
    names = {"s": "Swall", "n": "NWall", "e": "EWall", "w": "WWall"}
    if layout == "channel":
        names = {"s": "Swall", "n": "NWall", "e": "Master", "w": "Slave"}
    elif layout != "cavity":
        raise ValueError(f"unknown layout {layout!r}")

    def nid(i, j):
        return i * nn + j + 1

    xs = X.reshape(-1)
    ys = Y.reshape(-1)
    lines = []
    lines += [(11, 5, nid(i, 0), nid(i + 1, 0)) for i in range(n)]
    lines += [(12, 6, nid(i, n), nid(i - 1, n)) for i in range(n, 0, -1)]
    lines += [(13, 7, nid(n, j), nid(n, j + 1)) for j in range(n)]
    lines += [(14, 8, nid(0, j), nid(0, j - 1)) for j in range(n, 0, -1)]
    flip_on_x = layout != "channel"
    tris = []
    for i in range(n):
        for j in range(n):
            if ((2 * i >= n) if flip_on_x else False) == (2 * j >= n):
                tris.append((nid(i, j), nid(i + 1, j), nid(i + 1, j + 1)))
                tris.append((nid(i, j), nid(i + 1, j + 1), nid(i, j + 1)))
            else:
                tris.append((nid(i, j), nid(i + 1, j), nid(i, j + 1)))
                tris.append((nid(i, j + 1), nid(i + 1, j), nid(i + 1, j + 1)))
    header = _HEADER.format(**names)
    if hole is not None:
        x0, y0, r = (float(v) for v in hole)
        t = np.asarray(tris, dtype=np.int64) - 1
        gx = xs[t].mean(axis=1)
        gy = ys[t].mean(axis=1)
        t = t[(gx - x0) ** 2 + (gy - y0) ** 2 > r * r]
        # edges used by one kept triangle and not on the outer boundary
        outer = {tuple(sorted((a - 1, b - 1))) for _, _, a, b in lines}
        count = {}
        for tri in t:
            for k in range(3):
                e = (int(tri[k]), int(tri[(k + 1) % 3]))
                key = (min(e), max(e))
                count[key] = count.get(key, 0) + 1
        hole_lines = []
        for tri in t:
            for k in range(3):
                a0, b0 = int(tri[k]), int(tri[(k + 1) % 3])
                key = (min(a0, b0), max(a0, b0))
                if count[key] == 1 and key not in outer:
                    hole_lines.append((a0, b0))
        used = np.zeros(nn * nn, dtype=bool)
        used[t.reshape(-1)] = True
        new = np.cumsum(used) * used                     # 1-based new ids, 0 = dropped
        xs, ys = xs[used], ys[used]
        lines = [(p, g, int(new[a - 1]), int(new[b - 1])) for p, g, a, b in lines]
        lines += [(16, 9, int(new[a]), int(new[b])) for a, b in hole_lines]
        tris = [tuple(int(new[v]) for v in tri) for tri in t]
        header = header.replace("$PhysicalNames\n5\n", "$PhysicalNames\n6\n").replace(
            "$EndPhysicalNames", '1 16 "Hole (physical id 16)"\n$EndPhysicalNames')
    out = [header, "$Nodes\n", f"{xs.size:6d}\n"]
    ids = np.arange(1, xs.size + 1)
    out.extend(f"{k:6d} {x:.16f} {y:.16f} 0\n" for k, x, y in zip(ids, xs, ys))
    out.append("$EndNodes\n")
    out.append("$Elements\n")
    out.append(f"{len(lines) + len(tris):6d}\n")
    k = 0
    for phy, geo, a, b in lines:
        k += 1
        out.append(f"{k:6d} 1 2 {phy} {geo} {a:5d} {b:5d}\n")
    for a, b, c in tris:
        k += 1
        out.append(f"{k:6d} 2 2 15 10 {a:5d} {b:5d} {c:5d}\n")
    out.append("$EndElements\n")
    return "".join(out)


def write_square_mesh(path, n: int, **kw) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(square_mesh_text(n, **kw))
    return path
