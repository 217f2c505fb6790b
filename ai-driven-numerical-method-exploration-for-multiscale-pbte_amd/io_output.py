"""``RunRecord``: the machine-readable result of a run, and the run's files.

Field names, types and the derived properties follow the reference's
verifier contract (``solver-python/pybte/io_output.py:30-164``): verifiers read
these names, so they are kept as they are.  The field sampling grid, the
point location rule, the Tecplot layout and the ``RunTime.txt`` line follow
``io_output.py:167-308`` (file formats read by the reference's tools).
"""
from __future__ import annotations

import json
import platform
from dataclasses import dataclass, field, fields
from pathlib import Path

import numpy as np

__all__ = ["RunRecord", "sampling_grid", "locate_points", "sample_dofs", "sample_fields",
           "write_tecplot", "append_runtime_log"]


@dataclass
class RunRecord:
    iterations: int = 0
    converged: bool = False
    residual_history: np.ndarray = field(default_factory=lambda: np.zeros(0))
    residual_true: np.ndarray | None = None
    mass_history: np.ndarray = field(default_factory=lambda: np.zeros(0))
    temp: np.ndarray = field(default_factory=lambda: np.zeros(0))
    qx: np.ndarray = field(default_factory=lambda: np.zeros(0))
    qy: np.ndarray = field(default_factory=lambda: np.zeros(0))
    temp_dofs: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    qx_dofs: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    qy_dofs: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    sweep_count: int = 0
    factorisation_count: int = 0
    wall_clock: float = 0.0
    setup_wall_clock: float = 0.0
    config_hash: str = ""
    scheme: str = "CIS"
    acc_variant: str | None = None
    tol: float = 0.0
    tmax: int = 0
    tau_r: float = 0.0
    tau_n: float = 0.0
    tau_c: float = 0.0
    deg: int = 0
    npole: int = 0
    nazim: int = 0
    n_tris: int = 0
    n_faces: int = 0
    mesh_file: str = ""
    restart: dict = field(default_factory=dict)
    environment: dict = field(default_factory=dict)
    diagnostics: dict = field(default_factory=dict)

    @property
    def final_residual(self) -> float:
        return float(self.residual_history[-1]) if self.residual_history.size else float("nan")

    @property
    def mass(self) -> float:
        return float(self.temp.sum()) if self.temp.size else float("nan")

    @property
    def contraction(self) -> float:
        """Geometric factor fitted to the second half of the residual history."""
        h = np.asarray(self.residual_history, dtype=float)
        if h.size < 8:
            return float("nan")
        tail = h[max(1, h.size // 2):]
        tail = tail[np.isfinite(tail) & (tail > 0)]
        if tail.size < 4 or tail[-1] >= tail[0]:
            return float("nan")
        rho = (tail[-1] / tail[0]) ** (1.0 / (tail.size - 1))
        return float(rho) if 0.0 < rho < 1.0 else float("nan")

    @property
    def error_estimate(self) -> float:
        """``r rho/(1 - rho)``: the sum of the remaining geometric steps."""
        rho = self.contraction
        if not np.isfinite(rho):
            return float("nan")
        return float(self.final_residual * rho / (1.0 - rho))

    # properties that the record's serialised forms carry beside its fields
    DERIVED = ("final_residual", "mass", "contraction", "error_estimate")

    def _arrays(self) -> dict:
        return {f.name: getattr(self, f.name) for f in fields(self)
                if isinstance(getattr(self, f.name), np.ndarray)}

    def to_dict(self, include_fields: bool = True) -> dict:
        """Every field as plain Python values (arrays as lists; the 2-D DOF
        arrays ``None`` when ``include_fields`` is false), then the derived
        quantities of ``DERIVED``."""
        out = {}
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, np.ndarray):
                v = v.tolist() if (include_fields or v.ndim <= 1) else None
            elif isinstance(v, dict):
                v = json.loads(json.dumps(v, default=_json_default))
            out[f.name] = v
        out.update({name: getattr(self, name) for name in self.DERIVED})
        return out

    def to_json(self, path, include_fields: bool = True) -> Path:
        """``run_record.json``: ``to_dict`` as indented JSON."""
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(self.to_dict(include_fields), indent=2, default=_json_default))
        return path

    def to_npz(self, path) -> Path:
        """The arrays of the record as they are, and everything else as the
        JSON string ``meta``."""
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        meta = json.dumps(self.to_dict(include_fields=False), default=_json_default)
        np.savez_compressed(path, meta=np.array(meta), **self._arrays())
        return path

    @staticmethod
    def environment_fingerprint() -> dict:
        """Interpreter, the numerical packages' versions (None when absent;
        numba is never used here) and the machine."""
        def version(module):
            try:
                return __import__(module).__version__
            except ImportError:
                return None
        env = {"python": platform.python_version()}
        env.update({m: version(m) for m in ("numpy", "scipy")})
        env["numba"] = None
        env["torch"] = version("torch")
        env.update(platform=platform.platform(), machine=platform.machine(), backend="hip-gfx950")
        return env


def _json_default(v):
    """JSON form of the numpy values a record's dicts may hold."""
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, np.generic):
        return v.item()
    raise TypeError(f"{type(v).__name__} is not JSON serialisable")


# ---------------------------------------------------------------------------
# field output (io_output.py:167-308)
def sampling_grid(nx: int = 109, ny: int = 109):
    """Sampling coordinates ``(px, py)`` (``io_output.py:167-184``).  At
    109 x 109 both are one set: five points in each outer 2 % of [0, 1] on
    the smootherstep ramp ``0.02 a^3 (10 - 15 a + 6 a^2)``, a = 0, 1/8 ..
    4/8 (mirrored as ``1 - p`` at the far end), and 99 uniform points
    ``0.01 + 0.0098 k`` between, which resolve the kinetic boundary layer;
    other sizes are uniform on [0, 1]."""
    if (nx, ny) != (109, 109):
        return np.linspace(0.0, 1.0, nx), np.linspace(0.0, 1.0, ny)
    a = np.arange(5) / 8.0
    edge = a * a * a * (10.0 - 15.0 * a + 6.0 * a * a) * 0.02
    mid = (np.arange(6, 105) - 5.0) * (1.0 - 0.02) / 100.0 + 0.01
    p = np.concatenate([edge, mid, 1.0 - edge[::-1]])
    return p.copy(), p


def _inside(px, py, xs, ys):
    """The reference's containment test (``io_output.py:200-215``): every
    signed sub-area of the point against the element's edges (vertex
    order) at least -1e-12."""
    ok = np.ones(px.shape, dtype=bool)
    for il in range(3):
        x0, y0 = xs[:, il], ys[:, il]
        x1, y1 = xs[:, (il + 1) % 3], ys[:, (il + 1) % 3]
        area = 0.5 * (px * (y0 - y1) - py * (x0 - x1) + (x0 * y1 - x1 * y0))
        ok &= area >= -1e-12
    return ok


def locate_points(mesh, px, py) -> np.ndarray:
    """``(nx, ny)`` element ids of the grid points ``(px[i], py[j])``, -1
    outside the mesh.  The reference scans the elements in index order and
    takes the first that contains a point (``io_output.py:187-217``); the
    same answer, the smallest containing index, comes here from a bucket
    grid over the elements' bounding boxes (the scan is O(points x
    elements) in Python)."""
    xs = mesh.nodes[mesh.tri_nodes, 0]
    ys = mesh.nodes[mesh.tri_nodes, 1]
    X, Y = np.meshgrid(np.asarray(px, float), np.asarray(py, float), indexing="ij")
    qx, qy = X.ravel(), Y.ravel()
    n = xs.shape[0]
    out = np.full(qx.size, -1, dtype=np.int64)
    if n == 0 or qx.size == 0:
        return out.reshape(X.shape).astype(np.int32)
    # bounding boxes widened well past the test's slack (1e-12 in area)
    pad = 1e-8 * max(1.0, float(np.ptp(mesh.nodes)))
    bx0, bx1 = xs.min(1) - pad, xs.max(1) + pad
    by0, by1 = ys.min(1) - pad, ys.max(1) + pad
    lox, loy = bx0.min(), by0.min()
    g = max(1, int(np.sqrt(n / 2.0)))
    wx = max((bx1.max() - lox) / g, 1e-300)
    wy = max((by1.max() - loy) / g, 1e-300)

    def cell(v, lo, w):
        return np.clip(np.floor((v - lo) / w).astype(np.int64), 0, g - 1)

    i0, i1 = cell(bx0, lox, wx), cell(bx1, lox, wx)
    j0, j1 = cell(by0, loy, wy), cell(by1, loy, wy)
    ni, nj = i1 - i0 + 1, j1 - j0 + 1
    cnt = ni * nj
    tri = np.repeat(np.arange(n), cnt)
    k = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    bins = (np.repeat(i0, cnt) + k // np.repeat(nj, cnt)) * g + np.repeat(j0, cnt) + k % np.repeat(nj, cnt)
    order = np.lexsort((tri, bins))
    bins, tri = bins[order], tri[order]
    starts = np.searchsorted(bins, np.arange(g * g), side="left")
    ends = np.searchsorted(bins, np.arange(g * g), side="right")
    inb = (qx >= lox) & (qx <= lox + g * wx) & (qy >= loy) & (qy <= loy + g * wy)
    pb = cell(qx, lox, wx) * g + cell(qy, loy, wy)
    pts = np.flatnonzero(inb)
    ln = ends[pb[pts]] - starts[pb[pts]]
    pt = np.repeat(pts, ln)
    cand = tri[np.repeat(starts[pb[pts]], ln) + (np.arange(ln.sum()) - np.repeat(np.cumsum(ln) - ln, ln))]
    hit = _inside(qx[pt], qy[pt], xs[cand], ys[cand])
    best = np.full(qx.size, n, dtype=np.int64)
    np.minimum.at(best, pt[hit], cand[hit])
    out = np.where(best < n, best, -1)
    return out.reshape(X.shape).astype(np.int32)


def sample_dofs(mesh, basis, dofs, px, py, triid) -> list:
    """Each ``[P, N]`` DOF array of ``dofs`` evaluated at the grid points:
    the point mapped to its element's reference coordinates (the affine map
    of the integration tensors), the element basis evaluated there; 0
    outside the mesh.  Returns ``(nx, ny)`` arrays."""
    from .integration import _affine
    nxp, nyp = np.size(px), np.size(py)
    xs = mesh.nodes[mesh.tri_nodes, 0]
    ys = mesh.nodes[mesh.tri_nodes, 1]
    (bA, cA, dA), (fE, gE, hE) = _affine(xs, ys)
    X = np.repeat(np.asarray(px, float), nyp)
    Y = np.tile(np.asarray(py, float), nxp)
    k = np.asarray(triid).reshape(-1)
    hit = k >= 0
    kh = k[hit]
    xi = bA[kh] * X[hit] + cA[kh] * Y[hit] + dA[kh]
    eta = fE[kh] * X[hit] + gE[kh] * Y[hit] + hE[kh]
    phi = basis.eval_tri(xi, eta).reshape(kh.size, -1)
    res = []
    for c in dofs:
        c = np.asarray(c)
        v = np.zeros(nxp * nyp)
        v[hit] = np.sum(phi * c[:, kh].T, axis=1)
        res.append(v.reshape(nxp, nyp))
    return res


def sample_fields(mesh, basis, vdf, cxv, cyv, domega, cv, px, py, triid, uq=None, ndof_tri=None):
    """T, qx, qy and the GSIS stress trio (Nxx, Nxy, Nyy) on the grid
    (``io_output.py:220-276``) from a host distribution ``vdf[D, N, P]``:
    the angular moments per element DOF, then ``sample_dofs``; the stress
    rows of ``uq[7P, N]`` as ``-4/3 Lxx + 2/3 Lyy``, ``-(Lxy + Lyx)``,
    ``2/3 Lxx - 4/3 Lyy``; zeros without ``uq``."""
    nd = basis.ndof_tri if ndof_tri is None else ndof_tri
    w = np.asarray(domega, float)
    m0 = np.tensordot(w, vdf, axes=(0, 0)).T / cv
    mx = np.tensordot(np.asarray(cxv) * w, vdf, axes=(0, 0)).T
    my = np.tensordot(np.asarray(cyv) * w, vdf, axes=(0, 0)).T
    return _fields(mesh, basis, m0, mx, my, px, py, triid, uq, nd)


def _fields(mesh, basis, ts, qxs, qys, px, py, triid, uq, nd):
    pT, pqx, pqy = sample_dofs(mesh, basis, (ts, qxs, qys), px, py, triid)
    if uq is None:
        z = np.zeros_like(pT)
        return pT, pqx, pqy, z, z.copy(), z.copy()
    lxx, lxy, lyx, lyy = (uq[r * nd:(r + 1) * nd] for r in (3, 4, 5, 6))
    pxx, pxy, pyy = sample_dofs(mesh, basis, (-4.0 / 3.0 * lxx + 2.0 / 3.0 * lyy, -lxy - lyx,
                                              2.0 / 3.0 * lxx - 4.0 / 3.0 * lyy), px, py, triid)
    return pT, pqx, pqy, pxx, pxy, pyy


def write_tecplot(path, px, py, arrays, names) -> Path:
    """Point-ordered Tecplot ASCII of ``Out_Put_Conduction``
    (``io_output.py:279-291``): a VARIABLES line, ``ZONE I = nx J = ny``,
    then per point (x fastest) the coordinates and values as 16.6E."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    nxp, nyp = np.size(px), np.size(py)
    cols = [np.repeat(np.asarray(px, float)[None, :], nyp, axis=0).ravel(),
            np.repeat(np.asarray(py, float)[:, None], nxp, axis=1).ravel()]
    cols += [np.asarray(a, float).T.ravel() for a in arrays]
    lines = [" VARIABLES=" + ",".join(f'"{n}"' for n in ["x", "y", *names]),
             f" ZONE I = {nxp} J = {nyp}"]
    lines += [" " + "".join(f"{v:16.6E}" for v in row) for row in zip(*cols)]
    path.write_text("\n".join(lines) + "\n")
    return path


def append_runtime_log(path, scheme, tau_r, tau_n, n_tris, npole, nazim, step, seconds) -> Path:
    """One ``RunTime.txt`` line in the Fortran layout
    ``(1X,A4,2ES10.2,3I6,I8,F20.1)`` (``io_output.py:294-308``)."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("a") as fh:
        fh.write(f" {scheme:>4s}{tau_r:10.2E}{tau_n:10.2E}{n_tris:6d}{npole:6d}{nazim:6d}"
                 f"{step:8d}{seconds:20.1f}\n")
    return path
