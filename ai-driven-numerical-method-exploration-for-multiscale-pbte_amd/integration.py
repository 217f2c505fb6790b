"""Precomputed element and face integrals of the nodal DG discretisation.

Index orders are the reference's (``solver-python/pybte/integration.py:159-196``,
``docs/INDEXING.md:41-56``)::

    int_tri[L, I]               int_tri_tri[M, L, I]
    int_tri_tri_x[I, M, L]      int_tri_tri_y[I, M, L]
    int_tri_fc[I, IL, M, L]     int_tri_fc_fc[M, L, I, IL]
    int_fc_fc[M, L, F]          int_fc[M, F]
    int_tri_tri_fc[I, IL, M, L]

Two techniques, as in the reference (``integration.py:199-404``):

* analytic monomial moments ``int_T xi^a eta^b = a! b!/(a+b+2)!`` (and edge
  analogues) contracted with the basis coefficients -- mass, derivative and
  same-element edge tensors;
* a 15-point Gauss-Legendre rule along each face -- the element x face-basis
  tensor and the element x neighbour tensor (``int_tri_tri_fc``); on a wall
  face the latter holds ``int phi_M ds`` broadcast over ``L``.  These face
  loops run in ``pbte_face_integrals`` (``csrc/host_setup.cpp``) on host
  threads, element by element in the reference's expression order.

The contraction loops keep the reference's summation orders (they matter in
the last bit).  Monomial powers x^e are 1, x and x*x for e <= 2 (as numpy's
power gives them) and libm's ``pow`` for e >= 3; numpy's own power of an
array is a vector routine on CPUs with AVX-512 that differs from libm's by
one unit in the last place in a few per cent of cases, so for degrees 3 and
4 the reference's own tensors depend on the CPU it ran on to that extent.
"""
from __future__ import annotations

import ctypes as C_
from dataclasses import dataclass

import numpy as np

from . import _native
from .basis import Basis
from .constants import BC_PERIODIC
from .mesh import Mesh
from .quadrature import gauss_legendre_core, tri_quadrature

__all__ = ["Integrals", "build_integrals", "face_partners", "fac_div"]


def fac_div(ns: int, ne: int, ds: int, de: int) -> float:
    """``prod(ns..ne) / prod(ds..de)`` evaluated term by term."""
    numn = ne - ns + 1
    numd = de - ds + 1
    ans = 1.0
    for i in range(1, min(numn, numd) + 1):
        ans = ans * float(ns + i - 1) / float(ds + i - 1)
    for i in range(numd + 1, numn + 1):
        ans = ans * float(ns + i - 1)
    for i in range(numn + 1, numd + 1):
        ans = ans / float(ds + i - 1)
    return ans


def _affine(x, y):
    """Physical -> reference map ``xi = b x + c y + d``, ``eta = f x + g y + h``."""
    x1, x2, x3 = x[..., 0], x[..., 1], x[..., 2]
    y1, y2, y3 = y[..., 0], y[..., 1], y[..., 2]
    A = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)
    B = y3 - y1
    C = x1 - x3
    D = (x3 - x1) * y1 - (y3 - y1) * x1
    E = (x3 - x1) * (y2 - y1) - (x2 - x1) * (y3 - y1)
    F = y2 - y1
    G = x1 - x2
    H = (x2 - x1) * y1 - (y2 - y1) * x1
    return (B / A, C / A, D / A), (F / E, G / E, H / E)


def _pow(x, e: int):
    if e == 0:
        return np.ones_like(x)
    if e == 1:
        return x.copy()
    return np.power(x, float(e))


def _eval_mono(C, mono, xi, eta):
    """``sum_k C[m,k] xi^a_k eta^b_k`` accumulated in ascending ``k``."""
    out = np.zeros(np.broadcast(xi, eta).shape + (C.shape[0],))
    for k, (a, b) in enumerate(mono):
        out += (C[:, k] * _pow(xi, int(a))[..., None]) * _pow(eta, int(b))[..., None]
    return out


def _contract2(C, E, il_outer: bool):
    n, m = C.shape[0], E.shape[0]
    out = np.zeros((n, n))
    if il_outer:
        for il in range(m):
            for im in range(m):
                out += (C[:, il][None, :] * C[:, im][:, None]) * E[im, il]
    else:
        for im in range(m):
            for il in range(m):
                out += (C[:, im][:, None] * C[:, il][None, :]) * E[im, il]
    return out


def _contract1(C, e):
    out = np.zeros(C.shape[0])
    for im in range(e.size):
        out += C[:, im] * e[im]
    return out


@dataclass
class Integrals:
    int_tri: np.ndarray
    int_tri_tri: np.ndarray
    int_tri_tri_x: np.ndarray
    int_tri_tri_y: np.ndarray
    int_tri_fc: np.ndarray
    int_tri_fc_fc: np.ndarray
    int_fc_fc: np.ndarray
    int_fc: np.ndarray
    int_tri_tri_fc: np.ndarray
    qua_abs_x: np.ndarray
    qua_abs_y: np.ndarray
    qua_wei: np.ndarray
    nodfun_qua_p: np.ndarray
    ref_tri: np.ndarray
    ref_tri_tri: np.ndarray
    # the element operators' factors: [6][P][P] reference matrices (mass,
    # d/dxi, d/deta, edges 1-3) and [N][8] cell scalars (2 area, gradient
    # coefficients, edge lengths as the edge tensors scale them)
    ref_ops: np.ndarray | None = None
    cell_ops: np.ndarray | None = None


def face_partners(mesh: Mesh, bc_codes=None, bc_xoff=None, bc_yoff=None):
    """Per (element, local face) the element across -- the interior
    neighbour, or for a periodic face the element behind the partner face --
    and the shift of its nodes (the partner boundary's offset); ``-1`` and
    zero on a wall (``integration.py:321-340``)."""
    n_tris = mesh.n_tris
    bc_codes = np.asarray(bc_codes) if bc_codes is not None else np.zeros(0, dtype=int)
    nbc = bc_codes.size
    bc_xoff = np.asarray(bc_xoff, dtype=float) if bc_xoff is not None else np.zeros(nbc)
    bc_yoff = np.asarray(bc_yoff, dtype=float) if bc_yoff is not None else np.zeros(nbc)
    adj_tri, _ = mesh.adjacent_arrays()
    other = np.full((n_tris, 3), -1, dtype=np.int32)
    off_x = np.zeros((n_tris, 3))
    off_y = np.zeros((n_tris, 3))
    for j2 in range(3):
        fcid = mesh.tri_faces[:, j2].astype(np.int64)
        fbc = mesh.face_bc[fcid].astype(np.int64)
        per = np.zeros(n_tris, dtype=bool)
        if bc_codes.size:
            per = (fbc >= 0) & (bc_codes[np.clip(fbc, 0, None)] == BC_PERIODIC)
        interior = fbc < 0
        t1 = mesh.face_tri[fcid, 0]
        t2 = mesh.face_tri[fcid, 1]
        o = np.where(t1 == np.arange(n_tris), t2, t1).astype(np.int64)
        if per.any():
            pair = mesh.face_pair[fcid[per]].astype(np.int64)
            o[per] = adj_tri[pair]
            bcp = mesh.face_bc[pair].astype(np.int64)
            off_x[per, j2] = bc_xoff[bcp]
            off_y[per, j2] = bc_yoff[bcp]
        other[:, j2] = np.where(interior | per, o, -1)
    return other, off_x, off_y


def reference_element(basis: Basis):
    """The reference-element integrals every element tensor scales:
    ``(ref_tri, ref_tri_tri, ref_dx, ref_dy, r1, r2, r3)`` -- the basis
    integrals, the mass matrix, the two derivative matrices and the three
    edge matrices (``integration.py:199-260``)."""
    P = basis.ndof_tri
    mono = basis.mono
    C = basis.nodfun_tri
    ix, iy = mono[:, 0], mono[:, 1]

    e_vec = np.array([fac_div(1, int(iy[m]), int(ix[m]) + 1, int(ix[m] + iy[m]) + 2)
                      for m in range(P)])
    e_mass = np.empty((P, P))
    e_dx = np.zeros((P, P))
    e_dy = np.zeros((P, P))
    e_e1 = np.zeros((P, P))
    e_e2 = np.empty((P, P))
    e_e3 = np.zeros((P, P))
    sqrt2 = np.sqrt(2.0)
    for m in range(P):
        im, jm = int(ix[m]), int(iy[m])
        for l in range(P):
            il, jl = int(ix[l]), int(iy[l])
            e_mass[m, l] = fac_div(1, jl + jm, il + im + 1, il + im + jl + jm + 2)
            if im > 0:
                e_dx[m, l] = im * fac_div(1, jl + jm, il + im, il + im - 1 + jl + jm + 2)
            if jm > 0:
                e_dy[m, l] = jm * fac_div(1, jl + jm - 1, il + im + 1,
                                          il + im - 1 + jl + jm + 2)
            if jm == 0 and jl == 0:
                e_e1[m, l] = 1.0 / float(im + il + 1)
            e_e2[m, l] = sqrt2 * fac_div(1, jm + jl, im + il + 1, im + il + jm + jl + 1)
            if im == 0 and il == 0:
                e_e3[m, l] = 1.0 / float(jm + jl + 1)

    return (_contract1(C, e_vec), _contract2(C, e_mass, True), _contract2(C, e_dx, True),
            _contract2(C, e_dy, True), _contract2(C, e_e1, False), _contract2(C, e_e2, False),
            _contract2(C, e_e3, False))


def build_integrals(mesh: Mesh, basis: Basis, np_tri: int, np_fc: int,
                    bc_codes=None, bc_xoff=None, bc_yoff=None, threads: int = 0) -> Integrals:
    P = basis.ndof_tri
    Q = basis.ndof_fc
    deg = basis.deg
    mono = basis.mono
    C = basis.nodfun_tri
    Cf = basis.nodfun_fc
    nodes = mesh.nodes
    n_tris = mesh.n_tris
    area2 = 2.0 * mesh.tri_area
    sqrt2 = np.sqrt(2.0)
    ref_tri, ref_tri_tri, ref_dx, ref_dy, r1, r2, r3 = reference_element(basis)
    int_tri = ref_tri[:, None] * area2[None, :]
    int_tri_tri = ref_tri_tri[:, :, None] * area2[None, None, :]

    x = nodes[mesh.tri_nodes, 0]
    y = nodes[mesh.tri_nodes, 1]
    x1, x2, x3 = x[:, 0], x[:, 1], x[:, 2]
    y1, y2, y3 = y[:, 0], y[:, 1], y[:, 2]
    ss = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)
    int_tri_tri_x = (((y3 - y1) / ss)[:, None, None] * ref_dx[None]
                     + ((y1 - y2) / ss)[:, None, None] * ref_dy[None]) * area2[:, None, None]
    int_tri_tri_y = (((x1 - x3) / ss)[:, None, None] * ref_dx[None]
                     + ((x2 - x1) / ss)[:, None, None] * ref_dy[None]) * area2[:, None, None]

    len12 = np.sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1))
    len23 = np.sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2))
    len31 = np.sqrt((x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1))
    int_tri_fc = np.empty((n_tris, 3, P, P))
    int_tri_fc[:, 0] = len12[:, None, None] * r1[None]
    int_tri_fc[:, 1] = (len23 / sqrt2)[:, None, None] * r2[None]
    int_tri_fc[:, 2] = len31[:, None, None] * r3[None]

    e_line = np.zeros((Q, Q))
    for m in range(deg + 1):
        for l in range(deg + 1):
            if (m + l) % 2 == 0:
                e_line[m, l] = 2.0 / float(m + l + 1)
    ref_fc_fc = _contract2(Cf, e_line, False)
    ref_fc = _contract1(Cf, e_line[:, 0])
    half_len = mesh.face_len * 0.5
    int_fc_fc = ref_fc_fc[:, :, None] * half_len[None, None, :]
    int_fc = ref_fc[:, None] * half_len[None, :]

    # -- face integrals (pbte_face_integrals, host threads) ------------------
    yq, lpq = gauss_legendre_core(np_fc)
    other, off_x, off_y = face_partners(mesh, bc_codes, bc_xoff, bc_yoff)
    int_tri_fc_fc = np.empty((P, Q, n_tris, 3))
    int_tri_tri_fc = np.empty((n_tris, 3, P, P))
    keep = [np.ascontiguousarray(nodes, dtype=np.float64),
            np.ascontiguousarray(mesh.tri_nodes, dtype=np.int32),
            np.ascontiguousarray(mesh.tri_faces, dtype=np.int32),
            np.ascontiguousarray(mesh.face_nodes, dtype=np.int32),
            np.ascontiguousarray(mesh.face_len, dtype=np.float64),
            other, off_x, off_y,
            np.ascontiguousarray(C, dtype=np.float64), np.ascontiguousarray(Cf, dtype=np.float64),
            np.ascontiguousarray(mono, dtype=np.int32),
            np.ascontiguousarray(yq, dtype=np.float64), np.ascontiguousarray(lpq, dtype=np.float64)]
    d = _native.PbteFaceIntegralsDesc()
    d.n_tris, d.nd, d.nf, d.np_fc = n_tris, P, Q, int(np_fc)
    d.nodes, d.tri_nodes, d.tri_faces, d.face_nodes, d.face_len = (
        _native.f64(keep[0]), _native.i32(keep[1]), _native.i32(keep[2]), _native.i32(keep[3]),
        _native.f64(keep[4]))
    d.other, d.off_x, d.off_y = _native.i32(keep[5]), _native.f64(keep[6]), _native.f64(keep[7])
    d.C, d.Cf, d.mono = _native.f64(keep[8]), _native.f64(keep[9]), _native.i32(keep[10])
    d.yq, d.lpq = _native.f64(keep[11]), _native.f64(keep[12])
    if _native.lib().pbte_face_integrals(C_.byref(d), int_tri_fc_fc.ctypes.data, int_tri_tri_fc.ctypes.data,
                                         threads) != 0:
        raise ValueError("pbte_face_integrals: bad arguments")

    qax, qay, qw = tri_quadrature(np_tri)
    nodfun_qua_p = _eval_mono(C, mono, qax, qay).T
    return Integrals(int_tri=int_tri, int_tri_tri=int_tri_tri,
                     int_tri_tri_x=int_tri_tri_x, int_tri_tri_y=int_tri_tri_y,
                     int_tri_fc=int_tri_fc, int_tri_fc_fc=int_tri_fc_fc,
                     int_fc_fc=int_fc_fc, int_fc=int_fc, int_tri_tri_fc=int_tri_tri_fc,
                     qua_abs_x=qax, qua_abs_y=qay, qua_wei=qw, nodfun_qua_p=nodfun_qua_p,
                     ref_tri=ref_tri, ref_tri_tri=ref_tri_tri,
                     ref_ops=np.ascontiguousarray(np.stack([ref_tri_tri, ref_dx, ref_dy, r1, r2, r3])),
                     cell_ops=np.ascontiguousarray(np.stack(
                         [area2, (y3 - y1) / ss, (y1 - y2) / ss, (x1 - x3) / ss, (x2 - x1) / ss,
                          len12, len23 / sqrt2, len31], axis=1)))
