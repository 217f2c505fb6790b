"""Host restatement of the face integrals of the nodal DG basis: the
element x face-basis tensor ``int_tri_fc_fc [P, Q, N, 3]`` and the element
x neighbour tensor ``int_tri_tri_fc [N, 3, P, P]``.

TEST INFRASTRUCTURE ONLY.  The product computes them in
``pbte_face_integrals`` (``csrc/host_setup.cpp``); the tests compare the two.

Follows ``solver-python/pybte/integration.py:321-404``: a Gauss-Legendre
rule of ``np_fc`` points along each face, the element's basis at the
points through its inverse affine map, the face basis in the edge
parameter, and for an interior (or periodic, with the partner's shift)
face the neighbour's basis through its own map; on a wall the element's
basis integrated over the face, broadcast over the second index.  numpy,
vectorised over elements (element-wise arithmetic in the reference's
expression order).

Pinned by: the exact integrals of polynomials along straight edges (a
face of length L: int_tri_fc_fc summed over the face basis reproduces
int phi_m ds, and on the reference triangle the monomial edge moments of
``fac_div``), checked in tests/test_face_integrals_cpu.py, and by the
reference's FROZEN GSIS and CIS runs reached through the product.
"""
from __future__ import annotations

import numpy as np


def affine(x, y):
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


def eval_mono(C, mono, xi, eta):
    out = np.zeros(np.broadcast(xi, eta).shape + (C.shape[0],))
    for k, (a, b) in enumerate(mono):
        out += (C[:, k] * _pow(xi, int(a))[..., None]) * _pow(eta, int(b))[..., None]
    return out


def face_integrals(mesh, C, Cf, mono, yq, lpq, other, off_x, off_y):
    """``(int_tri_fc_fc, int_tri_tri_fc)``; ``other``, ``off_x``, ``off_y``
    ``[N, 3]`` as ``integration.face_partners`` gives them."""
    P, Q, np_fc = C.shape[0], Cf.shape[0], yq.size
    nodes = mesh.nodes
    n_tris = mesh.n_tris
    x = nodes[mesh.tri_nodes, 0]
    y = nodes[mesh.tri_nodes, 1]
    n1, n2 = np_fc, np_fc + 1
    out_fc = np.zeros((P, Q, n_tris, 3))
    out_tt = np.zeros((n_tris, 3, P, P))
    for j2 in range(3):
        f = mesh.tri_faces[:, j2].astype(np.int64)
        ln = mesh.face_len[f]
        lnc = ln[:, None]
        absc = (0.0 * (1.0 - yq) + lnc * (1.0 + yq)) / 2.0
        weig = n2 * n2 * (lnc - 0.0) / ((1.0 - yq * yq) * lpq * lpq) / n1 / n1
        fx1 = nodes[mesh.face_nodes[f, 0], 0][:, None]
        fy1 = nodes[mesh.face_nodes[f, 0], 1][:, None]
        fx2 = nodes[mesh.face_nodes[f, 1], 0][:, None]
        fy2 = nodes[mesh.face_nodes[f, 1], 1][:, None]
        xp = (fx2 - fx1) / lnc * absc + fx1
        yp = (fy2 - fy1) / lnc * absc + fy1
        (bA, cA, dA), (fE, gE, hE) = affine(x, y)
        xi = bA[:, None] * xp + cA[:, None] * yp + dA[:, None]
        eta = fE[:, None] * xp + gE[:, None] * yp + hE[:, None]
        pm = eval_mono(C, mono, xi, eta)
        t = 2.0 * absc / lnc - 1.0
        pl = np.zeros(t.shape + (Q,))
        for k in range(Q):
            pl += Cf[:, k] * _pow(t, k)[..., None]
        acc = np.zeros((n_tris, P, Q))
        for q in range(np_fc):
            acc += (pm[:, q, :, None] * pl[:, q, None, :]) * weig[:, q, None, None]
        out_fc[:, :, :, j2] = np.transpose(acc, (1, 2, 0))
        two = other[:, j2] >= 0
        o = np.where(two, other[:, j2], 0)
        xo = nodes[mesh.tri_nodes[o], 0] + off_x[:, j2][:, None]
        yo = nodes[mesh.tri_nodes[o], 1] + off_y[:, j2][:, None]
        (bB, cB, dB), (fF, gF, hF) = affine(xo, yo)
        xi2 = bB[:, None] * xp + cB[:, None] * yp + dB[:, None]
        eta2 = fF[:, None] * xp + gF[:, None] * yp + hF[:, None]
        pn = eval_mono(C, mono, xi2, eta2)
        acc2 = np.zeros((n_tris, P, P))
        for q in range(np_fc):
            acc2 += (pm[:, q, :, None] * pn[:, q, None, :]) * weig[:, q, None, None]
        accw = np.zeros((n_tris, P))
        for q in range(np_fc):
            accw += pm[:, q, :] * weig[:, q, None]
        out_tt[:, j2] = np.where(two[:, None, None], acc2, accw[:, :, None])
    return out_fc, out_tt
