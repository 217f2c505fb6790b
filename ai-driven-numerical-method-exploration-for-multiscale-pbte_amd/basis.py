"""Nodal Lagrange bases stored as monomial coefficients.

Face basis on [-1, 1] with equispaced nodes ``-1 + 2i/(Q-1)``; element basis on
the reference triangle (0,0)-(1,0)-(0,1) with nodes ``(i/deg, j/deg)`` listed
``j``-major.  Monomials are ordered by total degree, then by the eta exponent:
``xi^a eta^b`` has index ``(a+b)(a+b+1)/2 + b``
(``solver-python/pybte/basis.py:1-71``).

The coefficient tables are produced by expanding the Lagrange products one
linear factor at a time, in the same factor order as the reference
(``basis.py:46-160``), so that the coefficients agree to the last bit on the
shipped degrees; the tests check the Kronecker property and the partition of
unity, which any wrong coefficient breaks.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["Basis", "build_basis", "monomial_exponents", "basis_function_fc",
           "basis_function_tri", "eval_tri_basis"]


def monomial_exponents(deg: int) -> np.ndarray:
    return np.asarray([(k - j, j) for k in range(deg + 1) for j in range(k + 1)],
                      dtype=np.int64)


def _mono_index(a: int, b: int) -> int:
    return (a + b) * (a + b + 1) // 2 + b


def basis_function_fc(nn: int):
    """1-D Lagrange basis with ``nn`` equispaced nodes on [-1, 1].

    The product over ``j != i`` of ``(t - t_j)/(t_i - t_j)`` is accumulated
    factor by factor as ``b0 + b1 t`` with ``b0 = (2(j+1)-nn-1)/(2(j-i))`` and
    ``b1 = -(nn-1)/(2(j-i))``.
    """
    if nn < 2:
        raise ValueError("the face basis needs at least two nodes")
    node = np.array([-1.0 + 2.0 * i / (nn - 1) for i in range(nn)])
    coef = np.zeros((nn, nn))
    for i in range(nn):
        poly = [1.0]
        for j in range(nn):
            if j == i:
                continue
            b0 = float(2 * (j + 1) - nn - 1) / float(2 * (j - i))
            b1 = -float(nn - 1) / float(2 * (j - i))
            new = [0.0] * (len(poly) + 1)
            new[0] = poly[0] * b0
            new[len(poly)] = poly[-1] * b1
            for e in range(1, len(poly)):
                new[e] = poly[e] * b0 + poly[e - 1] * b1
            poly = new
        coef[i, :len(poly)] = poly
    return node, coef


def _univariate(rr: int, count: int):
    """Coefficients (by power) of ``prod_{l<count} (rr*s - l)/(count - l)``."""
    if count == 0:
        return [1.0]
    poly = [0.0, float(rr) / float(count)]        # the l = 0 factor rr*s/count
    for ell in range(1, count):
        d0 = float(ell) / float(ell - count)
        d1 = -float(rr) / float(ell - count)
        new = [0.0] * (len(poly) + 1)
        new[0] = poly[0] * d0
        new[len(poly)] = poly[-1] * d1
        for m in range(1, len(poly)):
            new[m] = poly[m] * d0 + poly[m - 1] * d1
        poly = new
    return poly


def basis_function_tri(rr: int):
    """Lagrange basis of degree ``rr`` on the reference triangle.

    Node ``(i/rr, j/rr)`` gets
    ``prod_{l<i}(rr xi - l)/(i-l) * prod_{m<j}(rr eta - m)/(j-m)
    * prod_{n<k}(rr(1-xi-eta) - n)/(k-n)`` with ``k = rr - i - j``.
    """
    nn = (rr + 1) * (rr + 2) // 2
    node = np.zeros((nn, 2))
    coef = np.zeros((nn, nn))
    c = 0
    for j in range(rr + 1):
        for i in range(rr - j + 1):
            node[c] = (float(i) / float(rr), float(j) / float(rr))
            k = rr - i - j
            px = _univariate(rr, i)           # powers of xi
            py = _univariate(rr, j)           # powers of eta
            row = np.zeros(nn)
            for a in range(i + 1):
                for b in range(j + 1):
                    row[_mono_index(a, b)] += px[a] * py[b]
            for n in range(k):
                prev = row.copy()
                d0 = float(n - rr) / float(n - k)
                d1 = float(rr) / float(n - k)
                for tot in range(rr + 1):
                    for b in range(tot + 1):
                        a = tot - b
                        pos = _mono_index(a, b)
                        row[pos] = prev[pos] * d0
                        if a > 0:
                            row[pos] += prev[_mono_index(a - 1, b)] * d1
                        if b > 0:
                            row[pos] += prev[_mono_index(a, b - 1)] * d1
            coef[c] = row
            c += 1
    return node, coef


@dataclass(frozen=True)
class Basis:
    deg: int
    node_fc: np.ndarray
    nodfun_fc: np.ndarray
    node_tri: np.ndarray
    nodfun_tri: np.ndarray
    mono: np.ndarray

    @property
    def ndof_tri(self) -> int:
        return self.nodfun_tri.shape[0]

    @property
    def ndof_fc(self) -> int:
        return self.nodfun_fc.shape[0]

    def eval_tri(self, xi, eta) -> np.ndarray:
        return eval_tri_basis(self.nodfun_tri, self.mono, xi, eta)

    def eval_fc(self, t) -> np.ndarray:
        t = np.atleast_1d(np.asarray(t, dtype=float))
        powers = np.stack([np.power(t, float(k)) for k in range(self.ndof_fc)], axis=-1)
        return powers @ self.nodfun_fc.T


def build_basis(deg: int) -> Basis:
    node_fc, nodfun_fc = basis_function_fc(deg + 1)
    node_tri, nodfun_tri = basis_function_tri(deg)
    return Basis(deg=deg, node_fc=node_fc, nodfun_fc=nodfun_fc, node_tri=node_tri,
                 nodfun_tri=nodfun_tri, mono=monomial_exponents(deg))


def eval_tri_basis(nodfun_tri, mono, xi, eta) -> np.ndarray:
    xi = np.asarray(xi, dtype=float)
    eta = np.asarray(eta, dtype=float)
    scalar = xi.ndim == 0 and eta.ndim == 0
    xi, eta = np.broadcast_arrays(np.atleast_1d(xi), np.atleast_1d(eta))
    terms = np.empty(xi.shape + (mono.shape[0],))
    for k, (a, b) in enumerate(mono):
        terms[..., k] = np.power(xi, float(a)) * np.power(eta, float(b))
    out = terms @ nodfun_tri.T
    return out[0] if scalar else out
