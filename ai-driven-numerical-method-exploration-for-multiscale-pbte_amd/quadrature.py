"""Quadrature rules consumed by the setup: Gauss-Legendre lines and the
symmetric triangle rules.

``gauss_legendre`` reproduces the reference's Newton iteration
(``solver-python/pybte/quadrature.py:44-90``): the initial guess
``cos((2i-1)pi/(2n)) + 0.27/n sin(...)`` and a three-term Legendre recurrence,
stopped when the largest root update is below ``1e-13``.  The iteration is kept
rather than replaced by ``numpy.polynomial.legendre.leggauss`` because its
roots are short by ~1e-13/n, and every angular sum downstream inherits that
(the reference's ``LIMITATIONS.md:167-193``); the tests pin the nodes to
``leggauss`` at 1e-13 and the weights to the exact moments.

``tri_quadrature`` holds the tabulated symmetric rules on the reference
triangle (weights sum to one) for NP = 3, 4, 6, 7, 12, 16
(``quadrature.py:103-195``); the point order is part of the state.
"""
from __future__ import annotations

import functools

import numpy as np

from .constants import PI

__all__ = ["gauss_legendre", "gauss_legendre_core", "tri_quadrature"]


@functools.lru_cache(maxsize=None)
def _newton_roots(n: int, tol: float = 1.0e-13, max_sweeps: int = 200):
    """Roots ``y`` of P_n on [-1, 1] and the scaled derivative ``lp``."""
    if n < 2:
        raise ValueError("the initial guess divides by (n-1); n >= 2 required")
    np1 = n + 1
    k = np.arange(1, n + 1, dtype=np.float64)
    y = (np.cos((2.0 * (k - 1.0) + 1.0) * PI / (2.0 * (n - 1.0) + 2.0))
         + 0.27 / n * np.sin(PI * (-1.0 + k * 2.0 / (n - 1.0)) * (n - 1.0) / np1))
    leg = np.zeros((n, np1))
    for _ in range(max_sweeps):
        leg[:, 0] = 1.0
        leg[:, 1] = y
        for j in range(2, n + 1):
            leg[:, j] = ((2.0 * j - 1.0) * y * leg[:, j - 1] - (j - 1) * leg[:, j - 2]) / j
        lp = np1 * (leg[:, n - 1] - y * leg[:, n]) / (1.0 - y * y)
        y_prev = y
        y = y_prev - leg[:, n] / lp
        if np.max(np.abs(y - y_prev)) < tol:
            break
    else:
        raise RuntimeError("Gauss-Legendre Newton iteration did not converge")
    y.setflags(write=False)
    lp.setflags(write=False)
    return y, lp


def gauss_legendre_core(n: int):
    """``(y, lp)``: interval-independent roots and derivative factors."""
    return _newton_roots(int(n))


def gauss_legendre(n: int, a: float, b: float):
    """Abscissae and weights on ``[a, b]`` (mapping as ``quadrature.py:33-41``)."""
    y, lp = _newton_roots(int(n))
    n1, n2 = n, n + 1
    absc = (a * (1.0 - y) + b * (1.0 + y)) / 2.0
    weig = n2 * n2 * (b - a) / ((1.0 - y * y) * lp * lp) / n1 / n1
    return absc, weig


def tri_quadrature(np_tri: int):
    """``(x, y, w)`` of the symmetric rule with ``np_tri`` points."""
    x = np.zeros(np_tri)
    y = np.zeros(np_tri)
    w = np.zeros(np_tri)

    def orbit3(i0, a):
        x[i0], y[i0] = a, 1.0 - 2.0 * a
        x[i0 + 1], y[i0 + 1] = a, a
        x[i0 + 2], y[i0 + 2] = 1.0 - 2.0 * a, a

    def orbit6(i0, a, b):
        c = 1.0 - a - b
        pts = [(b, c), (c, b), (a, b), (b, a), (c, a), (a, c)]
        for k, (px, py) in enumerate(pts):
            x[i0 + k], y[i0 + k] = px, py

    if np_tri == 3:
        w[:] = 1.0 / 3.0
        x[:], y[:] = [0.5, 0.5, 0.0], [0.0, 0.5, 0.5]
    elif np_tri == 4:
        w[0] = -27.0 / 48.0
        w[1:4] = 25.0 / 48.0
        x[:], y[:] = [1.0 / 3.0, 0.2, 0.6, 0.2], [1.0 / 3.0, 0.2, 0.2, 0.6]
    elif np_tri == 6:
        w[0:3] = 0.109951743655322
        w[3:6] = 0.223381589678011
        a, b = 0.091576213509771, 0.816847572980459
        c, e = 0.445948490915965, 0.108103018168070
        x[:] = [a, b, a, c, e, c]
        y[:] = [a, a, b, c, c, e]
    elif np_tri == 7:
        w[0] = 0.225
        w[1:4] = 0.125939180544827
        w[4:7] = 0.132394152788506
        a, b = 0.101286507323456, 0.797426985353087
        c, e = 0.470142064105115, 0.059715871789770
        x[:] = [1.0 / 3.0, a, b, a, c, c, e]
        y[:] = [1.0 / 3.0, a, a, b, e, c, c]
    elif np_tri == 12:
        w[0:3] = 0.050844906370207
        w[3:6] = 0.116786275726379
        w[6:12] = 0.082851075618374
        orbit3(0, 0.063089014491502)
        orbit3(3, 0.249286745170911)
        orbit6(6, 0.636502499121399, 0.310352451033785)
    elif np_tri == 16:
        w[0] = 0.144315607677787
        w[1:4] = 0.103217370534718
        w[4:7] = 0.032458497623198
        w[7:10] = 0.095091634267285
        w[10:16] = 0.027230314174435
        x[0], y[0] = 1.0 / 3.0, 1.0 / 3.0
        orbit3(1, 0.170569307751760)
        orbit3(4, 0.050547228317031)
        orbit3(7, 0.459292588292723)
        orbit6(10, 0.263112829634638, 0.008394777409958)
    else:
        raise ValueError(f"no triangle rule for NP_TRI={np_tri}")
    return x, y, w
