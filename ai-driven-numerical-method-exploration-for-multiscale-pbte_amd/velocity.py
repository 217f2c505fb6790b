"""Discrete ordinates: polar Gauss-Legendre rule on [0, pi] and an azimuthal
rule made of two Gauss-Legendre halves on [0, pi] and [pi, 2pi], every rule
stored in reversed order (``solver-python/pybte/velocity.py:59-86``).

The in-plane group velocity of ordinate ``(j1, j2)`` is
``(vg cos the_j1, vg sin the_j1 cos phi_j2)`` and its weight
``sin the_j1 wthe_j1 wphi_j2``; directions are flattened Fortran-style,
``d = j1 + npole * j2`` (``velocity.py:52-56``).  The reversal fixes the
direction numbering and hence the sweep orders, so it is kept.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .constants import PI
from .quadrature import gauss_legendre

__all__ = ["VelocityMesh", "build_velocity_mesh"]


@dataclass(frozen=True)
class VelocityMesh:
    the: np.ndarray
    wthe: np.ndarray
    phi: np.ndarray
    wphi: np.ndarray
    domega: np.ndarray      # (npole, nazim)
    cx: np.ndarray          # (npole, nazim)
    cy: np.ndarray          # (npole, nazim)

    @property
    def npole(self) -> int:
        return self.the.size

    @property
    def nazim(self) -> int:
        return self.phi.size

    @property
    def ndir(self) -> int:
        return self.npole * self.nazim

    def flat(self):
        """``(cx, cy, domega)`` with ``d = j1 + npole * j2``."""
        return (np.ascontiguousarray(self.cx.T.reshape(-1)),
                np.ascontiguousarray(self.cy.T.reshape(-1)),
                np.ascontiguousarray(self.domega.T.reshape(-1)))


def build_velocity_mesh(npole: int, nazim: int, vg: float = 1.0) -> VelocityMesh:
    nazim = (int(nazim) // 2) * 2
    if nazim < 2:
        raise ValueError("nazim must be >= 2")
    half = nazim // 2
    a, w = gauss_legendre(npole, 0.0, PI)
    the, wthe = a[::-1].copy(), w[::-1].copy()
    a1, w1 = gauss_legendre(half, 0.0, PI)
    a2, w2 = gauss_legendre(half, PI, 2.0 * PI)
    phi = np.concatenate([a1[::-1], a2[::-1]])
    wphi = np.concatenate([w1[::-1], w2[::-1]])
    sin_the = np.sin(the)
    domega = (sin_the * wthe)[:, None] * wphi[None, :]
    cx = np.repeat((vg * np.cos(the))[:, None], nazim, axis=1)
    cy = (vg * sin_the)[:, None] * np.cos(phi)[None, :]
    return VelocityMesh(the=the, wthe=wthe, phi=phi, wphi=wphi,
                        domega=domega, cx=cx, cy=cy)
