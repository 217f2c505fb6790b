"""Macroscopic moments on the device (``solver-python/pybte/moments.py:29-72``).

``Moments`` keeps ``ts/qxs/qys`` ``[P, N]`` and the element integrals
``temp/qx/qy`` ``[N]`` as binary64 device tensors; the reference's attribute
names return host copies.  ``temp``/``qx``/``qy`` are element *integrals*,
not averages.
"""
from __future__ import annotations

import numpy as np

from .device import dev_zeros, ptr, torch

__all__ = ["Moments", "compute_moments", "residual_iterate"]


class Moments:
    def __init__(self, ndof_tri: int, n_tris: int):
        self.nd, self.nt = ndof_tri, n_tris
        self.ts_d = dev_zeros((ndof_tri, n_tris))
        self.qxs_d = dev_zeros((ndof_tri, n_tris))
        self.qys_d = dev_zeros((ndof_tri, n_tris))
        self.temp_d = dev_zeros(n_tris)
        self.qx_d = dev_zeros(n_tris)
        self.qy_d = dev_zeros(n_tris)
        self._ctx = None

    @classmethod
    def zeros(cls, ndof_tri: int, n_tris: int) -> "Moments":
        return cls(ndof_tri, n_tris)

    def _host(self, t):
        return t.detach().cpu().numpy()

    @property
    def ts(self) -> np.ndarray:
        return self._host(self.ts_d)

    @property
    def qxs(self) -> np.ndarray:
        return self._host(self.qxs_d)

    @property
    def qys(self) -> np.ndarray:
        return self._host(self.qys_d)

    @property
    def temp(self) -> np.ndarray:
        return self._host(self.temp_d)

    @property
    def qx(self) -> np.ndarray:
        return self._host(self.qx_d)

    @property
    def qy(self) -> np.ndarray:
        return self._host(self.qy_d)

    def set_dofs(self, ts, qxs, qys) -> None:
        for dst, src in ((self.ts_d, ts), (self.qxs_d, qxs), (self.qys_d, qys)):
            dst.copy_(torch.as_tensor(np.ascontiguousarray(src), dtype=torch.float64))

    @property
    def mass(self) -> float:
        """Sum of ``temp`` (``_kernels.py:340-347``): on the device in a
        fixed two-pass order (per-workgroup partial sums over a grid-stride
        partition, then the partials in block order), so a run repeats bit
        for bit but the order is not the reference's sequential one; on the
        host (no context) in index order."""
        if self._ctx is not None:
            return self._ctx.mass(self)
        return float(sum(self.temp.tolist()))


def compute_moments(vdf, cxv, cyv, domega, int_tri, cv, mom: Moments, ctx=None) -> Moments:
    """In-place moments of ``vdf``; ``ctx`` is the solver's ``SweepContext``
    (``vdf`` may be the device tensor or the host mirror of ``Solver.vdf``)."""
    ctx = ctx or getattr(vdf, "_sweep_ctx", None) or mom._ctx
    if ctx is None:
        raise TypeError("compute_moments needs the solver's SweepContext (pass ctx=...)")
    ctx.moments(vdf, mom)
    return mom


def residual_iterate(temp, t_old, ctx=None) -> float:
    """``sqrt(sum (T - T_old)^2 / sum T^2)`` (``_kernels.py:340-361``),
    summed on the device in the fixed two-pass order of ``Moments.mass``."""
    if ctx is None:
        raise TypeError("residual_iterate needs the solver's SweepContext")
    return ctx.residual_iterate(temp, t_old)[0]
