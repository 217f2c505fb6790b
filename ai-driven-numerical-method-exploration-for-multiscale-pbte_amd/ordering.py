"""Per-direction sweep order, bit-exact with the reference's Fortran-scan
topological sort (``solver-python/pybte/ordering.py:39-200``).

The order is computed by the native host scheduler (``pbte_build_order``):
elements sorted by (pass number, index) with the pass numbers obtained in
O(N + E) from ``pass(i) = max(1, max_upwind_j pass(j) + [j > i])``; a
direction whose graph has a cycle falls back to the literal scan with the
reference's cycle-breaking rule.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native

__all__ = ["SweepOrder", "SweepCycleError", "build_sweep_order", "validate_sweep_order"]


class SweepCycleError(RuntimeError):
    """A direction's element dependency graph contains a cycle."""


@dataclass
class SweepOrder:
    order: np.ndarray            # (ndir, n_tris) int32
    npole: int
    nazim: int
    cycles_broken: np.ndarray    # (ndir,) int32

    def fortran_tri_order(self) -> np.ndarray:
        n_tris = self.order.shape[1]
        return (self.order.reshape(self.nazim, self.npole, n_tris).transpose(2, 1, 0) + 1
                ).astype(np.int32)

    def __getitem__(self, d):
        return self.order[d]


def build_sweep_order(mesh, vel, on_cycle: str = "raise") -> SweepOrder:
    if on_cycle not in ("raise", "break"):
        raise ValueError("on_cycle must be 'raise' or 'break'")
    cx, cy, _ = vel.flat()
    ndir, n_tris = cx.size, mesh.n_tris
    order = np.full((ndir, n_tris), -1, dtype=np.int32)
    cycles = np.zeros(ndir, dtype=np.int32)
    nx = np.ascontiguousarray(mesh.tri_normal[:, :, 0])
    ny = np.ascontiguousarray(mesh.tri_normal[:, :, 1])
    nb = np.ascontiguousarray(mesh.tri_neighbour, dtype=np.int32)
    st = _native.lib().pbte_build_order(ndir, n_tris, _native.i32(nb), _native.f64(nx),
                                        _native.f64(ny), _native.f64(np.ascontiguousarray(cx)),
                                        _native.f64(np.ascontiguousarray(cy)),
                                        1 if on_cycle == "break" else 0,
                                        _native.i32(order), _native.i32(cycles))
    if st != 0:
        bad = int(np.flatnonzero((order < 0).any(axis=1))[0])
        raise SweepCycleError(
            f"direction {bad} (j1={bad % vel.npole}, j2={bad // vel.npole}) has a cyclic "
            "element dependency; pass scheme.on_cycle='break' to lag the inflow instead.")
    return SweepOrder(order=order, npole=vel.npole, nazim=vel.nazim, cycles_broken=cycles)


def validate_sweep_order(mesh, vel, sweep: SweepOrder) -> None:
    """Every element's upwind neighbours precede it (``ordering.py:178-200``)."""
    cx, cy, _ = vel.flat()
    for d in range(cx.size):
        if sweep.cycles_broken[d]:
            continue
        pos = np.empty(mesh.n_tris, dtype=np.int64)
        pos[sweep.order[d]] = np.arange(mesh.n_tris)
        for il in range(3):
            nb = mesh.tri_neighbour[:, il]
            idx = np.flatnonzero(nb >= 0)
            flux = cx[d] * mesh.tri_normal[idx, il, 0] + cy[d] * mesh.tri_normal[idx, il, 1]
            down = idx[flux > 0]
            up = idx[flux < 0]
            if np.any(pos[nb[down]] < pos[down]):
                raise SweepCycleError(f"outflow neighbour precedes element, direction {d}")
            if np.any(pos[nb[up]] > pos[up]):
                raise SweepCycleError(f"inflow neighbour follows element, direction {d}")
