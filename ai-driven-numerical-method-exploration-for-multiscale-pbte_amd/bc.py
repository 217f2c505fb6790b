"""Boundary data: type codes, wall temperatures, periodic partners.

Follows ``solver-python/pybte/bc.py:120-170``: symmetry walls are refused,
periodic faces get the element behind the paired face and whether the pair
is traversed in reverse; the adiabatic emission ``flux_wall[F, P]`` lives on
the device and is refreshed by ``pbte_wall_flux`` (``bc.py:47-58``).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .constants import BC_NONTHERMALISING, BC_PERIODIC, BC_SYMMETRY

__all__ = ["BoundaryData", "build_boundary_data"]


@dataclass
class BoundaryData:
    bc_type: np.ndarray
    bc_temp: np.ndarray
    periodic_tri: np.ndarray
    periodic_flip: np.ndarray
    has_nonthermalising: bool
    has_periodic: bool
    n_faces: int
    ndof: int
    _ctx: object = None          # the owning solver's SweepContext

    @property
    def flux_wall(self) -> np.ndarray:
        """Host copy of the adiabatic emission ``flux_wall[F, P]``."""
        if self._ctx is None:
            return np.zeros((self.n_faces, self.ndof))
        return self._ctx.flux_wall.detach().cpu().numpy()

    def update_wall_flux(self, vdf, cxv=None, cyv=None, domega=None, int_tri_fc=None, mesh=None) -> None:
        """Diffuse emission of the adiabatic walls from ``vdf``
        (``bc.py:47-58``, ``_kernels.py:364-397``), into the device
        ``flux_wall`` the sweeps read.  ``vdf`` is the device tensor or the
        solver's host array; the ordinates, face integrals and mesh are the
        solver's own (the reference's argument list is accepted)."""
        if self._ctx is None:
            raise TypeError("update_wall_flux needs the boundary data of a Solver")
        if self.has_nonthermalising:
            self._ctx.wall_flux(vdf)


def build_boundary_data(mesh, boundaries, ndof_tri: int) -> BoundaryData:
    bc_type = np.array([b.code for b in boundaries], dtype=np.int32)
    bc_temp = np.array([float(b.temp) for b in boundaries], dtype=np.float64)
    bc_xoff = np.array([float(b.xoff) for b in boundaries], dtype=np.float64)
    bc_yoff = np.array([float(b.yoff) for b in boundaries], dtype=np.float64)
    if np.any(bc_type == BC_SYMMETRY):
        raise NotImplementedError(
            "BC_TYP=4 (symmetry) is carried by the mesh but not implemented in the "
            "transport solve; use a periodic pair or an adiabatic wall instead.")
    n_faces = mesh.n_faces
    periodic_tri = np.full(n_faces, -1, dtype=np.int32)
    periodic_flip = np.zeros(n_faces, dtype=bool)
    has_periodic = bool(np.any(bc_type == BC_PERIODIC))
    if has_periodic:
        for f in range(n_faces):
            bc = int(mesh.face_bc[f])
            if bc < 0 or bc_type[bc] != BC_PERIODIC:
                continue
            pair = int(mesh.face_pair[f])
            if pair < 0:
                raise ValueError(f"periodic face {f} has no partner")
            tri, _ = mesh.adjacent(pair)
            periodic_tri[f] = tri
            a = mesh.nodes[mesh.face_nodes[f, 0]] + np.array([bc_xoff[bc], bc_yoff[bc]])
            b1 = mesh.nodes[mesh.face_nodes[pair, 0]]
            b2 = mesh.nodes[mesh.face_nodes[pair, 1]]
            if np.allclose(a, b2) and not np.allclose(a, b1):
                periodic_flip[f] = True
            elif not np.allclose(a, b1):
                raise ValueError(f"periodic faces {f} and {pair} do not line up under the "
                                 f"declared offset ({bc_xoff[bc]}, {bc_yoff[bc]})")
    return BoundaryData(bc_type=bc_type, bc_temp=bc_temp, periodic_tri=periodic_tri,
                        periodic_flip=periodic_flip,
                        has_nonthermalising=bool(np.any(bc_type == BC_NONTHERMALISING)),
                        has_periodic=has_periodic, n_faces=n_faces, ndof=ndof_tri)
