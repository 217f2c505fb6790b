"""Boundary data: type codes, wall temperatures, periodic partners.

The data of ``solver-python/pybte/bc.py:120-170`` (the same fields and
exceptions), with the periodic pairing vectorised over faces; the adiabatic
emission ``flux_wall[F, P]`` lives on the device and is refreshed by
``pbte_wall_flux`` (``bc.py:47-58``).
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


def _periodic_partners(mesh, bc_type, offsets):
    """For every periodic face: the element behind its partner face and
    whether the pair runs in opposite senses.  A face's trace nodes run from
    its first node to its second; shifted by its boundary's offset, the first
    node lands on the partner's first node (same sense) or on its second
    (reversed: the partner's nodes are mirrored, which the HDG assembly
    needs)."""
    F = mesh.n_faces
    partner_tri = np.full(F, -1, dtype=np.int32)
    reversed_pair = np.zeros(F, dtype=bool)
    bc = mesh.face_bc
    code = np.where(bc >= 0, bc_type[np.clip(bc, 0, None)], -1)
    faces = np.flatnonzero(code == BC_PERIODIC)
    if faces.size == 0:
        return partner_tri, reversed_pair
    pair = mesh.face_pair[faces].astype(np.int64)
    lonely = faces[pair < 0]
    if lonely.size:
        raise ValueError(f"periodic face {int(lonely[0])} has no partner face")
    own_tri, _ = mesh.adjacent_arrays()
    partner_tri[faces] = own_tri[pair]
    shifted = mesh.nodes[mesh.face_nodes[faces, 0]] + offsets[bc[faces]]
    same = np.isclose(shifted, mesh.nodes[mesh.face_nodes[pair, 0]]).all(axis=1)
    mirrored = np.isclose(shifted, mesh.nodes[mesh.face_nodes[pair, 1]]).all(axis=1)
    bad = ~(same | mirrored)
    if bad.any():
        k = int(np.flatnonzero(bad)[0])
        f, g, b = int(faces[k]), int(pair[k]), int(bc[faces[k]])
        raise ValueError(f"periodic face {f} does not meet its partner {g} under the offset "
                         f"({offsets[b, 0]}, {offsets[b, 1]}) of its boundary")
    reversed_pair[faces] = mirrored & ~same
    return partner_tri, reversed_pair


def build_boundary_data(mesh, boundaries, ndof_tri: int) -> BoundaryData:
    """Per boundary its type code and wall temperature; per periodic face
    its partner element and orientation (``bc.py:120-170``).  Symmetry walls
    (type 4) are refused with ``NotImplementedError``: the reference carries
    them through its mesh reader but its transport solve has none."""
    bc_type = np.array([b.code for b in boundaries], dtype=np.int32)
    if (bc_type == BC_SYMMETRY).any():
        raise NotImplementedError("symmetry walls (type 4) have no transport boundary condition; "
                                  "a periodic pair or an adiabatic wall can stand in for one")
    bc_temp = np.array([float(b.temp) for b in boundaries], dtype=np.float64)
    offsets = np.array([[float(b.xoff), float(b.yoff)] for b in boundaries],
                       dtype=np.float64).reshape(-1, 2)
    partner_tri, reversed_pair = _periodic_partners(mesh, bc_type, offsets)
    return BoundaryData(bc_type=bc_type, bc_temp=bc_temp, periodic_tri=partner_tri,
                        periodic_flip=reversed_pair,
                        has_nonthermalising=bool((bc_type == BC_NONTHERMALISING).any()),
                        has_periodic=bool((bc_type == BC_PERIODIC).any()),
                        n_faces=mesh.n_faces, ndof=ndof_tri)
