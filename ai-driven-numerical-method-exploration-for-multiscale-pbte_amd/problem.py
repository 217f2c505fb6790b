"""Host-side setup chain of a case (``solver-python/pybte/driver.py:61-89``):
mesh, bases, ordinates, integral tensors, sweep order, boundary data.
Needs no device."""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

from .basis import Basis, build_basis
from .bc import BoundaryData, build_boundary_data
from .integration import Integrals, build_integrals
from .mesh import Mesh, build_mesh
from .ordering import SweepOrder, build_sweep_order, validate_sweep_order
from .velocity import VelocityMesh, build_velocity_mesh

__all__ = ["Problem", "build_problem"]


@dataclass
class Problem:
    case: object
    mesh: Mesh
    basis: Basis
    vel: VelocityMesh
    integrals: Integrals
    order: SweepOrder
    bcdata: BoundaryData


def build_problem(case, on_cycle: str | None = None, validate_order: bool = False) -> Problem:
    c = case
    if not c.boundaries:
        raise ValueError("the case declares no boundaries")
    mesh = build_mesh(c.mesh_path(), [b.phyid for b in c.boundaries], c.bc_codes(),
                      [b.name for b in c.boundaries], [b.xoff for b in c.boundaries],
                      [b.yoff for b in c.boundaries])
    basis = build_basis(c.dg.deg)
    vel = build_velocity_mesh(c.velmesh.npole, c.velmesh.nazim, c.flow.vg)
    # the sweep orders (native, threaded) are built on a worker thread while
    # this one builds the integral tensors
    with ThreadPoolExecutor(max_workers=1) as pool:
        fut = pool.submit(build_sweep_order, mesh, vel, on_cycle=on_cycle or c.scheme.on_cycle)
        integrals = build_integrals(mesh, basis, c.np_tri, c.np_fc, bc_codes=c.bc_codes(),
                                    bc_xoff=[b.xoff for b in c.boundaries],
                                    bc_yoff=[b.yoff for b in c.boundaries])
        order = fut.result()
    if validate_order:
        validate_sweep_order(mesh, vel, order)
    bcdata = build_boundary_data(mesh, c.boundaries, c.ndof_tri)
    return Problem(case=c, mesh=mesh, basis=basis, vel=vel, integrals=integrals, order=order,
                   bcdata=bcdata)
