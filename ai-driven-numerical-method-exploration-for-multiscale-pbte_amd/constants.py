"""Physical and bookkeeping constants of the gray linear Callaway solver.

Boundary codes and the ``NP_TRI`` table follow the reference's
``solver-python/pybte/constants.py:23-49``; ``PI`` is ``4*atan(1)`` in binary64,
which equals ``math.pi`` bit for bit.
"""
from __future__ import annotations

import math

PI: float = 4.0 * math.atan(1.0)

BC_THERMALISING = 1      # diffusely emitting wall at a prescribed temperature
BC_NONTHERMALISING = 2   # adiabatic, diffusely reflecting wall
BC_PERIODIC = 3          # periodic face pair
BC_SYMMETRY = 4          # specular plane: carried by the mesh, refused by the solve

BC_TYPE_NAMES = {
    "thermalising": BC_THERMALISING,
    "thermalizing": BC_THERMALISING,
    "isothermal": BC_THERMALISING,
    "nonthermalising": BC_NONTHERMALISING,
    "nonthermalizing": BC_NONTHERMALISING,
    "adiabatic": BC_NONTHERMALISING,
    "periodic": BC_PERIODIC,
    "symmetry": BC_SYMMETRY,
}
BC_NAME_BY_CODE = {
    BC_THERMALISING: "thermalising",
    BC_NONTHERMALISING: "nonthermalising",
    BC_PERIODIC: "periodic",
    BC_SYMMETRY: "symmetry",
}

#: quadrature points of the element rule per polynomial degree
NP_TRI_BY_DEG = {1: 3, 2: 6, 3: 12, 4: 16}

#: Gauss-Legendre points of the face rule (fixed at 15 in the reference)
NP_FC_DEFAULT = 15
