"""Restart files ``VDF*.out`` (``solver-python/pybte/driver.py:100-139``).

The file is the raw little-endian binary64 distribution in the Fortran
array ``VDF(NDOF_TRI, N_TRIS, NPOLE, NAZIM)``, column-major.  With the
ordinate numbering ``d = j1 + npole j2`` that is, byte for byte, the
C-ordered ``[D, N, P]`` array the solver keeps (offset ``l + P (i + N d)``
both ways), so reading and writing are plain copies.  The name carries the
degree, element count and ordinate counts, with the blanks of the Fortran
``I3`` / ``I2`` edit descriptors squeezed out.
"""
from __future__ import annotations

import hashlib
from pathlib import Path

import numpy as np

__all__ = ["restart_name", "read_restart", "write_restart"]


def restart_name(deg: int, n_tris: int, npole: int, nazim: int) -> str:
    return f"VDFP{deg}T{n_tris:3d}NP{npole:2d}NA{nazim:2d}.out".replace(" ", "")


def read_restart(path, shape) -> tuple[np.ndarray, str]:
    """``(vdf [D, N, P], sha256 of the file)``; ``ValueError`` when the
    file does not hold ``D N P`` values."""
    raw = Path(path).read_bytes()
    count = int(np.prod(shape))
    if len(raw) != 8 * count:
        raise ValueError(f"restart file {path} holds {len(raw)} bytes, not the {8 * count} of a "
                         f"{shape[2]}-DOF x {shape[1]}-element x {shape[0]}-ordinate distribution")
    return np.frombuffer(raw, dtype="<f8").reshape(shape).copy(), hashlib.sha256(raw).hexdigest()


def write_restart(path, vdf: np.ndarray) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_bytes(np.ascontiguousarray(vdf, dtype="<f8").tobytes())
    return path
