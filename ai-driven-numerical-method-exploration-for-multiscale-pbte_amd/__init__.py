"""MI355X-native gray linear Callaway DG discrete-ordinates solver.

Drop-in for the reference's ``pybte`` entry points::

    from pbte_amd import Case, Solver
    record = Solver(Case.from_yaml("case.yaml")).run()

Hot path: hand-written HIP kernels for gfx950 (``csrc/``) behind the C-ABI of
``include/pbte_hip.h``, loaded with ctypes; PyTorch-ROCm provides device
buffers and the stream.
"""
from .config import (DG, Boundary, Case, Flow, Iteration, MeshCfg, Output, Performance,
                     Restart, Scheme, VelMesh)
from .io_output import RunRecord

__all__ = ["Case", "Solver", "RunRecord", "Boundary", "Iteration", "Scheme", "VelMesh", "DG",
           "Flow", "MeshCfg", "Restart", "Output", "Performance"]

__version__ = "0.1.0"


def __getattr__(name):
    # the solver pulls in torch and the HIP library; import it on first use
    if name == "Solver":
        from .driver import Solver
        return Solver
    raise AttributeError(name)
