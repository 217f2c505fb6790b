"""``RunRecord``: the machine-readable result of a run.

Field names, types and the derived properties follow the reference's
verifier contract (``solver-python/pybte/io_output.py:30-164``): verifiers read
these names, so they are kept as they are.
"""
from __future__ import annotations

import json
import platform
import sys
from dataclasses import asdict, dataclass, field
from pathlib import Path

import numpy as np

__all__ = ["RunRecord"]


@dataclass
class RunRecord:
    iterations: int = 0
    converged: bool = False
    residual_history: np.ndarray = field(default_factory=lambda: np.zeros(0))
    residual_true: np.ndarray | None = None
    mass_history: np.ndarray = field(default_factory=lambda: np.zeros(0))
    temp: np.ndarray = field(default_factory=lambda: np.zeros(0))
    qx: np.ndarray = field(default_factory=lambda: np.zeros(0))
    qy: np.ndarray = field(default_factory=lambda: np.zeros(0))
    temp_dofs: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    qx_dofs: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    qy_dofs: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    sweep_count: int = 0
    factorisation_count: int = 0
    wall_clock: float = 0.0
    setup_wall_clock: float = 0.0
    config_hash: str = ""
    scheme: str = "CIS"
    acc_variant: str | None = None
    tol: float = 0.0
    tmax: int = 0
    tau_r: float = 0.0
    tau_n: float = 0.0
    tau_c: float = 0.0
    deg: int = 0
    npole: int = 0
    nazim: int = 0
    n_tris: int = 0
    n_faces: int = 0
    mesh_file: str = ""
    restart: dict = field(default_factory=dict)
    environment: dict = field(default_factory=dict)
    diagnostics: dict = field(default_factory=dict)

    @property
    def final_residual(self) -> float:
        return float(self.residual_history[-1]) if self.residual_history.size else float("nan")

    @property
    def mass(self) -> float:
        return float(self.temp.sum()) if self.temp.size else float("nan")

    @property
    def contraction(self) -> float:
        """Geometric factor fitted to the second half of the residual history."""
        h = np.asarray(self.residual_history, dtype=float)
        if h.size < 8:
            return float("nan")
        tail = h[max(1, h.size // 2):]
        tail = tail[np.isfinite(tail) & (tail > 0)]
        if tail.size < 4 or tail[-1] >= tail[0]:
            return float("nan")
        rho = (tail[-1] / tail[0]) ** (1.0 / (tail.size - 1))
        return float(rho) if 0.0 < rho < 1.0 else float("nan")

    @property
    def error_estimate(self) -> float:
        """``r rho/(1 - rho)``: the sum of the remaining geometric steps."""
        rho = self.contraction
        if not np.isfinite(rho):
            return float("nan")
        return float(self.final_residual * rho / (1.0 - rho))

    def to_dict(self, include_fields: bool = True) -> dict:
        d = asdict(self)
        for k, v in list(d.items()):
            if isinstance(v, np.ndarray):
                d[k] = None if (not include_fields and v.ndim > 1) else v.tolist()
        d["final_residual"] = self.final_residual
        d["mass"] = self.mass
        d["contraction"] = self.contraction
        d["error_estimate"] = self.error_estimate
        return d

    def to_json(self, path, include_fields: bool = True) -> Path:
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(self.to_dict(include_fields), indent=2))
        return path

    def to_npz(self, path) -> Path:
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        payload = {k: v for k, v in asdict(self).items() if isinstance(v, np.ndarray)}
        payload["meta"] = np.array(json.dumps(self.to_dict(include_fields=False)))
        np.savez_compressed(path, **payload)
        return path

    @staticmethod
    def environment_fingerprint() -> dict:
        import numpy as _np
        try:
            import scipy
            scipy_v = scipy.__version__
        except ImportError:  # pragma: no cover
            scipy_v = None
        try:
            import torch
            torch_v = torch.__version__
        except ImportError:  # pragma: no cover
            torch_v = None
        return {"python": sys.version.split()[0], "numpy": _np.__version__, "scipy": scipy_v,
                "numba": None, "torch": torch_v, "platform": platform.platform(),
                "machine": platform.machine(), "backend": "hip-gfx950"}
