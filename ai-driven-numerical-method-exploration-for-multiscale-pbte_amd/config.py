"""YAML case description (``Case.from_yaml``), the reference's schema.

A case file of the reference (``solver-python/pybte/config.py:35-355``) loads
unchanged: the same sections, keys, defaults, accepted values and
``ValueError`` on an invalid one.  The structure here is a small schema
layer: every section is a dataclass whose fields carry their coercion
(YAML 1.1 reads ``1.0e5`` as a string, so every scalar is converted to its
declared type) and whose rules are a table of (test, message) pairs checked
after construction.

Additions of this build, none of which changes the discrete problem:

* ``performance.kernel`` -- ``"hip"``; the reference's ``"numba"`` and
  ``"numpy"`` are accepted as aliases so that shipped files keep working;
* ``performance.precision`` -- ``"fp64"`` (every sweep in binary64) or
  ``"fp32"`` (binary32 sweeps inside a binary64 correction);
* ``performance.refine_inner``, ``refine_switch``, ``refine_tol``, ``autotune``,
  ``autotune_cache``, ``trace_leaf`` -- documented at their fields;
* ``scheme.krylov_restart`` -- the GMRES basis size (the reference derives
  it from a byte budget with a floor of 32, ``krylov.py:249-251``).
"""
from __future__ import annotations

import hashlib
import json
from dataclasses import asdict, dataclass, field, fields, replace
from pathlib import Path

from .constants import BC_TYPE_NAMES, NP_FC_DEFAULT, NP_TRI_BY_DEG

__all__ = ["Case", "Boundary", "Iteration", "Scheme", "VelMesh", "DG", "Flow",
           "MeshCfg", "Restart", "Output", "Performance"]


# ---------------------------------------------------------------------------
# schema helpers
# ---------------------------------------------------------------------------
def _key(default, conv, factory=None):
    """A section field with its coercion."""
    if factory is not None:
        return field(default_factory=factory, metadata={"conv": conv})
    return field(default=default, metadata={"conv": conv})


def _lower(v) -> str:
    return str(v).lower()


def _upper(v) -> str:
    return str(v).upper()


def _floats(v) -> list:
    return [float(x) for x in v]


def _opt_str(v):
    return None if v is None else str(v)


_KERNEL_ALIASES = {"hip": "hip", "numba": "hip", "numpy": "hip"}


def _kernel(v) -> str:
    k = str(v).lower()
    if k not in _KERNEL_ALIASES:
        raise ValueError(f"performance.kernel {v!r} is not one of 'hip', 'numba', 'numpy'")
    return _KERNEL_ALIASES[k]


class _Section:
    """Coerces every field, then checks the section's rules in order."""

    RULES: tuple = ()

    def __post_init__(self):
        for f in fields(self):
            conv = f.metadata.get("conv")
            if conv is not None:
                setattr(self, f.name, conv(getattr(self, f.name)))
        for test, message in self.RULES:
            if not test(self):
                raise ValueError(message)


# ---------------------------------------------------------------------------
# sections
# ---------------------------------------------------------------------------
@dataclass
class Iteration(_Section):
    tol: float = _key(1.0e-8, float)
    tmax: int = _key(10_000, int)
    # also evaluate the transport residual of the discrete system (never the
    # stopping test), every true_residual_every iterations
    true_residual: bool = _key(False, bool)
    true_residual_every: int = _key(1, int)


@dataclass
class Scheme(_Section):
    accflag: int = _key(0, int)                          # 0 CIS, 1 GSIS
    acc_variant: str = _key("A", _upper)
    stabilisation: list = _key(None, _floats, factory=lambda: [1.0, 1.0, 1.0])
    scale_by_hmin: bool = _key(False, bool)
    on_cycle: str = _key("raise", str)
    defect_omega: float = _key(0.0, float)
    defect_every: int = _key(1, int)
    defect_anderson: int = _key(0, int)
    method: str = _key("source", _lower)
    krylov_tol_factor: float = _key(1.0e-6, float)
    krylov_max_bytes: float = _key(5.0e8, float)
    krylov_max_rounds: int = _key(3, int)
    krylov_precond: bool = _key(False, bool)
    # GMRES basis size; 0 keeps the reference's byte-budget rule
    krylov_restart: int = _key(0, int)

    RULES = (
        (lambda s: s.method in ("source", "krylov"), "scheme.method: expected 'source' or 'krylov'"),
        (lambda s: s.krylov_max_bytes > 0 and s.krylov_max_rounds >= 1,
         "scheme.krylov_max_bytes must be positive and scheme.krylov_max_rounds at least 1"),
        (lambda s: s.krylov_restart >= 0, "scheme.krylov_restart must not be negative"),
        (lambda s: s.defect_every >= 0, "scheme.defect_every must not be negative"),
        (lambda s: s.defect_anderson >= 0, "scheme.defect_anderson must not be negative"),
        (lambda s: 0.0 <= s.defect_omega <= 1.0, "scheme.defect_omega lies outside [0, 1]"),
        (lambda s: s.on_cycle in ("raise", "break"), "scheme.on_cycle: expected 'raise' or 'break'"),
    )


@dataclass
class VelMesh(_Section):
    npole: int = _key(20, int)
    nazim: int = _key(40, int)


@dataclass
class DG(_Section):
    deg: int = _key(3, int)


@dataclass
class Flow(_Section):
    cv: float = _key(1.0, float)
    vg: float = _key(1.0, float)
    tau_r: float = _key(1.0e-3, float)
    tau_n: float = _key(1.0e5, float)
    tau_thr: float = _key(1.0, float)

    RULES = (
        (lambda s: s.tau_r > 0.0 and s.tau_n > 0.0, "flow.tau_r and flow.tau_n must be positive"),
        (lambda s: s.vg > 0.0, "flow.vg must be positive"),
    )


@dataclass
class MeshCfg(_Section):
    file: str = _key("meshes/A1_Nx11_Ny11.msh", str)


@dataclass
class Boundary(_Section):
    name: str
    phyid: int = field(metadata={"conv": int})
    type: str = "thermalising"
    temp: float = _key(0.0, float)
    xoff: float = _key(0.0, float)
    yoff: float = _key(0.0, float)

    @property
    def code(self) -> int:
        """The boundary-condition code of ``type`` (constants.py)."""
        code = BC_TYPE_NAMES.get(str(self.type).strip().lower())
        if code is None:
            raise ValueError(f"boundary {self.name!r}: type {self.type!r} is not one of "
                             f"{sorted(set(BC_TYPE_NAMES))}")
        return code


@dataclass
class Restart(_Section):
    enabled: bool = _key(False, bool)
    path: str | None = _key(None, _opt_str)
    # with restart off, refuse to run over a VDF*.out left in the output dir
    error_on_stale: bool = _key(True, bool)


@dataclass
class Output(_Section):
    dir: str = _key("out", str)
    field: bool = _key(True, bool)
    runtime_log: bool = _key(True, bool)
    run_record: bool = _key(True, bool)
    tecplot: bool = _key(False, bool)
    npz: bool = _key(True, bool)
    field_nx: int = _key(109, int)
    field_ny: int = _key(109, int)


@dataclass
class Performance(_Section):
    """How the device computes; nothing here changes the discrete problem."""
    precompute_inverse: bool = _key(True, bool)
    kernel: str = _key("hip", _kernel)
    threads: int = _key(0, int)
    precision: str = _key("fp64", _lower)
    # fp32 source iteration (CIS / GSIS): binary32 iterations solve for the
    # correction of a binary64 iterate; each inner solve ends when its step
    # is this fraction of the binary64 step it corrects
    refine_inner: float = _key(1.0e-3, float)
    # fp32 source iteration where the correction does not apply (broken
    # sweep cycles, the defect repair on): binary32 sweeps until the iterate
    # step falls below this (or stalls), then binary64 sweeps to the tolerance
    refine_switch: float = _key(1.0e-6, float)
    # fp32 Krylov: 0 decides convergence as the binary64 driver does (the
    # transport residual below iteration.tol); > 0 stops the refinement once
    # the estimated relative L2 error of the temperature and heat-flux
    # moments is below this
    refine_tol: float = _key(0.0, float)
    # time the sweep's tile shapes on the case's own mesh and ordinates and
    # sweep with the fastest (kept per process and, with autotune_cache
    # naming a JSON file, across runs)
    autotune: bool = _key(False, bool)
    autotune_cache: str = _key("", lambda v: str(v or ""))
    # elements per leaf of the trace solve's nested dissection, 0 = 16
    trace_leaf: int = _key(0, int)
    # binary32 sweeps: "fused" builds and factorises each element operator
    # in binary32 inside the sweep (no binary32 factors stored or read, the
    # reference's on-the-fly mode); "stored" reads binary32 factors formed
    # once in binary64.  Fused falls back to stored where it cannot run.
    fp32_operator: str = _key("fused", _lower)

    RULES = (
        (lambda s: s.threads >= 0, "performance.threads must not be negative"),
        (lambda s: s.precision in ("fp64", "fp32"), "performance.precision: expected 'fp64' or 'fp32'"),
        (lambda s: s.refine_switch > 0.0, "performance.refine_switch must be positive"),
        (lambda s: 0.0 < s.refine_inner < 1.0, "performance.refine_inner must lie in (0, 1)"),
        (lambda s: s.refine_tol >= 0.0, "performance.refine_tol must not be negative"),
        (lambda s: s.trace_leaf >= 0, "performance.trace_leaf must not be negative"),
        (lambda s: s.fp32_operator in ("fused", "stored"),
         "performance.fp32_operator: expected 'fused' or 'stored'"),
    )


# section name in the YAML -> its class (``boundaries`` is a list)
_SECTIONS = {"iteration": Iteration, "scheme": Scheme, "velmesh": VelMesh, "dg": DG,
             "flow": Flow, "mesh": MeshCfg, "restart": Restart, "output": Output,
             "performance": Performance}

# what the numbers depend on: everything except where they are written and
# how the device computes them
_NOT_NUMERICS = ("output", "performance")


@dataclass
class Case:
    iteration: Iteration = field(default_factory=Iteration)
    scheme: Scheme = field(default_factory=Scheme)
    velmesh: VelMesh = field(default_factory=VelMesh)
    dg: DG = field(default_factory=DG)
    flow: Flow = field(default_factory=Flow)
    mesh: MeshCfg = field(default_factory=MeshCfg)
    boundaries: list = field(default_factory=list)
    restart: Restart = field(default_factory=Restart)
    output: Output = field(default_factory=Output)
    performance: Performance = field(default_factory=Performance)
    basedir: str = "."            # relative paths of the case resolve here

    def __post_init__(self):
        # the azimuthal count is made even (config.py:252-255 of the reference)
        self.velmesh.nazim -= self.velmesh.nazim % 2
        checks = (
            (self.velmesh.nazim >= 2, "velmesh.nazim is below 2 once made even"),
            (self.dg.deg in NP_TRI_BY_DEG, f"dg.deg must be one of {sorted(NP_TRI_BY_DEG)}"),
            (self.scheme.acc_variant in ("A", "B"), "scheme.acc_variant: expected 'A' or 'B'"),
            (len(self.scheme.stabilisation) == 3, "scheme.stabilisation needs exactly three values"),
        )
        for ok, message in checks:
            if not ok:
                raise ValueError(message)

    # -- derived sizes and constants ------------------------------------------
    @property
    def ndof_tri(self) -> int:
        d = self.dg.deg
        return (d + 1) * (d + 2) // 2

    @property
    def ndof_fc(self) -> int:
        return self.dg.deg + 1

    @property
    def np_tri(self) -> int:
        return NP_TRI_BY_DEG[self.dg.deg]

    @property
    def np_fc(self) -> int:
        return NP_FC_DEFAULT

    @property
    def tau_c(self) -> float:
        """Combined relaxation time, 1 / (1/tau_R + 1/tau_N)."""
        return 1.0 / (1.0 / self.flow.tau_r + 1.0 / self.flow.tau_n)

    @property
    def nbc(self) -> int:
        return len(self.boundaries)

    def bc_codes(self) -> list:
        return [b.code for b in self.boundaries]

    def _resolve(self, p) -> Path:
        p = Path(p)
        return p if p.is_absolute() else Path(self.basedir) / p

    def mesh_path(self) -> Path:
        return self._resolve(self.mesh.file)

    def output_path(self) -> Path:
        return self._resolve(self.output.dir)

    # -- serialisation -------------------------------------------------------
    def to_dict(self) -> dict:
        out = {name: asdict(getattr(self, name)) for name in _SECTIONS}
        out["boundaries"] = [asdict(b) for b in self.boundaries]
        return out

    def canonical_json(self) -> str:
        """The inputs that determine the discrete problem, as sorted compact
        JSON."""
        d = {k: v for k, v in self.to_dict().items() if k not in _NOT_NUMERICS}
        return json.dumps(d, sort_keys=True, separators=(",", ":"))

    @property
    def config_hash(self) -> str:
        """SHA-256 over the canonical JSON followed by the mesh file's bytes
        (when the file exists)."""
        h = hashlib.sha256(self.canonical_json().encode())
        mesh = self.mesh_path()
        if mesh.exists():
            h.update(mesh.read_bytes())
        return h.hexdigest()

    def replace(self, **kw) -> "Case":
        return replace(self, **kw)

    @classmethod
    def from_dict(cls, data: dict, basedir: str = ".") -> "Case":
        kw = {name: klass(**(data.get(name) or {})) for name, klass in _SECTIONS.items()}
        kw["boundaries"] = [Boundary(**b) for b in data.get("boundaries") or []]
        return cls(basedir=basedir, **kw)

    @classmethod
    def from_yaml(cls, path) -> "Case":
        import yaml

        path = Path(path)
        return cls.from_dict(yaml.safe_load(path.read_text()) or {}, basedir=str(path.parent))

    def to_yaml(self, path) -> None:
        import yaml

        Path(path).write_text(yaml.safe_dump(self.to_dict(), sort_keys=False))
