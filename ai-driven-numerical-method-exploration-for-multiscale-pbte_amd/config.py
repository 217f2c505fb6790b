"""YAML case description: the same schema as the reference's ``Case``.

Field names, defaults and validation follow ``solver-python/pybte/config.py:35-355``
so that an existing case file loads unchanged.  Two ``performance`` fields are
new and change only how the device computes, never the discrete problem:

* ``performance.kernel`` -- ``"hip"``; the reference's ``"numba"``/``"numpy"``
  are accepted as aliases so that shipped YAML keeps working.
* ``performance.precision`` -- ``"fp64"`` (default: every sweep in binary64) or
  ``"fp32"`` (sweeps and element LU stored and accumulated in binary32 inside
  an FP64 defect correction).

``scheme.krylov_restart`` exposes the GMRES basis size directly (the
reference derives it from a byte budget and floors it at 32,
``krylov.py:249-251``).
"""
from __future__ import annotations

import hashlib
import json
from dataclasses import asdict, dataclass, field, replace
from pathlib import Path

from .constants import BC_TYPE_NAMES, NP_FC_DEFAULT, NP_TRI_BY_DEG

__all__ = ["Case", "Boundary", "Iteration", "Scheme", "VelMesh", "DG", "Flow",
           "MeshCfg", "Restart", "Output", "Performance"]


@dataclass
class Iteration:
    tol: float = 1.0e-8
    tmax: int = 10_000
    true_residual: bool = False
    true_residual_every: int = 1

    def __post_init__(self):
        self.tol = float(self.tol)
        self.tmax = int(self.tmax)
        self.true_residual = bool(self.true_residual)
        self.true_residual_every = int(self.true_residual_every)


@dataclass
class Scheme:
    accflag: int = 0                     # 0 = CIS, 1 = GSIS
    acc_variant: str = "A"
    stabilisation: list = field(default_factory=lambda: [1.0, 1.0, 1.0])
    scale_by_hmin: bool = False
    on_cycle: str = "raise"
    defect_omega: float = 0.0
    defect_every: int = 1
    defect_anderson: int = 0
    method: str = "source"
    krylov_tol_factor: float = 1.0e-6
    krylov_max_bytes: float = 5.0e8
    krylov_max_rounds: int = 3
    krylov_precond: bool = False
    #: GMRES basis size; 0 keeps the reference's byte-budget rule
    krylov_restart: int = 0

    def __post_init__(self):
        self.method = str(self.method).lower()
        if self.method not in ("source", "krylov"):
            raise ValueError("scheme.method must be 'source' or 'krylov'")
        self.krylov_tol_factor = float(self.krylov_tol_factor)
        self.krylov_max_bytes = float(self.krylov_max_bytes)
        self.krylov_max_rounds = int(self.krylov_max_rounds)
        self.krylov_precond = bool(self.krylov_precond)
        self.krylov_restart = int(self.krylov_restart)
        if self.krylov_max_bytes <= 0 or self.krylov_max_rounds < 1:
            raise ValueError("scheme.krylov_max_bytes must be > 0 and "
                             "krylov_max_rounds >= 1")
        if self.krylov_restart < 0:
            raise ValueError("scheme.krylov_restart must be >= 0")
        self.accflag = int(self.accflag)
        self.acc_variant = str(self.acc_variant).upper()
        self.stabilisation = [float(v) for v in self.stabilisation]
        self.scale_by_hmin = bool(self.scale_by_hmin)
        self.defect_omega = float(self.defect_omega)
        self.defect_every = int(self.defect_every)
        if self.defect_every < 0:
            raise ValueError("scheme.defect_every must be >= 0")
        self.defect_anderson = int(self.defect_anderson)
        if self.defect_anderson < 0:
            raise ValueError("scheme.defect_anderson must be >= 0")
        if not 0.0 <= self.defect_omega <= 1.0:
            raise ValueError("scheme.defect_omega must be in [0, 1]")
        if self.on_cycle not in ("raise", "break"):
            raise ValueError("scheme.on_cycle must be 'raise' or 'break'")


@dataclass
class VelMesh:
    npole: int = 20
    nazim: int = 40

    def __post_init__(self):
        self.npole = int(self.npole)
        self.nazim = int(self.nazim)


@dataclass
class DG:
    deg: int = 3

    def __post_init__(self):
        self.deg = int(self.deg)


@dataclass
class Flow:
    cv: float = 1.0
    vg: float = 1.0
    tau_r: float = 1.0e-3
    tau_n: float = 1.0e5
    tau_thr: float = 1.0

    def __post_init__(self):
        self.cv = float(self.cv)
        self.vg = float(self.vg)
        self.tau_r = float(self.tau_r)
        self.tau_n = float(self.tau_n)
        self.tau_thr = float(self.tau_thr)
        if self.tau_r <= 0.0 or self.tau_n <= 0.0:
            raise ValueError("tau_r and tau_n must be positive")
        if self.vg <= 0.0:
            raise ValueError("vg must be positive")


@dataclass
class MeshCfg:
    file: str = "meshes/A1_Nx11_Ny11.msh"


@dataclass
class Boundary:
    name: str
    phyid: int
    type: str = "thermalising"
    temp: float = 0.0
    xoff: float = 0.0
    yoff: float = 0.0

    def __post_init__(self):
        self.phyid = int(self.phyid)
        self.temp = float(self.temp)
        self.xoff = float(self.xoff)
        self.yoff = float(self.yoff)

    @property
    def code(self) -> int:
        key = str(self.type).strip().lower()
        if key not in BC_TYPE_NAMES:
            raise ValueError(f"unknown boundary type {self.type!r}; "
                             f"expected one of {sorted(set(BC_TYPE_NAMES))}")
        return BC_TYPE_NAMES[key]


@dataclass
class Restart:
    enabled: bool = False
    path: str | None = None
    error_on_stale: bool = True


@dataclass
class Output:
    dir: str = "out"
    field: bool = True
    runtime_log: bool = True
    run_record: bool = True
    tecplot: bool = False
    npz: bool = True
    field_nx: int = 109
    field_ny: int = 109


_KERNEL_ALIASES = {"hip": "hip", "numba": "hip", "numpy": "hip"}


@dataclass
class Performance:
    """Storage and kernel choices; none of them changes the discrete problem."""
    precompute_inverse: bool = True
    kernel: str = "hip"
    threads: int = 0
    precision: str = "fp64"
    # fp32 source iteration (CIS / GSIS): binary32 sweeps until the iterate
    # step falls below this (or stalls), then binary64 sweeps to the
    # tolerance; an addition of this build
    refine_switch: float = 1.0e-6
    # fp32 Krylov: 0 decides convergence as the binary64 driver does (the
    # transport residual below iteration.tol); > 0 stops the refinement
    # once the estimated relative L2 error of the temperature and heat-flux
    # moments is below this, the accuracy asked of a mixed-precision solve;
    # an addition of this build
    refine_tol: float = 0.0
    # time the sweep's tile shapes (directions per workgroup, workgroup
    # width, pipeline depth) on the case's own mesh and ordinates and sweep
    # with the fastest; results kept per process and, when autotune_cache
    # names a JSON file, across runs; an addition of this build
    autotune: bool = False
    autotune_cache: str = ""

    def __post_init__(self):
        self.precompute_inverse = bool(self.precompute_inverse)
        self.autotune = bool(self.autotune)
        self.autotune_cache = str(self.autotune_cache or "")
        key = str(self.kernel).lower()
        if key not in _KERNEL_ALIASES:
            raise ValueError("performance.kernel must be 'hip' "
                             "(or the reference names 'numba'/'numpy')")
        self.kernel = _KERNEL_ALIASES[key]
        self.threads = int(self.threads)
        if self.threads < 0:
            raise ValueError("performance.threads must be >= 0")
        self.precision = str(self.precision).lower()
        if self.precision not in ("fp64", "fp32"):
            raise ValueError("performance.precision must be 'fp64' or 'fp32'")
        self.refine_switch = float(self.refine_switch)
        if not self.refine_switch > 0.0:
            raise ValueError("performance.refine_switch must be > 0")
        self.refine_tol = float(self.refine_tol)
        if not self.refine_tol >= 0.0:
            raise ValueError("performance.refine_tol must be >= 0")


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
    basedir: str = "."

    def __post_init__(self):
        # the reference forces an even azimuthal count (config.py:252-255)
        self.velmesh.nazim = (self.velmesh.nazim // 2) * 2
        if self.velmesh.nazim < 2:
            raise ValueError("nazim must be >= 2 after the even-ing step")
        if self.dg.deg not in NP_TRI_BY_DEG:
            raise ValueError(f"deg must be one of {sorted(NP_TRI_BY_DEG)}")
        if self.scheme.acc_variant not in ("A", "B"):
            raise ValueError("scheme.acc_variant must be 'A' or 'B'")
        if len(self.scheme.stabilisation) != 3:
            raise ValueError("scheme.stabilisation must have three entries")

    @property
    def ndof_tri(self) -> int:
        return (self.dg.deg + 1) * (self.dg.deg + 2) // 2

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
        return 1.0 / (1.0 / self.flow.tau_r + 1.0 / self.flow.tau_n)

    @property
    def nbc(self) -> int:
        return len(self.boundaries)

    def bc_codes(self) -> list:
        return [b.code for b in self.boundaries]

    def mesh_path(self) -> Path:
        p = Path(self.mesh.file)
        return p if p.is_absolute() else (Path(self.basedir) / p)

    def output_path(self) -> Path:
        p = Path(self.output.dir)
        return p if p.is_absolute() else (Path(self.basedir) / p)

    def to_dict(self) -> dict:
        d = asdict(self)
        d.pop("basedir", None)
        return d

    def canonical_json(self) -> str:
        """Everything that can change the numbers, and nothing else."""
        d = self.to_dict()
        d.pop("output", None)
        d.pop("performance", None)
        return json.dumps(d, sort_keys=True, separators=(",", ":"))

    @property
    def config_hash(self) -> str:
        h = hashlib.sha256(self.canonical_json().encode()).hexdigest()
        mesh = self.mesh_path()
        if mesh.exists():
            h2 = hashlib.sha256(mesh.read_bytes()).hexdigest()
            h = hashlib.sha256((h + h2).encode()).hexdigest()
        return h

    def replace(self, **kw) -> "Case":
        return replace(self, **kw)

    @classmethod
    def from_dict(cls, data: dict, basedir: str = ".") -> "Case":
        def sub(klass, key):
            return klass(**(data.get(key) or {}))

        bnds = [Boundary(**b) for b in (data.get("boundaries") or [])]
        return cls(iteration=sub(Iteration, "iteration"), scheme=sub(Scheme, "scheme"),
                   velmesh=sub(VelMesh, "velmesh"), dg=sub(DG, "dg"),
                   flow=sub(Flow, "flow"), mesh=sub(MeshCfg, "mesh"),
                   boundaries=bnds, restart=sub(Restart, "restart"),
                   output=sub(Output, "output"),
                   performance=sub(Performance, "performance"), basedir=basedir)

    @classmethod
    def from_yaml(cls, path) -> "Case":
        import yaml

        path = Path(path)
        data = yaml.safe_load(path.read_text()) or {}
        return cls.from_dict(data, basedir=str(path.parent))

    def to_yaml(self, path) -> None:
        import yaml

        Path(path).write_text(yaml.safe_dump(self.to_dict(), sort_keys=False))
