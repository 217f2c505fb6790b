"""Vectors for GSIS with a finite normal-process relaxation time, which no
fixture of the reference pins: its frozen GSIS rows all use tau_N = 1e5
(tests/integration/test_regression.py:29-48), where the drift channel
tau_C / tau_N of the correction is negligible
(acceleration/correction.py:13-17).

Imports the reference's pure-Python solver (solver-python/pybte; numba is
absent here, so its kernels run as plain Python) and runs small cases at
degree 2 (degree 1 for the adiabatic cavity), 4 x 8 ordinates:

* c0, c1  cavity A1_Nx6_Ny6 (50 triangles), thermalising walls, north T = 1,
          tau_R / tau_N = 1e-2 / 1e-3 and 1e-3 / 1e-4;
* c2, c3  periodic channel C1_Nx6_Ny6 (east / west periodic), the same taus;
* c4      jittered 6 x 6 square (meshes/J_Nx6_jit25_s0.msh, written by this
          package's own mesh writer, jitter 0.25 h, seed 0): west wall T = 1,
          east 0, north / south adiabatic, 1e-3 / 1e-4;
* c5      cavity A1_Nx11_Ny11 with adiabatic east / west walls, degree 1,
          1e-2 / 1e-3.

Each case runs to the tolerance 1e-8 or to its tmax, and the iteration
count, residual and mass histories and the final element integrals of T, qx,
qy go to gsis_nprocess.npz.  Cases run in parallel processes; the file is
written once all are done.

Run from the repository root: python tests/golden/make_gsis_nprocess.py
"""
from __future__ import annotations

import sys
import tempfile
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

REF = Path("/root/reference/solver-python")
HERE = Path(__file__).resolve().parent
OUT = HERE / "gsis_nprocess.npz"
JIT_MESH = HERE / "meshes" / "J_Nx6_jit25_s0.msh"

THERMAL = """  - {name: SWall, phyid: 11, type: thermalising, temp: 0.0}
  - {name: NWall, phyid: 12, type: thermalising, temp: 1.0}
  - {name: EWall, phyid: 13, type: thermalising, temp: 0.0}
  - {name: WWall, phyid: 14, type: thermalising, temp: 0.0}
"""
PERIODIC = """  - {name: SWall,  phyid: 11, type: thermalising, temp: 0.0}
  - {name: NWall,  phyid: 12, type: thermalising, temp: 1.0}
  - {name: Master, phyid: 13, type: periodic, temp: 0.0, xoff: -1.0}
  - {name: Slave,  phyid: 14, type: periodic, temp: 0.0, xoff:  1.0}
"""
LEFT_HOT = """  - {name: SWall, phyid: 11, type: nonthermalising, temp: 0.0}
  - {name: NWall, phyid: 12, type: nonthermalising, temp: 0.0}
  - {name: EWall, phyid: 13, type: thermalising,    temp: 0.0}
  - {name: WWall, phyid: 14, type: thermalising,    temp: 1.0}
"""
ADIABATIC_EW = """  - {name: SWall, phyid: 11, type: thermalising,    temp: 0.0}
  - {name: NWall, phyid: 12, type: thermalising,    temp: 1.0}
  - {name: EWall, phyid: 13, type: nonthermalising, temp: 0.0}
  - {name: WWall, phyid: 14, type: nonthermalising, temp: 0.0}
"""

# (mesh, boundaries, deg, tau_r, tau_n, tmax)
CASES = [
    (REF / "meshes" / "A1_Nx6_Ny6.msh", THERMAL, 2, 1e-2, 1e-3, 3000),
    (REF / "meshes" / "A1_Nx6_Ny6.msh", THERMAL, 2, 1e-3, 1e-4, 3000),
    (REF / "meshes" / "C1_Nx6_Ny6.msh", PERIODIC, 2, 1e-2, 1e-3, 3000),
    (REF / "meshes" / "C1_Nx6_Ny6.msh", PERIODIC, 2, 1e-3, 1e-4, 3000),
    (JIT_MESH, LEFT_HOT, 2, 1e-3, 1e-4, 3000),
    (REF / "meshes" / "A1_Nx11_Ny11.msh", ADIABATIC_EW, 1, 1e-2, 1e-3, 3000),
]


def case_yaml(k, mesh, bnd, deg, tau_r, tau_n, tmax, workdir):
    text = f"""iteration:  {{tol: 1.0e-8, tmax: {tmax}}}
scheme:     {{accflag: 1}}
velmesh:    {{npole: 4, nazim: 8}}
dg:         {{deg: {deg}}}
flow:       {{cv: 1.0, vg: 1.0, tau_r: {tau_r!r}, tau_n: {tau_n!r}, tau_thr: 1.0}}
mesh:       {{file: {mesh}}}
boundaries:
{bnd}restart:    {{enabled: false}}
output:     {{dir: {workdir}/out{k}, field: false, runtime_log: false, run_record: false}}
"""
    p = Path(workdir) / f"gsis_nprocess_{k}.yaml"
    p.write_text(text)
    return p


def run_case(k):
    sys.path.insert(0, str(REF))
    from pybte.config import Case
    from pybte.driver import Solver
    work = tempfile.mkdtemp(prefix=f"gsis_np{k}_")
    t0 = time.perf_counter()
    rec = Solver(Case.from_yaml(case_yaml(k, *CASES[k], work))).run()
    out = {
        f"c{k}_tau": np.array(CASES[k][3:5]),
        f"c{k}_deg": np.array(CASES[k][2]),
        f"c{k}_tmax": np.array(CASES[k][5]),
        f"c{k}_iterations": np.array(rec.iterations),
        f"c{k}_converged": np.array(bool(rec.converged)),
        f"c{k}_residual_history": np.asarray(rec.residual_history),
        f"c{k}_mass_history": np.asarray(rec.mass_history),
        f"c{k}_temp": np.asarray(rec.temp),
        f"c{k}_qx": np.asarray(rec.qx),
        f"c{k}_qy": np.asarray(rec.qy),
    }
    print(f"case {k}: {rec.iterations} iterations, converged {rec.converged}, "
          f"mass {rec.mass!r} ({time.perf_counter() - t0:.0f} s)", flush=True)
    return out


def main():
    if not JIT_MESH.exists():
        sys.path.insert(0, str(HERE.parent.parent))
        from __graft_entry__ import load_package
        load_package()
        from pbte_amd.mesh import write_square_mesh
        write_square_mesh(JIT_MESH, 6, jitter=0.25, seed=0)
    only = [int(a) for a in sys.argv[1:]] or list(range(len(CASES)))
    out = {}
    if OUT.exists():
        with np.load(OUT) as z:
            out.update({k: z[k] for k in z.files})
    with ProcessPoolExecutor(max_workers=len(only)) as ex:
        for res in ex.map(run_case, only):
            out.update(res)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, flush=True)


if __name__ == "__main__":
    main()
