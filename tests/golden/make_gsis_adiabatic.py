"""Vectors for GSIS with diffusely reflecting (adiabatic) walls, which no
fixture of the reference pins: its frozen GSIS rows all use thermalising
walls (tests/integration/test_regression.py:29-48) and its adiabatic
cavity test checks convergence and int T dA only
(tests/integration/test_deviations.py:35-65).

Imports the reference's pure-Python solver (solver-python/pybte; numba is
absent here, so its kernels run as plain Python) and runs small cases
(10 x 10 cavity mesh A1_Nx11_Ny11, degree 1, 4 x 8 ordinates):

* hot north wall, cold south wall, adiabatic east and west walls (the
  layout of cases/cavity_adiabatic_gsis.yaml), GSIS at tau_N = 1e5 and
  tau_R = 1e-2 (to convergence, tol 1e-8) and tau_R = 1e-3 (the first 600
  iterations: it converges far more slowly);

and stores the iteration count, residual and mass histories and the final
element integrals of T, qx, qy in gsis_adiabatic.npz (rewritten after each
case).

Run from the repository root: python tests/golden/make_gsis_adiabatic.py
"""
from __future__ import annotations

import sys
import tempfile
import time
from pathlib import Path

import numpy as np

REF = Path("/root/reference/solver-python")
OUT = Path(__file__).resolve().parent / "gsis_adiabatic.npz"
CASES = [(1e-2, 1e5, 20000), (1e-3, 1e5, 600)]


def case_yaml(tau_r, tau_n, tmax, workdir):
    text = f"""iteration:  {{tol: 1.0e-8, tmax: {tmax}}}
scheme:     {{accflag: 1}}
velmesh:    {{npole: 4, nazim: 8}}
dg:         {{deg: 1}}
flow:       {{cv: 1.0, vg: 1.0, tau_r: {tau_r!r}, tau_n: {tau_n!r}, tau_thr: 1.0}}
mesh:       {{file: {REF / 'meshes' / 'A1_Nx11_Ny11.msh'}}}
boundaries:
  - {{name: SWall, phyid: 11, type: thermalising,    temp: 0.0}}
  - {{name: NWall, phyid: 12, type: thermalising,    temp: 1.0}}
  - {{name: EWall, phyid: 13, type: nonthermalising, temp: 0.0}}
  - {{name: WWall, phyid: 14, type: nonthermalising, temp: 0.0}}
restart:    {{enabled: false}}
output:     {{dir: {workdir}/out, field: false, runtime_log: false, run_record: false}}
"""
    p = Path(workdir) / f"gsis_adiabatic_{tau_r:g}.yaml"
    p.write_text(text)
    return p


def main():
    sys.path.insert(0, str(REF))
    from pybte.config import Case
    from pybte.driver import Solver
    work = tempfile.mkdtemp(prefix="gsis_adiabatic_")
    out = {}
    for k, (tau_r, tau_n, tmax) in enumerate(CASES):
        t0 = time.perf_counter()
        rec = Solver(Case.from_yaml(case_yaml(tau_r, tau_n, tmax, work))).run()
        print(f"tau_r {tau_r:g}: {rec.iterations} iterations, converged {rec.converged}, "
              f"mass {rec.mass!r} ({time.perf_counter() - t0:.0f} s)", flush=True)
        out[f"c{k}_tau"] = np.array([tau_r, tau_n])
        out[f"c{k}_tmax"] = np.array(tmax)
        out[f"c{k}_iterations"] = np.array(rec.iterations)
        out[f"c{k}_converged"] = np.array(bool(rec.converged))
        out[f"c{k}_residual_history"] = np.asarray(rec.residual_history)
        out[f"c{k}_mass_history"] = np.asarray(rec.mass_history)
        out[f"c{k}_temp"] = np.asarray(rec.temp)
        out[f"c{k}_qx"] = np.asarray(rec.qx)
        out[f"c{k}_qy"] = np.asarray(rec.qy)
        np.savez_compressed(OUT, **out)
        print("wrote", OUT, flush=True)


if __name__ == "__main__":
    main()
