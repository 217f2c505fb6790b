"""Extract the reference's own golden data into small committed fixtures.

Reads (never imports or runs) files under /root/reference:

* solver-python/meshes/{A1,C1}_Nx{6,11}_Ny{6,11}.msh          -> meshes/
* tasks/t01-square/verifier/cases/F*.yaml                      -> cases/
* docs/refactor/evidence/smoke_cis.tar.gz :: F{1,2,3}_1_1.npz  -> cis_F*_1_1.npz
  (converged source-iteration vdf and residual / mass histories); the vdf is
  kept for 8 directions only, the moments and histories in full
* tasks/t01-square/reference/*.npz + index.json                 -> krylov/
  (certified Krylov fixed points: temp, qx, qy, temp_dofs, mass)
* docs/refactor/evidence/numerical_baseline.json                -> copied

Run from the repository root: python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import io
import json
import shutil
import tarfile
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent
KEEP_DIRS = np.array([0, 1, 17, 42, 99, 123, 150, 199])


def main():
    (OUT / "meshes").mkdir(exist_ok=True)
    for name in ("A1_Nx6_Ny6", "A1_Nx11_Ny11", "C1_Nx6_Ny6", "C1_Nx11_Ny11"):
        shutil.copy(REF / "solver-python/meshes" / f"{name}.msh", OUT / "meshes" / f"{name}.msh")
    (OUT / "cases").mkdir(exist_ok=True)
    for p in sorted((REF / "tasks/t01-square/verifier/cases").glob("F*.yaml")):
        text = p.read_text().replace("../meshes/", "../meshes/")
        (OUT / "cases" / p.name).write_text(text)
    with tarfile.open(REF / "docs/refactor/evidence/smoke_cis.tar.gz") as tar:
        for m in tar.getmembers():
            if m.name.endswith("_1_1.npz") and "/evaluations/" in "/" + m.name:
                z = np.load(io.BytesIO(tar.extractfile(m).read()), allow_pickle=False)
                tag = Path(m.name).stem
                np.savez_compressed(OUT / f"cis_{tag}.npz", vdf_dirs=KEEP_DIRS,
                                    vdf=z["vdf"][KEEP_DIRS], temp=z["temp"], qx=z["qx"],
                                    qy=z["qy"], residual_history=z["residual_history"],
                                    mass_history=z["mass_history"])
    (OUT / "krylov").mkdir(exist_ok=True)
    src = REF / "tasks/t01-square/reference"
    for p in sorted(src.glob("*.npz")):
        z = np.load(p, allow_pickle=False)
        np.savez_compressed(OUT / "krylov" / p.name, **{k: z[k] for k in
                            ("temp", "qx", "qy", "temp_dofs", "iterations",
                             "transport_residual", "mass")})
    shutil.copy(src / "index.json", OUT / "krylov" / "index.json")
    shutil.copy(REF / "docs/refactor/evidence/numerical_baseline.json",
                OUT / "numerical_baseline.json")
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
