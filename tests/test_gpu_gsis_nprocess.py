"""GSIS with a finite normal-process relaxation time against the
reference's own runs (tests/golden/gsis_nprocess.npz, written by
tests/golden/make_gsis_nprocess.py, which imports the reference's
pure-Python solver): tau_R / tau_N = 1e-2 / 1e-3 and 1e-3 / 1e-4 on the
cavity and the periodic channel, a jittered square with adiabatic walls and
the adiabatic cavity.  This is the regime where the drift term of the
correction (correction.py:29-75) carries weight, which the reference's
frozen rows (tau_N = 1e5) leave out.

FP64: the iteration count, the residual history (rtol 1e-6), the mass
history and the final element integrals (1e-10).  FP32 sweeps with the
binary64 finish: the converged cases reach the binary64 fixed point within
1e-6 relative L2 in T and in q."""
import numpy as np
import pytest

from conftest import GOLDEN, MESHES

pytestmark = pytest.mark.gpu

Z = np.load(GOLDEN / "gsis_nprocess.npz")
REF_MESH = {0: "A1_Nx6_Ny6", 1: "A1_Nx6_Ny6", 2: "C1_Nx6_Ny6", 3: "C1_Nx6_Ny6",
            4: "J_Nx6_jit25_s0", 5: "A1_Nx11_Ny11"}
WALLS = {
    "thermal": [("SWall", 11, "thermalising", 0.0), ("NWall", 12, "thermalising", 1.0),
                ("EWall", 13, "thermalising", 0.0), ("WWall", 14, "thermalising", 0.0)],
    "periodic": [("SWall", 11, "thermalising", 0.0), ("NWall", 12, "thermalising", 1.0),
                 ("Master", 13, "periodic", 0.0), ("Slave", 14, "periodic", 0.0)],
    "left_hot": [("SWall", 11, "nonthermalising", 0.0), ("NWall", 12, "nonthermalising", 0.0),
                 ("EWall", 13, "thermalising", 0.0), ("WWall", 14, "thermalising", 1.0)],
    "adiabatic_ew": [("SWall", 11, "thermalising", 0.0), ("NWall", 12, "thermalising", 1.0),
                     ("EWall", 13, "nonthermalising", 0.0), ("WWall", 14, "nonthermalising", 0.0)],
}
LAYOUT = {0: "thermal", 1: "thermal", 2: "periodic", 3: "periodic", 4: "left_hot", 5: "adiabatic_ew"}


def _case(pkg, k, precision="fp64"):
    tau_r, tau_n = (float(v) for v in Z[f"c{k}_tau"])
    bnd = []
    for name, phy, typ, temp in WALLS[LAYOUT[k]]:
        b = {"name": name, "phyid": phy, "type": typ, "temp": temp}
        if typ == "periodic":
            b["xoff"] = -1.0 if name == "Master" else 1.0
        bnd.append(b)
    return pkg.Case.from_dict({
        "iteration": {"tol": 1e-8, "tmax": int(Z[f"c{k}_tmax"])},
        "scheme": {"accflag": 1},
        "velmesh": {"npole": 4, "nazim": 8},
        "dg": {"deg": int(Z[f"c{k}_deg"])},
        "flow": {"cv": 1.0, "vg": 1.0, "tau_r": tau_r, "tau_n": tau_n, "tau_thr": 1.0},
        "mesh": {"file": str(MESHES / f"{REF_MESH[k]}.msh")},
        "boundaries": bnd,
        "output": {"field": False, "runtime_log": False, "run_record": False},
        "performance": {"precision": precision},
    })


@pytest.mark.parametrize("k", range(6), ids=lambda k: f"c{k}_{LAYOUT[k]}_tau{float(Z[f'c{k}_tau'][0]):g}")
def test_gsis_nprocess_matches_reference_run(pkg, k):
    print(f"case GSIS finite tau_N vs reference run c{k}", flush=True)
    rec = pkg.Solver(_case(pkg, k)).run()
    want = Z[f"c{k}_residual_history"]
    assert rec.converged == bool(Z[f"c{k}_converged"])
    assert rec.iterations == int(Z[f"c{k}_iterations"])
    np.testing.assert_allclose(rec.residual_history, want, rtol=1e-6, atol=0)
    np.testing.assert_allclose(rec.mass_history, Z[f"c{k}_mass_history"], rtol=0, atol=1e-10)
    for f in ("temp", "qx", "qy"):
        np.testing.assert_allclose(getattr(rec, f), Z[f"c{k}_{f}"], rtol=0, atol=1e-10)


@pytest.mark.parametrize("k", [0, 2, 5], ids=lambda k: f"c{k}_{LAYOUT[k]}")
def test_gsis_nprocess_fp32_reaches_fp64_fixed_point(pkg, k):
    print(f"case GSIS finite tau_N fp32 c{k}", flush=True)
    r32 = pkg.Solver(_case(pkg, k, "fp32")).run()
    assert r32.converged
    T = Z[f"c{k}_temp"]
    q = np.concatenate([Z[f"c{k}_qx"], Z[f"c{k}_qy"]])
    assert np.linalg.norm(r32.temp - T) / np.linalg.norm(T) <= 1e-6
    assert np.linalg.norm(np.concatenate([r32.qx, r32.qy]) - q) / np.linalg.norm(q) <= 1e-6
