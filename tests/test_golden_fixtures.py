"""The committed reference-run fixtures are well formed (CPU only)."""
import numpy as np

from conftest import GOLDEN


def test_gsis_adiabatic_fixture():
    g = np.load(GOLDEN / "gsis_adiabatic.npz")
    assert bool(g["c0_converged"])
    h = g["c0_residual_history"]
    assert len(h) == int(g["c0_iterations"]) and h[-1] < 1e-8 <= h[-2]
    # the adiabatic side walls leave int T dA = 1/2 (test_deviations.py:63-65)
    assert abs(float(g["c0_mass_history"][-1]) - 0.5) < 1e-3
    assert g["c0_temp"].shape == g["c0_qx"].shape == g["c0_qy"].shape == (200,)
