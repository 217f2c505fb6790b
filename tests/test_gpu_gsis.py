"""GPU GSIS (HDG synthetic acceleration) against the reference's FROZEN
regression rows (tests/integration/test_regression.py:29-48): iteration
count exact and converged mass to 1e-12, on the shipped 200-triangle cavity
(thermalising walls, north wall T = 1, tol 1e-8)."""
import numpy as np
import pytest

from conftest import CASES, GOLDEN

pytestmark = pytest.mark.gpu

FROZEN_GSIS = {
    (1.0, 1e5, 3, 20, 40): (22, 0.2496873776739861),
    (1e-1, 1e5, 3, 20, 40): (30, 0.24999647389562388),
    (1e-2, 1e5, 1, 20, 40): (35, 0.24995968584823433),
    (1e-2, 1e5, 2, 20, 40): (29, 0.24995682731416324),
    (1e-2, 1e5, 3, 10, 20): (26, 0.24996286908141993),
    (1e-3, 1e5, 3, 20, 40): (50, 0.24999752787448726),
}


def _shipped(pkg, tau_r, tau_n, deg, npole, nazim, precision="fp64"):
    c = pkg.Case.from_yaml(CASES / "F1_1_1.yaml")
    c.scheme.accflag = 1
    c.flow.tau_r, c.flow.tau_n = tau_r, tau_n
    c.dg.deg = deg
    c.velmesh.npole, c.velmesh.nazim = npole, nazim
    c.iteration.tol, c.iteration.tmax = 1e-8, 200000
    c.performance.precision = precision
    return c


@pytest.mark.parametrize("key", list(FROZEN_GSIS), ids=lambda k: f"tau{k[0]:g}_deg{k[2]}_{k[3]}x{k[4]}")
def test_frozen_gsis(pkg, key):
    want_it, want_mass = FROZEN_GSIS[key]
    print(f"case frozen GSIS {key}", flush=True)
    rec = pkg.Solver(_shipped(pkg, *key)).run()
    assert rec.converged
    assert rec.iterations == want_it
    assert abs(rec.mass - want_mass) <= 1e-12


def test_gsis_adiabatic_cavity_conserves_energy(pkg):
    """Adiabatic side walls: the converged cavity holds int T dA = 0.5
    (test_deviations.py:63-65; global_solve.py comment: 30 iterations)."""
    c = pkg.Case.from_yaml(CASES / "F2_1_1.yaml")
    c.scheme.accflag = 1
    c.flow.tau_r, c.flow.tau_n = 1e-2, 1e5
    print("case GSIS adiabatic cavity", flush=True)
    rec = pkg.Solver(c).run()
    assert rec.converged
    assert abs(rec.mass - 0.5) < 1e-6


def test_gsis_periodic_channel_is_x_independent(pkg):
    """As test_deviations.py:69-87: x-spread of T below 1e-3 of its range,
    int T dA = 1/2 within 1e-5, a mirrored periodic pair present."""
    c = pkg.Case.from_yaml(CASES / "F3_1_1.yaml")
    c.scheme.accflag = 1
    c.flow.tau_r, c.flow.tau_n = 1e-2, 1e5
    print("case GSIS periodic channel", flush=True)
    s = pkg.Solver(c)
    rec = s.run()
    assert rec.converged
    # element mean temperature depends on y only: group elements by centroid y
    xy = s.mesh.nodes[s.mesh.tri_nodes].mean(axis=1)
    tbar = rec.temp / s.mesh.tri_area
    key = np.round(xy[:, 1], 9)
    spread = max(np.ptp(tbar[key == k]) for k in np.unique(key))
    assert int(s.bcdata.periodic_flip.sum()) > 0
    assert spread < 1e-3 * np.ptp(tbar)
    assert abs(rec.mass - 0.5) < 1e-5


@pytest.mark.parametrize("key", [(1e-2, 1e5, 2, 20, 40), (1e-3, 1e5, 3, 20, 40)],
                         ids=lambda k: f"tau{k[0]:g}_deg{k[2]}")
def test_gsis_fp32_refined_reaches_fp64_fixed_point(pkg, key):
    """fp32 GSIS: binary32 GSIS iterations solve for the corrections of a
    binary64 iterate; the result is the binary64 GSIS fixed point within
    the tolerance, with a handful of binary64 steps."""
    want_it, want_mass = FROZEN_GSIS[key]
    print(f"case GSIS fp32 refined {key}", flush=True)
    r32 = pkg.Solver(_shipped(pkg, *key, precision="fp32")).run()
    assert r32.converged
    assert r32.diagnostics["precision_scheme"] == "defect-correction"
    assert r32.diagnostics["steps_fp32"] > 0
    assert len(r32.diagnostics["fp64_steps"]) <= 6
    assert abs(r32.mass - want_mass) <= 1e-7
    r64 = pkg.Solver(_shipped(pkg, *key)).run()
    assert np.linalg.norm(r32.temp - r64.temp) / np.linalg.norm(r64.temp) <= 1e-6
    q32 = np.concatenate([r32.qx, r32.qy])
    q64 = np.concatenate([r64.qx, r64.qy])
    assert np.linalg.norm(q32 - q64) / np.linalg.norm(q64) <= 1e-6


@pytest.mark.parametrize("k", [0, 1], ids=["tau1e-2_converged", "tau1e-3_600_steps"])
def test_gsis_adiabatic_cavity_matches_reference_run(pkg, k):
    """Adiabatic east and west walls, GSIS (degree 1, 4 x 8 ordinates): the
    reference's own runs of this case (written by
    tests/golden/make_gsis_adiabatic.py, which imports the reference's
    pure-Python solver) -- iteration count, residual and mass histories,
    final moments.  At tau_R = 1e-3 the reference itself is still far from
    converged after 600 steps; the first 600 steps are compared."""
    g = np.load(GOLDEN / "gsis_adiabatic.npz")
    c = pkg.Case.from_yaml(CASES / "F1_1_1.yaml")
    c.scheme.accflag = 1
    c.dg.deg = 1
    c.velmesh.npole, c.velmesh.nazim = 4, 8
    c.flow.tau_r, c.flow.tau_n = (float(v) for v in g[f"c{k}_tau"])
    c.flow.tau_thr = 1.0
    c.iteration.tol, c.iteration.tmax = 1e-8, int(g[f"c{k}_tmax"])
    for b in c.boundaries:
        if b.name in ("EWall", "WWall"):
            b.type = "nonthermalising"
    print(f"case GSIS adiabatic cavity vs reference run {k}", flush=True)
    rec = pkg.Solver(c).run()
    want = g[f"c{k}_residual_history"]
    assert rec.converged == bool(g[f"c{k}_converged"])
    assert abs(rec.iterations - int(g[f"c{k}_iterations"])) <= 1
    n = min(len(want), len(rec.residual_history))
    np.testing.assert_allclose(rec.residual_history[:n], want[:n], rtol=1e-6, atol=0)
    np.testing.assert_allclose(rec.mass_history[:n], g[f"c{k}_mass_history"][:n], rtol=0, atol=1e-10)
    np.testing.assert_allclose(rec.temp, g[f"c{k}_temp"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(rec.qx, g[f"c{k}_qx"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(rec.qy, g[f"c{k}_qy"], rtol=0, atol=1e-10)


@pytest.mark.parametrize("tag,prec", [("F2_1_1", "fp64"), ("F3_1_1", "fp32"), ("F1_1_1", "fp32")])
def test_gsis_pending_correction_gives_the_eager_bits(pkg, tag, prec):
    """The accelerated step leaves the correction of the distribution pending
    (the next sweep rewrites the distribution; only the elements behind
    adiabatic walls are read before it).  Against the correction applied
    every step: the same residual and mass histories, moments and final
    distribution, bit for bit -- adiabatic walls, a periodic pair, the fp32
    phase and its switch to binary64."""
    c = pkg.Case.from_yaml(CASES / f"{tag}.yaml")
    c.scheme.accflag = 1
    c.flow.tau_r, c.flow.tau_n = 1e-2, 1e-3
    c.velmesh.npole, c.velmesh.nazim = 4, 8
    c.iteration.tol, c.iteration.tmax = 1e-8, 60
    c.performance.precision = prec
    c.performance.refine_switch = 1e-3
    print(f"case GSIS pending correction {tag} {prec}", flush=True)
    eager = pkg.Solver(c)
    assert eager.acc.lazy
    eager.acc.lazy = False
    r1 = eager._run_source()          # the source iteration (fp32: with its switch to binary64)
    lazy = pkg.Solver(c)
    r2 = lazy._run_source()
    assert r1.iterations == r2.iterations
    np.testing.assert_array_equal(r1.residual_history, r2.residual_history)
    np.testing.assert_array_equal(r1.mass_history, r2.mass_history)
    np.testing.assert_array_equal(r1.temp_dofs, r2.temp_dofs)
    np.testing.assert_array_equal(np.asarray(eager.vdf), np.asarray(lazy.vdf))
