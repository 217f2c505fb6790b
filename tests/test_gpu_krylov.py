"""GPU Krylov solves against the reference's certified fixed points
(tasks/t01-square/reference/*.npz: transport residual ~1e-16) and sweep
counts (F1/F2/F3_1_1: 21 / 23 / 27 sweeps, numerical_baseline.json)."""
import json

import numpy as np
import pytest

from conftest import CASES, GOLDEN

pytestmark = pytest.mark.gpu

SWEEPS = {"F1_1_1": 21, "F2_1_1": 23, "F3_1_1": 27}


def _run(pkg, tag, precision="fp64", **kw):
    c = pkg.Case.from_yaml(CASES / f"{tag}.yaml")
    c.scheme.method = "krylov"
    c.performance.precision = precision
    for k, v in kw.items():
        obj, attr = k.split("__")
        setattr(getattr(c, obj), attr, v)
    print(f"case krylov {tag} {precision} {kw}", flush=True)
    return pkg.Solver(c).run()


# F*_10_1e-2: tau_R = 10, tau_N = 1e-2, the normal-process-dominated
# certified fixed points (tasks/t01-square/reference)
@pytest.mark.parametrize("tag", ["F1_1_1", "F2_1_1", "F3_1_1", "F1_1e-1",
                                 "F1_10_1e-2", "F2_10_1e-2", "F3_10_1e-2"])
def test_krylov_fp64_reaches_certified_fixed_point(pkg, tag):
    rec = _run(pkg, tag)
    z = np.load(GOLDEN / "krylov" / f"{tag}.npz")
    idx = json.loads((GOLDEN / "krylov" / "index.json").read_text())[tag]
    assert rec.converged
    assert rec.diagnostics["transport_residual"] < 1e-8
    # Short runs take the certified run's product count to one product.  A
    # long run stops where its GMRES residual estimate meets the round-off
    # target (1e-14 relative) or the extension past tol 1e-2 ends
    # (krylov.py:321-345), which depends on rounding near that floor:
    # measured F1 / F2 / F3_10_1e-2 574 / 688 / 775 products against the
    # certified runs' 578 / 728 / 914.  Never more than the certified run.
    print(f"products {rec.iterations} certified run {idx['iterations']}", flush=True)
    if idx["iterations"] < 100:
        assert abs(rec.iterations - idx["iterations"]) <= 1
    else:
        assert 0.8 * idx["iterations"] <= rec.iterations <= idx["iterations"] + 1
    np.testing.assert_allclose(rec.temp_dofs, z["temp_dofs"], rtol=0, atol=1e-9)
    assert abs(rec.mass - float(z["mass"])) <= 1e-9


@pytest.mark.parametrize("refine_tol", [0.0, 1e-6], ids=["residual", "tq1e-6"])
@pytest.mark.parametrize("tag", ["F1_1_1", "F2_1_1", "F3_1_1", "F1_1e-2", "F1_1e-3",
                                 "F1_10_1e-2", "F2_10_1e-2", "F3_10_1e-2"])
def test_krylov_fp32_refined_matches_fp64(pkg, tag, refine_tol):
    """FP32 sweeps inside the FP64 defect correction reach the FP64 fixed
    point within 1e-6 relative L2 on T and on the heat flux, whether the
    refinement stops on the transport residual or on its own estimate of
    the T and q error (performance.refine_tol)."""
    rec = _run(pkg, tag, "fp32", performance__refine_tol=refine_tol)
    z = np.load(GOLDEN / "krylov" / f"{tag}.npz")
    assert rec.converged
    rel_t = np.linalg.norm(rec.temp - z["temp"]) / np.linalg.norm(z["temp"])
    q = np.hypot(rec.qx, rec.qy)
    qz = np.hypot(z["qx"], z["qy"])
    rel_q = np.linalg.norm(q - qz) / np.linalg.norm(qz)
    print(f"rel L2 T {rel_t:.2e} q {rel_q:.2e} sweeps fp64 {rec.diagnostics['sweeps_fp64']} "
          f"fp32 {rec.diagnostics['sweeps_fp32']}", flush=True)
    assert rel_t <= 1e-6 and rel_q <= 1e-6


@pytest.mark.parametrize("tag", ["F1_1e-2", "F1_1e-3"])
def test_krylov_moment_preconditioner_reaches_certified_fixed_point(pkg, tag):
    """scheme.krylov_precond with accflag 1: GMRES left-preconditioned by
    the HDG moment solve (krylov.py:369-414) converges to the same
    certified transport fixed point."""
    rec = _run(pkg, tag, scheme__accflag=1, scheme__krylov_precond=True)
    z = np.load(GOLDEN / "krylov" / f"{tag}.npz")
    assert rec.diagnostics["krylov_precond"] and rec.diagnostics["precond_calls"] > 0
    assert rec.converged
    assert rec.diagnostics["transport_residual"] < 1e-8
    np.testing.assert_allclose(rec.temp_dofs, z["temp_dofs"], rtol=0, atol=1e-9)
    assert abs(rec.mass - float(z["mass"])) <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_krylov_lookahead_product_changes_nothing(pkg, precision):
    """With no user callback GMRES queues the next product before it reads
    the Gram-Schmidt coefficients back; with a callback it does not.  Both
    runs give the same iterates bit for bit, the same histories and the
    same sweep count (a queued product the loop never reaches is
    discarded and uncounted)."""
    c = pkg.Case.from_yaml(CASES / "F1_1e-1.yaml")
    c.scheme.method = "krylov"
    c.performance.precision = precision
    print(f"case krylov lookahead F1_1e-1 {precision}", flush=True)
    fast = pkg.Solver(c).run()
    seen = []
    plain = pkg.Solver(c).run(callback=lambda step, res, s: seen.append(res))
    assert fast.converged and plain.converged
    assert fast.iterations == plain.iterations == len(seen)
    assert fast.sweep_count == plain.sweep_count
    np.testing.assert_array_equal(fast.temp_dofs, plain.temp_dofs)
    np.testing.assert_array_equal(np.asarray(fast.residual_history), np.asarray(plain.residual_history))
    np.testing.assert_array_equal(np.asarray(fast.mass_history), np.asarray(plain.mass_history))
    assert fast.diagnostics["sweeps_discarded"] > 0 and plain.diagnostics["sweeps_discarded"] == 0


def test_cgs2_matches_multidot_update_sequence(pkg):
    """pbte_vec_cgs2 (update and next dot product fused in one pass over
    the basis for up to 32 rows) gives the bits of two multidot / update
    pairs and a norm on the same vectors."""
    import torch
    c = pkg.Case.from_yaml(CASES / "F1_1_1.yaml")
    print("case cgs2 F1_1_1", flush=True)
    ctx = pkg.Solver(c).ctx
    rng = np.random.default_rng(5)
    n = 100003
    for m in (1, 7, 30, 33):
        Q = torch.tensor(rng.standard_normal((m, n)), device="cuda")
        w0 = torch.tensor(rng.standard_normal(n), device="cuda")
        w1 = w0.clone()
        h, hn = ctx.cgs2(Q, m, w1)
        w2 = w0.clone()
        c1 = ctx.multidot(Q, m, w2)
        ctx.update(Q, m, c1, w2)
        c2 = ctx.multidot(Q, m, w2)
        ctx.update(Q, m, c2, w2)
        assert np.array_equal(h, c1 + c2), m
        assert hn == ctx.norm(w2), m
        assert torch.equal(w1, w2), m
