"""GSIS fixed-point repair (``scheme.defect_omega``, relaxation and Anderson)
on the device, with the properties the reference's own tests state
(solver-python/tests/unit/test_defect_correction.py:41-98): omega = 0 is
the published scheme bit for bit; with the repair, GSIS lands on CIS's
discrete fixed point (temperature gap < 1e-8 of its scale, transport
residual < 1e-10) where the published scheme sits > 1e-3 away; Anderson
takes fewer outer refreshes than relaxation.  The device Anderson step is
also checked against numpy's least squares on a window of random pairs."""
import numpy as np
import pytest

from conftest import CASES, MESHES

pytestmark = pytest.mark.gpu


def _case(pkg, accflag, *, omega=0.0, every=1, anderson=0):
    # the reference test's case: the cavity (north wall T = 1), tau_R = 1,
    # tau_N = 1e5, 20 x 40 ordinates, degree 1 on the 6 x 6 mesh, tol 1e-12
    c = pkg.Case.from_yaml(CASES / "F1_1_1.yaml")
    c.scheme.accflag = accflag
    c.scheme.defect_omega = omega
    c.scheme.defect_every = every
    c.scheme.defect_anderson = anderson
    c.flow.tau_r, c.flow.tau_n = 1.0, 1e5
    c.dg.deg = 1
    c.velmesh.npole, c.velmesh.nazim = 20, 40
    c.mesh.file = str(MESHES / "A1_Nx6_Ny6.msh")
    c.iteration.tol, c.iteration.tmax = 1e-12, 20000
    c.output.field = c.output.run_record = c.output.runtime_log = False
    return c


def _run(pkg, **kw):
    print(f"case defect {kw}", flush=True)
    s = pkg.Solver(_case(pkg, **kw))
    return s, s.run()


def test_omega_zero_is_the_published_scheme(pkg):
    _, a = _run(pkg, accflag=1)
    _, b = _run(pkg, accflag=1, omega=0.0, every=0, anderson=8)
    assert a.iterations == b.iterations
    assert np.array_equal(a.temp, b.temp)
    assert a.mass == b.mass


def test_defect_recovers_the_cis_fixed_point(pkg):
    _, rc = _run(pkg, accflag=0)
    sp, rp = _run(pkg, accflag=1)
    sd, rd = _run(pkg, accflag=1, omega=1.0, every=0, anderson=8)
    scale = np.abs(rc.temp).max()
    gap_pub = np.abs(rc.temp - rp.temp).max() / scale
    gap_fix = np.abs(rc.temp - rd.temp).max() / scale
    print(f"gap published {gap_pub:.2e} repaired {gap_fix:.2e}", flush=True)
    assert rd.converged
    assert gap_pub > 1e-3
    assert gap_fix < 1e-8
    assert gap_fix < gap_pub / 1e4
    assert sd.ctx.true_residual(sd.mom, sd.vdf) < 1e-10
    assert sp.ctx.true_residual(sp.mom, sp.vdf) > 1e-6


def test_anderson_beats_plain_relaxation_on_the_outer_loop(pkg):
    sp, rp = _run(pkg, accflag=1, omega=1.0, every=0, anderson=0)
    sa, ra = _run(pkg, accflag=1, omega=1.0, every=0, anderson=8)
    print(f"outer refreshes relaxation {len(sp.acc.defect_steps)} anderson {len(sa.acc.defect_steps)}; "
          f"iterations {rp.iterations} / {ra.iterations}", flush=True)
    assert rp.converged and ra.converged
    assert len(sa.acc.defect_steps) < len(sp.acc.defect_steps)
    assert ra.iterations < rp.iterations


def test_diagnostics_report_the_defect(pkg):
    _, r = _run(pkg, accflag=1, omega=1.0, every=0, anderson=4)
    d = r.diagnostics
    assert d["defect_omega"] == 1.0
    assert d["defect_every"] == 0
    assert d["defect_anderson"] == 4
    assert d["defect_outer"] >= 1
    assert d["defect_norm"] > 0.0


def _anderson_numpy(D, G):
    """The Anderson type-II step of acceleration/__init__.py:160-172 in
    numpy: the least-squares problem on the whole difference matrix."""
    R = G - D
    dR = (R[1:] - R[:-1]).T
    gamma, *_ = np.linalg.lstsq(dR, R[-1], rcond=None)
    return G[-1] - (G[1:] - G[:-1]).T @ gamma


@pytest.mark.parametrize("window", [2, 4, 9])
def test_device_anderson_step_matches_numpy_least_squares(pkg, window):
    """Feed the device Anderson step a sequence of random (defect, offset)
    pairs; after each, the new defect equals numpy's lstsq-based step over
    the same window within 1e-12 of its scale, and the window slides once
    it holds ``anderson + 1`` pairs."""
    import torch
    s = pkg.Solver(_case(pkg, 1, omega=1.0, every=0, anderson=window - 1))
    acc = s.acc
    n = acc.defect_d.numel()
    rng = np.random.default_rng(window)
    D_hist, G_hist = [], []
    for step in range(window + 3):
        d = rng.standard_normal(n)
        # offsets near an affine map of d, so the window is well conditioned
        g = 0.5 * d + rng.standard_normal(n) * 0.1 + 1.0
        acc.defect_d.copy_(torch.from_numpy(d.reshape(acc.defect_d.shape)))
        o = torch.from_numpy(g).to(acc.defect_d.device)
        acc._anderson_step(o)
        D_hist.append(d)
        G_hist.append(g)
        D_hist, G_hist = D_hist[-window:], G_hist[-window:]
        got = acc.defect_d.reshape(-1).cpu().numpy()
        if len(D_hist) == 1:
            want = d + 1.0 * (g - d)
            assert np.array_equal(got, want)
        else:
            want = _anderson_numpy(np.asarray(D_hist), np.asarray(G_hist))
            err = np.abs(got - want).max() / np.abs(want).max()
            print(f"window {window} step {step} rel err {err:.2e}", flush=True)
            assert err < 1e-12


def test_device_relaxation_is_numpy_bit_for_bit(pkg):
    """refresh_defect's relaxation and its step measure: the numpy
    expressions of acceleration/__init__.py:143-147, bitwise."""
    import torch
    s = pkg.Solver(_case(pkg, 1, omega=0.7, every=0))
    acc = s.acc
    rng = np.random.default_rng(3)
    shape = acc.defect_d.shape
    d = rng.standard_normal(shape)
    uq = rng.standard_normal(acc.last["uq"].shape)
    mk = rng.standard_normal(shape)
    acc.defect_d.copy_(torch.from_numpy(d))
    acc.last["uq"].copy_(torch.from_numpy(uq))
    acc.mkin_d.copy_(torch.from_numpy(mk))
    step = acc.refresh_defect()
    o = uq.reshape(shape) - mk
    want_step = float(np.abs(o - d).max())
    want = d.copy()
    want += 0.7 * (o - want)
    assert step == want_step
    assert np.array_equal(acc.defect_d.cpu().numpy(), want)
    # a NaN in the offset is reported, as numpy's max reports it
    uq[0, 0] = np.nan
    acc.last["uq"].copy_(torch.from_numpy(uq))
    assert np.isnan(acc.refresh_defect())
