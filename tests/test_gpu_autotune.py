"""The sweep tile-shape autotuner: a tuned run is the untuned run, bit for
bit (tile shapes change where items are swept, never their arithmetic), and
the chosen shape is kept in the cache file and reused."""
import json

import numpy as np
import pytest

from conftest import CASES

pytestmark = pytest.mark.gpu


def _case(pkg, tag, **kw):
    c = pkg.Case.from_yaml(CASES / f"{tag}.yaml")
    for k, v in kw.items():
        obj, attr = k.split("__")
        setattr(getattr(c, obj), attr, v)
    return c


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_autotuned_run_is_the_untuned_run(pkg, tmp_path, prec):
    from pbte_amd import autotune
    cache = tmp_path / "tiles.json"
    tuned = _case(pkg, "F2_1_1", performance__precision=prec, performance__autotune=True,
                  performance__autotune_cache=str(cache), iteration__tmax=30)
    plain = _case(pkg, "F2_1_1", performance__precision=prec, iteration__tmax=30)
    print(f"case autotune F2_1_1 {prec}", flush=True)
    s = pkg.Solver(tuned)
    shape = s.tile_shape
    assert shape is not None and shape["ms"] > 0
    assert len(shape["timed"]) >= 2                  # at least two shapes built and timed
    assert (shape["max_bundle"], shape["block"], shape["variant"]) in autotune.CANDIDATES
    r1 = s.run()
    r0 = pkg.Solver(plain).run()
    assert r1.iterations == r0.iterations
    np.testing.assert_array_equal(r1.residual_history, r0.residual_history)
    np.testing.assert_array_equal(r1.temp_dofs, r0.temp_dofs)
    assert r1.diagnostics["tile_shape"]["max_bundle"] == shape["max_bundle"]
    table = json.loads(cache.read_text())
    key = autotune.signature(tuned, s.problem)
    assert table[key]["max_bundle"] == shape["max_bundle"]
    # a new process would read the file: drop the in-process memo and
    # check that the entry comes back without timing anything
    autotune._memo.clear()
    again = autotune.tune(tuned, s.problem)
    assert again == table[key]


def test_mixed_mode_autotune_is_bitwise_untuned(pkg, tmp_path):
    """An fp32 case whose binary64 factors do not fit ("mixed" factor
    mode): the tuner times its binary32 sweeps, and the tuned run is the
    untuned run bit for bit."""
    from pbte_amd.problem import build_problem
    base = _case(pkg, "F2_1_1", performance__precision="fp32", iteration__tmax=20)
    pb = build_problem(base)
    nbytes = float(pb.vel.ndir) * pb.mesh.n_tris * base.ndof_tri ** 2 * 8.0
    tuned = _case(pkg, "F2_1_1", performance__precision="fp32", performance__autotune=True,
                  iteration__tmax=20)
    s = pkg.Solver(tuned, max_bytes=0.9 * nbytes)
    assert s.ctx.mode == "mixed"
    assert s.tile_shape is not None
    r1 = s.run()
    r0 = pkg.Solver(base, max_bytes=0.9 * nbytes).run()
    np.testing.assert_array_equal(r1.residual_history, r0.residual_history)
    np.testing.assert_array_equal(r1.temp_dofs, r0.temp_dofs)


def test_untimeable_case_keeps_default_schedule_once(pkg, monkeypatch):
    """An fp64 on-the-fly case stores no factors: nothing can be timed, the
    tuner returns None (the default schedule is kept), and a second solver
    of the same case does not build the candidates again."""
    from pbte_amd import autotune, sweep
    c = _case(pkg, "F1_1_1", performance__autotune=True, performance__precompute_inverse=False,
              iteration__tmax=3)
    autotune._memo.clear()
    s = pkg.Solver(c)
    assert s.tile_shape is None and s.ctx.mode == "onthefly"
    built = []
    orig = sweep.SweepContext.build.__func__

    def counting(cls, *a, **kw):
        built.append(kw.get("max_bundle"))
        return orig(cls, *a, **kw)
    monkeypatch.setattr(sweep.SweepContext, "build", classmethod(counting))
    s2 = pkg.Solver(c)
    assert s2.tile_shape is None
    assert len(built) == 1                   # the solver's own context only
    assert s2.run().iterations == 3
