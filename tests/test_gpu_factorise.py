"""The device factorisation of the element operators (pbte_factorise).

Singular element operators are reported as the reference reports them
(solver-python/pybte/sweep.py:122-128): every (direction, element) operator
with a zero pivot is counted and context set-up raises
``numpy.linalg.LinAlgError`` naming that count.  The factors built from the
reference-element matrices and cell scalars equal those built from the
stored element tensors bit for bit (the sweeps they give are compared), in
both precisions and from the one-pass binary64 + binary32 factorisation."""
import dataclasses

import numpy as np
import pytest

from conftest import CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_singular_element_operators_raise_linalgerror(pkg, precision):
    from importlib import import_module
    sweep = import_module(pkg.__name__ + ".sweep")
    c = pkg.Case.from_yaml(CASES / "F1_1_1.yaml")
    c.performance.precision = precision
    s = pkg.Solver(c)
    # zero the element matrices of the first k cells: their operators
    # (collision mass, streaming and face terms) are the zero matrix, whose
    # first pivot is exactly zero in every direction
    k = 5
    ig = s.integrals
    zeroed = {}
    for name in ("int_tri_tri", "int_tri_tri_x", "int_tri_tri_y", "int_tri_fc"):
        a = getattr(ig, name).copy()
        if name == "int_tri_tri":
            a[..., :k] = 0.0
        else:
            a[:k] = 0.0
        zeroed[name] = a
    # and the cell scalars the operators are rebuilt from
    co = ig.cell_ops.copy()
    co[:k] = 0.0
    zeroed["cell_ops"] = co
    bad_ig = dataclasses.replace(ig, **zeroed)
    want = k * s.ndir
    print(f"case singular {precision}: {k} cells x {s.ndir} directions", flush=True)
    with pytest.raises(np.linalg.LinAlgError, match=f"^{want} element operators are singular"):
        sweep.SweepContext.build(c, s.mesh, bad_ig, s.vel, s.order, s.bcdata)
    # the solver's own context is untouched and still factorises cleanly
    s.ctx.factorise(precision)
    rec = s.run()
    assert rec.converged


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_reference_matrix_factors_equal_tensor_factors(pkg, deg):
    """A context without the reference-element matrices factorises from the
    stored tensors; one sweep from the same moments gives the same bits in
    binary64 and binary32, the binary32 factors taken from the one-pass
    factorisation on one side and from a binary32-only pass on the other."""
    import torch
    from importlib import import_module
    sweep = import_module(pkg.__name__ + ".sweep")
    c = pkg.Case.from_yaml(CASES / "F2_1_1.yaml")
    c.dg.deg = deg
    c.performance.precision = "fp32"
    c.performance.fp32_operator = "stored"
    s = pkg.Solver(c)
    assert s.ctx.mode == "precomputed" and not s.ctx.fused
    plain = dataclasses.replace(s.integrals, ref_ops=None, cell_ops=None)
    print(f"case factor paths deg {deg}", flush=True)
    old = sweep.SweepContext.build(c, s.mesh, plain, s.vel, s.order, s.bcdata, precompute=False)
    old.factorise("fp64")
    old.factorise("fp32")
    rng = np.random.default_rng(deg)
    for m in (s.mom.ts_d, s.mom.qxs_d, s.mom.qys_d):
        m.copy_(torch.from_numpy(rng.standard_normal(tuple(m.shape))))
    for dtype in (torch.float64, torch.float32):
        a = torch.zeros((s.ndir, s.n_tris, s.ndof), dtype=dtype, device="cuda")
        b = torch.zeros_like(a)
        s.ctx.wall_flux(a)
        old.flux_wall.copy_(s.ctx.flux_wall)
        s.ctx.sweep(s.mom, a)
        old.sweep(s.mom, b)
        torch.cuda.synchronize()
        assert torch.equal(a, b), dtype
