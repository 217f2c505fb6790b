"""Singular element operators are reported as the reference reports them
(solver-python/pybte/sweep.py:122-128): the device factorisation counts
every (direction, element) operator with a zero pivot and context set-up
raises ``numpy.linalg.LinAlgError`` naming that count."""
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
        a[:k] = 0.0
        zeroed[name] = a
    bad_ig = dataclasses.replace(ig, **zeroed)
    want = k * s.ndir
    print(f"case singular {precision}: {k} cells x {s.ndir} directions", flush=True)
    with pytest.raises(np.linalg.LinAlgError, match=f"^{want} element operators are singular"):
        sweep.SweepContext.build(c, s.mesh, bad_ig, s.vel, s.order, s.bcdata)
    # the solver's own context is untouched and still factorises cleanly
    s.ctx.factorise(precision)
    rec = s.run()
    assert rec.converged
