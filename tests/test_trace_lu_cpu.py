"""Symbolic phase of the device trace solve (host only, no device): the
dissection owns every unknown once, keeps every block of K, and rejects a
pattern whose block couples faces that share no element."""
import numpy as np
import pytest

from trace_helpers import trace_system


@pytest.mark.parametrize("case,deg", [("F1_1_1.yaml", 2), ("F3_1_1.yaml", 1)])
def test_symbolic_phase_covers_the_system(pkg, case, deg):
    from pbte_amd.trace_solve import TraceSolver, block_pattern
    pb, K, t1, t2, blk = trace_system(pkg, case, deg)
    ptr, col, val = block_pattern(K, blk)
    ts = TraceSolver(pb.mesh, t1, t2, ptr, col, blk, leaf_cells=8)
    info = ts.info()
    assert info["unknowns"] == K.shape[0]
    assert info["blocks"] == col.size
    # a binary tree: every internal node has two children
    assert info["nodes"] % 2 == 1 and info["levels"] >= 4
    # the fronts hold at least the matrix itself
    assert info["factor_entries"] >= K.nnz


def test_symbolic_phase_rejects_unrelated_block(pkg):
    from pbte_amd import _native
    from pbte_amd.trace_solve import TraceSolver, block_pattern
    pb, K, t1, t2, blk = trace_system(pkg, "F1_1_1.yaml", 1)
    ptr, col, _ = block_pattern(K, blk)
    # face 0 (a south-wall face) coupled to the last face (far away)
    col = col.copy()
    col[ptr[0]] = pb.mesh.n_faces - 1
    with pytest.raises(_native.NativeError, match="share no cell"):
        TraceSolver(pb.mesh, t1, t2, ptr, col, blk, leaf_cells=8)
