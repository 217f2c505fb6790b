"""Device trace solve (multifrontal, binary64) against SciPy's SuperLU with
the reference's own call (global_solve.py:120-124, COLAMD) on the trace
matrices of the test meshes: thermalising, adiabatic and periodic walls,
degrees 1-3, a jittered mesh, two dissection leaf sizes."""
import numpy as np
import pytest

from conftest import MESHES
from trace_helpers import trace_system

pytestmark = pytest.mark.gpu

CASES_TL = [("F1_1_1.yaml", 1, None), ("F1_1_1.yaml", 2, None), ("F1_1_1.yaml", 3, None),
            ("F2_1_1.yaml", 2, None), ("F3_1_1.yaml", 2, None), ("F3_1_1.yaml", 3, None),
            ("F1_1_1.yaml", 2, MESHES / "J_Nx6_jit25_s0.msh")]


@pytest.mark.parametrize("case,deg,mesh", CASES_TL,
                         ids=lambda v: str(v if not hasattr(v, "name") else v.name))
@pytest.mark.parametrize("leaf", [4, 16])
@pytest.mark.parametrize("pivoting", [0, 1], ids=["block_pivots", "partial_pivots"])
def test_trace_solve_matches_superlu(pkg, case, deg, mesh, leaf, pivoting):
    import torch
    from scipy.sparse.linalg import splu

    from pbte_amd.trace_solve import TraceSolver, block_pattern
    print(f"case trace solve {case} deg {deg} mesh {mesh} leaf {leaf} pivoting {pivoting}", flush=True)
    pb, K, t1, t2, blk = trace_system(pkg, case, deg, mesh=mesh)
    ptr, col, val = block_pattern(K, blk)
    ts = TraceSolver(pb.mesh, t1, t2, ptr, col, blk, leaf_cells=leaf)
    ts.set_pivoting(pivoting)
    ts.factor(val)
    q = ts.quality()
    assert q["pivoting"] == ("block" if pivoting == 0 else "partial")
    assert q["backward_error"] <= 1e-14
    lu = splu(K.tocsc(), permc_spec="COLAMD")
    rng = np.random.default_rng(7)
    for _ in range(3):
        b = rng.standard_normal(K.shape[0])
        bd = torch.from_numpy(b).cuda()
        xd = torch.empty_like(bd)
        ts.solve(bd, xd)
        x = xd.cpu().numpy()
        xr = lu.solve(b)
        err = np.abs(x - xr).max() / np.abs(xr).max()
        res = np.abs(K @ x - b).max() / np.abs(b).max()
        print(f"  rel diff {err:.2e} residual {res:.2e}", flush=True)
        assert err <= 1e-12
        assert res <= 1e-13


def test_trace_solve_singular_raises(pkg):
    from pbte_amd.trace_solve import TraceSolver, block_pattern
    pb, K, t1, t2, blk = trace_system(pkg, "F1_1_1.yaml", 1)
    ptr, col, val = block_pattern(K, blk)
    val = val.copy()
    # zero the block row of one face: its unknowns have no pivot
    val[ptr[5]:ptr[6]] = 0.0
    ts = TraceSolver(pb.mesh, t1, t2, ptr, col, blk, leaf_cells=8)
    with pytest.raises(np.linalg.LinAlgError):     # both pivotings fail
        ts.factor(val)


def test_graph_replays_are_the_direct_solve(pkg):
    """The solve recorded as a HIP graph and replayed (several times, with
    other device work and waits between the replays, in both precisions,
    again after a new factorisation) gives the directly launched solve's
    bits on every replay."""
    import torch

    from pbte_amd.trace_solve import TraceSolver, block_pattern
    pb, K, t1, t2, blk = trace_system(pkg, "F3_1_1.yaml", 2)
    ptr, col, val = block_pattern(K, blk)
    direct = TraceSolver(pb.mesh, t1, t2, ptr, col, blk, leaf_cells=4)
    direct.set_graphs(False)
    graphed = TraceSolver(pb.mesh, t1, t2, ptr, col, blk, leaf_cells=4)
    rng = np.random.default_rng(11)
    bd = torch.empty(K.shape[0], dtype=torch.float64, device="cuda")
    xa, xb = torch.empty_like(bd), torch.empty_like(bd)
    for scale in (1.0, 3.0):
        direct.factor(val * scale)
        graphed.factor(val * scale)
        for rep in range(4):
            bd.copy_(torch.from_numpy(rng.standard_normal(K.shape[0])))
            for prec in ("fp64", "fp32"):
                direct.solve(bd, xa, prec)
                graphed.solve(bd, xb, prec)
                junk = torch.randn(4096, device="cuda", dtype=torch.float64).sum()   # other work between replays
                torch.cuda.synchronize()
                assert torch.equal(xa, xb), (scale, rep, prec, float(junk))
