"""HDG operators and trace matrix built on the device (pbte_hdg_local,
pbte_hdg_assemble) against the host restatement of assembly.py:44-268
(oracle/hdg.py): inv(AA_SOL), AA_TRACE, BA_SOL and every entry of K, on
thermalising, adiabatic and periodic (mirrored pair) walls, degrees 1-3,
variants A and B, a jittered mesh."""
import ctypes as C

import numpy as np
import pytest

from conftest import CASES, MESHES

pytestmark = pytest.mark.gpu

SETUPS = [("F1_1_1.yaml", 1, "A", None), ("F1_1_1.yaml", 2, "A", None), ("F1_1_1.yaml", 3, "A", None),
          ("F2_1_1.yaml", 2, "A", None), ("F3_1_1.yaml", 2, "A", None), ("F3_1_1.yaml", 1, "B", None),
          ("F1_1_1.yaml", 2, "B", None), ("F1_1_1.yaml", 2, "A", MESHES / "J_Nx6_jit25_s0.msh")]


@pytest.mark.parametrize("case,deg,variant,mesh", SETUPS,
                         ids=lambda v: str(v if not hasattr(v, "name") else v.name))
def test_device_hdg_build_matches_host(pkg, case, deg, variant, mesh):
    from scipy.sparse import bsr_matrix

    from oracle.hdg import assemble_global_matrix, build_ba_sol, build_local_operators
    from pbte_amd.acceleration import build_hdg_device, face_sides
    from pbte_amd.device import stream_handle
    from pbte_amd.problem import build_problem
    print(f"case hdg build {case} deg {deg} variant {variant} mesh {mesh}", flush=True)
    c = pkg.Case.from_yaml(CASES / case)
    c.scheme.accflag = 1
    c.scheme.acc_variant = variant
    c.dg.deg = deg
    c.flow.tau_r, c.flow.tau_n = 1e-2, 1e-3
    if mesh is not None:
        c.mesh.file = str(mesh)
    pb = build_problem(c)
    st = np.asarray(c.scheme.stabilisation, dtype=float)
    bc = pb.bcdata.bc_type
    inv_h, aat_h = build_local_operators(pb.mesh, pb.integrals, c.ndof_tri, c.ndof_fc, c.flow.cv,
                                         c.flow.tau_r, c.tau_c, st, variant)
    ba_h = build_ba_sol(pb.mesh, pb.integrals, c.ndof_tri, c.ndof_fc, st, bc, variant)
    K_h, _ = assemble_global_matrix(pb.mesh, pb.integrals, ba_h, aat_h, c.ndof_fc, bc,
                                    drop_tol=1e-30 if variant == "A" else 0.0,
                                    periodic_flip=pb.bcdata.periodic_flip)
    ops = build_hdg_device(c, pb.mesh, pb.integrals, pb.bcdata, st, face_sides(pb.mesh, bc),
                           C.c_void_p(stream_handle()))
    inv_d = ops["inv_aa_sol"].cpu().numpy()
    aat_d = ops["aa_trace"].cpu().numpy()
    ba_d = ops["ba_sol"].cpu().numpy()
    # the inverse to a relative 1e-13 of each element's largest entry
    scale = np.abs(inv_h).max(axis=(1, 2))[:, None, None]
    assert (np.abs(inv_d - inv_h) <= 1e-13 * scale).all()
    ascale = np.abs(aat_h).max(axis=(1, 2, 3))[:, None, None, None]
    assert (np.abs(aat_d - aat_h) <= 1e-13 * ascale).all()
    assert np.array_equal(ba_d, ba_h)
    blk = 3 * c.ndof_fc
    K_d = bsr_matrix((ops["bsr_val"].cpu().numpy(), ops["bsr_col"], ops["bsr_ptr"]),
                     shape=K_h.shape).tocsr()
    diff = abs(K_d - K_h)
    assert diff.max() <= 1e-13 * abs(K_h).max()
    assert ops["nnz"] == K_d.count_nonzero()
    # the same sparsity up to entries that cancel to within rounding
    assert abs(ops["nnz"] - K_h.nnz) <= 0.01 * K_h.nnz
