"""pbte_hdg_assemble_static (no allocation: scratch in a file-scope device
array) against pbte_hdg_assemble, through the C-ABI: the same block values
and nonzero count, bit for bit, on two calls in a row, the second with a
smaller system (the scratch keeps the first call's contents); a system
beyond the scratch is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from conftest import CASES

pytestmark = pytest.mark.gpu


def _operators(pkg, deg):
    from pbte_amd import _native
    from pbte_amd.acceleration import build_hdg_device, face_sides
    from pbte_amd.device import ptr, stream_handle, to_dev
    from pbte_amd.problem import build_problem
    c = pkg.Case.from_yaml(CASES / "F3_1_1.yaml")
    c.scheme.accflag = 1
    c.dg.deg = deg
    c.flow.tau_r, c.flow.tau_n = 1e-2, 1e-3
    pb = build_problem(c)
    st = np.asarray(c.scheme.stabilisation, dtype=float)
    sides = face_sides(pb.mesh, pb.bcdata.bc_type)
    stream = C.c_void_p(stream_handle())
    ops = build_hdg_device(c, pb.mesh, pb.integrals, pb.bcdata, st, sides, stream)
    # the descriptor and device arrays for a direct call
    ig, mesh = pb.integrals, pb.mesh
    code = np.where(mesh.face_bc >= 0, pb.bcdata.bc_type[np.clip(mesh.face_bc, 0, None)], 0)
    kind = np.where((mesh.face_bc[mesh.tri_faces] < 0) | (code[mesh.tri_faces] == 3), 0,
                    np.where(code[mesh.tri_faces] == 1, 1, 2)).astype(np.int32)
    keep = [to_dev(np.ascontiguousarray(a, dtype=np.float64)) for a in
            (ig.int_tri_tri, ig.int_tri_tri_x, ig.int_tri_tri_y, ig.int_tri_fc, ig.int_tri_fc_fc,
             ig.int_fc_fc, mesh.tri_normal)]
    keep.append(to_dev(kind))
    hd = _native.PbteHdgDesc()
    hd.n_tris, hd.nd, hd.nf, hd.n_faces = mesh.n_tris, c.ndof_tri, c.ndof_fc, mesh.n_faces
    hd.variant_b = 0
    hd.cv, hd.tau_r, hd.tau_c = c.flow.cv, c.flow.tau_r, c.tau_c
    hd.st[:] = [float(v) for v in st]
    hd.drop_tol = 1e-30
    (hd.int_tri_tri, hd.int_tri_tri_x, hd.int_tri_tri_y, hd.int_tri_fc, hd.int_tri_fc_fc, hd.int_fc_fc,
     hd.tri_normal, hd.face_kind) = [ptr(t) for t in keep]
    t1, l1, t2, l2, periodic = sides
    mirror = (pb.bcdata.periodic_flip & periodic).astype(np.uint8)
    dev = [to_dev(np.ascontiguousarray(a, dtype=np.int32)) for a in (t1, l1, t2, l2)]
    dev += [to_dev(mirror), to_dev(np.ascontiguousarray(mesh.tri_faces, dtype=np.int32)),
            to_dev(ops["bsr_ptr"]), to_dev(ops["bsr_col"])]
    return hd, keep, dev, ops, stream


def _call(fn, hd, dev, ops, stream, nnzb=None):
    import torch

    from pbte_amd.device import ptr
    out = torch.full_like(ops["bsr_val"], float("nan"))
    nnz = C.c_int64(-1)
    err = C.create_string_buffer(512)
    nb = int(ops["bsr_col"].size) if nnzb is None else nnzb
    st = fn(C.byref(hd), *[ptr(t) for t in dev], nb, ptr(ops["aa_trace"]), ptr(ops["ba_sol"]), ptr(out),
            C.byref(nnz), stream, err, 512)
    return st, out, nnz.value, err.value.decode()


def test_static_assembly_matches_allocating_assembly(pkg):
    from pbte_amd import _native
    lib = _native.lib()
    for deg in (2, 1):              # the second call is the smaller system
        print(f"case no-alloc assembly F3_1_1 deg {deg}", flush=True)
        hd, keep, dev, ops, stream = _operators(pkg, deg)
        s1, v1, n1, e1 = _call(lib.pbte_hdg_assemble, hd, dev, ops, stream)
        s2, v2, n2, e2 = _call(lib.pbte_hdg_assemble_static, hd, dev, ops, stream)
        assert s1 == 0 and s2 == 0, (e1, e2)
        assert n1 == n2 == ops["nnz"]
        assert np.array_equal(v1.cpu().numpy(), v2.cpu().numpy())
    st, _, _, msg = _call(lib.pbte_hdg_assemble_static, hd, dev, ops, stream, nnzb=5_000_000)
    assert st < 0 and "scratch" in msg
