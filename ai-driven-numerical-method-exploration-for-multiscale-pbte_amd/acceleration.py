"""GSIS: general synthetic iterative scheme with the HDG macroscopic system
(``solver-python/pybte/acceleration/__init__.py:34-217``).

After each sweep, on the device:

1. ``pbte_gsis_moments``  one pass over ``vdf``: the CIS moments and the four
   third-order angular moments of the higher-order-term source;
2. ``pbte_gsis_source``   HoT stress rows, then through one operator per
   element the T, qx, qy rows of ``AA_SRC = inv(AA_SOL) work`` and the
   element sides' sums ``BA_SOL AA_SRC`` of the trace right-hand side;
3. ``pbte_gsis_rhs``      trace right-hand side from those sums, wall rows
   from the kinetic traces (adiabatic flux projected on the tangent);
4. trace solve            ``K u = ffa`` with ``K`` factorised once on the
   device (``trace_solve.TraceSolver``: nested-dissection multifrontal, in
   place of the reference's SuperLU, ``global_solve.py:120-124, 190-195``);
5. ``pbte_gsis_local``    ``UQ = AA_SRC + AA_TRACE U_TRACE``, the T, qx, qy
   rows (``pbte_gsis_uq`` gives every row, for the field output);
6. ``pbte_gsis_correct``  damped blend of ``UQ`` into ``vdf`` and the moments.

On a binary32 distribution (the binary32 iterations of an fp32 run) steps
2, 4 and 5 read binary32 copies of the element operators and of the trace
factors; sums and every vector stay binary64.

``scheme.defect_omega`` (the fixed-point repair of the reference) keeps the
reference's semantics: with ``omega = 0`` the defect is identically zero.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .constants import BC_NONTHERMALISING, BC_PERIODIC, BC_THERMALISING, PI
from .device import dev_zeros, ptr, stream_handle, to_dev, torch
from .trace_solve import TraceSolver

__all__ = ["Acceleration", "build_hdg_device", "face_sides", "prepare", "trace_pattern"]


def face_sides(mesh, bc_type):
    """Per face the (element, local face) of each side that contributes to
    its trace row, ``-1`` where absent (``global_solve.py:89-106``): an
    interior face has its two elements; a periodic face its own element
    and its partner's; a wall face its own element only."""
    bc = mesh.face_bc
    code = np.where(bc >= 0, np.asarray(bc_type)[np.clip(bc, 0, None)], 0)
    interior = bc < 0
    periodic = ~interior & (code == BC_PERIODIC)
    own_t, own_l = mesh.adjacent_arrays()
    t1 = np.where(interior, mesh.face_tri[:, 0], own_t).astype(np.int64)
    l1 = np.where(interior, mesh.face_lfc[:, 0], own_l).astype(np.int64)
    t2 = np.where(interior, mesh.face_tri[:, 1], -1).astype(np.int64)
    l2 = np.where(interior, mesh.face_lfc[:, 1], -1).astype(np.int64)
    partner = np.where(periodic, mesh.face_pair, 0).astype(np.int64)
    t2 = np.where(periodic, own_t[partner], t2)
    l2 = np.where(periodic, own_l[partner], l2)
    return t1, l1, t2, l2, periodic


def trace_pattern(tri_faces, t1, t2):
    """Block pattern of the trace matrix: row f couples to the faces of the
    elements on its sides (its own among them), ascending."""
    F = t1.size
    cols = np.concatenate([tri_faces[t1], np.where(t2[:, None] >= 0, tri_faces[np.maximum(t2, 0)], -1)],
                          axis=1)
    cols = np.sort(cols, axis=1)
    keep = cols >= 0
    keep[:, 1:] &= cols[:, 1:] != cols[:, :-1]
    counts = keep.sum(axis=1)
    bptr = np.zeros(F + 1, dtype=np.int32)
    np.cumsum(counts, out=bptr[1:])
    return bptr, np.ascontiguousarray(cols[keep], dtype=np.int32)


def build_hdg_device(c, mesh, ig, bcdata, st, sides, stream):
    """The HDG operators and the trace matrix built on the device
    (``pbte_hdg_local``, ``pbte_hdg_assemble``; ``assembly.py:44-268``):
    a dict of device tensors ``inv_aa_sol [N,7P,7P]``, ``aa_trace
    [N,3,7P,3Q]``, ``ba_sol [N,3,3Q,7P]``, ``bsr_val [nnzb,3Q,3Q]``, the host
    block pattern ``bsr_ptr``, ``bsr_col`` and the count ``nnz`` of nonzero
    entries of K."""
    t1, l1, t2, l2, periodic = sides
    P, Q, N, F = c.ndof_tri, c.ndof_fc, mesh.n_tris, mesh.n_faces
    bc_type = bcdata.bc_type
    variant = c.scheme.acc_variant
    code = np.where(mesh.face_bc >= 0, bc_type[np.clip(mesh.face_bc, 0, None)], 0)
    fcode = code[mesh.tri_faces]
    interior = mesh.face_bc[mesh.tri_faces] < 0
    kind = np.where(interior | (fcode == BC_PERIODIC), 0,
                    np.where(fcode == BC_THERMALISING, 1, 2)).astype(np.int32)
    tens = [to_dev(np.ascontiguousarray(a, dtype=np.float64)) for a in
            (ig.int_tri_tri, ig.int_tri_tri_x, ig.int_tri_tri_y, ig.int_tri_fc, ig.int_tri_fc_fc,
             ig.int_fc_fc, mesh.tri_normal)]
    kind_d = to_dev(kind)
    hd = _native.PbteHdgDesc()
    hd.n_tris, hd.nd, hd.nf, hd.n_faces = N, P, Q, F
    hd.variant_b = 1 if variant == "B" else 0
    hd.cv, hd.tau_r, hd.tau_c = c.flow.cv, c.flow.tau_r, c.tau_c
    hd.st[:] = [float(v) for v in st]
    hd.drop_tol = 1e-30 if variant == "A" else 0.0
    (hd.int_tri_tri, hd.int_tri_tri_x, hd.int_tri_tri_y, hd.int_tri_fc, hd.int_tri_fc_fc,
     hd.int_fc_fc, hd.tri_normal) = [ptr(t) for t in tens]
    hd.face_kind = ptr(kind_d)
    inv_aa_sol = dev_zeros((N, 7 * P, 7 * P))
    aa_trace = dev_zeros((N, 3, 7 * P, 3 * Q))
    ba_sol = dev_zeros((N, 3, 3 * Q, 7 * P))
    err = C.create_string_buffer(512)
    if _native.lib().pbte_hdg_local(C.byref(hd), ptr(inv_aa_sol), ptr(aa_trace), ptr(ba_sol), stream,
                                    err, 512) < 0:
        raise np.linalg.LinAlgError(err.value.decode())
    mirror = (bcdata.periodic_flip & periodic).astype(np.uint8)
    bptr, bcol = trace_pattern(mesh.tri_faces.astype(np.int64), t1, t2)
    side_d = [to_dev(np.ascontiguousarray(a, dtype=np.int32)) for a in (t1, l1, t2, l2, mesh.tri_faces)]
    mirror_d, bptr_d, bcol_d = to_dev(mirror), to_dev(bptr), to_dev(bcol)
    blk = 3 * Q
    bval = dev_zeros((bcol.size, blk, blk))
    nnz = C.c_int64(0)
    if _native.lib().pbte_hdg_assemble(C.byref(hd), *[ptr(t) for t in side_d[:4]], ptr(mirror_d),
                                       ptr(side_d[4]), ptr(bptr_d), ptr(bcol_d), int(bcol.size),
                                       ptr(aa_trace), ptr(ba_sol), ptr(bval), C.byref(nnz), stream,
                                       err, 512) < 0:
        raise _native.NativeError(f"hdg_assemble: {err.value.decode()}")
    return {"inv_aa_sol": inv_aa_sol, "aa_trace": aa_trace, "ba_sol": ba_sol, "bsr_val": bval,
            "bsr_ptr": bptr, "bsr_col": bcol, "nnz": int(nnz.value)}


def prepare(case, problem, stream) -> dict:
    """The part of the set-up that needs no sweep context: the HDG operators
    on the device and the factorised trace matrix.  ``Solver`` runs it on a
    worker thread while the sweep context is built (the host schedule and
    this device work overlap; every native call releases the GIL)."""
    c = case
    mesh, ig, bcdata = problem.mesh, problem.integrals, problem.bcdata
    st = np.asarray(c.scheme.stabilisation, dtype=float).copy()
    if c.scheme.scale_by_hmin:
        st[1] /= mesh.hmin
        st[2] /= mesh.hmin
    stream = C.c_void_p(stream)
    sides = face_sides(mesh, bcdata.bc_type)
    ops = build_hdg_device(c, mesh, ig, bcdata, st, sides, stream)
    trace = TraceSolver(mesh, sides[0], sides[2], ops["bsr_ptr"], ops["bsr_col"], 3 * c.ndof_fc, stream=stream,
                        leaf_cells=c.performance.trace_leaf)
    trace.factor(ops.pop("bsr_val"))
    return {"st": st, "sides": sides, "ops": ops, "trace": trace}


class Acceleration:
    def __init__(self, solver, prepared: dict | None = None):
        c = solver.case
        self.solver = s = solver
        self.variant = c.scheme.acc_variant
        self.nd, self.nf = c.ndof_tri, c.ndof_fc
        if prepared is None:
            prepared = prepare(c, s.problem, stream_handle())
        self.st = st = prepared["st"]
        mesh, ig = s.mesh, s.integrals
        bc_type = s.bcdata.bc_type
        P, N, F, Q = self.nd, s.n_tris, mesh.n_faces, self.nf
        t1, l1, t2, l2, periodic = prepared["sides"]
        ops = prepared.pop("ops")
        inv_aa_sol, aa_trace, ba_sol = ops["inv_aa_sol"], ops["aa_trace"], ops["ba_sol"]
        blk = 3 * Q
        self.n_trace, self.nnz = F * blk, ops["nnz"]
        self.trace = prepared["trace"]
        self.trace_info = self.trace.info()
        del ops
        # the T, qx, qy columns of inv(AA_SOL) serve the moment preconditioner
        self.inv_aa_sol = inv_aa_sol if c.scheme.krylov_precond else None
        flip = (s.bcdata.periodic_flip & periodic).astype(np.uint8)
        code = np.where(mesh.face_bc >= 0, bc_type[np.clip(mesh.face_bc, 0, None)], 0)
        wall_th = np.flatnonzero((mesh.face_bc >= 0) & (code == BC_THERMALISING))
        wall_ad = np.flatnonzero((mesh.face_bc >= 0) & (code == BC_NONTHERMALISING))
        wall = np.concatenate([wall_th, wall_ad]).astype(np.int32)
        wtri, wlfc = t1[wall].astype(np.int32), l1[wall].astype(np.int32)
        scale = c.flow.tau_r if self.variant == "B" else 1.0
        keep = []

        def f64(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return _native.f64(a)

        def i32(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(a)
            return _native.i32(a)

        cxv, cyv, dw = s.vel.flat()
        g = _native.PbteGsisData()
        g.cxv, g.cyv, g.domega = f64(cxv), f64(cyv), f64(dw)
        g.inv_aa_sol = C.cast(C.c_void_p(ptr(inv_aa_sol)), C.POINTER(C.c_double))
        g.aa_trace = C.cast(C.c_void_p(ptr(aa_trace)), C.POINTER(C.c_double))
        g.ba_sol = C.cast(C.c_void_p(ptr(ba_sol)), C.POINTER(C.c_double))
        g.tri_fc_fc, g.tri_hmin = f64(ig.int_tri_fc_fc), f64(mesh.tri_hmin)
        g.tri_faces = i32(mesh.tri_faces)
        g.t1, g.l1, g.t2, g.l2 = i32(t1), i32(l1), i32(t2), i32(l2)
        flip = np.ascontiguousarray(flip)
        keep.append(flip)
        g.flip = flip.ctypes.data_as(C.POINTER(C.c_uint8))
        g.n_wall, g.n_wall_th = int(wall.size), int(wall_th.size)
        g.wall_face, g.wall_tri, g.wall_lfc = i32(wall), i32(wtri), i32(wlfc)
        g.wall_nx = f64(mesh.tri_normal[wtri, wlfc, 0])
        g.wall_ny = f64(mesh.tri_normal[wtri, wlfc, 1])
        g.tc5 = c.tau_c * scale / 5.0
        g.tau_thr = c.flow.tau_thr
        g.drift = c.tau_c / c.flow.tau_n * 3.0 / 4.0 / PI / c.flow.vg / c.flow.vg
        g.variant_b = 1 if self.variant == "B" else 0
        g.use_hmin = 1 if self.variant == "A" else 0
        ctx = s.ctx
        _native.check(ctx.handle, _native.lib().pbte_gsis_setup(ctx.handle, C.byref(g)), "gsis_setup")
        torch.cuda.current_stream().synchronize()      # the operators are built from these buffers
        # BA_SOL also serves the moment preconditioner's operator
        self.ba_sol = ba_sol if c.scheme.krylov_precond else None
        del aa_trace, ba_sol

        # the step's buffers, one set per distribution precision (the
        # refined run alternates binary32 and binary64 steps; the field
        # output reads the last binary64 step's)
        self._sets: dict = {}
        self.last = self._set(torch.float64)
        self.mkin_d = dev_zeros((3, P, N))
        self.defect_d = dev_zeros((3, P, N))
        # The correction of the distribution is left pending: the next sweep
        # writes every value of the distribution from the corrected moments,
        # so only the values read before it is written need the correction --
        # the elements behind adiabatic walls (their emission is computed
        # first) and behind periodic faces (the sweep reads the partner's
        # value of the previous iterate where the partner comes later in
        # its order) -- and the distribution the run hands back.  Not with
        # cycles the sweep breaks: those read the previous iterate anywhere.
        self.lazy = int(s.order.cycles_broken.sum()) == 0
        self.pending = False
        self.coef_d = dev_zeros((3, P, N)) if self.lazy else None
        read = np.flatnonzero((mesh.face_bc >= 0) & ((code == BC_NONTHERMALISING) | (code == BC_PERIODIC)))
        bcells = np.unique(t1[read]).astype(np.int32)
        self.boundary_cells_d = to_dev(bcells) if read.size else None
        rest = np.setdiff1d(np.arange(N, dtype=np.int32), bcells).astype(np.int32)
        self.rest_cells_d = to_dev(rest) if self.lazy and rest.size else None
        self._state_coef: dict = {}       # per distribution buffer: its correction, pending or None
        self.omega = c.scheme.defect_omega
        self.defect_every = c.scheme.defect_every
        self.anderson = c.scheme.defect_anderson
        self._ncall = 0
        self.defect_steps: list[float] = []
        self._o_d = None                  # the offset o = UQ[:3P] - M_kin
        self._and_pairs = None            # Anderson window on the device
        self._and_work = None
        self._and_rows: list[int] = []    # window rows, oldest first

    # -- one acceleration cycle -------------------------------------------
    def _set(self, dtype) -> dict:
        st = self._sets.get(dtype)
        if st is None:
            P, N = self.nd, self.solver.n_tris
            F, Q = self.solver.mesh.n_faces, self.nf
            st = self._sets[dtype] = {"mom7": dev_zeros((7, P, N)), "src": dev_zeros((3 * P, N)),
                                      "ffa": dev_zeros((F, 3 * Q)), "u": dev_zeros((F, 3 * Q)),
                                      "uq": dev_zeros((3 * P, N))}
        return st

    def _stage(self, vdf) -> dict:
        """Steps 1-5 on a swept distribution: its moments, the element pass,
        the trace right-hand side, the trace solve and the local recovery
        of T, qx, qy -- with the binary32 operators and trace factors when
        the distribution is binary32."""
        h = self.solver.ctx.handle
        lib = _native.lib()
        fp32 = vdf.dtype == torch.float32
        prec = _native.PBTE_FP32 if fp32 else _native.PBTE_FP64
        st = self.last = self._set(vdf.dtype)
        _native.check(h, lib.pbte_gsis_moments(h, prec, ptr(vdf), ptr(st["mom7"])), "gsis_moments")
        _native.check(h, lib.pbte_gsis_source(h, prec, ptr(st["mom7"]), ptr(st["src"])), "gsis_source")
        _native.check(h, lib.pbte_gsis_rhs(h, prec, ptr(vdf), ptr(st["ffa"])), "gsis_rhs")
        self.trace.solve(st["ffa"], st["u"], "fp32" if fp32 else "fp64")
        _native.check(h, lib.pbte_gsis_local(h, prec, ptr(st["src"]), ptr(st["u"]), ptr(st["uq"])),
                      "gsis_local")
        return st

    def apply(self) -> None:
        s = self.solver
        h = s.ctx.handle
        lib = _native.lib()
        vdf = s.sweep_vdf
        prec = _native.PBTE_FP64 if vdf.dtype == torch.float64 else _native.PBTE_FP32
        st = self._stage(vdf)
        m = s.mom
        defect = ptr(self.defect_d) if self.omega else 0
        if self.lazy:
            _native.check(h, lib.pbte_gsis_correct_moments(h, ptr(st["mom7"]), ptr(st["uq"]), defect,
                                                           ptr(m.ts_d), ptr(m.qxs_d), ptr(m.qys_d),
                                                           ptr(m.temp_d), ptr(m.qx_d), ptr(m.qy_d),
                                                           ptr(self.mkin_d), ptr(self.coef_d)),
                          "gsis_correct_moments")
            self.pending = True
        else:
            _native.check(h, lib.pbte_gsis_correct(h, prec, ptr(vdf), ptr(st["mom7"]), ptr(st["uq"]),
                                                   defect, ptr(m.ts_d), ptr(m.qxs_d), ptr(m.qys_d),
                                                   ptr(m.temp_d), ptr(m.qx_d), ptr(m.qy_d),
                                                   ptr(self.mkin_d)), "gsis_correct")
        self._ncall += 1
        if self.omega and self.defect_every and self._ncall % self.defect_every == 0:
            self.refresh_defect()

    def full_uq(self):
        """UQ [7P, N] of the last binary64 step (of the last step when none
        ran in binary64), every row: the field output's stress rows."""
        st = self._sets.get(torch.float64) if torch.float64 in self._sets else self.last
        uq = dev_zeros((7 * self.nd, self.solver.n_tris))
        h = self.solver.ctx.handle
        _native.check(h, _native.lib().pbte_gsis_uq(h, ptr(st["mom7"]), ptr(st["u"]), ptr(uq)), "gsis_uq")
        return uq

    def settle(self, boundary_only: bool = False) -> None:
        """Apply the pending correction to the distribution: to every element,
        or (``boundary_only``) to the elements behind adiabatic walls and
        periodic faces, the only ones the next step reads before its sweep
        rewrites them (the correction stays pending for the rest)."""
        if not self.pending:
            return
        s = self.solver
        vdf = s.sweep_vdf
        prec = _native.PBTE_FP64 if vdf.dtype == torch.float64 else _native.PBTE_FP32
        h = s.ctx.handle
        if boundary_only:
            cells = self.boundary_cells_d
            if cells is not None:
                _native.check(h, _native.lib().pbte_gsis_apply_correction(
                    h, prec, ptr(vdf), ptr(self.coef_d), ptr(cells), int(cells.numel())), "gsis_apply_correction")
            return
        _native.check(h, _native.lib().pbte_gsis_apply_correction(h, prec, ptr(vdf), ptr(self.coef_d), None, 0),
                      "gsis_apply_correction")
        self.pending = False

    def correct_state(self, vdf, out) -> None:
        """The correction of a swept distribution for the fixed-point
        operator of the refined iteration: the corrected moments into
        ``out = (ts, qxs, qys, temp, qx, qy)``, the correction applied to
        the boundary elements of ``vdf`` (those the state's emission and
        periodic blocks are read from) and left pending for the rest
        (``finish_state``)."""
        h = self.solver.ctx.handle
        lib = _native.lib()
        prec = _native.PBTE_FP64 if vdf.dtype == torch.float64 else _native.PBTE_FP32
        st = self._stage(vdf)
        key = vdf.data_ptr()
        coef = self._state_coef.get(key)
        if coef is None:
            coef = self._state_coef[key] = [dev_zeros(self.coef_d.shape), False]
        _native.check(h, lib.pbte_gsis_correct_moments(h, ptr(st["mom7"]), ptr(st["uq"]), None,
                                                       *(ptr(t) for t in out), None, ptr(coef[0])),
                      "gsis_correct_moments")
        cells = self.boundary_cells_d
        if cells is not None:
            _native.check(h, lib.pbte_gsis_apply_correction(h, prec, ptr(vdf), ptr(coef[0]), ptr(cells),
                                                            int(cells.numel())), "gsis_apply_correction")
        coef[1] = True
        self._ncall += 1

    def finish_state(self, vdf) -> None:
        """Apply the correction ``correct_state`` left pending to the other
        elements of ``vdf``."""
        coef = self._state_coef.get(vdf.data_ptr())
        if coef is None or not coef[1]:
            return
        h = self.solver.ctx.handle
        prec = _native.PBTE_FP64 if vdf.dtype == torch.float64 else _native.PBTE_FP32
        cells = self.rest_cells_d
        if cells is not None:
            _native.check(h, _native.lib().pbte_gsis_apply_correction(
                h, prec, ptr(vdf), ptr(coef[0]), ptr(cells), int(cells.numel())), "gsis_apply_correction")
        coef[1] = False

    def drop_state(self) -> None:
        """Release the refined iteration's correction buffers."""
        self._state_coef.clear()

    def swept(self) -> None:
        """The sweep has rewritten the distribution: nothing is pending."""
        self.pending = False

    def trace_solve(self, ffa, out, precision: str = "fp64") -> None:
        """``K u = ffa`` on the device (``global_solve.py:190-195``)."""
        self.trace.solve(ffa, out, precision)

    @property
    def solve_count(self) -> int:
        return self.trace.solve_count

    # -- the fixed-point repair (acceleration/__init__.py:118-172) ----------
    def refresh_defect(self) -> float:
        """Move the defect toward the current offset ``o = UQ[:3P] - M_kin``
        and return ``max |o - defect|`` before the move: relaxation
        ``defect += omega (o - defect)``, or an Anderson step on the map
        ``defect -> o`` (``scheme.defect_anderson``).  Every vector
        operation runs on the device (``pbte_vec_*``)."""
        ctx = self.solver.ctx
        if self._o_d is None:
            self._o_d = dev_zeros(self.defect_d.numel())
        o = self._o_d
        ctx.defect(self.last["uq"], self.mkin_d, None, o)
        if self.anderson:
            step = ctx.relax(0.0, o, self.defect_d)
            self._anderson_step(o)
        else:
            step = ctx.relax(self.omega, o, self.defect_d)
        self.defect_steps.append(step)
        return step

    def _anderson_step(self, o) -> None:
        """Anderson (type II) on ``d -> o(d)`` over the last
        ``defect_anderson + 1`` pairs (acceleration/__init__.py:149-172).
        The pairs stay on the device.  ``min |dR gamma - R_last|`` (``R =
        G - D``, ``dR`` its successive differences) is solved through a thin
        QR of ``dR`` -- Gram-Schmidt twice on the device -- and numpy's
        least squares on the small triangle, with the singular-value cut-off
        numpy would use for ``dR`` itself; then ``d = G_last - dG gamma`` on
        the device.  A non-finite result empties the window and relaxes."""
        ctx = self.solver.ctx
        n = o.numel()
        W = self.anderson + 1
        if self._and_pairs is None:
            self._and_pairs = dev_zeros((2, W, n))     # defect rows, offset rows
            self._and_work = dev_zeros((3, W, n))      # R, differences of G, Q
        Dw, Gw = self._and_pairs[0], self._and_pairs[1]
        rows = self._and_rows
        free = rows.pop(0) if len(rows) == W else min(set(range(W)) - set(rows))
        Dw[free].copy_(self.defect_d.view(-1))
        Gw[free].copy_(o.view(-1))
        rows.append(free)
        m = len(rows)
        if m == 1:
            ctx.relax(self.omega, o, self.defect_d)
            return
        R, dG, Q = self._and_work[0], self._and_work[1], self._and_work[2]
        for j, r in enumerate(rows):
            ctx.defect(Gw[r], Dw[r], None, R[j])
        k = m - 1
        Rk = np.zeros((k, k))
        for j in range(k):
            ctx.defect(R[j + 1], R[j], None, Q[j])
            if j == 0:
                nrm = ctx.norm(Q[0])
            else:
                Rk[:j, j], nrm = ctx.cgs2(Q, j, Q[j])
            Rk[j, j] = nrm
            if nrm > 0.0:
                ctx.axpby(1.0 / nrm, Q[j], 0.0, Q[j])
        b = ctx.multidot(Q, k, R[k])
        gamma = np.linalg.lstsq(Rk, b, rcond=np.finfo(np.float64).eps * max(n, k))[0]
        for j in range(k):
            ctx.defect(Gw[rows[j + 1]], Gw[rows[j]], None, dG[j])
        new = R[0]
        new.copy_(Gw[rows[-1]])
        ctx.update(dG, k, gamma, new)
        if not (np.all(np.isfinite(gamma)) and np.isfinite(ctx.relax(0.0, new, new))):
            rows.clear()
            ctx.relax(self.omega, o, self.defect_d)
            return
        self.defect_d.view(-1).copy_(new)

    def _defect_norm(self) -> float:
        """``max |defect|`` on the device (the offset buffer, cleared, as
        the zero it is measured against)."""
        if self._o_d is None:
            self._o_d = dev_zeros(self.defect_d.numel())
        self._o_d.zero_()
        return self.solver.ctx.relax(0.0, self.defect_d, self._o_d)

    def diagnostics(self) -> dict:
        return {"acc_variant": self.variant, "global_matrix_n": int(self.n_trace),
                "global_matrix_nnz": int(self.nnz), "global_solve_count": int(self.solve_count),
                "stabilisation": [float(v) for v in self.st], "defect_omega": float(self.omega),
                "defect_every": int(self.defect_every), "defect_anderson": int(self.anderson),
                "defect_outer": len(self.defect_steps),
                "defect_norm": self._defect_norm() if self.omega else 0.0,
                "trace_solve": "device-multifrontal", "trace_ordering": "nested-dissection",
                "trace_factor_entries": self.trace_info["factor_entries"]}
