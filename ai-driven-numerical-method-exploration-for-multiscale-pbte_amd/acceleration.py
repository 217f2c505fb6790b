"""GSIS: general synthetic iterative scheme with the HDG macroscopic system
(``solver-python/pybte/acceleration/__init__.py:34-217``).

After each sweep, on the device:

1. ``pbte_gsis_moments``  one pass over ``vdf``: the CIS moments and the four
   third-order angular moments of the higher-order-term source;
2. ``pbte_gsis_source``   HoT stress rows and ``AA_SRC = inv(AA_SOL) work``;
3. ``pbte_gsis_rhs``      trace right-hand side, wall rows from the kinetic
   traces (adiabatic flux projected on the tangent);
4. trace solve            ``K u = ffa`` with ``K`` factorised once on the
   device (``trace_solve.TraceSolver``: nested-dissection multifrontal, in
   place of the reference's SuperLU, ``global_solve.py:120-124, 190-195``);
5. ``pbte_gsis_local``    ``UQ = AA_SRC + AA_TRACE U_TRACE``;
6. ``pbte_gsis_correct``  damped blend of ``UQ`` into ``vdf`` and the moments.

``scheme.defect_omega`` (the fixed-point repair of the reference) keeps the
reference's semantics: with ``omega = 0`` the defect is identically zero.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .constants import BC_NONTHERMALISING, BC_THERMALISING, PI
from .device import dev_zeros, ptr, stream_handle, torch
from .hdg import assemble_global_matrix, build_ba_sol, build_local_operators, face_sides
from .trace_solve import TraceSolver, block_pattern

__all__ = ["Acceleration"]

class Acceleration:
    def __init__(self, solver):
        c = solver.case
        self.solver = s = solver
        self.variant = c.scheme.acc_variant
        self.nd, self.nf = c.ndof_tri, c.ndof_fc
        st = np.asarray(c.scheme.stabilisation, dtype=float).copy()
        if c.scheme.scale_by_hmin:
            st[1] /= s.mesh.hmin
            st[2] /= s.mesh.hmin
        self.st = st
        mesh, ig = s.mesh, s.integrals
        bc_type = s.bcdata.bc_type
        inv_aa_sol, aa_trace = build_local_operators(mesh, ig, self.nd, self.nf, c.flow.cv,
                                                     c.flow.tau_r, c.tau_c, st, self.variant)
        # kept on the host for the Krylov moment preconditioner only
        self.inv_aa_sol = inv_aa_sol if c.scheme.krylov_precond else None
        ba_sol = build_ba_sol(mesh, ig, self.nd, self.nf, st, bc_type, self.variant)
        drop = 1e-30 if self.variant == "A" else 0.0
        self.K, self.nnz = assemble_global_matrix(mesh, ig, ba_sol, aa_trace, self.nf, bc_type,
                                                  drop_tol=drop,
                                                  periodic_flip=s.bcdata.periodic_flip)
        t1, l1, t2, l2, periodic = face_sides(mesh, bc_type)
        bptr, bcol, bval = block_pattern(self.K, 3 * self.nf)
        self.trace = TraceSolver(mesh, t1, t2, bptr, bcol, 3 * self.nf, stream=C.c_void_p(stream_handle()),
                                 leaf_cells=c.performance.trace_leaf)
        self.trace.factor(bval)
        self.trace_info = self.trace.info()
        del bval
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
        g.inv_aa_sol, g.aa_trace, g.ba_sol = f64(inv_aa_sol), f64(aa_trace), f64(ba_sol)
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

        P, N, F, Q = self.nd, s.n_tris, mesh.n_faces, self.nf
        self.mom7 = dev_zeros((7, P, N))
        self.aa_src_d = dev_zeros((7 * P, N))
        self.ffa_d = dev_zeros((F, 3 * Q))
        self.u_d = dev_zeros((F, 3 * Q))
        self.uq_d = dev_zeros((7 * P, N))
        self.mkin_d = dev_zeros((3, P, N))
        self.defect_d = dev_zeros((3, P, N))
        self.omega = c.scheme.defect_omega
        self.defect_every = c.scheme.defect_every
        self.anderson = c.scheme.defect_anderson
        self._ncall = 0
        self.defect_steps: list[float] = []
        self._and_d: list = []
        self._and_g: list = []

    # -- one acceleration cycle -------------------------------------------
    def apply(self) -> None:
        s = self.solver
        ctx = s.ctx
        lib = _native.lib()
        vdf = s.sweep_vdf
        prec = _native.PBTE_FP64 if vdf.dtype == torch.float64 else _native.PBTE_FP32
        h = ctx.handle
        _native.check(h, lib.pbte_gsis_moments(h, prec, ptr(vdf), ptr(self.mom7)), "gsis_moments")
        _native.check(h, lib.pbte_gsis_source(h, ptr(self.mom7), ptr(self.aa_src_d)), "gsis_source")
        _native.check(h, lib.pbte_gsis_rhs(h, prec, ptr(vdf), ptr(self.aa_src_d), ptr(self.ffa_d)),
                      "gsis_rhs")
        self.trace_solve()
        _native.check(h, lib.pbte_gsis_local(h, ptr(self.aa_src_d), ptr(self.u_d), ptr(self.uq_d)),
                      "gsis_local")
        m = s.mom
        defect = ptr(self.defect_d) if self.omega else 0
        _native.check(h, lib.pbte_gsis_correct(h, prec, ptr(vdf), ptr(self.mom7), ptr(self.uq_d),
                                               defect, ptr(m.ts_d), ptr(m.qxs_d), ptr(m.qys_d),
                                               ptr(m.temp_d), ptr(m.qx_d), ptr(m.qy_d),
                                               ptr(self.mkin_d)), "gsis_correct")
        self._ncall += 1
        if self.omega and self.defect_every and self._ncall % self.defect_every == 0:
            self.refresh_defect()

    def trace_solve(self, ffa=None, out=None) -> None:
        """``K u = ffa`` on the device (``global_solve.py:190-195``); by
        default from ``ffa_d`` into ``u_d``."""
        ffa = self.ffa_d if ffa is None else ffa
        out = self.u_d if out is None else out
        self.trace.solve(ffa, out)

    @property
    def solve_count(self) -> int:
        return self.trace.solve_count

    # -- the fixed-point repair (acceleration/__init__.py:118-172) ----------
    def refresh_defect(self) -> float:
        P = self.nd
        o = self.uq_d[:3 * P].view(3, P, -1) - self.mkin_d
        step = float((o - self.defect_d).abs().max())
        if self.anderson:
            self._anderson_step(o)
        else:
            self.defect_d.add_(self.omega * (o - self.defect_d))
        self.defect_steps.append(step)
        return step

    def _anderson_step(self, o) -> None:
        self._and_d.append(self.defect_d.reshape(-1).cpu().numpy().copy())
        self._and_g.append(o.reshape(-1).cpu().numpy().copy())
        if len(self._and_d) > self.anderson + 1:
            self._and_d.pop(0)
            self._and_g.pop(0)
        if len(self._and_d) == 1:
            self.defect_d.add_(self.omega * (o - self.defect_d))
            return
        D = np.asarray(self._and_d)
        G = np.asarray(self._and_g)
        R = G - D
        dR = (R[1:] - R[:-1]).T
        gamma, *_ = np.linalg.lstsq(dR, R[-1], rcond=None)
        new = G[-1] - (G[1:] - G[:-1]).T @ gamma
        if not np.all(np.isfinite(new)):
            self._and_d.clear()
            self._and_g.clear()
            self.defect_d.add_(self.omega * (o - self.defect_d))
            return
        self.defect_d.copy_(torch.from_numpy(new.reshape(self.defect_d.shape)))

    def diagnostics(self) -> dict:
        return {"acc_variant": self.variant, "global_matrix_n": int(self.K.shape[0]),
                "global_matrix_nnz": int(self.K.nnz), "global_solve_count": int(self.solve_count),
                "stabilisation": [float(v) for v in self.st], "defect_omega": float(self.omega),
                "defect_every": int(self.defect_every), "defect_anderson": int(self.anderson),
                "defect_outer": len(self.defect_steps),
                "defect_norm": float(self.defect_d.abs().max()) if self.omega else 0.0,
                "trace_solve": "device-multifrontal", "trace_ordering": "nested-dissection",
                "trace_factor_entries": self.trace_info["factor_entries"]}

    def global_matrix(self):
        return self.K
