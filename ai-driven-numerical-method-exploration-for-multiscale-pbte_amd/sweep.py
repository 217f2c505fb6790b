"""``SweepContext``: the operator boundary of the solve on the device.

Replaces the reference's ``SweepContext`` (``solver-python/pybte/sweep.py:37-197``)
with the same method names and side effects:

* ``sweep(mom, vdf)``       one transport sweep, in place on ``vdf``,
                            ``sweep_count += 1`` (``sweep.py:132-158``);
* ``true_residual(mom, vdf)`` the dw-weighted ``||A f - b|| / ||b||`` with no
                            sweep (``sweep.py:161-197``);
* ``factorise()``           LU of every (direction, element) operator, once
                            (``sweep.py:116-129``); singular blocks raise
                            ``numpy.linalg.LinAlgError``.

``vdf`` is a device tensor ``[D, N, P]``; its dtype selects the sweep
precision (float64: binary64 LU and accumulation; float32: binary32 LU and
accumulation, moments still summed in binary64).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .constants import BC_PERIODIC, PI
from .device import dev_zeros, ptr, stream_handle, torch

__all__ = ["SweepContext"]

_PREC = {torch.float64: _native.PBTE_FP64, torch.float32: _native.PBTE_FP32}


def _other_lfc(mesh, bcdata) -> np.ndarray:
    """Local face index, in the element across, of each (element, face)."""
    n = mesh.n_tris
    out = np.full((n, 3), -1, dtype=np.int32)
    f = mesh.tri_faces
    t0 = mesh.face_tri[f, 0]
    own = np.arange(n)[:, None]
    interior = mesh.face_bc[f] < 0
    lfc_other = np.where(t0 == own, mesh.face_lfc[f, 1], mesh.face_lfc[f, 0])
    out[interior] = lfc_other[interior]
    if bcdata.has_periodic:
        _, adj_lfc = mesh.adjacent_arrays()
        bc = mesh.face_bc[f]
        per = (bc >= 0) & (np.asarray(bcdata.bc_type)[np.clip(bc, 0, None)] == BC_PERIODIC)
        out[per] = adj_lfc[mesh.face_pair[f[per]]]
    return out


class SweepContext:
    def __init__(self):
        self.handle = None
        self.sweep_count = 0
        self.factorisations = 0                 # element LUs of on-the-fly sweeps
        self.mode = "precomputed"
        self._precomputed = set()               # precisions with stored factors
        self.kernel = "hip"
        self.use_pbuf = False
        self.pbuf = None
        self.sweep_variant = 0
        self.fused = False                      # binary32 sweeps build their element operators
        self.one_dir = True
        self._solver = None                     # weak reference to the Solver that owns this context

    # -- construction ------------------------------------------------------
    @classmethod
    def build(cls, case, mesh, integrals, vel, sweep_order, bcdata, precompute=None,
              max_bytes: float | None = None, max_bundle: int = 0, block: int = 0,
              variant: int = 0, lds_slots_cap: int = 0):
        """``precompute`` (default ``performance.precompute_inverse``) and
        ``max_bytes`` choose the mode as the reference does (sweep.py:104-110):
        precomputed factors when wanted and their bytes fit, else "onthefly".
        The default cap is half of the device memory (the reference's 8e9
        is sized for host RAM).  An fp32 run stores factors in both
        precisions; when only the binary32 ones fit, the mode is "mixed":
        binary32 sweeps use stored factors, binary64 sweeps build theirs on
        the fly."""
        self = cls()
        cxv, cyv, domega = vel.flat()
        self.cxv, self.cyv, self.domega = cxv, cyv, domega
        self.ndir, self.n_tris, self.nd = cxv.size, mesh.n_tris, case.ndof_tri
        self.n_faces = mesh.n_faces
        self.cv, self.vg = float(case.flow.cv), float(case.flow.vg)
        self.tau_r, self.tau_n, self.tau_c = float(case.flow.tau_r), float(case.flow.tau_n), case.tau_c
        self.has_walls = bool(bcdata.has_nonthermalising)
        self.has_periodic = bool(bcdata.has_periodic)
        keep = {}

        def f64(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep[id(a)] = a
            return _native.f64(a)

        def i32(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep[id(a)] = a
            return _native.i32(a)

        dims = _native.PbteDims(ndir=self.ndir, n_tris=self.n_tris, nd=self.nd, nf=case.ndof_fc,
                                n_faces=self.n_faces, nbc=len(bcdata.bc_type))
        su = _native.PbteSetup()
        su.cxv, su.cyv, su.domega = f64(cxv), f64(cyv), f64(domega)
        su.cv, su.vg, su.tau_r, su.tau_n, su.tau_c, su.pi = (self.cv, self.vg, self.tau_r,
                                                             self.tau_n, self.tau_c, PI)
        su.mass = f64(integrals.int_tri_tri)
        su.gx, su.gy = f64(integrals.int_tri_tri_x), f64(integrals.int_tri_tri_y)
        su.fcm, su.ttfc = f64(integrals.int_tri_fc), f64(integrals.int_tri_tri_fc)
        su.nx, su.ny = f64(mesh.tri_normal[:, :, 0]), f64(mesh.tri_normal[:, :, 1])
        su.int_tri = f64(integrals.int_tri)
        su.tri_faces, su.face_bc = i32(mesh.tri_faces), i32(mesh.face_bc)
        su.neighbour, su.periodic_tri = i32(mesh.tri_neighbour), i32(bcdata.periodic_tri)
        su.other_lfc = i32(_other_lfc(mesh, bcdata))
        su.bc_type, su.bc_temp = i32(bcdata.bc_type), f64(bcdata.bc_temp)
        su.face_tri, su.face_lfc = i32(mesh.face_tri), i32(mesh.face_lfc)
        su.order = i32(sweep_order.order)
        su.block, su.sweep_variant = int(block), int(variant)
        self.sweep_variant = int(variant)
        su.lds_slots_cap = int(lds_slots_cap)
        su.lds_value_bytes = 4 if case.performance.precision == "fp32" else 8
        # max_bundle 0: with at least two directions per CU, pairs of
        # directions with the same order share a workgroup (and each cell
        # record load) when their trace pool fits the LDS without any bundle
        # falling back to global memory; else one direction per workgroup
        tries = [int(max_bundle)]
        if max_bundle == 0:
            ncu = torch.cuda.get_device_properties(0).multi_processor_count
            tries = [2, 1] if self.ndir >= 2 * ncu else [1]
        err = C.create_string_buffer(512)
        torch.cuda.synchronize()
        h = None
        for k, mb in enumerate(tries):
            su.max_bundle = mb
            su.pairs_or_fail = 1 if k < len(tries) - 1 else 0
            h = _native.lib().pbte_ctx_create(C.byref(dims), C.byref(su), C.c_void_p(stream_handle()),
                                              err, 512)
            last = k == len(tries) - 1
            if not h:
                if last:
                    raise _native.NativeError(f"pbte_ctx_create: {err.value.decode()}")
                continue
            if not last:
                out = (C.c_int64 * 6)()
                _native.check(h, _native.lib().pbte_schedule_info(h, out), "schedule")
                if out[4] > 0:                  # a bundle would sweep through global memory
                    _native.lib().pbte_ctx_destroy(h)
                    h = None
                    continue
            break
        self.handle = h
        if getattr(integrals, "ref_ops", None) is not None:
            _native.check(h, _native.lib().pbte_set_cell_operators(
                h, f64(integrals.ref_ops), f64(integrals.cell_ops)), "set_cell_operators")
        # every bundle one direction: the sweep runs the kernel instance
        # without bundle arithmetic (a batched launch runs one instance)
        self.one_dir = self.schedule_info()["bundles"] == self.ndir
        self.flux_wall = dev_zeros((max(self.n_faces, 1), self.nd))
        self.precision = case.performance.precision
        if precompute is None:
            precompute = case.performance.precompute_inverse
        if max_bytes is None:
            max_bytes = 0.5 * torch.cuda.get_device_properties(0).total_memory
        nbytes = self.operator_bytes(self.n_tris, self.nd)
        self.mode = "onthefly"
        self._precomputed = set()
        # binary32 sweeps fused: each element operator built and factorised
        # in the sweep, so only the binary64 factors are stored
        self.fused = (self.precision == "fp32" and case.performance.fp32_operator == "fused"
                      and self.set_fused(True))
        # bytes of the stored factors: binary64, plus binary32 for fp32 runs
        # that read stored binary32 factors
        need = nbytes * (1.5 if self.precision == "fp32" and not self.fused else 1.0)
        if precompute and need > max_bytes >= 0.5 * nbytes:
            # binary32 factors fit, binary64 ones (beside them) do not: the
            # binary32 sweeps use stored (or fused) factors, the binary64
            # sweeps refine from the binary32 ones (fp64 runs included:
            # binary64 results without binary64 factors)
            self.factorise("fp32")
            self.mode = "mixed"
        elif precompute and need <= max_bytes:
            self.factorise("fp64+fp32" if self.precision == "fp32" and not self.fused else "fp64")
        return self

    _PATHS = {1: "plain", 2: "pipelined", 3: "fused", 4: "fused-split", 5: "onthefly", 6: "refined"}

    def sweep_path(self, precision: str = "fp32") -> str:
        """The kernel a sweep of this precision runs: "plain", "pipelined"
        (stored factors), "fused", "fused-split", "onthefly" or "refined"."""
        code = _native.PBTE_FP64 if precision == "fp64" else _native.PBTE_FP32
        return self._PATHS[_native.check(self.handle, _native.lib().pbte_sweep_path(self.handle, code), "sweep_path")]

    def staging(self, on=None) -> dict:
        """LDS staging of the pipelined sweeps' previous-sweep and explicit
        periodic-buffer reads: switch it on or off (``on`` True/False; None
        leaves it) and return the setting with the schedule's stage sizes."""
        out = (C.c_int64 * 4)()
        st = _native.check(self.handle, _native.lib().pbte_staging(
            self.handle, -1 if on is None else (1 if on else 0), out), "staging")
        return {"on": bool(st), "slots": int(out[0]), "slots_pbuf": int(out[1]),
                "staged_old_reads": int(out[2]), "staged_periodic_reads": int(out[3])}

    def set_fused(self, on: bool) -> bool:
        """Switch the fused binary32 sweep on (True when it runs here: the
        context's elements and schedule allow it; "split": in the split
        kernel where it fits) or off."""
        h = self.handle
        if not on:
            _native.check(h, _native.lib().pbte_set_fused(h, 0), "set_fused")
            self.fused = False
            return False
        self.fused = _native.lib().pbte_set_fused(h, 2 if on == "split" else 1) == 0
        return self.fused

    def __del__(self):
        self.close()

    def close(self):
        h = getattr(self, "handle", None)
        self.handle = None
        if h:
            try:
                _native.lib().pbte_ctx_destroy(h)
            except Exception:  # interpreter shutdown: modules already torn down
                pass

    def schedule_info(self) -> dict:
        out = (C.c_int64 * 6)()
        _native.check(self.handle, _native.lib().pbte_schedule_info(self.handle, out), "schedule")
        return {"levels": int(out[0]), "bundles": int(out[1]), "max_width": int(out[2]),
                "pool_peak": int(out[3]), "global_bundles": int(out[4]),
                "pool_slots": int(out[5])}

    def timing(self, capacity: int) -> None:
        """Arm HIP events around the next ``capacity`` sweep-kernel launches."""
        _native.check(self.handle, _native.lib().pbte_timing(self.handle, int(capacity)), "timing")

    def timing_read(self, n: int) -> np.ndarray:
        """Durations (ms) of the recorded sweep-kernel launches."""
        out = (C.c_float * max(int(n), 1))()
        m = _native.lib().pbte_timing_read(self.handle, out, int(n))
        _native.check(self.handle, m, "timing_read")
        return np.array(out[:m], dtype=np.float64)

    def operator_bytes(self, n_tris: int, ndof: int) -> float:
        return float(self.ndir) * n_tris * ndof * ndof * 8.0

    def factorise(self, precision: str = "fp64") -> None:
        """Factors of every element operator in ``precision``: "fp64",
        "fp32", or "fp64+fp32" (both from one binary64 pass)."""
        code = {"fp64": _native.PBTE_FP64, "fp32": _native.PBTE_FP32,
                "fp64+fp32": _native.PBTE_FP64_FP32}[precision]
        bad = _native.check(self.handle, _native.lib().pbte_factorise(self.handle, code), "factorise")
        if bad:
            raise np.linalg.LinAlgError(
                f"{bad} element operators are singular; check TAU_C and the mesh")
        self._precomputed.update(precision.split("+"))
        self.mode = "precomputed"

    def ensure_pbuf(self, dtype):
        if self.pbuf is None or self.pbuf.dtype != dtype:
            self.pbuf = dev_zeros((self.ndir, max(self.n_faces, 1), self.nd), dtype=dtype)
        return self.pbuf

    # -- operators ---------------------------------------------------------
    def _device(self, vdf):
        """``(device tensor, hand-back)`` for a distribution argument: a
        device tensor as it is; a host array (the owning solver's ``vdf``
        mirror, or any ``[D, N, P]`` array) through ``Solver.device_vdf``."""
        if isinstance(vdf, torch.Tensor):
            return vdf, None
        solver = self._solver() if self._solver is not None else None
        if solver is None:
            raise TypeError("a host distribution needs the Solver that owns this context")
        return solver.device_vdf(vdf)

    def sweep(self, mom, vdf, homogeneous: bool = False, flux_wall=None, src=None) -> None:
        """One transport sweep of ``vdf`` in place: the device tensor, or a
        host array (``solver.vdf``) written back after the sweep.

        ``src`` optionally overrides the moment source with a tuple of device
        ``[P, N]`` binary64 tensors ``(ts, qxs, qys)``."""
        vdf, back = self._device(vdf)
        solver = self._solver() if back is None and self._solver is not None else None
        if solver is not None and vdf is solver.sweep_vdf:
            solver.device_changed()             # the owner's device state moves on
        self._sweep(mom, vdf, homogeneous, flux_wall, src)
        if back is not None:
            back()

    def _sweep(self, mom, vdf, homogeneous, flux_wall, src) -> None:
        prec = _PREC[vdf.dtype]
        ts, qxs, qys = src if src is not None else (mom.ts_d, mom.qxs_d, mom.qys_d)
        fw = self.flux_wall if flux_wall is None else flux_wall
        pb = None
        if self.use_pbuf:
            pb = self.ensure_pbuf(vdf.dtype)
        fused = self.fused and prec == _native.PBTE_FP32
        stored = ("fp64" if prec == _native.PBTE_FP64 else "fp32") in self._precomputed and not fused
        fn = _native.lib().pbte_sweep if stored else _native.lib().pbte_sweep_onthefly
        st = fn(self.handle, prec, ptr(ts), ptr(qxs), ptr(qys), ptr(vdf), ptr(fw), ptr(pb),
                1 if self.use_pbuf else 0, 1 if homogeneous else 0)
        _native.check(self.handle, st, "sweep")
        self.sweep_count += 1
        if self.mode == "onthefly" or fused:
            # every element operator was built and LU-factorised for this sweep
            self.factorisations += self.ndir * self.n_tris

    def moments(self, vdf, mom, out=None) -> None:
        """Moments of ``vdf`` into ``mom`` (or into ``out`` = 6 device tensors)."""
        vdf, _ = self._device(vdf)
        prec = _PREC[vdf.dtype]
        o = out if out is not None else (mom.ts_d, mom.qxs_d, mom.qys_d, mom.temp_d, mom.qx_d, mom.qy_d)
        st = _native.lib().pbte_moments(self.handle, prec, ptr(vdf), *(ptr(t) for t in o))
        _native.check(self.handle, st, "moments")

    def wall_flux(self, vdf, out=None) -> None:
        vdf, _ = self._device(vdf)
        prec = _PREC[vdf.dtype]
        st = _native.lib().pbte_wall_flux(self.handle, prec, ptr(vdf),
                                          ptr(self.flux_wall if out is None else out))
        _native.check(self.handle, st, "wall_flux")

    def transport_residual_parts(self, mom, vdf):
        vdf, _ = self._device(vdf)
        if vdf.dtype != torch.float64:
            vdf = vdf.to(torch.float64)
        out = np.zeros(2 * self.ndir)
        st = _native.lib().pbte_transport_residual(
            self.handle, ptr(mom.ts_d), ptr(mom.qxs_d), ptr(mom.qys_d), ptr(vdf),
            ptr(self.flux_wall), out.ctypes.data_as(C.POINTER(C.c_double)))
        _native.check(self.handle, st, "transport_residual")
        return out[:self.ndir], out[self.ndir:]

    def true_residual(self, mom, vdf) -> float:
        num, den = self.transport_residual_parts(mom, vdf)
        d = float(den.sum())
        if d <= 0.0:
            return float("nan")
        return float(np.sqrt(num.sum() / d))

    def residual_iterate(self, temp_d, t_old_d):
        res = C.c_double()
        mass = C.c_double()
        st = _native.lib().pbte_residual_iterate(self.handle, ptr(temp_d), ptr(t_old_d),
                                                 C.byref(res), C.byref(mass))
        _native.check(self.handle, st, "residual_iterate")
        return res.value, mass.value

    def mass(self, mom) -> float:
        return self.residual_iterate(mom.temp_d, mom.temp_d)[1]

    # -- Krylov vector kernels --------------------------------------------
    # one device call takes at most _ROWS basis rows; larger bases go in
    # chunks (every row's dot product is still one call's)
    _ROWS = 512

    def multidot(self, Q, m: int, w) -> np.ndarray:
        out = np.zeros(max(m, 1))
        for k0 in range(0, m, self._ROWS):
            k = min(self._ROWS, m - k0)
            st = _native.lib().pbte_vec_multidot(
                self.handle, ptr(Q[k0]) if Q.dim() > 1 else ptr(Q), Q.stride(0), k, ptr(w), w.numel(),
                out[k0:].ctypes.data_as(C.POINTER(C.c_double)))
            _native.check(self.handle, st, "multidot")
        return out[:m]

    def update(self, Q, m: int, c: np.ndarray, w) -> None:
        c = np.ascontiguousarray(c, dtype=np.float64)
        for k0 in range(0, m, self._ROWS):
            k = min(self._ROWS, m - k0)
            st = _native.lib().pbte_vec_update(
                self.handle, ptr(Q[k0]) if Q.dim() > 1 else ptr(Q), Q.stride(0), k,
                c[k0:].ctypes.data_as(C.POINTER(C.c_double)), ptr(w), w.numel())
            _native.check(self.handle, st, "update")

    def cgs2(self, Q, m: int, w):
        """Gram-Schmidt of ``w`` against ``Q[:m]`` applied twice, on the device
        (``krylov.py:203-207``); returns ``(c1 + c2, |w|)`` with ``w``
        orthogonalised in place.  Bitwise the multidot/update sequence."""
        if m > self._ROWS:
            c1 = self.multidot(Q, m, w)
            self.update(Q, m, c1, w)
            c2 = self.multidot(Q, m, w)
            self.update(Q, m, c2, w)
            return c1 + c2, self.norm(w)
        out = np.zeros(2 * m + 1)
        st = _native.lib().pbte_vec_cgs2(self.handle, ptr(Q), Q.stride(0), m, ptr(w), w.numel(),
                                         out.ctypes.data_as(C.POINTER(C.c_double)))
        _native.check(self.handle, st, "cgs2")
        return out[:m] + out[m:2 * m], float(np.sqrt(max(out[2 * m], 0.0)))

    def cgs2_begin(self, Q, m: int, w, next_row=None) -> None:
        """Queue :meth:`cgs2` (``m <= 512``) and, with ``next_row``,
        ``next_row = w / |w|`` on the device; :meth:`cgs2_end` collects."""
        st = _native.lib().pbte_vec_cgs2_begin(self.handle, ptr(Q), Q.stride(0), m, ptr(w), w.numel(),
                                               None if next_row is None else ptr(next_row))
        _native.check(self.handle, st, "cgs2_begin")
        self._cg_m = m

    def cgs2_end(self):
        m = self._cg_m
        out = np.zeros(2 * m + 1)
        st = _native.lib().pbte_vec_cgs2_end(self.handle, out.ctypes.data_as(C.POINTER(C.c_double)))
        _native.check(self.handle, st, "cgs2_end")
        return out[:m] + out[m:2 * m], float(np.sqrt(max(out[2 * m], 0.0)))

    def defect(self, v, t, g, out) -> None:
        """``out = (v - t) + g`` (``g`` None: ``v - t``)."""
        st = _native.lib().pbte_vec_defect(self.handle, ptr(v), ptr(t), None if g is None else ptr(g),
                                           ptr(out), out.numel())
        _native.check(self.handle, st, "defect")

    def relax(self, omega: float, o, d) -> float:
        """``max |o - d|`` (NaN when an entry is NaN), then
        ``d += omega (o - d)`` unless ``omega`` is 0; waits for the stream."""
        step = C.c_double()
        st = _native.lib().pbte_vec_relax(self.handle, float(omega), ptr(o), ptr(d), d.numel(), C.byref(step))
        _native.check(self.handle, st, "relax")
        return step.value

    def residual_iterate_async(self, temp_d, t_old_d, out_d) -> None:
        """Residual and mass into the device pair ``out_d`` without waiting."""
        st = _native.lib().pbte_residual_iterate_async(self.handle, ptr(temp_d), ptr(t_old_d), ptr(out_d))
        _native.check(self.handle, st, "residual_iterate_async")

    def axpby(self, a: float, x, b: float, y) -> None:
        st = _native.lib().pbte_vec_axpby(self.handle, float(a), ptr(x), float(b), ptr(y), y.numel())
        _native.check(self.handle, st, "axpby")

    def norm(self, w) -> float:
        return float(np.sqrt(max(self.multidot(w.view(1, -1), 1, w)[0], 0.0)))
