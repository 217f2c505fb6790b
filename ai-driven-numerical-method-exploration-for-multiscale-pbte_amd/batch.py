"""Batched solves: several independent problems advanced in lockstep, the
transport sweeps of all of them in one launch of the pipelined kernel
(``pbte_batch_sweep``).  No counterpart in the reference, which solves one
case per process; each problem's iterates are the ones its source
iteration alone (``Solver._run_source``: for fp32 the binary32 sweeps with
the switch to binary64) gives it, bit for bit (the same kernels in the same
order, only the sweep launches are merged).

For many small problems: one small problem's sweep fills a few dozen of
the GPU's 256 CUs (a workgroup per direction), a batch of them fills all.
"""
from __future__ import annotations

import ctypes as C
import time

from . import _native
from .device import dev_zeros, ptr, torch

__all__ = ["SweepBatch", "launch_key", "run_source_batch"]

_PREC = {torch.float64: _native.PBTE_FP64, torch.float32: _native.PBTE_FP32}


class SweepBatch:
    """The sweeps of several ``SweepContext`` objects (one element degree,
    one stream) in one kernel launch."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        n = len(self.contexts)
        arr = (C.c_void_p * n)(*[c.handle for c in self.contexts])
        err = C.create_string_buffer(512)
        h = _native.lib().pbte_batch_create(arr, n, err, 512)
        if not h:
            raise _native.NativeError(f"pbte_batch_create: {err.value.decode()}")
        self.handle = h

    def close(self) -> None:
        h, self.handle = self.handle, None
        if h:
            try:
                _native.lib().pbte_batch_destroy(h)
            except Exception:  # interpreter shutdown
                pass

    def __del__(self):
        self.close()

    def pipelined(self, prec) -> bool:
        """Whether every context runs the pipelined kernel in ``prec``
        (``PBTE_FP64`` / ``PBTE_FP32``), so one launch sweeps them all."""
        st = _native.lib().pbte_batch_pipelined(self.handle, prec)
        if st < 0:
            raise _native.NativeError("batch_pipelined: null batch")
        return st == 1

    def sweep(self, moms, vdfs, homogeneous: bool = False) -> None:
        """One sweep of every context: ``vdfs[k]`` in place from ``moms[k]``
        (all distributions of one precision), ``sweep_count`` += 1 each.
        Contexts that do not run the pipelined kernel in this precision
        (factors built on the fly, the plain level kernel) are swept one by
        one, each as ``SweepContext.sweep`` does."""
        n = len(self.contexts)
        prec = _PREC[vdfs[0].dtype]
        if any(_PREC[v.dtype] != prec for v in vdfs):
            raise ValueError("a batched sweep runs one precision")
        if not self.pipelined(prec):
            for c, m, v in zip(self.contexts, moms, vdfs):
                c.sweep(m, v, homogeneous=homogeneous)
            return

        def arr(vals):
            return (C.c_void_p * n)(*vals)

        ts = arr([ptr(m.ts_d) for m in moms])
        qx = arr([ptr(m.qxs_d) for m in moms])
        qy = arr([ptr(m.qys_d) for m in moms])
        vd = arr([ptr(v) for v in vdfs])
        fw = arr([ptr(c.flux_wall) for c in self.contexts])
        st = _native.lib().pbte_batch_sweep(self.handle, prec, ts, qx, qy, vd, fw, 1 if homogeneous else 0)
        if st < 0:
            msg = _native.lib().pbte_batch_last_error(self.handle)
            raise _native.NativeError(f"batch_sweep: {msg.decode() if msg else 'unknown error'}")
        for c in self.contexts:
            c.sweep_count += 1


def launch_key(ctx) -> tuple:
    """Contexts with equal keys share a launch: one kernel instance
    (element degree, pipeline depth, and whether every bundle holds one
    direction, whose instance drops the bundle arithmetic)."""
    return (ctx.nd, 2 if ctx.sweep_variant == 2 else 3, ctx.one_dir)


class _Track:
    """One problem's histories and switch state (``Solver._run_source``)."""

    def __init__(self, solver):
        c = solver.case
        self.s = solver
        self.tol, self.tmax = c.iteration.tol, c.iteration.tmax
        self.switch = max(self.tol, c.performance.refine_switch)
        self.residuals, self.masses = [], []
        self.step = 0
        self.step32 = 0
        self.converged = False
        self.done = False
        self.wall = 0.0
        self.slot = dev_zeros(2)


def run_source_batch(solvers, callback=None):
    """Source iteration (CIS, or GSIS with ``accflag: 1``) of every solver
    in lockstep; returns their ``RunRecord`` objects in order.

    Per step: the wall emission of each unfinished problem, one batched
    sweep per group of one precision and one ``launch_key``, each
    problem's moments (or GSIS correction), the residual and mass of every
    problem summed on the device, one wait, then the stop tests of
    ``Solver._run_source`` (the fp32 switch included).
    ``callback(k, step, res, solver)`` after each problem's step.  Options
    that need the transport residual per step (``iteration.true_residual``)
    are not supported here."""
    tracks = [_Track(s) for s in solvers]
    for t in tracks:
        if t.s.case.iteration.true_residual:
            raise ValueError("run_source_batch: iteration.true_residual is not supported")
        if t.s.case.scheme.method == "krylov":
            raise ValueError("run_source_batch runs source iteration (scheme.method: source)")
    batches = {}             # one per set of problems swept together

    def batch_of(group):
        key = tuple(group)
        if key not in batches:
            batches[key] = SweepBatch([tracks[k].s.ctx for k in group])
        return batches[key]

    t0 = time.perf_counter()
    try:
        while True:
            active = [k for k, t in enumerate(tracks) if not t.done]
            if not active:
                break
            for k in active:
                s = tracks[k].s
                if s.acc is not None:
                    s.acc.settle(boundary_only=True)
                if s.bcdata.has_nonthermalising:
                    s.ctx.wall_flux(s.sweep_vdf)
            groups = {}
            for k in active:
                s = tracks[k].s
                groups.setdefault((s.sweep_precision,) + launch_key(s.ctx), []).append(k)
            for group in groups.values():
                batch_of(group).sweep([tracks[k].s.mom for k in group],
                                      [tracks[k].s.sweep_vdf for k in group])
                for k in group:
                    if tracks[k].s.acc is not None:
                        tracks[k].s.acc.swept()
            for k in active:
                t = tracks[k]
                s = t.s
                if s.acc is None:
                    s.ctx.moments(s.sweep_vdf, s.mom)
                else:
                    s.acc.apply()
                s.ctx.residual_iterate_async(s.mom.temp_d, s.t_old_d, t.slot)
                s.t_old_d.copy_(s.mom.temp_d)
            vals = torch.stack([tracks[k].slot for k in active]).cpu().numpy()
            now = time.perf_counter() - t0
            for (k, (res, mass)) in zip(active, vals):
                t = tracks[k]
                s = t.s
                t.step += 1
                res, mass = float(res), float(mass)
                t.residuals.append(res)
                t.masses.append(mass)
                if callback is not None:
                    callback(k, t.step, res, s)
                if s.sweep_precision == "fp32":
                    t.step32 += 1
                    stalled = t.step32 > 8 and res >= min(t.residuals[-9:-1])
                    if res < t.switch or stalled:
                        s.promote()
                        t.step32 = t.step
                    if t.step >= t.tmax:
                        t.done, t.wall = True, now
                    continue
                if res < t.tol:
                    if (s.acc is not None and s.acc.omega and not s.acc.defect_every
                            and s.acc.refresh_defect() > t.tol):
                        continue
                    t.converged, t.done, t.wall = True, True, now
                elif t.step >= t.tmax:
                    t.done, t.wall = True, now
        for t in tracks:
            t.s.settle()
        torch.cuda.synchronize()
    finally:
        for b in batches.values():
            b.close()
    recs = []
    for t in tracks:
        s = t.s
        extra = {"batched": len(tracks)}
        if s.precision == "fp32":
            extra.update({"steps_fp32": t.step32 if s.sweep_precision == "fp64" else t.step,
                          "refine_switch": t.switch})
        recs.append(s._record(t.step, t.converged, t.residuals, t.masses, None, t.wall,
                              scheme="GSIS" if s.case.scheme.accflag == 1 else "CIS", extra=extra))
    return recs
