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
from .krylov import SweepRequest

__all__ = ["SweepBatch", "build_solvers", "launch_key", "run_refined_batch", "run_source_batch"]

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

    def join(self) -> None:
        """The contexts' stream waits for this batch's last launch."""
        _join_all([self])

    def sweep(self, moms, vdfs, homogeneous: bool = False, join: bool = True) -> None:
        """One sweep of every context: ``vdfs[k]`` in place from ``moms[k]``
        (all distributions of one precision), ``sweep_count`` += 1 each.
        Contexts that do not run the pipelined kernel in this precision
        (factors built on the fly, the plain level kernel) are swept one by
        one, each as ``SweepContext.sweep`` does."""
        self.sweep_requests([SweepRequest(c, v, homogeneous, None, None)
                             for c, v in zip(self.contexts, vdfs)], moms, join=join)

    def sweep_requests(self, reqs, moms=None, join: bool = True) -> None:
        """The sweeps ``reqs[k]`` (``krylov.SweepRequest`` of context k: its
        distribution, moment source -- ``src`` or ``moms[k]`` --, flux_wall,
        and the explicit periodic buffer when the context's ``use_pbuf`` is
        set), all of one precision and one ``homogeneous`` flag, in one
        launch when every context runs the pipelined kernel; else one by
        one.  Each context's result is its ``SweepContext.sweep``, bit for
        bit.  ``join=False`` leaves the launch running beside the contexts'
        stream (launches of several batches then overlap): call ``join``
        before the stream uses the results."""
        n = len(self.contexts)
        prec = _PREC[reqs[0].vdf.dtype]
        hom = reqs[0].homogeneous
        if any(_PREC[r.vdf.dtype] != prec or r.homogeneous != hom for r in reqs):
            raise ValueError("a batched sweep runs one precision and one source flag")
        if not self.pipelined(prec):
            for k, r in enumerate(reqs):
                r.ctx.sweep(None if moms is None else moms[k], r.vdf, homogeneous=hom,
                            flux_wall=r.flux_wall, src=r.src)
            return

        def arr(vals):
            return (C.c_void_p * n)(*vals)

        srcs = [r.src if r.src is not None else (moms[k].ts_d, moms[k].qxs_d, moms[k].qys_d)
                for k, r in enumerate(reqs)]
        ts = arr([ptr(s[0]) for s in srcs])
        qx = arr([ptr(s[1]) for s in srcs])
        qy = arr([ptr(s[2]) for s in srcs])
        vd = arr([ptr(r.vdf) for r in reqs])
        fw = arr([ptr(r.ctx.flux_wall if r.flux_wall is None else r.flux_wall) for r in reqs])
        pb = arr([ptr(r.ctx.ensure_pbuf(r.vdf.dtype)) if r.ctx.use_pbuf else None for r in reqs])
        for r in reqs:
            solver = r.ctx._solver() if r.ctx._solver is not None else None
            if solver is not None and r.vdf is solver.sweep_vdf:
                solver.device_changed()
        fn = _native.lib().pbte_batch_sweep if join else _native.lib().pbte_batch_sweep_async
        st = fn(self.handle, prec, ts, qx, qy, vd, fw, pb, 1 if hom else 0)
        if st < 0:
            msg = _native.lib().pbte_batch_last_error(self.handle)
            raise _native.NativeError(f"batch_sweep: {msg.decode() if msg else 'unknown error'}")
        for r in reqs:
            r.ctx.sweep_count += 1
            if r.ctx.fused and prec == _native.PBTE_FP32:
                r.ctx.factorisations += r.ctx.ndir * r.ctx.n_tris


def build_solvers(cases, workers: int = 4, **kw):
    """``Solver(case, **kw)`` of every case, ``workers`` of them at a time
    on host threads (the native set-up, the device work it queues and
    numpy's array passes release the interpreter lock), in case order.  The
    solvers are those built one by one: each set-up is independent of the
    others and every context keeps the creating thread's current stream,
    which for threads that never set one is the same default stream."""
    from concurrent.futures import ThreadPoolExecutor

    from .driver import Solver
    cases = list(cases)
    if workers <= 1 or len(cases) <= 1:
        return [Solver(c, **kw) for c in cases]
    with ThreadPoolExecutor(max_workers=min(workers, len(cases))) as pool:
        return list(pool.map(lambda c: Solver(c, **kw), cases))


def _join_all(batches) -> None:
    for b in batches:
        if _native.lib().pbte_batch_join(b.handle) < 0:
            msg = _native.lib().pbte_batch_last_error(b.handle)
            raise _native.NativeError(f"batch_join: {msg.decode() if msg else 'unknown error'}")


def launch_key(ctx) -> tuple:
    """Contexts with equal keys share a launch: one kernel instance
    (element degree, pipeline depth, and whether every bundle holds one
    direction, whose instance drops the bundle arithmetic, and whether its
    binary32 sweeps are fused)."""
    return (ctx.nd, ctx.sweep_variant if ctx.sweep_variant in (2, 4) else 3, ctx.one_dir, ctx.fused)


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
            # the groups' launches overlap on the device; the stream joins
            # them all before the per-problem work
            launched = []
            for group in groups.values():
                b = batch_of(group)
                b.sweep([tracks[k].s.mom for k in group], [tracks[k].s.sweep_vdf for k in group], join=False)
                launched.append(b)
            _join_all(launched)
            for group in groups.values():
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


def run_refined_batch(solvers, callback=None):
    """The fp32 source iteration as a binary64 defect correction
    (``refine.refined_steps``, CIS or GSIS) of every solver in lockstep;
    returns their ``RunRecord`` objects in order.

    Every problem advances to its next request: the sweeps asked for are
    grouped by precision, source flag and ``launch_key`` and each group is
    swept in one launch; the residual reads of all problems share one
    wait.  Each problem's iterates are those of its refined iteration alone,
    bit for bit.  ``callback(k, step, res, solver)`` after each recorded
    step of problem k."""
    from .krylov import SweepRequest as _Sweep
    from .refine import ReadRequest, applicable, refined_steps
    for s in solvers:
        if s.precision != "fp32" or s.sweep_precision != "fp32" or s.case.scheme.method == "krylov":
            raise ValueError("run_refined_batch runs fp32 source iteration (refined)")
        ok, why = applicable(s)
        if not ok:
            raise ValueError(f"run_refined_batch: {why}")
    gens = {}
    for k, s in enumerate(solvers):
        cb = None if callback is None else (lambda step, res, solver, k=k: callback(k, step, res, solver))
        gens[k] = refined_steps(s, cb)
    msgs = {k: None for k in gens}
    recs = {}
    batches = {}

    def batch_of(group):
        key = tuple(group)
        if key not in batches:
            batches[key] = SweepBatch([solvers[k].ctx for k in group])
        return batches[key]

    try:
        while gens:
            reqs = {}
            for k, g in list(gens.items()):
                try:
                    reqs[k] = g.send(msgs[k])
                except StopIteration as stop:
                    recs[k] = stop.value
                    del gens[k]
            groups = {}
            reads = []
            for k, r in reqs.items():
                if isinstance(r, _Sweep):
                    key = (r.vdf.dtype, r.homogeneous) + launch_key(r.ctx)
                    groups.setdefault(key, []).append(k)
                    msgs[k] = None
                elif isinstance(r, ReadRequest):
                    reads.append(k)
                else:
                    raise TypeError(f"unknown request {type(r).__name__}")
            launched = []
            for group in groups.values():
                b = batch_of(group)
                b.sweep_requests([reqs[k] for k in group], join=False)
                launched.append(b)
            _join_all(launched)
            if reads:
                vals = torch.stack([reqs[k].slot for k in reads]).cpu().numpy()
                for k, (res, mass) in zip(reads, vals):
                    msgs[k] = (float(res), float(mass))
    finally:
        for g in gens.values():
            g.close()
        for b in batches.values():
            b.close()
    return [recs[k] for k in range(len(solvers))]
