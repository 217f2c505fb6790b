"""Binary32 source iteration inside a binary64 defect correction.

The outer loop of the reference (``driver.py:184-211``; GSIS: the same loop
with ``acceleration/__init__.py:95-115`` after each sweep) is the fixed-point
iteration ``x <- F(x)`` of one step ``F(u) = A u + g`` on the state ``u``
that a step reads: the moment source ``[ts, qxs, qys]``, the adiabatic
walls' emission and the periodic partners' previous values (the state of
``krylov.FixedPointOperator``).  With GSIS, ``F`` is the sweep followed by
the macroscopic correction, and the state holds the corrected moments and
the emission and partner values of the corrected distribution.  ``A`` is
linear: the wall temperatures enter only through ``g`` (the correction has
no source of its own when the defect repair is off).

With ``performance.precision: fp32`` the iteration runs as

1. ``fx = F64(x)``, ``r = fx - x`` -- one binary64 step;
2. if the reference's residual of that step, ``|T(fx) - T(x)| / |T(fx)|``
   over the element temperatures, is below ``iteration.tol``: ``fx`` is
   the solution (the reference stops on the same test of the same map);
3. else solve ``(I - A) e = r`` by the homogeneous iteration
   ``e <- A32 e + r`` (binary32 sweeps without the wall sources; the
   correction's moments, trace solve and state in binary64), whose step
   ``e_{j+1} - e_j`` is the defect ``r`` would leave at ``x + e_j``; it is
   measured as the reference measures a step, on ``T(x) + T(e_j)``.  The
   inner loop stops when that step is ``performance.refine_inner`` times
   the outer one (or 0.8 of the tolerance, or it stops falling: the
   binary32 rounding of ``e``); then ``x += e`` and back to 1.

Rounding of the binary32 products is relative to ``e``, which shrinks
with every refinement, so the fixed point is the binary64 one.  Every
sweep is one entry of the residual history; the inner entries are the
inner steps, the outer entries the binary64 steps.
"""
from __future__ import annotations

import time

import numpy as np

from .device import dev_zeros, torch
from .krylov import FixedPointOperator

__all__ = ["ReadRequest", "applicable", "drive", "refined_steps", "run_refined"]

# inner steps without a 10 % improvement that end an inner solve (a
# contraction slower than 0.9 per this many steps is taken as the floor)
_STALL_WINDOW = 25


def applicable(solver) -> tuple[bool, str]:
    """Whether the defect correction applies (else the fp32 run switches
    from binary32 to binary64 sweeps)."""
    if int(solver.order.cycles_broken.sum()):
        return False, "the sweep ordering has broken cycles: one sweep is not a function of the state"
    acc = solver.acc
    if acc is not None and acc.omega:
        return False, "the defect repair (scheme.defect_omega) makes the step depend on the iteration count"
    if acc is not None and not acc.lazy:
        return False, "the correction is applied eagerly"
    return True, ""


class ReadRequest:
    """A device pair ``(residual, mass)`` the iteration needs on the host:
    the driver sends back the two values (one wait for every problem of a
    batch)."""

    __slots__ = ("slot",)

    def __init__(self, slot):
        self.slot = slot


def drive(steps):
    """Run a step generator alone: each sweep as it is asked for, each
    read at once; returns the generator's result."""
    msg = None
    try:
        while True:
            req = steps.send(msg)
            if isinstance(req, ReadRequest):
                v = req.slot.cpu().numpy()
                msg = (float(v[0]), float(v[1]))
            else:
                req.run()
                msg = None
    except StopIteration as stop:
        return stop.value


def run_refined(solver, callback=None, progress_every: int = 0):
    return drive(refined_steps(solver, callback, progress_every))


def refined_steps(solver, callback=None, progress_every: int = 0):
    """The refined iteration as a generator: it yields the sweeps it needs
    (``krylov.SweepRequest``: sweep, then resume) and the residual reads
    (``ReadRequest``: resume with ``(residual, mass)``), and returns the
    ``RunRecord``.  ``drive`` runs it alone; ``batch.run_refined_batch``
    runs several in lockstep with their sweeps batched."""
    c = solver.case
    tol, tmax = c.iteration.tol, c.iteration.tmax
    eta = c.performance.refine_inner
    want_true = c.iteration.true_residual
    ctx = solver.ctx
    acc = solver.acc
    op = FixedPointOperator(solver, acc=acc)
    residuals, masses, true_res, fp64_steps = [], [], [], []
    refinements = 0
    inner_steps = []
    t0 = time.perf_counter()
    try:
        n = op.n
        x = op.state() if solver.restart_info.get("enabled") else dev_zeros(n)
        tx = op._temp.clone() if solver.restart_info.get("enabled") else dev_zeros(solver.n_tris)
        fx, r, e, tmp = dev_zeros(n), dev_zeros(n), dev_zeros(n), dev_zeros(n)
        tfx, qxfx, qyfx = (dev_zeros(solver.n_tris) for _ in range(3))
        cur, prev = dev_zeros(solver.n_tris), dev_zeros(solver.n_tris)
        slot = dev_zeros(2)
        converged = False

        def record(res, mass, fp64):
            residuals.append(res)
            masses.append(mass)
            step = len(residuals)
            if fp64:
                fp64_steps.append(step)
            if want_true:
                true_res.append(np.nan)
            if progress_every and step % progress_every == 0:
                print(f"  step {step:8d}  residual {res:.6e}  mass {mass:.10f}", flush=True)
            if callback is not None:
                callback(step, res, solver)

        while True:
            # binary64 step from x: fx = F(x), and the reference's residual
            yield op.begin(x, "fp64")
            op.end(fx, "fp64")
            tfx.copy_(op._temp)
            qxfx.copy_(op._qx)
            qyfx.copy_(op._qy)
            ctx.residual_iterate_async(tfx, tx, slot)
            res, mass = yield ReadRequest(slot)
            record(res, mass, True)
            if want_true:
                op.publish(fx)
                if acc is not None:
                    acc.finish_state(solver.vdf_d)
                true_res[-1] = ctx.true_residual(solver.mom, solver.vdf_d)
            if res < tol:
                converged = True
                break
            if len(residuals) >= tmax:
                break
            # inner: e_{j+1} = A32 e_j + r from e_0 = 0 (so e_1 = r), with
            # cur = T(x) + T(e_j) = T(fx) + T(A e_{j-1})
            ctx.defect(fx, x, None, r)
            e.copy_(r)
            cur.copy_(tfx)
            target = max(eta * res, 0.8 * tol)
            anchor, since, k = np.inf, 0, 0
            while len(residuals) < tmax:
                yield op.begin(e, "fp32", homogeneous=True)
                op.end(tmp, "fp32")
                ctx.axpby(1.0, r, 1.0, tmp)
                prev.copy_(cur)
                cur.copy_(tfx)
                ctx.axpby(1.0, op._temp, 1.0, cur)
                e, tmp = tmp, e
                k += 1
                ctx.residual_iterate_async(cur, prev, slot)
                rin, mass = yield ReadRequest(slot)
                record(rin, mass, False)
                if rin < target:
                    break
                # at the binary32 floor: no step 10 % below the last
                # such step within the window
                if rin < 0.9 * anchor:
                    anchor, since = rin, 0
                else:
                    since += 1
                    if since >= _STALL_WINDOW:
                        break
            inner_steps.append(k)
            if len(residuals) >= tmax:
                break                                       # the last binary64 state is the answer
            ctx.axpby(1.0, e, 1.0, x)                       # x += e
            tx.copy_(cur)
            refinements += 1
        # fx = F64(x) of the last outer evaluation is the solver's state
        op.publish(fx, cells=(tfx, qxfx, qyfx))
        solver.t_old_d.copy_(tx)
        if acc is not None:
            acc.finish_state(solver.vdf_d)
    finally:
        ctx.use_pbuf = False
        solver.sweep_precision = "fp64"
        if acc is not None:
            acc.drop_state()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    extra = {"precision_scheme": "defect-correction", "refinements": refinements,
             "refine_inner": eta, "inner_steps": inner_steps, "steps_fp32": len(residuals) - len(fp64_steps),
             "fp64_steps": fp64_steps, "sweeps_fp64": op.sweeps64, "sweeps_fp32": op.sweeps - op.sweeps64}
    return solver._record(len(residuals), converged, residuals, masses, true_res if want_true else None, wall,
                          scheme="GSIS" if acc is not None else "CIS", extra=extra)
