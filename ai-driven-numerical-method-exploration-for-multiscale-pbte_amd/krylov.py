"""Krylov solution of the outer iteration on the device
(``solver-python/pybte/krylov.py:55-363``).

One source-iteration step is the affine map ``u -> F(u) = T u + g`` on the
state ``u = [ts, qxs, qys (, flux_wall) (, periodic partner DOFs)]``; GMRES
solves ``(I - T) u = g`` with one sweep per product.  Everything of size
``n`` stays in HBM: the basis ``Q`` (binary64 tensor), the products, and the
orthogonalisation -- classical Gram-Schmidt applied twice through the
``pbte_vec_multidot`` / ``pbte_vec_update`` kernels, whose reductions have a
fixed order.  Only the ``j+2`` Hessenberg entries per product and the
Givens recurrence run on the host, as scalars.

Precision (``performance.precision``):

``fp64``  the reference's algorithm: ``A v = v - (F(v) - g)`` with binary64
          sweeps, staged targets, a closing sweep and the true transport
          residual as the convergence decision (``krylov.py:285-363``).
``fp32``  FP64 defect correction around FP32 sweeps.  The outer loop keeps
          ``x`` in binary64 and evaluates ``r = F64(x) - x`` (one binary64
          sweep, ``= g - (I - T) x``); the correction ``(I - T) e = r`` is
          solved by GMRES whose products are homogeneous binary32 sweeps
          (boundary sources removed, moments summed in binary64); ``x += e``.
          The last ``F64(x)`` is the closing sweep.  Convergence is decided
          by the binary64 transport residual, as in the reference, or, with
          ``performance.refine_tol > 0``, by the estimated relative error of
          the temperature and heat-flux moments (the accuracy a
          mixed-precision solve is asked for).
"""
from __future__ import annotations

import time

import numpy as np

from . import _native
from .device import dev_zeros, ptr, torch

__all__ = ["FixedPointOperator", "MomentPreconditioner", "SweepRequest", "applicable", "gmres_full",
           "run_krylov", "krylov_basis_cap"]

_ENOUGH_FACTOR = 1.0e-2
_FLOOR = 1.0e-14


def applicable(solver) -> tuple[bool, str]:
    if int(solver.order.cycles_broken.sum()):
        return False, ("the sweep ordering has broken cycles: part of the inflow is lagged, "
                       "so one sweep is not a function of the state alone")
    return True, ""


class SweepRequest:
    """One sweep a fixed-point product asks for: ``vdf`` in place from the
    moment source ``src`` (and ``flux_wall``; the context's explicit
    periodic buffer when its ``use_pbuf`` is set), with or without the wall
    sources.  ``run()`` sweeps it alone; a batch driver may sweep several
    problems' requests in one launch instead (the same bits)."""

    __slots__ = ("ctx", "vdf", "homogeneous", "flux_wall", "src")

    def __init__(self, ctx, vdf, homogeneous, flux_wall, src):
        self.ctx, self.vdf, self.homogeneous, self.flux_wall, self.src = ctx, vdf, homogeneous, flux_wall, src

    def run(self) -> None:
        self.ctx.sweep(None, self.vdf, homogeneous=self.homogeneous, flux_wall=self.flux_wall, src=self.src)


class FixedPointOperator:
    """``u -> F(u)`` on device vectors, one sweep apiece; with ``acc`` (an
    ``acceleration.Acceleration``) each sweep is followed by the GSIS
    correction, and the state is that of the corrected distribution."""

    def __init__(self, solver, acc=None):
        self.s = solver
        self.acc = acc
        self.ctx = solver.ctx
        self.nd, self.nt = solver.ndof, solver.n_tris
        self.nmom = 3 * self.nd * self.nt
        self.wall = bool(solver.bcdata.has_nonthermalising)
        self.nwall = solver.mesh.n_faces * self.nd if self.wall else 0
        self.per = bool(solver.bcdata.has_periodic)
        self.npf = int(np.count_nonzero(solver.bcdata.periodic_tri >= 0)) if self.per else 0
        self.ndir = solver.ndir
        self.nper = self.ndir * self.npf * self.nd if self.per else 0
        self.n = self.nmom + self.nwall + self.nper
        self.sweeps = 0
        self.sweeps64 = 0
        self.discarded = 0
        self._last = "fp64"
        solver.ctx.use_pbuf = self.per
        self._temp = dev_zeros(self.nt)
        self._qx = dev_zeros(self.nt)
        self._qy = dev_zeros(self.nt)

    def _blocks(self, u):
        k = self.nd * self.nt
        ts = u[:k]
        qxs = u[k:2 * k]
        qys = u[2 * k:3 * k]
        fw = u[self.nmom:self.nmom + self.nwall].view(-1, self.nd) if self.wall else None
        pb = u[self.nmom + self.nwall:] if self.per else None
        return ts, qxs, qys, fw, pb

    def apply(self, u, out, precision: str = "fp64", homogeneous: bool = False) -> None:
        """``out = F(u)`` (``homogeneous``: without the wall sources, ``T u``)."""
        self.begin(u, precision, homogeneous).run()
        self.end(out, precision)

    def begin(self, u, precision: str = "fp64", homogeneous: bool = False) -> "SweepRequest":
        """First half of :meth:`apply`: the periodic partners of ``u`` into
        the explicit buffer; returns the sweep to run (alone, or batched
        with other problems' sweeps), after which :meth:`end` finishes."""
        s, ctx = self.s, self.ctx
        vdf = s.vdf_d if precision == "fp64" else s.vdf32_d
        code = _native.PBTE_FP64 if precision == "fp64" else _native.PBTE_FP32
        ts, qxs, qys, fw, pb = self._blocks(u)
        if self.per:
            buf = ctx.ensure_pbuf(vdf.dtype)
            _native.check(ctx.handle, _native.lib().pbte_periodic_scatter(
                ctx.handle, code, ptr(pb), ptr(buf)), "periodic_scatter")
        return SweepRequest(ctx, vdf, homogeneous, fw, (ts, qxs, qys))

    def end(self, out, precision: str = "fp64") -> None:
        """Second half of :meth:`apply`, after the sweep: the moments (or the
        GSIS correction), the wall emission and the periodic partners of
        the swept distribution into ``out``."""
        s, ctx = self.s, self.ctx
        vdf = s.vdf_d if precision == "fp64" else s.vdf32_d
        code = _native.PBTE_FP64 if precision == "fp64" else _native.PBTE_FP32
        self.sweeps += 1
        self._last = precision
        if precision == "fp64":
            self.sweeps64 += 1
        ots, oqxs, oqys, ofw, opb = self._blocks(out)
        if self.acc is not None:
            self.acc.correct_state(vdf, (ots, oqxs, oqys, self._temp, self._qx, self._qy))
        else:
            ctx.moments(vdf, None, out=(ots, oqxs, oqys, self._temp, self._qx, self._qy))
        if self.wall:
            ctx.wall_flux(vdf, out=ofw)
        if self.per:
            _native.check(ctx.handle, _native.lib().pbte_periodic_gather(
                ctx.handle, code, ptr(vdf), ptr(opb)), "periodic_gather")

    def discard(self) -> None:
        """Uncount the last product: it was computed ahead and not used."""
        self.sweeps -= 1
        self.discarded += 1
        self.ctx.sweep_count -= 1
        if self._last == "fp64":
            self.sweeps64 -= 1

    def publish(self, u, cells=None) -> None:
        """Copy a state's moment block (and emission) into the solver; the
        element moments are the last product's, or ``cells`` = (temp, qx,
        qy) when ``u`` is not that product."""
        s = self.s
        ts, qxs, qys, fw, _ = self._blocks(u)
        s.mom.ts_d.copy_(ts.view(self.nd, self.nt))
        s.mom.qxs_d.copy_(qxs.view(self.nd, self.nt))
        s.mom.qys_d.copy_(qys.view(self.nd, self.nt))
        temp, qx, qy = cells if cells is not None else (self._temp, self._qx, self._qy)
        s.mom.temp_d.copy_(temp)
        s.mom.qx_d.copy_(qx)
        s.mom.qy_d.copy_(qy)
        if self.wall:
            s.ctx.flux_wall.copy_(fw)

    def state(self):
        """The state of the solver's current distribution, without sweeping."""
        s = self.s
        u = dev_zeros(self.n)
        ts, qxs, qys, fw, pb = self._blocks(u)
        s.ctx.moments(s.vdf_d, None, out=(ts, qxs, qys, self._temp, self._qx, self._qy))
        if self.wall:
            s.ctx.wall_flux(s.vdf_d, out=fw)
        if self.per:
            _native.check(s.ctx.handle, _native.lib().pbte_periodic_gather(
                s.ctx.handle, _native.PBTE_FP64, ptr(s.vdf_d), ptr(pb)), "periodic_gather")
        return u


def _back_substitute(r, y):
    out = np.array(y, dtype=float)
    for i in range(out.size - 1, -1, -1):
        s = out[i] - r[i, i + 1:] @ out[i + 1:]
        out[i] = s / r[i, i] if r[i, i] != 0.0 else 0.0
    return out


def gmres_full(ctx, matvec, b, x=None, tol_abs=0.0, maxiter=None, max_basis=None, callback=None,
               precond=None, overlap=False):
    """GMRES with a device basis; ``matvec(v, out)`` writes ``A v``.

    Follows ``krylov.py:150-246``: restarts at ``max_basis``, CGS2, Givens
    rotations, ``callback(res)`` once per product (true stops), and the
    optional left preconditioner ``precond(v)`` (in place) applied to every
    residual and product.  Returns ``(x, converged, n_matvec)``; ``x`` is
    updated in place.

    ``overlap`` (no preconditioner; ``matvec`` takes ``record=False`` and
    has ``matvec.commit()`` and ``matvec.discard()``): the next basis vector
    is normalised on the device and its product queued before the host
    waits for the Gram-Schmidt coefficients, so the Givens work and the
    launches overlap the sweep.  A queued product is committed (counted,
    its history entry made) when the loop reaches it, and discarded when
    the loop stops first; the arithmetic is that of the plain loop."""
    n = b.numel()
    x = dev_zeros(n) if x is None else x
    maxiter = int(n if maxiter is None else maxiter)
    m_cap = maxiter if max_basis is None else max(1, min(int(max_basis), maxiter))
    r = dev_zeros(n)
    w = dev_zeros(n)
    Q = None
    nmv = 0
    converged = halted = False
    xnz = ctx.norm(x) > 0.0
    while nmv < maxiter and not converged and not halted:
        if xnz:
            matvec(x, r)
            ctx.axpby(1.0, b, -1.0, r)          # r = b - A x
            if precond is not None:
                precond(r)
            nmv += 1
            if callback is not None and callback(ctx.norm(r)):
                halted = True
                break
        else:
            r.copy_(b)
            if precond is not None:
                precond(r)
        beta = ctx.norm(r)
        if beta <= tol_abs:
            converged = True
            break
        m = min(m_cap, maxiter - nmv)
        if m <= 0:
            break
        if Q is None or Q.shape[0] < m + 1:
            Q = dev_zeros((m + 1, n))
        ctx.axpby(1.0 / beta, r, 0.0, Q[0])
        cols, cs, sn, gam = [], [], [], [beta]
        ahead = overlap and precond is None
        pending = False                         # w holds the queued product of Q[j]
        for j in range(m):
            if pending:
                matvec.commit()
                pending = False
            else:
                matvec(Q[j], w)
            if precond is not None:
                precond(w)
            nmv += 1
            if ahead and j + 1 < m and j + 1 <= ctx._ROWS:
                ctx.cgs2_begin(Q, j + 1, w, next_row=Q[j + 1])
                matvec(Q[j + 1], w, record=False)
                pending = True
                hcol, hn = ctx.cgs2_end()
            else:
                hcol, hn = ctx.cgs2(Q, j + 1, w)
            col = np.empty(j + 2)
            col[:j + 1] = hcol
            col[j + 1] = hn
            for i in range(j):
                t = cs[i] * col[i] + sn[i] * col[i + 1]
                col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1]
                col[i] = t
            d = float(np.hypot(col[j], col[j + 1]))
            c_, s_ = (1.0, 0.0) if d == 0.0 else (col[j] / d, col[j + 1] / d)
            cs.append(c_)
            sn.append(s_)
            col[j] = d
            cols.append(col[:j + 1].copy())
            gam.append(-s_ * gam[j])
            gam[j] = c_ * gam[j]
            res = abs(gam[j + 1])
            stop = bool(callback(res)) if callback is not None else False
            if res <= tol_abs:
                converged = True
                break
            if stop:
                halted = True
                break
            if hn == 0.0:
                break
            if j + 1 < m and not pending:
                ctx.axpby(1.0 / hn, w, 0.0, Q[j + 1])
        if pending:
            matvec.discard()
            pending = False
        k = len(cols)
        if k:
            rmat = np.zeros((k, k))
            for j, col in enumerate(cols):
                rmat[:j + 1, j] = col
            y = _back_substitute(rmat, np.asarray(gam[:k]))
            ctx.update(Q, k, -y, x)                # x += Q[:k]^T y
            xnz = True
    return x, converged, nmv


class MomentPreconditioner:
    """Approximate ``(I - T)^-1`` by the macroscopic moment system
    (``krylov.py:369-414``; experimental there and here, switched on by
    ``scheme.krylov_precond`` with ``scheme.accflag: 1``).

    For a state vector ``r``: the T, qx, qy blocks times ``sign`` become the
    element source through ``inv(AA_SOL)``, the trace system is solved with
    a zero distribution on the walls, and the local recovery's T, qx, qy
    replace those blocks; wall and periodic parts pass through."""

    def __init__(self, solver, acc, sign: float = -1.0):
        self.s, self.acc, self.sign = solver, acc, float(sign)
        self.nd, self.nt = solver.ndof, solver.n_tris
        self.nmom = 3 * self.nd * self.nt
        ctx = solver.ctx
        _native.check(ctx.handle, _native.lib().pbte_gsis_precond_setup(
            ctx.handle, ptr(acc.inv_aa_sol), ptr(acc.ba_sol)), "gsis_precond_setup")
        self.src = dev_zeros((3 * self.nd, self.nt))
        st = acc.last
        self.ffa = dev_zeros(st["ffa"].shape)
        self.u = dev_zeros(st["u"].shape)
        self.uq = dev_zeros((3 * self.nd, self.nt))
        self.calls = 0

    def __call__(self, v) -> None:
        h = self.s.ctx.handle
        lib = _native.lib()
        fp64 = _native.PBTE_FP64
        _native.check(h, lib.pbte_gsis_precond_src(h, ptr(v), self.sign, ptr(self.src)), "gsis_precond_src")
        _native.check(h, lib.pbte_gsis_rhs(h, fp64, None, ptr(self.ffa)), "gsis_rhs")
        self.acc.trace_solve(self.ffa, self.u)
        _native.check(h, lib.pbte_gsis_local(h, fp64, ptr(self.src), ptr(self.u), ptr(self.uq)), "gsis_local")
        v[:self.nmom].copy_(self.uq.reshape(-1))
        self.calls += 1


def _precond(solver):
    c = solver.case
    if c.scheme.krylov_precond and solver.acc is not None:
        return MomentPreconditioner(solver, solver.acc)
    return None


def krylov_basis_cap(size: int, tmax: int, max_bytes: float) -> int:
    per = max(1.0, 8.0 * float(size))
    return int(max(32, min(int(tmax), int(max_bytes / per))))


def run_krylov(solver, callback=None, progress_every: int = 0):
    ok, why = applicable(solver)
    if not ok:
        raise ValueError(f"Krylov acceleration does not apply: {why}")
    op = FixedPointOperator(solver)
    try:
        if solver.precision == "fp32":
            return _run_mixed(solver, op, callback, progress_every)
        return _run_fp64(solver, op, callback, progress_every)
    finally:
        solver.ctx.use_pbuf = False


class _History:
    """Residual and mass histories (``krylov.py:289-308``).  The mass of
    every sweep is summed on the device into a slot of a device buffer
    without waiting for it; ``finish`` reads the buffer back once."""

    _CHUNK = 4096

    def __init__(self, solver, callback, progress_every, want_true):
        self.s, self.cb, self.every = solver, callback, progress_every
        self.residuals, self.masses, self.true_res = [], [], []
        self.want_true = want_true
        self._chunks = []
        self._ctx = None

    def _slot(self, k):
        return self._chunks[k // self._CHUNK][k % self._CHUNK]

    def sweep_done(self, op):
        k = len(self.masses)
        if k // self._CHUNK >= len(self._chunks):
            self._chunks.append(dev_zeros((self._CHUNK, 2)))
        self._ctx = op.ctx
        op.ctx.residual_iterate_async(op._temp, op._temp, self._slot(k))
        self.masses.append(np.nan)
        self.residuals.append(np.nan)
        if self.want_true:
            self.true_res.append(np.nan)

    def _mass(self, k):
        return float(self._slot(k)[1].item())

    def close(self, res):
        self.residuals[-1] = res
        step = len(self.residuals)
        if self.every and step % self.every == 0:
            print(f"  step {step:8d}  residual {res:.6e}  mass {self._mass(len(self.masses) - 1):.10f}",
                  flush=True)
        if self.cb is not None:
            self.cb(step, res, self.s)

    def finish(self):
        """Read every sweep's mass back into ``masses``."""
        n = len(self.masses)
        if n:
            host = torch.cat(self._chunks).cpu().numpy()
            self.masses = [float(v) for v in host[:n, 1]]
        return self.masses


def _cap(solver, op):
    c = solver.case
    if c.scheme.krylov_restart:
        return int(c.scheme.krylov_restart)
    return krylov_basis_cap(op.n, c.iteration.tmax, c.scheme.krylov_max_bytes)


def _run_fp64(solver, op, callback, progress_every):
    """The reference driver (``krylov.py:285-363``) on device vectors."""
    c = solver.case
    tol, tmax = c.iteration.tol, c.iteration.tmax
    ctx = solver.ctx
    h = _History(solver, callback, progress_every, c.iteration.true_residual)
    t0 = time.perf_counter()
    x = op.state() if solver.restart_info.get("enabled") else dev_zeros(op.n)
    zero = dev_zeros(op.n)
    g = dev_zeros(op.n)
    op.apply(zero, g)
    h.sweep_done(op)
    gnorm = ctx.norm(g)
    scale = gnorm or 1.0
    h.close(1.0)
    tmp = dev_zeros(op.n)

    def matvec(v, out, record=True):
        op.apply(v, tmp)
        if record:
            h.sweep_done(op)
        ctx.defect(v, tmp, g, out)              # out = v - (F(v) - g)

    matvec.commit = lambda: h.sweep_done(op)
    matvec.discard = op.discard

    enough = max(tol * _ENOUGH_FACTOR, _FLOOR) * scale
    target = max(tol * c.scheme.krylov_tol_factor, _FLOOR) * scale
    cap = _cap(solver, op)
    reached = -1

    def on_residual(res):
        nonlocal reached
        h.close(res / scale)
        n = len(h.residuals)
        if reached < 0 and res <= enough:
            reached = n
        return reached >= 0 and n - reached > max(64, reached // 4)

    converged = False
    final = float("nan")
    fx = dev_zeros(op.n)
    precond = _precond(solver)
    for _ in range(c.scheme.krylov_max_rounds):
        budget = tmax - len(h.residuals) - 1
        if budget <= 0:
            break
        reached = -1
        x, _, _ = gmres_full(ctx, matvec, g, x=x, tol_abs=target, maxiter=budget, max_basis=cap,
                             callback=on_residual, precond=precond, overlap=callback is None)
        op.apply(x, fx)                                     # closing sweep
        h.sweep_done(op)
        op.publish(fx)
        final = float(ctx.true_residual(solver.mom, solver.vdf_d))
        if h.want_true:
            h.true_res[-1] = final
        h.close(final)
        converged = bool(final < tol) if np.isfinite(final) else gnorm == 0.0
        if converged or len(h.residuals) >= tmax:
            break
        target = max(target * 1.0e-3, _FLOOR * scale)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    h.finish()
    return solver._record(len(h.residuals), converged, h.residuals, h.masses,
                          h.true_res if h.want_true else None, wall, scheme="KRYLOV",
                          extra={"krylov_size": op.n, "krylov_periodic_state": op.nper,
                                 "krylov_basis_cap": cap, "krylov_precond": precond is not None,
                                 "precond_calls": precond.calls if precond is not None else 0,
                                 "transport_residual": final, "sweeps_fp64": op.sweeps64,
                                 "sweeps_fp32": op.sweeps - op.sweeps64,
                                 "sweeps_discarded": op.discarded})


def _run_mixed(solver, op, callback, progress_every, inner_reduction: float = 1.0e-2,
               max_refine: int = 64, stall_products: int = 300):
    """FP64 defect correction around FP32-sweep GMRES (module docstring).

    ``inner_reduction``: each inner solve reduces the defect by this
    factor.  Two digits: a binary64 refinement sweep costs two binary32
    products with stored binary64 factors and about four when they are
    refined from the binary32 inverse (the "mixed" mode), so deeper inner
    solves do not pay (on a 362^2 P3 square with a hole and 1 024
    ordinates, inner solves to 1e-4 took 36 more binary32 products to save
    two binary64 sweeps)."""
    c = solver.case
    tol, tmax = c.iteration.tol, c.iteration.tmax
    ctx = solver.ctx
    h = _History(solver, callback, progress_every, c.iteration.true_residual)
    t0 = time.perf_counter()
    x = op.state() if solver.restart_info.get("enabled") else dev_zeros(op.n)
    fx = dev_zeros(op.n)
    r = dev_zeros(op.n)
    tmp = dev_zeros(op.n)

    def residual():
        op.apply(x, fx, "fp64")                 # F64(x) = T x + g
        h.sweep_done(op)
        r.copy_(fx)
        ctx.axpby(-1.0, x, 1.0, r)              # r = F(x) - x
        return ctx.norm(r)

    rnorm = residual()
    scale = rnorm or 1.0                        # = ||g|| when x starts at 0

    def matvec(v, out, record=True):            # (I - T32) v, homogeneous
        op.apply(v, tmp, "fp32", homogeneous=True)
        if record:
            h.sweep_done(op)
        ctx.defect(v, tmp, None, out)

    matvec.commit = lambda: h.sweep_done(op)
    matvec.discard = op.discard

    # the defect the refinement drives to: the level at which the binary64
    # driver counts its GMRES residual as enough (two digits below the
    # tolerance); the decision is the transport residual, as there
    target = max(tol * _ENOUGH_FACTOR, _FLOOR) * scale
    cap = _cap(solver, op)
    converged = False
    final = float("nan")
    refinements = 0
    e = dev_zeros(op.n)
    precond = _precond(solver)
    # performance.refine_tol > 0: the error of x after a refinement is about
    # rho / (1 - rho) times that refinement's correction, rho the defect
    # reduction it achieved (the next corrections shrink by rho each); the
    # refinement stops when that estimate, relative to x, is a tenth of
    # refine_tol for the temperature moments and for the heat-flux moments
    refine_tol = c.performance.refine_tol
    kt = op.nd * op.nt
    est = np.inf
    while True:
        op.publish(fx)
        final = float(ctx.true_residual(solver.mom, solver.vdf_d))
        if h.want_true:
            h.true_res[-1] = final
        h.close(final if refinements else 1.0)
        if refine_tol > 0.0:
            converged = rnorm == 0.0 or est <= 0.1 * refine_tol
        else:
            converged = bool(final < tol) and rnorm <= target
        if converged or refinements >= max_refine or len(h.residuals) >= tmax:
            break
        budget = tmax - len(h.residuals) - 1
        if budget <= 0:
            break
        e.zero_()
        # two digits per refinement: binary32 products lose about
        # eps32 / (smallest eigenvalue of I - T) of accuracy, which near the
        # diffusive limit stops the inner residual well above eps32; an inner
        # solve that no longer improves by 10 % within stall_products
        # products hands back to the binary64 refinement
        best = [np.inf, 0]

        def inner(res):
            h.close(res / scale)
            best[1] += 1
            if res < 0.9 * best[0]:
                best[0], best[1] = res, 0
            return best[1] > stall_products

        gmres_full(ctx, matvec, r, x=e, tol_abs=max(inner_reduction * rnorm, 0.1 * target),
                   maxiter=budget, max_basis=cap, callback=inner, precond=precond, overlap=callback is None)
        ctx.axpby(1.0, e, 1.0, x)
        refinements += 1
        rnorm_old, rnorm = rnorm, residual()
        if refine_tol > 0.0:
            rho = rnorm / rnorm_old if rnorm_old > 0.0 else 0.0
            est = np.inf
            if rho < 1.0:
                grow = rho / (1.0 - rho)
                parts = ((e[:kt], x[:kt]), (e[kt:3 * kt], x[kt:3 * kt]))
                est = max(grow * ctx.norm(de) / max(ctx.norm(xx), np.finfo(float).tiny) for de, xx in parts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    h.finish()
    return solver._record(len(h.residuals), converged, h.residuals, h.masses,
                          h.true_res if h.want_true else None, wall, scheme="KRYLOV",
                          extra={"krylov_size": op.n, "krylov_periodic_state": op.nper,
                                 "krylov_basis_cap": cap, "krylov_precond": precond is not None,
                                 "precond_calls": precond.calls if precond is not None else 0,
                                 "transport_residual": final, "refinements": refinements,
                                 "final_defect": rnorm / scale, "refine_tol": refine_tol,
                                 "refine_error_estimate": est, "sweeps_fp64": op.sweeps64,
                                 "sweeps_fp32": op.sweeps - op.sweeps64,
                                 "sweeps_discarded": op.discarded})
