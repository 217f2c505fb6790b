"""``Solver``: set up a case and run the outer iteration on the device.

Same constructor, methods and record as the reference's
``solver-python/pybte/driver.py:41-290``::

    STEP = 1
    DO
        adiabatic walls: refresh the diffuse emission      (driver.py:150-153)
        transport sweep                                    (ctx.sweep)
        CIS: moments | GSIS: HoT source, trace solve, local recovery, correction
        iterate residual, mass
        EXIT when residual < tol or STEP >= tmax
    END DO

The distribution and moments stay in HBM for the whole run; ``vdf`` and the
``Moments`` attributes copy to the host on access.
"""
from __future__ import annotations

import time
import warnings
import weakref
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

from .config import Case
from .restart import read_restart, restart_name, write_restart
from .device import dev_zeros, require_gpu, stream_handle, torch
from .io_output import RunRecord, _fields, append_runtime_log, locate_points, sampling_grid, write_tecplot
from .moments import Moments
from .problem import build_problem
from .sweep import SweepContext

__all__ = ["Solver", "HostVdf"]


class _Mirror:
    """Which copy of a solver's distribution is newer: the host array
    (``dirty``, written through ``Solver.vdf``) or the device tensor
    (``stale``: a device step changed it since the host array was read)."""

    def __init__(self, solver):
        self._solver = weakref.ref(solver)     # no cycle: the solver frees its HBM on del
        self.array = None
        self.dirty = False
        self.stale = False

    @property
    def solver(self):
        return self._solver()

    def push(self) -> None:
        """Host writes to the device (when there are any)."""
        if self.dirty:
            s = self.solver
            s.sweep_vdf.copy_(torch.from_numpy(np.asarray(self.array)).to(s.sweep_vdf.dtype))
            self.dirty = False

    def pull(self) -> None:
        """The device state into the host array, in place."""
        np.asarray(self.array)[...] = self.solver.sweep_vdf.detach().to(torch.float64).cpu().numpy()
        self.dirty = False
        self.stale = False

    def refresh(self) -> None:
        """Before a host write: a held array that device steps have left
        behind is brought up to date first, so the write lands on the
        current distribution (as on the reference's live array)."""
        if self.stale and self.array is not None and not self.dirty:
            self.pull()


class HostVdf(np.ndarray):
    """``Solver.vdf``: the host ``[D, N, P]`` binary64 mirror of the
    distribution in HBM, as callers of the reference use ``solver.vdf``
    (``run_agent_case.py``, ``evaluate.py``): element writes
    (``solver.vdf[:] = f``, also through views and in-place operators) make
    it newer than the device copy; the solver's device operations given it
    (``compute_moments``, ``update_wall_flux``, ``ctx.true_residual``,
    ``ctx.sweep``) upload it first, and ``ctx.sweep`` writes its result
    back into it, in place."""

    def __array_finalize__(self, obj):
        shared = obj is not None and np.may_share_memory(self, obj)
        self._mirror = getattr(obj, "_mirror", None) if shared else None

    def _touch(self):
        if self._mirror is not None:
            self._mirror.dirty = True

    def _refresh(self):
        if self._mirror is not None:
            self._mirror.refresh()

    def __setitem__(self, key, value):
        self._refresh()
        super().__setitem__(key, value)
        self._touch()


def _inplace(name):
    base = getattr(np.ndarray, name)

    def op(self, other):
        self._refresh()
        out = base(self, other)
        self._touch()
        return out
    op.__name__ = name
    return op


for _name in ("__iadd__", "__isub__", "__imul__", "__itruediv__", "__ipow__"):
    setattr(HostVdf, _name, _inplace(_name))


class Solver:
    def __init__(self, case: Case, validate_order: bool = False, on_cycle: str | None = None,
                 max_bundle: int = 0, block: int = 0, sweep_variant: int = 0,
                 lds_slots_cap: int = 0, max_bytes: float | None = None):
        t0 = time.perf_counter()
        require_gpu()
        self.case = c = case
        pb = build_problem(c, on_cycle=on_cycle, validate_order=validate_order)
        self.problem = pb
        self.mesh, self.basis, self.vel = pb.mesh, pb.basis, pb.vel
        self.integrals, self.order, self.bcdata = pb.integrals, pb.order, pb.bcdata
        self.ndir, self.n_tris, self.ndof = self.vel.ndir, self.mesh.n_tris, c.ndof_tri
        self.tile_shape = None
        if c.performance.autotune and not (max_bundle or block or sweep_variant):
            from .autotune import tune
            self.tile_shape = tune(c, pb, max_bytes=max_bytes)
            if self.tile_shape is None:
                warnings.warn("performance.autotune: no tile shape could be timed (no stored factors of "
                              "the run's precision); the default schedule is used")
            else:
                max_bundle, block, sweep_variant = (self.tile_shape[k] for k in ("max_bundle", "block", "variant"))
        # GSIS: the HDG operators and the trace factorisation (device work
        # and the trace solver's host symbolic phase) are built on a worker
        # thread while this thread builds the sweep context (host schedule)
        prep = None
        if c.scheme.accflag == 1:
            from .acceleration import prepare
            pool = ThreadPoolExecutor(max_workers=1)
            prep = pool.submit(prepare, c, pb, stream_handle())
            pool.shutdown(wait=False)
        try:
            self.ctx = SweepContext.build(c, self.mesh, self.integrals, self.vel, self.order,
                                          self.bcdata, max_bundle=max_bundle, block=block,
                                          variant=sweep_variant, lds_slots_cap=lds_slots_cap,
                                          max_bytes=max_bytes)
        except BaseException:
            if prep is not None:          # let the worker finish its device work first
                try:
                    prep.result()
                except Exception:
                    pass
            raise
        prepared = prep.result() if prep is not None else None
        # element operators factorised (driver.py:84-85, 155-156; krylov.py:
        # 119-120): D N once when stored ("precomputed", and the binary32 ones
        # of "mixed", whose binary64 sweeps refine against the stored binary32
        # inverse and factorise nothing), plus D N per "onthefly" sweep
        self._stored_factorisations = self.ndir * self.n_tris if self.ctx.mode != "onthefly" else 0
        self.precision = c.performance.precision
        self.sweep_precision = self.precision     # fp32 runs finish in fp64
        self.vdf_d = dev_zeros((self.ndir, self.n_tris, self.ndof))
        self.vdf32_d = (dev_zeros((self.ndir, self.n_tris, self.ndof), dtype=torch.float32)
                        if self.precision == "fp32" else None)
        self.mom = Moments.zeros(self.ndof, self.n_tris)
        self.mom._ctx = self.ctx
        self._mirror = _Mirror(self)
        self.ctx._solver = weakref.ref(self)
        self.bcdata._ctx = self.ctx
        self.t_old_d = dev_zeros(self.n_tris)
        self.acc = None
        self.restart_info = self._apply_restart()
        if c.scheme.accflag == 1:
            from .acceleration import Acceleration
            self.acc = Acceleration(self, prepared=prepared)
        torch.cuda.synchronize()
        self.setup_seconds = time.perf_counter() - t0

    @property
    def factorisation_count(self) -> int:
        return self._stored_factorisations + self.ctx.factorisations

    # -- state views ---------------------------------------------------------
    @property
    def sweep_vdf(self):
        """The device tensor the sweeps run on (binary32 in the fp32 phase)."""
        return self.vdf32_d if self.sweep_precision == "fp32" else self.vdf_d

    def promote(self) -> None:
        """Continue with binary64 sweeps from the binary32 distribution."""
        if self.sweep_precision == "fp32":
            self.settle()
            self.vdf_d.copy_(self.vdf32_d)
            self.sweep_precision = "fp64"

    @property
    def vdf(self) -> np.ndarray:
        """The distribution ``[D, N, P]`` (binary64) as a host array
        (``HostVdf``): refreshed from HBM on access, or, when it was written
        on the host since, uploaded to HBM instead."""
        self.settle()
        m = self._mirror
        if m.array is None:
            m.array = np.empty((self.ndir, self.n_tris, self.ndof)).view(HostVdf)
            m.array._mirror = m
            m.pull()
        elif m.dirty:
            m.push()
        else:
            m.pull()
        return m.array

    @vdf.setter
    def vdf(self, value) -> None:
        if self.acc is not None:
            self.acc.swept()          # replaced: a pending correction no longer applies
        self.vdf_d.copy_(torch.as_tensor(np.ascontiguousarray(value, dtype=np.float64)))
        if self.sweep_precision == "fp32":
            self.vdf32_d.copy_(self.vdf_d)
        self._mirror.dirty = False
        self._mirror.stale = True

    def device_vdf(self, vdf):
        """The device tensor a device operation runs on for ``vdf``, and
        the function that hands a changed distribution back (None when
        there is nothing to hand back): a device tensor is itself; this
        solver's host mirror is the solver's distribution (host writes
        uploaded first); any other host array ``[D, N, P]`` is copied to a
        scratch binary64 tensor and copied back."""
        if isinstance(vdf, torch.Tensor):
            return vdf, None
        m = getattr(vdf, "_mirror", None)
        shape = (self.ndir, self.n_tris, self.ndof)
        if m is self._mirror and vdf.shape == shape and np.shares_memory(vdf, m.array):
            m.push()
            return self.sweep_vdf, m.pull
        arr = np.asarray(vdf)
        if arr.shape != shape:
            raise ValueError(f"vdf shape {arr.shape} does not match the discretisation {shape}")
        dev = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to(self.vdf_d.device)

        def back():
            arr[...] = dev.cpu().numpy()
        return dev, back

    def vdf_fortran(self) -> np.ndarray:
        npole, nazim = self.case.velmesh.npole, self.case.velmesh.nazim
        return np.transpose(self.vdf.reshape(nazim, npole, self.n_tris, self.ndof), (3, 2, 1, 0))

    # -- restart (driver.py:100-139) ------------------------------------------
    def _restart_default(self) -> Path:
        c = self.case
        return Path(c.basedir) / restart_name(c.dg.deg, self.n_tris, c.velmesh.npole, c.velmesh.nazim)

    def _apply_restart(self) -> dict:
        """Start from a ``VDF*.out`` file when ``restart.enabled``.  With
        restart off, a file of this run's name in the case directory is an
        error unless ``restart.error_on_stale`` is off: the Fortran code
        would have read it without a word."""
        r = self.case.restart
        off = {"enabled": False, "path": None, "sha256": None}
        default = self._restart_default()
        if not r.enabled:
            if r.error_on_stale and default.exists():
                raise FileExistsError(f"{default} exists while restart is off; remove it, or set "
                                      "restart.enabled or turn restart.error_on_stale off")
            return off
        path = default if not r.path else Path(r.path)
        path = path if path.is_absolute() else Path(self.case.basedir) / path
        if not path.exists():
            raise FileNotFoundError(f"restart.enabled is set and {path} is missing")
        arr, digest = read_restart(path, (self.ndir, self.n_tris, self.ndof))
        self.vdf = arr
        return {"enabled": True, "path": str(path), "sha256": digest}

    def save_restart(self, path=None) -> Path:
        """Write the distribution as a ``VDF*.out`` restart file."""
        return write_restart(Path(path) if path else self._restart_default(), self.vdf)

    # -- one outer iteration (driver.py:149-161) -------------------------------
    def settle(self) -> None:
        """Bring a pending GSIS correction into the distribution (the
        accelerated step leaves it pending; see ``Acceleration.settle``)."""
        if self.acc is not None:
            self.acc.settle()

    def device_changed(self) -> None:
        """Host writes through ``vdf`` reach the device first; the host
        array is behind the device after (``step``, ``run``, device sweeps)."""
        self._mirror.push()
        self._mirror.stale = True

    def step(self) -> None:
        self.device_changed()
        v = self.sweep_vdf
        if self.acc is not None:
            self.acc.settle(boundary_only=True)
        if self.bcdata.has_nonthermalising:
            self.ctx.wall_flux(v)
        self.ctx.sweep(self.mom, v)
        if self.acc is not None:
            self.acc.swept()
        if self.acc is None:
            self.ctx.moments(v, self.mom)
        else:
            self.acc.apply()

    def run(self, callback=None, progress_every: int = 0) -> RunRecord:
        self.device_changed()
        if self.case.scheme.method == "krylov":
            from .krylov import run_krylov
            return run_krylov(self, callback, progress_every)
        if self.precision == "fp32" and self.sweep_precision == "fp32":
            from .refine import applicable, run_refined
            if applicable(self)[0]:
                return run_refined(self, callback, progress_every)
        return self._run_source(callback, progress_every)

    def _run_source(self, callback=None, progress_every: int = 0) -> RunRecord:
        c = self.case
        tol, tmax = c.iteration.tol, c.iteration.tmax
        want_true = c.iteration.true_residual
        every = max(1, int(c.iteration.true_residual_every))
        residuals, masses, true_res = [], [], []
        t0 = time.perf_counter()
        step = 0
        converged = False
        # fp32: binary32 sweeps until the iterate step is below the switch
        # level or has stopped falling (it floors at the binary32 rounding
        # of the sweep), then binary64 sweeps to the tolerance
        switch = max(tol, c.performance.refine_switch)
        step32 = 0
        while True:
            step += 1
            self.step()
            res, mass = self.ctx.residual_iterate(self.mom.temp_d, self.t_old_d)
            self.t_old_d.copy_(self.mom.temp_d)
            residuals.append(res)
            masses.append(mass)
            if want_true:
                if step % every == 0:
                    self.settle()
                true_res.append(self.ctx.true_residual(self.mom, self.sweep_vdf)
                                if step % every == 0 else np.nan)
            if progress_every and step % progress_every == 0:
                print(f"  step {step:8d}  residual {res:.6e}  mass {mass:.10f}", flush=True)
            if callback is not None:
                callback(step, res, self)
            if self.sweep_precision == "fp32":
                step32 += 1
                # no new minimum over 8 steps: at the rounding floor (a slow
                # but steady decrease keeps making new minima)
                stalled = step32 > 8 and res >= min(residuals[-9:-1])
                if res < switch or stalled:
                    self.promote()
                    step32 = step
                if step >= tmax:
                    break
                continue
            if res < tol:
                if (self.acc is not None and self.acc.omega and not self.acc.defect_every
                        and self.acc.refresh_defect() > tol):
                    continue
                converged = True
                break
            if step >= tmax:
                break
        self.settle()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        extra = None
        if self.precision == "fp32":
            extra = {"steps_fp32": step32 if self.sweep_precision == "fp64" else step,
                     "refine_switch": switch}
        return self._record(step, converged, residuals, masses,
                            true_res if want_true else None, wall,
                            scheme="GSIS" if c.scheme.accflag == 1 else "CIS", extra=extra)

    def _record(self, iterations, converged, residuals, masses, true_res, wall, scheme,
                extra=None) -> RunRecord:
        c = self.case
        mom = self.mom
        rec = RunRecord(
            iterations=iterations, converged=converged,
            residual_history=np.asarray(residuals),
            residual_true=np.asarray(true_res) if true_res is not None else None,
            mass_history=np.asarray(masses), temp=mom.temp, qx=mom.qx, qy=mom.qy,
            temp_dofs=mom.ts, qx_dofs=mom.qxs, qy_dofs=mom.qys,
            sweep_count=self.ctx.sweep_count, factorisation_count=self.factorisation_count,
            wall_clock=wall, setup_wall_clock=self.setup_seconds, config_hash=c.config_hash,
            scheme=scheme, acc_variant=c.scheme.acc_variant if c.scheme.accflag == 1 else None,
            tol=c.iteration.tol, tmax=c.iteration.tmax, tau_r=c.flow.tau_r, tau_n=c.flow.tau_n,
            tau_c=c.tau_c, deg=c.dg.deg, npole=c.velmesh.npole, nazim=c.velmesh.nazim,
            n_tris=self.n_tris, n_faces=self.mesh.n_faces, mesh_file=str(c.mesh_path()),
            restart=self.restart_info, environment=RunRecord.environment_fingerprint(),
            diagnostics={"sweep_mode": self.ctx.mode, "precision": self.precision,
                         "cycles_broken": int(self.order.cycles_broken.sum()),
                         "hmin": float(self.mesh.hmin), "n_faces_b": int(self.mesh.n_faces_b),
                         "boundary_types": [b.type for b in c.boundaries],
                         **{f"schedule_{k}": v for k, v in self.ctx.schedule_info().items()}})
        if self.tile_shape is not None:
            rec.diagnostics["tile_shape"] = {k: self.tile_shape[k] for k in ("max_bundle", "block", "variant", "ms")}
        if self.acc is not None:
            rec.diagnostics.update(self.acc.diagnostics())
        if extra:
            rec.diagnostics.update(extra)
        return rec

    # -- output (driver.py:256-295) --------------------------------------------
    def write_outputs(self, rec: RunRecord) -> dict:
        """The run's files under ``output.dir``, as the reference writes
        them: ``run_record.json``, a ``RunTime.txt`` line, and the fields
        T, qx, qy (and the GSIS stress trio Nxx, Nxy, Nyy) sampled on the
        output grid into ``field.npz`` and/or ``field.dat`` (Tecplot).
        The fields are the angular moments of the distribution, as the
        reference samples them (after a GSIS correction the run's moments
        are not exactly those of the corrected distribution), summed in
        HBM by the moment kernel: one copy of ``3 P N`` values to the
        host, not of the distribution."""
        self.settle()
        c = self.case
        out = c.output_path()
        out.mkdir(parents=True, exist_ok=True)
        written = {}
        if c.output.run_record:
            written["run_record"] = str(rec.to_json(out / "run_record.json"))
        if c.output.runtime_log:
            written["runtime_log"] = str(append_runtime_log(
                out / "RunTime.txt", rec.scheme, c.flow.tau_r, c.flow.tau_n, self.n_tris,
                c.velmesh.npole, c.velmesh.nazim, rec.iterations, rec.wall_clock))
        if c.output.field:
            px, py = sampling_grid(c.output.field_nx, c.output.field_ny)
            triid = locate_points(self.mesh, px, py)
            missing = int(np.count_nonzero(triid < 0))
            if missing:
                warnings.warn(f"{missing} sampling points fell outside the mesh")
            uq = self.acc.full_uq().cpu().numpy() if self.acc is not None else None
            m = Moments.zeros(self.ndof, self.n_tris)
            self.ctx.moments(self.sweep_vdf, m)
            fields = _fields(self.mesh, self.basis, m.ts, m.qxs, m.qys, px, py, triid, uq, self.ndof)
            names = ["T", "qx", "qy", "Nxx", "Nxy", "Nyy"]
            if c.output.npz:
                np.savez_compressed(out / "field.npz", px=px, py=py, triid=triid,
                                    **dict(zip(names, fields)))
                written["field_npz"] = str(out / "field.npz")
            if c.output.tecplot:
                written["field_dat"] = str(write_tecplot(out / "field.dat", px, py, fields, names))
        return written
