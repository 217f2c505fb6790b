"""Benchmark of the gray Callaway DG discrete-ordinates hot path on one MI355X.

Default workload (--workload config2): the GSIS configuration -- a 256 x 256
triangulated unit square with its interior vertices moved by
U(-0.25h, 0.25h) (numpy default_rng(0)), 131 072 P2 triangles (Np = 6), 512
ordinates (16 Gauss-Legendre polar x 32 azimuth), Kn_R = 1e-3, Kn_N = 1e-4,
one wall at dT = +1 and the opposite wall at 0, the other two sides a
periodic pair (in place of the specular walls, which the reference does not
have; for this layout both give the same one-dimensional solution).

A step is one transport sweep over all 512 directions in binary32 (element
inverses and distribution; the sweeps of the FP32 GSIS solve), the moments
already resident in HBM.  One JSON line:

  value          sweep DOF-updates/s over the timed steps (wall clock);
  roofline       algorithmic bytes per sweep / mean sweep-kernel time, the
                 kernel timed with HIP events recorded on its launch stream
                 around each launch; `traffic` is the HBM bytes per launch
                 from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs
                 (profiles/sweep_traffic.json);
  solve          the FP64 GSIS solve and the FP32-sweep GSIS solve with the
                 binary64 finish of the same case, each to the iterate step
                 1e-8: time-to-solution (setup + run), solves/hour, iterations,
                 the relative L2 difference of T and of the heat flux between
                 them, and the time-to-solution speedup;
  cpu_baseline   the C restatement of the reference's sweep (binary64,
                 OpenMP over directions) on this host, on a bounded sample:
                 the same mesh with 16 of the ordinates, at the host's thread
                 count and at one thread.

Other workloads (each prints its own line):
--workload sweep   the Krylov configuration: 128 x 128 structured square,
                   32 768 P2 triangles, 256 ordinates (8 x 32), sweep
                   DOF-updates/s and FP64 vs FP32+refinement GMRES solves at
                   Kn_R in {0.01, 0.1, 1, 10} (Kn_N = 0.1 Kn_R).  With --n 362
                   --deg 3 --npole 16 --nazim 64 --jitter 0.25 --hole 0.2
                   --tau-r 0.05 it is the large single solve (P3, 1 024
                   ordinates, a central hot hole);
--workload cis     the plain-CIS configuration (16 x 16, 512 P1 triangles,
                   64 ordinates, Kn_R = Kn_N = 1) to 1e-8;
--workload gsis    the GSIS solves alone at --n;
--workload batch   64 independent problems drawn from default_rng(1)
                   (2k-65k jittered P1/P2 triangles, 64-512 ordinates, Kn_R
                   log-uniform on [1e-3, 10]); a step is one FP32 sweep of
                   every problem; a GSIS solve leg of the same draws.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12
VALU_F32_PEAK = 157.3e12        # MI355X published FP32 vector peak (packed; 78.6e12 unpacked)
METRIC = "solves/hour + sweep DOF-updates/s, FP32+refine vs FP64 GPU at 1e-6 rel L2"
TRAFFIC_FILE = ROOT / "profiles" / "sweep_traffic.json"
FP32_OPERATOR = "fused"         # --fp32-operator


def log(msg):
    """Progress on stderr (stdout carries the one JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def committed_traffic(workload):
    """HBM bytes per sweep launch measured by rocprofv3 (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration) for this workload."""
    try:
        table = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    row = table.get("workloads", {}).get(workload)
    return None if row is None else row.get("hbm_bytes_per_launch")


def make_case(pkg, n, deg, npole, nazim, tau_r, tau_n, precision, workdir, jitter=0.0, seed=0, hole=0.0,
              method="krylov", accflag=0, tmax=20000):
    """Diffuse isothermal walls, the north one at dT = +1; with ``hole`` > 0
    a central circular hole of that radius whose wall is at dT = +1 and
    the four outer walls at 0."""
    from pbte_amd.mesh import write_square_mesh
    mesh = Path(workdir) / f"square_{n}_{jitter}_{seed}_{hole}.msh"
    if not mesh.exists():
        write_square_mesh(mesh, n, jitter=jitter, seed=seed, hole=(0.5, 0.5, hole) if hole > 0 else None)
    walls = [
        {"name": "SWall", "phyid": 11, "type": "thermalising", "temp": 0.0},
        {"name": "NWall", "phyid": 12, "type": "thermalising", "temp": 0.0 if hole > 0 else 1.0},
        {"name": "EWall", "phyid": 13, "type": "thermalising", "temp": 0.0},
        {"name": "WWall", "phyid": 14, "type": "thermalising", "temp": 0.0},
    ]
    if hole > 0:
        walls.append({"name": "Hole", "phyid": 16, "type": "thermalising", "temp": 1.0})
    return pkg.Case.from_dict({
        "iteration": {"tol": 1e-8, "tmax": tmax},
        "scheme": {"accflag": accflag, "method": method, "krylov_restart": 30},
        "velmesh": {"npole": npole, "nazim": nazim},
        "dg": {"deg": deg},
        "flow": {"cv": 1.0, "vg": 1.0, "tau_r": tau_r, "tau_n": tau_n},
        "mesh": {"file": str(mesh)},
        "boundaries": walls,
        "restart": {"enabled": False, "error_on_stale": False},
        "output": {"field": False, "runtime_log": False, "run_record": False},
        "performance": {"precision": precision, "refine_tol": 1e-6 if precision == "fp32" else 0.0,
                        "fp32_operator": FP32_OPERATOR},
    }, basedir=str(workdir))


def gsis_case(pkg, n, deg, npole, nazim, tau_r, tau_n, precision, workdir, jitter=0.25, seed=0):
    """Hot wall (dT = +1) opposite a cold wall (0), the other two sides a
    periodic pair: for this layout a specular pair and a periodic pair give
    the same one-dimensional solution, and the reference has no specular
    wall.  The walls are north and south of the mesh writer's channel
    layout (the problem turned by 90 degrees)."""
    from pbte_amd.mesh import write_square_mesh
    mesh = Path(workdir) / f"channel_{n}_{jitter}_{seed}.msh"
    if not mesh.exists():
        write_square_mesh(mesh, n, layout="channel", jitter=jitter, seed=seed)
    return pkg.Case.from_dict({
        "iteration": {"tol": 1e-8, "tmax": 5000},
        "scheme": {"accflag": 1, "method": "source"},
        "velmesh": {"npole": npole, "nazim": nazim},
        "dg": {"deg": deg},
        "flow": {"cv": 1.0, "vg": 1.0, "tau_r": tau_r, "tau_n": tau_n},
        "mesh": {"file": str(mesh)},
        "boundaries": [
            {"name": "SWall", "phyid": 11, "type": "thermalising", "temp": 0.0},
            {"name": "NWall", "phyid": 12, "type": "thermalising", "temp": 1.0},
            {"name": "Master", "phyid": 13, "type": "periodic", "xoff": -1.0},
            {"name": "Slave", "phyid": 14, "type": "periodic", "xoff": 1.0},
        ],
        "restart": {"enabled": False, "error_on_stale": False},
        "output": {"field": False, "runtime_log": False, "run_record": False},
        "performance": {"precision": precision, "fp32_operator": FP32_OPERATOR},
    }, basedir=str(workdir))


def gsis_solves(pkg, args, work, n):
    """FP64 and FP32-sweep GSIS solves of the config-2 case on an n x n mesh;
    the sweep kernel is timed with HIP events over every sweep of each
    solve.  Returns the per-precision figures and the comparison."""
    import torch
    res, fields, shape = {}, {}, None
    for prec in ("fp64", "fp32"):
        case = gsis_case(pkg, n, args.deg, args.npole, args.nazim, 1e-3, 1e-4, prec, work,
                         jitter=args.jitter if args.jitter else 0.25)
        log(f"gsis {prec}: setup")
        t0 = time.perf_counter()
        s = pkg.Solver(case)
        setup = time.perf_counter() - t0
        s.ctx.timing(case.iteration.tmax)
        log(f"gsis {prec}: run")
        t1 = time.perf_counter()

        def progress(k, r, _s, prec=prec):
            if k % 50 == 0:
                log(f"gsis {prec}: step {k}, residual {r:.3e}")

        rec = s.run(callback=progress)
        torch.cuda.synchronize()
        run = time.perf_counter() - t1
        ms = s.ctx.timing_read(case.iteration.tmax)
        s.ctx.timing(0)
        fields[prec] = (rec.temp, np.concatenate([rec.qx, rec.qy]))
        shape = (s.ndir, s.n_tris, s.ndof)
        k32 = int(rec.diagnostics.get("steps_fp32", 0))
        # which sweeps ran in binary64: the refined iteration lists its
        # binary64 steps; the switch runs its binary32 sweeps first
        is64 = np.ones(len(ms), dtype=bool)
        if "fp64_steps" in rec.diagnostics:
            is64[:] = False
            is64[np.asarray(rec.diagnostics["fp64_steps"], dtype=np.int64) - 1] = True
        else:
            is64[:k32] = False
        res[prec] = {"time_to_solution_s": setup + run, "setup_s": setup, "run_s": run,
                     "solves_per_hour": 3600.0 / (setup + run), "iterations": rec.iterations,
                     "converged": rec.converged, "mass": rec.mass, "steps_fp32": k32,
                     "steps_fp64": int(is64.sum()), "sweeps": len(ms),
                     "precision_scheme": rec.diagnostics.get("precision_scheme", prec),
                     "sweep_kernel_ms_fp32_mean": float(np.mean(ms[~is64])) if (~is64).any() else None,
                     "sweep_kernel_ms_fp64_mean": float(np.mean(ms[is64])) if is64.any() else None,
                     "sweep_kernel_s_total": float(np.sum(ms)) * 1e-3,
                     "trace_unknowns": rec.diagnostics.get("global_matrix_n"),
                     "trace_solve": rec.diagnostics.get("trace_solve"),
                     "trace_ordering": rec.diagnostics.get("trace_ordering")}
        del s
        torch.cuda.empty_cache()
    (t64, q64), (t32, q32) = fields["fp64"], fields["fp32"]
    res["fp32"]["rel_l2_T_vs_fp64"] = float(np.linalg.norm(t32 - t64) / np.linalg.norm(t64))
    res["fp32"]["rel_l2_q_vs_fp64"] = float(np.linalg.norm(q32 - q64) / np.linalg.norm(q64))
    res["fp32"]["max_abs_T_vs_fp64"] = float(np.abs(t32 - t64).max())
    res["speedup_fp32_vs_fp64"] = res["fp64"]["time_to_solution_s"] / res["fp32"]["time_to_solution_s"]
    res["speedup_fp32_vs_fp64_run"] = res["fp64"]["run_s"] / res["fp32"]["run_s"]
    return res, shape


def gsis_workload(pkg, args, work):
    """The GSIS solves alone (value: solves/hour of the FP32 solve)."""
    n = args.n
    res, (D, N, P) = gsis_solves(pkg, args, work, n)
    t32 = res["fp32"]["sweep_kernel_ms_fp32_mean"] or res["fp32"]["sweep_kernel_ms_fp64_mean"]
    bytes_alg = b_alg(D, N, P, 4, 4)
    workload = (f"gsis jittered channel {n}x{n} -> {N} P{args.deg} triangles, {D} ordinates "
                f"({args.npole}x{args.nazim}), Kn_R 1e-3, Kn_N 1e-4, hot/cold walls, periodic sides")
    out = {
        "metric": METRIC, "value": res["fp32"]["solves_per_hour"], "unit": "solves/hour",
        "n_gpus": 1, "steps": res["fp32"]["iterations"], "warmup": 0,
        "ms_per_step": 1e3 * res["fp32"]["run_s"] / max(res["fp32"]["iterations"], 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: the workload's mesh, ordinates and boundary data",
        "config": {"workload": workload, "n_tris": N, "ndof": P, "ndir": D},
        "roofline": {"bound": "hbm", "achieved": bytes_alg / (t32 * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": bytes_alg / (t32 * 1e-3) / HBM_PEAK, "traffic": None,
                     "bytes_alg_per_launch": bytes_alg},
        "solve": res, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def trace_workload(pkg, args, work):
    """The GSIS trace solve of the config-2 mesh alone: its dissection per
    level and args.steps solves from the binary64 factors and from their
    binary32 copy (after args.warmup each), timed with HIP events on the
    solver's stream.  The trace system does not depend on the ordinates, so
    the solver is built with 4 x 4 of them."""
    import torch
    case = gsis_case(pkg, args.n, args.deg, 4, 4, 1e-3, 1e-4, "fp32", work,
                     jitter=args.jitter if args.jitter else 0.25)
    log(f"trace: jittered channel {args.n}x{args.n}, P{args.deg}")
    s = pkg.Solver(case)
    tr = s.acc.trace
    st = s.acc.last
    rhs = torch.randn(st["ffa"].shape, dtype=torch.float64, device="cuda",
                      generator=torch.Generator(device="cuda").manual_seed(0))
    u = torch.zeros_like(st["u"])
    times = {}
    for prec in ("fp64", "fp32"):
        for _ in range(args.warmup):
            tr.solve(rhs, u, prec)
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
        for a, b in ev:
            a.record()
            tr.solve(rhs, u, prec)
            b.record()
        torch.cuda.synchronize()
        times[prec] = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    info = tr.info()
    out = {"workload": f"gsis trace solve, jittered channel {args.n}x{args.n}, P{args.deg}",
           "unknowns": info["unknowns"], "fronts": info["nodes"], "factor_entries": info["factor_entries"],
           "solve_ms_fp64": times["fp64"], "solve_ms_fp32": times["fp32"],
           "factor_GBps_fp64": 8.0 * info["factor_entries"] / times["fp64"] / 1e6,
           "factor_GBps_fp32": 4.0 * info["factor_entries"] / times["fp32"] / 1e6,
           "levels": tr.levels(), "steps": args.steps, "warmup": args.warmup}
    print(json.dumps(out), flush=True)


def config2_workload(pkg, args, work):
    """The headline line: the config-2 FP32 sweep (value, roofline), the
    FP64 vs FP32 GSIS solves of the same case (solve) and the CPU sweep
    (cpu_baseline)."""
    import torch
    n = args.n
    log(f"config 2: jittered channel {n}x{n}, P{args.deg}, {args.npole}x{args.nazim} ordinates")
    case = gsis_case(pkg, n, args.deg, args.npole, args.nazim, 1e-3, 1e-4, "fp32", work,
                     jitter=args.jitter if args.jitter else 0.25)
    case.scheme.accflag = 0                  # the sweep alone: no HDG set-up
    t0 = time.perf_counter()
    s = pkg.Solver(case, max_bundle=args.bundle, block=args.block, sweep_variant=args.variant)
    setup_s = time.perf_counter() - t0
    if args.split and s.ctx.fused:
        s.ctx.set_fused("split")
    if args.no_staging:
        s.ctx.staging(False)
    D, N, P = s.ndir, s.n_tris, s.ndof
    log(f"solver set up in {setup_s:.1f} s ({N} triangles); sweeping")
    wl = (f"config 2: jittered square {n}x{n} (+-0.25h, seed 0) -> {N} P{args.deg} triangles, "
          f"{D} ordinates ({args.npole}x{args.nazim}), Kn_R 1e-3, Kn_N 1e-4, fp32 sweep")
    if not s.ctx.fused:
        wl += ", stored"
    out = sweep_line(s, args, D, N, P, "fp32", wl)
    out["setup_s"] = setup_s
    if s.ctx.fused and not args.no_stored:
        log("the same sweep with stored binary32 inverses")
        out["sweep_stored"] = stored_sweep(s, args, D, N, P, wl)
    del s
    import gc
    gc.collect()
    torch.cuda.empty_cache()
    if not args.no_solve:
        res, _ = gsis_solves(pkg, args, work, n)
        out["solve"] = res
    if not args.no_cpu:
        c64 = gsis_case(pkg, n, args.deg, 4, 4, 1e-3, 1e-4, "fp64", work,
                        jitter=args.jitter if args.jitter else 0.25)
        out["cpu_baseline"] = cpu_baseline(pkg, c64, args.cpu_seconds,
                                           sample=f"the {n}x{n} mesh with 16 ordinates (4x4) of the 512")
    print(json.dumps(out), flush=True)


def timed_sweeps(s, steps, warmup):
    """(wall seconds, per-launch kernel ms) of `steps` sweeps after `warmup`."""
    import torch
    vdf = s.sweep_vdf
    for _ in range(warmup):
        s.ctx.sweep(s.mom, vdf)
    s.ctx.timing(steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.ctx.sweep(s.mom, vdf)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = s.ctx.timing_read(steps)
    s.ctx.timing(0)
    return wall, ms


def roofline(s, D, N, P, precision, t_sweep, workload):
    """The sweep's roofline object.  `achieved` prices the workload at the
    survey's algorithmic bytes of the precomputed-LU sweep (B_alg) over the
    measured kernel time; a fused binary32 sweep (element operators built
    and factorised in registers) reads none of B_alg's factor bytes, so its
    own bytes and its VALU rate are given beside it."""
    s_lu = 4 if precision == "fp32" else 8
    bytes_alg = b_alg(D, N, P, s_lu, s_lu)
    out = {"bound": "hbm", "achieved": bytes_alg / t_sweep / 1e9, "peak": HBM_PEAK / 1e9,
           "unit": "GB/s", "frac": bytes_alg / t_sweep / HBM_PEAK,
           "traffic": committed_traffic(workload), "bytes_alg_per_launch": bytes_alg}
    if precision == "fp32" and s.ctx.fused:
        own = b_alg(D, N, P, 0, s_lu)
        flops = D * N * fused_flops({3: 1, 6: 2, 10: 3, 15: 4}[P])
        out.update({
            "basis": "B_alg of the precomputed-LU sweep of this workload over the fused sweep's time",
            "kernel_bytes_alg_per_launch": own, "kernel_bytes_frac": own / t_sweep / HBM_PEAK,
            "valu": {"flops_per_launch": flops, "achieved": flops / t_sweep / 1e12,
                     "peak": VALU_F32_PEAK / 1e12, "unit": "TFLOP/s", "frac": flops / t_sweep / VALU_F32_PEAK}})
    return out


def fused_flops(deg):
    """Floating-point operations of one fused binary32 element solve (item):
    the right-hand side (source, three inflow faces), A from the nonzero
    reference-element entries and the three outflow faces, elimination
    without pivoting and back substitution; an FMA counts two."""
    from pbte_amd.basis import build_basis
    from pbte_amd.integration import reference_element
    basis = build_basis(deg)
    P, Q = basis.ndof_tri, basis.ndof_fc
    _, m, dx, dy, _, _, _ = reference_element(basis)
    nz = sum(int(np.count_nonzero(np.abs(a) > 1e-14 * np.abs(a).max())) for a in (m, dx, dy))
    face = 5 + 2 * Q * Q + 2 * Q
    rhs = 5 * P + 3 * face
    build = 6 + 2 * nz + 3 * (5 + 2 * Q * Q)
    lu = sum(1 + (P - k - 1) * (3 + 2 * (P - k - 1)) for k in range(P))
    back = sum(2 * (P - 1 - i) + 1 for i in range(P))
    return rhs + build + lu + back


def sweep_line(s, args, D, N, P, precision, workload):
    """Time args.steps sweeps of a set-up solver (after args.warmup) and
    build the JSON line's sweep fields."""
    rng = np.random.default_rng(0)
    s.mom.set_dofs(rng.random((P, N)), 0.1 * rng.random((P, N)), 0.1 * rng.random((P, N)))
    wall, ms = timed_sweeps(s, args.steps, args.warmup)
    t_sweep = float(np.mean(ms)) * 1e-3
    dof = D * N * P
    return {
        "metric": METRIC,
        "value": dof * args.steps / wall,
        "unit": "DOF-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if precision == "fp32" else "f64",
        "data": "synthetic moments on the workload's mesh and ordinates",
        "config": {"workload": workload, "n_tris": N, "ndof": P, "ndir": D, "dof_per_sweep": dof},
        "roofline": roofline(s, D, N, P, precision, t_sweep, workload),
        "sweep_kernel_ms_mean": t_sweep * 1e3,
        "sweep_kernel_ms_min": float(np.min(ms)),
        "schedule": s.ctx.schedule_info(),
        "tile_shape": s.tile_shape,
        "factor_mode": s.ctx.mode,
        "fp32_operator": "fused" if precision == "fp32" and s.ctx.fused else "stored",
        "sweep_path": s.ctx.sweep_path(precision),
        "staging": s.ctx.staging(),
    }


def stored_sweep(s, args, D, N, P, workload):
    """The same binary32 sweep with stored binary32 inverses (formed in
    binary64, read by the kernel): its time and HBM roofline beside the
    fused sweep's."""
    s.ctx.set_fused(False)
    s.ctx.factorise("fp32")
    wall, ms = timed_sweeps(s, args.steps, args.warmup)
    t = float(np.mean(ms)) * 1e-3
    bytes_alg = b_alg(D, N, P, 4, 4)
    return {"sweep_kernel_ms_mean": t * 1e3, "dof_updates_per_s": D * N * P / t,
            "roofline": {"bound": "hbm", "achieved": bytes_alg / t / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": bytes_alg / t / HBM_PEAK, "traffic": committed_traffic(workload + ", stored"),
                         "bytes_alg_per_launch": bytes_alg}}


def cis_workload(pkg, args, work):
    """FP64 CIS to the iterate step 1e-8 on the GPU and with the C
    restatement of the reference on the host (same case, same schedule of
    operations), both timed from the built problem to the converged moments."""
    import torch
    from oracle import cpu
    from pbte_amd.problem import build_problem
    n, npole, nazim = 16, 4, 16
    case = make_case(pkg, n, 1, npole, nazim, 1.0, 1.0, "fp64", work)
    case.scheme.method = "source"
    t0 = time.perf_counter()
    s = pkg.Solver(case)
    setup = time.perf_counter() - t0
    s.ctx.timing(case.iteration.tmax)
    t1 = time.perf_counter()
    rec = s.run()
    torch.cuda.synchronize()
    run = time.perf_counter() - t1
    ms = s.ctx.timing_read(case.iteration.tmax)
    s.ctx.timing(0)
    cpu.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t2 = time.perf_counter()
    o = cpu.Oracle(build_problem(case))
    o.factorise()
    res, mass = o.run_cis(tol=case.iteration.tol, tmax=case.iteration.tmax)
    cpu_s = time.perf_counter() - t2
    D, N, P = s.ndir, s.n_tris, s.ndof
    t_sweep = float(np.mean(ms)) * 1e-3
    bytes_alg = b_alg(D, N, P, 8, 8)
    out = {
        "metric": METRIC, "value": 3600.0 / (setup + run), "unit": "solves/hour",
        "n_gpus": 1, "steps": rec.iterations, "warmup": 0,
        "ms_per_step": 1e3 * run / max(rec.iterations, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: the workload's mesh, ordinates and boundary data",
        "config": {"workload": f"cis square {n}x{n} -> {N} P1 triangles, {D} ordinates "
                               f"({npole}x{nazim}), Kn_R 1, Kn_N 1", "n_tris": N, "ndof": P, "ndir": D},
        "roofline": {"bound": "hbm", "achieved": bytes_alg / t_sweep / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": bytes_alg / t_sweep / HBM_PEAK, "traffic": None,
                     "bytes_alg_per_launch": bytes_alg, "note": "latency bound: 2 MB per sweep"},
        "cis": {"setup_s": setup, "run_s": run, "iterations": rec.iterations,
                "converged": rec.converged, "mass": rec.mass, "sweep_kernel_ms_mean": t_sweep * 1e3,
                "cpu_iterations": len(res), "cpu_mass": float(mass[-1]),
                "mass_abs_diff_vs_cpu": abs(rec.mass - float(mass[-1]))},
        "cpu_baseline": {"value": 3600.0 / cpu_s, "unit": "solves/hour", "cores": threads,
                         "kind": "port", "sample": f"one FP64 CIS solve of the same case "
                                                   f"({cpu_s:.2f} s incl. factorisation), "
                                                   f"OMP_NUM_THREADS={threads}"},
    }
    print(json.dumps(out), flush=True)


BATCH_CELLS = (2048, 8192, 32768, 65536)
BATCH_DIRS = {64: (4, 16), 128: (8, 16), 256: (8, 32), 512: (16, 32)}
BATCH_FACTOR_BYTES = 2.0e9      # per-problem cap on stored element factors in the batch solve leg
BATCH_SETUP_THREADS = 4         # problems of the batched solves set up at once (host threads)


def batch_draws(count=64, seed=1):
    """The batched workload's problems, drawn from default_rng(seed) in this
    order: cell counts from BATCH_CELLS, element degrees from {1, 2},
    ordinate counts from {64, 128, 256, 512} (polar x azimuth as in
    BATCH_DIRS), log10 Kn_R uniform on [-3, 1], Kn_N / Kn_R uniform on
    [0.01, 1].  A square of n x n cells split in two triangles each, with
    n = round(sqrt(cells / 2)) (65 536 cells: 181 x 181 = 65 522)."""
    rng = np.random.default_rng(seed)
    cells = rng.choice(BATCH_CELLS, count)
    degs = rng.choice([1, 2], count)
    dirs = rng.choice(sorted(BATCH_DIRS), count)
    kn_r = 10.0 ** rng.uniform(-3.0, 1.0, count)
    kn_n = kn_r * rng.uniform(0.01, 1.0, count)
    return [{"n": int(round(np.sqrt(c / 2))), "deg": int(g), "ndir": int(d), "tau_r": float(a),
             "tau_n": float(b)} for c, g, d, a, b in zip(cells, degs, dirs, kn_r, kn_n)]


def batch_workload(pkg, args, work):
    """The batched configuration: 64 independent problems (batch_draws),
    jittered meshes (interior vertices moved by U(-0.25h, 0.25h), seed 1),
    diffuse isothermal walls with the north one at dT = +1.

    Sweep leg, at the drawn sizes: one step is one FP32 sweep of every
    problem; the problems that run one kernel instance (element degree,
    pipeline depth, one or two directions per workgroup: batch.launch_key)
    go in one launch of the pipelined kernel (pbte_batch_sweep), plus each
    problem's source kernel.  value = total DOF-updates per step /
    wall time per step; the same sweeps launched problem by problem are
    timed beside it.

    Solve leg (--batch-solve-div k, default 1: the drawn sizes; 0 skips
    it): the same 64 draws (n // k cells a side, at least 4), GSIS with
    FP32 sweeps inside the binary64 defect correction to the iterate step
    1e-8, all 64 in lockstep (batch_solves); the aggregate solves/hour and
    each problem's iterations and time to tolerance."""
    import torch
    from pbte_amd.batch import SweepBatch, launch_key
    draws = batch_draws(args.batch_count, 1)
    log(f"batch: building {len(draws)} problems (sweep leg)")
    t0 = time.perf_counter()
    solvers = []
    rng = np.random.default_rng(0)
    for k, d in enumerate(draws):
        npole, nazim = BATCH_DIRS[d["ndir"]]
        case = make_case(pkg, d["n"], d["deg"], npole, nazim, d["tau_r"], d["tau_n"], "fp32", work,
                         jitter=0.25, seed=1, method="source")
        case.performance.autotune = args.autotune
        # binary32 sweeps only: fused, no element factors stored
        s = pkg.Solver(case, max_bundle=args.bundle, max_bytes=0.0 if FP32_OPERATOR == "fused" else None)
        P, N = s.ndof, s.n_tris
        s.mom.set_dofs(rng.random((P, N)), 0.1 * rng.random((P, N)), 0.1 * rng.random((P, N)))
        solvers.append(s)
        if k % 8 == 7:
            log(f"batch: {k + 1} problems set up ({time.perf_counter() - t0:.1f} s)")
    setup_s = time.perf_counter() - t0
    groups = {}
    for k, s in enumerate(solvers):
        groups.setdefault(launch_key(s.ctx), []).append(k)
    batches = [(SweepBatch([solvers[k].ctx for k in g]), g) for g in groups.values()]

    def step_batched():
        # the groups' launches overlap on the device, joined at the step's end
        for b, g in batches:
            b.sweep([solvers[k].mom for k in g], [solvers[k].sweep_vdf for k in g], join=False)
        for b, _ in batches:
            b.join()

    def step_alone():
        for s in solvers:
            s.ctx.sweep(s.mom, s.sweep_vdf)

    def timed(step):
        for _ in range(args.warmup):
            step()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        ev0.record()
        for _ in range(args.steps):
            step()
        ev1.record()
        torch.cuda.synchronize()
        return time.perf_counter() - t, ev0.elapsed_time(ev1) * 1e-3

    log("batch: batched sweeps")
    wall_b, dev_b = timed(step_batched)
    log("batch: the same sweeps problem by problem")
    wall_a, dev_a = timed(step_alone)
    for b, _ in batches:
        b.close()
    dof = sum(s.ndir * s.n_tris * s.ndof for s in solvers)
    bytes_alg = sum(b_alg(s.ndir, s.n_tris, s.ndof, 4, 4) for s in solvers)
    t_step = dev_b / args.steps
    cpu_case = None
    if not args.no_cpu:
        # the median problem by DOF count
        small = sorted(range(len(draws)), key=lambda k: solvers[k].ndir * solvers[k].n_tris * solvers[k].ndof)
        small = small[len(small) // 2]
        d = draws[small]
        npole, nazim = BATCH_DIRS[d["ndir"]]
        cpu_case = (small, make_case(pkg, d["n"], d["deg"], npole, nazim, d["tau_r"], d["tau_n"], "fp64",
                                     work, jitter=0.25, seed=1, method="source"))
    shapes = [{"n_tris": s.n_tris, "ndof": s.ndof, "ndir": s.ndir, "tau_r": d["tau_r"], "tau_n": d["tau_n"],
               "tile": None if s.tile_shape is None else
               [s.tile_shape[k] for k in ("max_bundle", "block", "variant")]}
              for s, d in zip(solvers, draws)]
    del solvers, batches
    torch.cuda.empty_cache()
    out = {
        "metric": METRIC, "value": dof * args.steps / wall_b, "unit": "DOF-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall_b / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic moments on the drawn problems' meshes and ordinates",
        "config": {"workload": f"batch of {len(draws)} jittered squares (seed 1), "
                               "2k-65k P1/P2 triangles, 64-512 ordinates, fp32 sweep",
                   "problems": len(draws), "dof_per_step": dof, "launches_per_step": len(groups),
                   "autotune": bool(args.autotune)},
        "roofline": {"bound": "hbm", "achieved": bytes_alg / t_step / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": bytes_alg / t_step / HBM_PEAK, "traffic": None,
                     "bytes_alg_per_launch": bytes_alg / len(groups),
                     "note": "one step's launches together (sweeps and source kernels), HIP events"},
        "batch": {"setup_s": setup_s, "step_ms_batched": 1e3 * t_step,
                  "step_ms_one_by_one": 1e3 * dev_a / args.steps,
                  "speedup_batched_vs_one_by_one": dev_a / dev_b,
                  "wall_ms_batched": 1e3 * wall_b / args.steps,
                  "wall_ms_one_by_one": 1e3 * wall_a / args.steps, "problems": shapes},
    }
    if args.batch_solve_div > 0:
        out["batch"]["solve"] = batch_solves(pkg, args, work, draws)
    if cpu_case is not None:
        k, case = cpu_case
        out["cpu_baseline"] = cpu_baseline(pkg, case, args.cpu_seconds, max_sweeps=1000)
        out["cpu_baseline"]["sample"] += f" (problem {k}, the median of the batch by DOF count)"
    print(json.dumps(out), flush=True)


def batch_solves(pkg, args, work, draws):
    """GSIS of every drawn problem (meshes args.batch_solve_div times
    coarser a side; 1: the drawn sizes), fp32 as the binary64 defect
    correction, in lockstep (batch.run_refined_batch: each problem's sweeps
    merged into the persistent work-queue launches); problems whose sweep
    ordering breaks cycles run the fp32 switch scheme in a second lockstep
    group (batch.run_source_batch).  Problems are admitted in draw order
    while the device memory in use stays below --batch-mem-frac of the
    total (the solve's own buffers, binary32 trace-factor copies among
    them, are allocated as it runs); each admitted group is solved, freed,
    and the next admitted.  Binary64 factors are stored for problems whose
    factors take at most BATCH_FACTOR_BYTES, binary32 ones (the refined
    binary64 sweep) below twice that, none above (the binary64 sweeps
    build their operators); binary32 sweeps are fused."""
    import gc

    import torch
    from pbte_amd.batch import build_solvers, run_refined_batch, run_source_batch
    from pbte_amd.refine import applicable
    div = args.batch_solve_div
    log(f"batch: {len(draws)} GSIS problems (meshes / {div} a side)")
    total = torch.cuda.mem_get_info()[1]
    per = [None] * len(draws)
    groups = []
    setup_all = run_all = 0.0
    k = 0
    # the case files (mesh written in gmsh format, YAML) are the solves'
    # input, written before the timed set-up as the config-2 solves' are;
    # each set-up reads its mesh file
    t0 = time.perf_counter()
    cases = []
    for d in draws:
        npole, nazim = BATCH_DIRS[d["ndir"]]
        case = make_case(pkg, max(4, d["n"] // div), d["deg"], npole, nazim, d["tau_r"], d["tau_n"],
                         "fp32", work, jitter=0.25, seed=1, method="source", accflag=1, tmax=args.batch_tmax)
        case.performance.autotune = args.autotune
        cases.append(case)
    case_files_s = time.perf_counter() - t0

    def progress(j, step, res, s):
        if step % 200 == 0 and j == 0:
            log(f"batch: first problem of the group at step {step}")

    while k < len(draws):
        t0 = time.perf_counter()
        solvers, idx = [], []
        while k < len(draws):
            # BATCH_SETUP_THREADS set-ups at a time; admission is checked
            # after each such chunk
            chunk = list(range(k, min(k + BATCH_SETUP_THREADS, len(draws))))
            solvers += build_solvers([cases[j] for j in chunk], workers=BATCH_SETUP_THREADS,
                                     max_bytes=BATCH_FACTOR_BYTES)
            idx += chunk
            k = chunk[-1] + 1
            free, _ = torch.cuda.mem_get_info()
            if total - free > args.batch_mem_frac * total:
                break
        setup = time.perf_counter() - t0
        used = (total - torch.cuda.mem_get_info()[0]) / 2**30
        refined = [j for j, s in enumerate(solvers) if applicable(s)[0]]
        switch = [j for j in range(len(solvers)) if j not in refined]
        log(f"batch: problems {idx[0]}-{idx[-1]} set up in {setup:.1f} s ({used:.1f} GiB in use); solving "
            f"({len(refined)} refined, {len(switch)} with the switch scheme)")
        t1 = time.perf_counter()
        recs = [None] * len(solvers)
        if refined:
            for j, r in zip(refined, run_refined_batch([solvers[j] for j in refined], callback=progress)):
                recs[j] = r
        if switch:
            for j, r in zip(switch, run_source_batch([solvers[j] for j in switch])):
                recs[j] = r
        torch.cuda.synchronize()
        run = time.perf_counter() - t1
        log(f"batch: group solved in {run:.1f} s")
        for j, (s, r) in enumerate(zip(solvers, recs)):
            per[idx[j]] = {"n_tris": s.n_tris, "ndof": s.ndof, "ndir": s.ndir, "tau_r": s.case.flow.tau_r,
                           "iterations": r.iterations, "converged": r.converged,
                           "time_to_tolerance_s": r.wall_clock,
                           "scheme": r.diagnostics.get("precision_scheme", "switch"),
                           "factor_mode": s.ctx.mode, "steps_fp32": r.diagnostics.get("steps_fp32"),
                           "group": len(groups)}
        groups.append({"problems": len(solvers), "setup_s": setup, "run_s": run, "device_gib": used})
        setup_all += setup
        run_all += run
        del solvers, recs
        gc.collect()
        torch.cuda.empty_cache()
    return {"mesh_div": div, "case_files_s": case_files_s, "setup_s": setup_all, "run_s": run_all,
            "solves_per_hour": 3600.0 * len(draws) / (setup_all + run_all),
            "converged": int(sum(p["converged"] for p in per)),
            "refined": int(sum(p["scheme"] == "defect-correction" for p in per)),
            "groups": groups, "problems": per}


def b_alg(D, N, P, s_lu, s_f):
    """Algorithmic HBM bytes of one precomputed-LU sweep: per (cell, dir) the
    LU (P^2 s_lu), pivots (P bytes), the distribution write and one upwind
    read (2 P s_f); per cell the direction-independent mass, face, source
    and moment data (8 (4 P^2 + 3 P + 6) bytes)."""
    return D * N * (P * P * s_lu + P + 2 * P * s_f) + N * 8 * (4 * P * P + 3 * P + 6)


def cpu_baseline(pkg, case, seconds=15.0, max_sweeps=1000, sample="the same workload"):
    """The C restatement of the reference sweep (binary64, OpenMP over
    directions, cells in the reference order, precomputed LU) on this host:
    about `seconds` of sweeps at the host's thread count (OMP_NUM_THREADS or
    the CPU count), then as long at one thread."""
    from oracle import cpu
    from pbte_amd.problem import build_problem
    cpu.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = cpu.Oracle(build_problem(case))
    o.factorise()
    rng = np.random.default_rng(0)
    o.ts[:] = rng.random(o.ts.shape)
    dof = o.D * o.N * o.P

    def timed(nthreads):
        cpu.lib().orc_set_threads(nthreads)
        o.sweep()                                   # warm
        n, t0 = 0, time.perf_counter()
        while True:
            o.sweep()
            n += 1
            el = time.perf_counter() - t0
            if el > seconds or n >= max_sweeps:
                break
        return dof * n / el, n, el

    v, n, el = timed(threads)
    v1, n1, el1 = timed(1)
    cpu.lib().orc_set_threads(threads)
    return {"value": v, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} FP64 sweeps ({el:.1f} s) of {sample} after factorisation, "
                      f"{threads} OpenMP threads; at 1 thread {n1} sweeps ({el1:.1f} s)",
            "value_1_thread": v1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=0, help="mesh size a side (default 256 for config2 and gsis, "
                                                      "128 for sweep)")
    ap.add_argument("--workload", default="config2", choices=["config2", "sweep", "cis", "gsis", "batch", "trace"])
    ap.add_argument("--batch-count", type=int, default=64, help="problems of the batch workload")
    ap.add_argument("--batch-solve-div", type=int, default=1,
                    help="batch workload: solve leg on meshes this many times coarser a side (0 skips)")
    ap.add_argument("--batch-tmax", type=int, default=5000, help="batch workload: GSIS iteration cap")
    ap.add_argument("--batch-mem-frac", type=float, default=0.55,
                    help="batch solve leg: admit problems while this fraction of device memory is in use")
    ap.add_argument("--deg", type=int, default=2)
    ap.add_argument("--npole", type=int, default=0, help="polar ordinates (default 16 for config2 and gsis, 8)")
    ap.add_argument("--nazim", type=int, default=32)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--bundle", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0,
                    help="sweep kernel: 0 default, 1 plain level sweep, 2 / 3 pipelined with the "
                         "gather 1 / 2 levels ahead")
    ap.add_argument("--jitter", type=float, default=0.0)
    ap.add_argument("--hole", type=float, default=0.0,
                    help="radius of a central circular hole (hot wall; the outer walls cold)")
    ap.add_argument("--tau-r", type=float, default=1.0, help="Kn_R of the sweep workload (Kn_N = 0.1 Kn_R)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the sweep's tile shapes by timing them (performance.autotune)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stored", action="store_true",
                    help="config2: skip timing the stored-inverse binary32 sweep beside the fused one")
    ap.add_argument("--fp32-operator", default="fused", choices=["fused", "stored"],
                    help="binary32 sweeps: element operators fused into the sweep or stored inverses")
    ap.add_argument("--no-staging", action="store_true",
                    help="read the previous-sweep and periodic-buffer upwind values from global memory "
                         "instead of staging them in LDS")
    ap.add_argument("--split", action="store_true",
                    help="fused binary32 sweeps in the split kernel (producer and consumer halves)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-solve", action="store_true", help="skip the Krylov solve leg")
    ap.add_argument("--solve-kn", default="0.01,0.1,1,10",
                    help="comma-separated Kn_R values of the Krylov solve leg")
    ap.add_argument("--solve-precisions", default="fp64,fp32",
                    help="precisions of the Krylov solve leg, in order")
    ap.add_argument("--restart", type=int, default=30, help="GMRES restart length of the solve leg")
    ap.add_argument("--precond", action="store_true",
                    help="solve leg: GMRES preconditioned by the GSIS moment solve (accflag 1, krylov_precond)")
    ap.add_argument("--solve-tmax", type=int, default=20000, help="product cap of the solve leg")
    ap.add_argument("--solve-max-bytes", type=float, default=0.0,
                    help="factor byte cap of the solve leg (0: half the device memory)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy, max(args.steps, 1))
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    from __graft_entry__ import load_package
    pkg = load_package()
    import torch
    global FP32_OPERATOR
    FP32_OPERATOR = args.fp32_operator

    work = Path(tempfile.mkdtemp(prefix="pbte_bench_"))
    big = args.workload in ("config2", "gsis", "trace")
    args.n = args.n or (256 if big else 128)
    args.npole = args.npole or (16 if big else 8)
    if args.workload == "config2":
        config2_workload(pkg, args, work)
        return
    if args.workload == "cis":
        cis_workload(pkg, args, work)
        return
    if args.workload == "batch":
        batch_workload(pkg, args, work)
        return
    if args.workload == "gsis":
        gsis_workload(pkg, args, work)
        return
    if args.workload == "trace":
        trace_workload(pkg, args, work)
        return
    log(f"mesh {args.n}x{args.n} (hole {args.hole:g}), P{args.deg}, {args.npole}x{args.nazim} ordinates")
    case = make_case(pkg, args.n, args.deg, args.npole, args.nazim, args.tau_r, 0.1 * args.tau_r,
                     args.precision, work, jitter=args.jitter, hole=args.hole)
    case.performance.autotune = args.autotune
    t0 = time.perf_counter()
    s = pkg.Solver(case, max_bundle=args.bundle, block=args.block, sweep_variant=args.variant)
    setup_s = time.perf_counter() - t0
    if args.split and s.ctx.fused:
        s.ctx.set_fused("split")
    if args.no_staging:
        s.ctx.staging(False)
    log(f"solver set up in {setup_s:.1f} s ({s.n_tris} triangles, factor mode {s.ctx.mode}); sweeping")
    D, N, P = s.ndir, s.n_tris, s.ndof
    vdf = s.sweep_vdf
    rng = np.random.default_rng(0)
    s.mom.set_dofs(rng.random((P, N)), 0.1 * rng.random((P, N)), 0.1 * rng.random((P, N)))
    for _ in range(args.warmup):
        s.ctx.sweep(s.mom, vdf)
    s.ctx.timing(args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.ctx.sweep(s.mom, vdf)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = s.ctx.timing_read(args.steps)
    s.ctx.timing(0)
    t_sweep = float(np.mean(ms)) * 1e-3
    dof = D * N * P
    s_lu = 4 if args.precision == "fp32" else 8
    bytes_alg = b_alg(D, N, P, s_lu, s_lu)
    hole = f" with a central hole r={args.hole:g}" if args.hole > 0 else ""
    workload = (f"square {args.n}x{args.n}{hole} -> {N} P{args.deg} triangles, "
                f"{D} ordinates ({args.npole}x{args.nazim}), {args.precision} sweep")
    if args.precision == "fp32" and not s.ctx.fused:
        workload += ", stored"
    out = {
        "metric": METRIC,
        "value": dof * args.steps / wall,
        "unit": "DOF-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic moments on the workload's mesh and ordinates",
        "config": {"workload": workload, "n_tris": N, "ndof": P, "ndir": D, "dof_per_sweep": dof,
                   "jitter": args.jitter},
        "roofline": roofline(s, D, N, P, args.precision, t_sweep, workload),
        "sweep_kernel_ms_mean": t_sweep * 1e3,
        "sweep_kernel_ms_min": float(np.min(ms)),
        "schedule": s.ctx.schedule_info(),
        "tile_shape": s.tile_shape,
        "setup_s": setup_s,
        "factor_mode": s.ctx.mode,
        "fp32_operator": "fused" if args.precision == "fp32" and s.ctx.fused else "stored",
        "sweep_path": s.ctx.sweep_path(args.precision),
        "staging": s.ctx.staging(),
    }
    if s.ctx.mode == "mixed":
        # binary64 sweeps of the refinement build their operators on the fly
        s.ctx.timing(2)
        for _ in range(2):
            s.ctx.sweep(s.mom, s.vdf_d)
        torch.cuda.synchronize()
        out["fp64_onthefly_sweep_ms"] = float(np.mean(s.ctx.timing_read(2)))
        s.ctx.timing(0)
    del s, vdf
    import gc
    gc.collect()
    torch.cuda.empty_cache()
    if not args.no_solve:
        out["solve"] = solve_times(pkg, args, work)
    if not args.no_cpu:
        c64 = make_case(pkg, args.n, args.deg, args.npole, args.nazim, args.tau_r, 0.1 * args.tau_r, "fp64",
                        work, jitter=args.jitter, hole=args.hole)
        out["cpu_baseline"] = cpu_baseline(pkg, c64, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def calibrate_copy(nbytes, reps):
    """A plain device-to-device copy of nbytes (reads nbytes, writes nbytes):
    the reference point that scales rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = nbytes // 4
    src = torch.rand(n, device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"copy_bytes": 4 * n, "reps": reps, "GB_per_s": 2 * 4 * n * reps / el / 1e9}),
          flush=True)


def solve_times(pkg, args, work):
    """Krylov solves of the workload at each Kn_R of --solve-kn (Kn_N =
    0.1 Kn_R), GMRES(--restart), in each precision of --solve-precisions;
    each solve's result is also logged as it finishes."""
    out = {}
    precs = [p.strip() for p in args.solve_precisions.split(",") if p.strip()]
    max_bytes = args.solve_max_bytes if args.solve_max_bytes > 0 else None
    for tau_r in (float(v) for v in args.solve_kn.split(",")):
        res = {}
        for prec in precs:
            case = make_case(pkg, args.n, args.deg, args.npole, args.nazim, tau_r, 0.1 * tau_r, prec,
                             work, jitter=args.jitter, hole=args.hole, tmax=args.solve_tmax,
                             accflag=1 if args.precond else 0)
            case.scheme.krylov_restart = args.restart
            case.scheme.krylov_precond = bool(args.precond)

            def progress(step, r, solver, prec=prec, tau_r=tau_r):
                if step % 200 == 0:
                    log(f"solve Kn_R {tau_r:g} {prec}: step {step}, residual {r:.3e}")

            t0 = time.perf_counter()
            s = pkg.Solver(case, max_bytes=max_bytes)
            t1 = time.perf_counter()
            rec = s.run(callback=progress)
            t2 = time.perf_counter()
            tts = t2 - t0
            d = rec.diagnostics
            res[prec] = {"time_to_solution_s": tts, "setup_s": s.setup_seconds, "run_s": t2 - t1,
                         "iterations": rec.iterations, "converged": rec.converged,
                         "sweeps_fp64": d.get("sweeps_fp64"), "sweeps_fp32": d.get("sweeps_fp32"),
                         "factor_mode": s.ctx.mode, "restart": args.restart,
                         "preconditioner": "GSIS moment solve" if args.precond else None,
                         "solves_per_hour": 3600.0 / tts, "temp": rec.temp,
                         "q": np.concatenate([rec.qx, rec.qy])}
            log(f"solve Kn_R {tau_r:g} {prec}: " + json.dumps(
                {k: v for k, v in res[prec].items() if k not in ("temp", "q")}))
            del s, rec
            import gc
            import torch
            gc.collect()
            torch.cuda.empty_cache()
        fields = {p: (res[p].pop("temp"), res[p].pop("q")) for p in precs}
        if "fp64" in res and "fp32" in res:
            (t64, q64), (t32, q32) = fields["fp64"], fields["fp32"]
            res["fp32"]["rel_l2_T_vs_fp64"] = float(np.linalg.norm(t32 - t64) / np.linalg.norm(t64))
            res["fp32"]["rel_l2_q_vs_fp64"] = float(np.linalg.norm(q32 - q64) / np.linalg.norm(q64))
            res["speedup_fp32_refine_vs_fp64"] = (res["fp64"]["time_to_solution_s"] /
                                                  res["fp32"]["time_to_solution_s"])
            res["speedup_run_only"] = res["fp64"]["run_s"] / res["fp32"]["run_s"]
        out[f"kn_r={tau_r:g}"] = res
    return out


if __name__ == "__main__":
    main()
