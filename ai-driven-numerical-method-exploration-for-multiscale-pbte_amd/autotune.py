"""Tile-shape autotuner of the pipelined transport sweep.

The sweep kernel has three shape parameters that change its speed and never
its numbers (every shape computes each item with the same arithmetic, bit
for bit):

* directions per workgroup (``max_bundle`` 1 or 2): pairs of directions
  with the same sweep order share a workgroup and each cell-record load, at
  the price of half as many workgroups;
* workgroup width (``block``: 0 sizes it to the widest level, rounded up to
  64 and at most 256; 64 / 128 / 192 / 256 sweep a wider level in chunks of
  that many items, one barrier per chunk);
* pipeline depth (``variant`` 0: the gather two levels ahead of the solve,
  2: one level ahead).

``tune`` builds the sweep context of each candidate on the case's own mesh
and ordinates, times a few sweeps of the run's precision with HIP events on
the launch stream, and returns the fastest.  Results are kept per process
and, when ``performance.autotune_cache`` names a JSON file, across runs,
keyed by the device, the sizes, the precision and a digest of the mesh and
ordinates.  No counterpart in the reference (its sweep runs one direction
per Numba thread, ``_kernels.py:24-27``).
"""
from __future__ import annotations

import hashlib
import json
import os
from pathlib import Path

import numpy as np

from . import _native
from .device import dev_zeros, torch

__all__ = ["CANDIDATES", "signature", "tune"]

# (max_bundle, block, variant)
CANDIDATES = tuple((b, blk, var) for var in (0, 2) for b in (1, 2) for blk in (0, 128))

_memo: dict = {}


def signature(case, problem) -> str:
    """Cache key: device, sizes, precision and a digest of the mesh
    geometry, the boundary layout and the ordinates."""
    mesh, vel = problem.mesh, problem.vel
    h = hashlib.sha256()
    for a in (mesh.nodes, mesh.tri_nodes, mesh.face_bc, *vel.flat()):
        h.update(np.ascontiguousarray(a).tobytes())
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    return (f"{dev}|N{mesh.n_tris}|P{case.ndof_tri}|D{vel.ndir}|{case.performance.precision}|"
            f"{h.hexdigest()[:20]}")


def _load(path: str) -> dict:
    try:
        return json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return {}


def _store(path: str, key: str, entry: dict) -> None:
    table = _load(path)
    table[key] = entry
    tmp = f"{path}.tmp{os.getpid()}"
    Path(tmp).write_text(json.dumps(table, indent=1, sort_keys=True))
    os.replace(tmp, path)


def tune(case, problem, candidates=CANDIDATES, reps: int = 4, max_bytes: float | None = None) -> dict | None:
    """The fastest tile shape of the case's sweep: ``{"max_bundle",
    "block", "variant", "ms", "timed": [...]}``, or None when no candidate
    can be timed (the caller keeps the default schedule).  Candidates that
    do not build, that store no factors of the run's precision, that fall
    back to global-memory bundles, or whose bundles would not pair, are
    skipped."""
    from .moments import Moments
    from .sweep import SweepContext
    key = signature(case, problem)
    cands = tuple(tuple(c) for c in candidates)
    memo_key = (key, cands)
    if memo_key in _memo:
        return _memo[memo_key]
    path = case.performance.autotune_cache
    if path:
        hit = _load(path).get(key)
        if hit is not None and tuple(tuple(c) for c in hit.get("candidates", ())) == cands:
            _memo[memo_key] = hit
            return hit
    pb = problem
    fp32 = case.performance.precision == "fp32"
    dtype = torch.float32 if fp32 else torch.float64
    P, N, D = case.ndof_tri, pb.mesh.n_tris, pb.vel.ndir
    timed = []
    for mb, blk, var in cands:
        if mb > 1 and D < 2:
            continue
        try:
            ctx = SweepContext.build(case, pb.mesh, pb.integrals, pb.vel, pb.order, pb.bcdata,
                                     max_bundle=mb, block=blk, variant=var, max_bytes=max_bytes)
        except _native.NativeError:
            continue
        try:
            info = ctx.schedule_info()
            # the timed sweeps need stored factors of the run's precision
            # (binary32 ones are stored in the "mixed" mode too)
            stored = ctx.mode == "precomputed" or (fp32 and ctx.mode == "mixed")
            if not stored or info["global_bundles"] > 0:
                continue
            if mb > 1 and info["bundles"] == D:
                continue                        # nothing paired: the same as one direction
            mom = Moments.zeros(P, N)
            mom._ctx = ctx
            vdf = dev_zeros((D, N, P), dtype=dtype)
            ctx.sweep(mom, vdf)                 # warm
            ctx.timing(reps)
            for _ in range(reps):
                ctx.sweep(mom, vdf)
            torch.cuda.synchronize()
            ms = float(np.min(ctx.timing_read(reps)))
            ctx.timing(0)
            timed.append({"max_bundle": mb, "block": blk, "variant": var, "ms": ms})
            del vdf, mom
        finally:
            ctx.close()
    if not timed:
        # remembered: every later solver of this signature keeps the
        # default schedule without building the candidates again
        _memo[memo_key] = None
        torch.cuda.empty_cache()
        return None
    best = min(timed, key=lambda r: r["ms"])
    entry = {**best, "timed": timed, "candidates": [list(c) for c in cands]}
    _memo[memo_key] = entry
    if path:
        _store(path, key, entry)
    torch.cuda.empty_cache()
    return entry
