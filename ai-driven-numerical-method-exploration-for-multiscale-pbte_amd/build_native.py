"""Build the native library in-tree: ``libpbte_hip.so`` (HIP kernels for
gfx950 plus the host-side scheduler, trace-solve symbolic phase, mesh reader
and face integrals) next to this file.

Each source compiles to an object of its own (in parallel, under
``_build/`` at the repository root), then one link.
``python build_native.py`` or ``build()`` from the package root entry point.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "libpbte_hip.so"
OBJDIR = ROOT / "_build"
SOURCES = [PKG / "csrc" / "pbte_kernels.hip", PKG / "csrc" / "trace_lu.hip", PKG / "csrc" / "hdg_setup.hip",
           PKG / "csrc" / "schedule.cpp", PKG / "csrc" / "host_setup.cpp"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # contraction only within one source expression: identical expressions give
    # identical arithmetic in every kernel variant
    "-ffp-contract=on",
    "-pthread",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}: {cmd[-1]}")


def _stale(src: Path, obj: Path) -> bool:
    if not obj.exists():
        return True
    newest = max(src.stat().st_mtime, (ROOT / "include" / "pbte_hip.h").stat().st_mtime,
                 Path(__file__).stat().st_mtime)
    return obj.stat().st_mtime < newest


def build(verbose: bool = False, force: bool = True) -> Path:
    """Compile every source (``force``) or only those newer than their
    object, then link."""
    OBJDIR.mkdir(exist_ok=True)
    inc = f"-I{ROOT / 'include'}"
    objs = [OBJDIR / (s.name + ".o") for s in SOURCES]
    cmds = [[hipcc(), *HIP_FLAGS, inc, "-c", str(s), "-o", str(o)] for s, o in zip(SOURCES, objs)
            if force or _stale(s, o)]
    with ThreadPoolExecutor(max_workers=max(1, len(cmds))) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-pthread", *map(str, objs), "-o", str(tmp)],
         verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, force="--incremental" not in sys.argv))
