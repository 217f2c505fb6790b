"""Device solve of the HDG trace system ``K u = f`` (binary64).

Replaces the SuperLU factorisation and solve of ``GlobalSolver``
(``solver-python/pybte/acceleration/global_solve.py:120-124`` factorise,
``:190-195`` solve) by the nested-dissection multifrontal solver of
``csrc/trace_lu.hip`` behind ``pbte_trace_*``:

* ``TraceSolver(mesh, bc_type, pattern, blk)`` runs the symbolic phase on
  the host: the dissection tree over the element centroids, the fronts and
  the assembly maps of the block pattern of ``K`` (``3Q x 3Q`` blocks, one
  block row per face);
* ``factor(values)`` assembles and eliminates the fronts on the device;
* ``solve(rhs, out)`` is one forward and one backward pass over the stored
  front rows, stream-ordered, no host round trip.

A zero pivot column of a front raises ``numpy.linalg.LinAlgError``, as
SuperLU's factorisation of a singular matrix raises ``RuntimeError`` in the
reference (the trace matrix is nonsingular for every admissible case).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .device import ptr, torch

__all__ = ["TraceSolver", "block_pattern"]


def block_pattern(K, blk: int):
    """``(ptr, col, values [nnzb, blk, blk])`` of ``K`` in block-row form:
    face-major unknowns, every block that holds a stored entry."""
    B = K.tobsr(blocksize=(blk, blk))
    B.sort_indices()
    return (np.ascontiguousarray(B.indptr, dtype=np.int32),
            np.ascontiguousarray(B.indices, dtype=np.int32),
            np.ascontiguousarray(B.data, dtype=np.float64))


class TraceSolver:
    def __init__(self, mesh, t1, t2, bsr_ptr, bsr_col, blk: int, stream=None, leaf_cells: int = 0):
        self.blk = int(blk)
        self.n = int(mesh.n_faces) * self.blk
        cent = np.ascontiguousarray(mesh.nodes[mesh.tri_nodes].mean(axis=1), dtype=np.float64)
        keep = [np.ascontiguousarray(t1, dtype=np.int32), np.ascontiguousarray(t2, dtype=np.int32),
                np.ascontiguousarray(mesh.tri_faces, dtype=np.int32), cent,
                np.ascontiguousarray(bsr_ptr, dtype=np.int32), np.ascontiguousarray(bsr_col, dtype=np.int32)]
        self.nnzb = int(keep[5].size)
        d = _native.PbteTraceDesc()
        d.n_faces, d.n_cells, d.blk, d.leaf_cells = int(mesh.n_faces), int(mesh.n_tris), self.blk, int(leaf_cells)
        d.face_t1, d.face_t2, d.tri_faces = _native.i32(keep[0]), _native.i32(keep[1]), _native.i32(keep[2])
        d.cell_xy = _native.f64(keep[3])
        d.bsr_ptr, d.bsr_col = _native.i32(keep[4]), _native.i32(keep[5])
        err = C.create_string_buffer(512)
        self._lib = _native.lib()
        self.handle = self._lib.pbte_trace_create(C.byref(d), stream, err, 512)
        if not self.handle:
            raise _native.NativeError(f"trace solver: {err.value.decode()}")
        self.factored = False
        self.solve_count = 0

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self._lib.pbte_trace_destroy(h)
            self.handle = None

    def _check(self, status: int, what: str) -> int:
        if status < 0:
            msg = self._lib.pbte_trace_last_error(self.handle)
            raise _native.NativeError(f"{what}: {msg.decode() if msg else 'unknown error'}")
        return status

    def info(self) -> dict:
        out = (C.c_int64 * 8)()
        self._check(self._lib.pbte_trace_info(self.handle, out), "trace_info")
        keys = ("nodes", "levels", "max_front", "factor_entries", "flops", "workspace_bytes",
                "unknowns", "blocks")
        return {k: int(v) for k, v in zip(keys, out)}

    def levels(self) -> list:
        """Per level of the dissection tree (root first): fronts, widest
        front, fully summed and contribution rows, stored entries of
        [Z V] and W, and the solve's forward and backward work items."""
        keys = ("fronts", "n_max", "rows", "contrib_rows", "zv_entries", "w_entries", "fw_items", "bw_items")
        out = []
        for lv in range(self.info()["levels"]):
            v = (C.c_int64 * 8)()
            self._check(self._lib.pbte_trace_level(self.handle, lv, v), "trace_level")
            out.append({k: int(x) for k, x in zip(keys, v)})
        return out

    def set_pivoting(self, mode: int) -> None:
        """-1 automatic (diagonal-block pivots, partial pivoting as the
        fallback), 0 diagonal-block pivots only, 1 partial pivoting only."""
        self._check(self._lib.pbte_trace_options(self.handle, int(mode)), "trace_options")

    def set_graphs(self, on: bool) -> None:
        """Record each (precision, rhs, x) solve as a HIP graph and replay
        it (the default), or launch every kernel directly."""
        self._check(self._lib.pbte_trace_graphs(self.handle, 1 if on else 0), "trace_graphs")

    def quality(self) -> dict:
        piv, berr = C.c_int32(), C.c_double()
        self._check(self._lib.pbte_trace_quality(self.handle, C.byref(piv), C.byref(berr)), "trace_quality")
        return {"pivoting": "block" if piv.value == 0 else "partial", "backward_error": berr.value}

    def factor(self, values) -> None:
        """Factorise from the block values (host ``[nnzb, blk, blk]`` array
        or a device tensor of that size)."""
        if isinstance(values, np.ndarray):
            values = torch.from_numpy(np.ascontiguousarray(values, dtype=np.float64)).to("cuda")
        if values.numel() != self.nnzb * self.blk * self.blk:
            raise ValueError("block values do not match the pattern")
        st = self._check(self._lib.pbte_trace_factor(self.handle, ptr(values)), "trace_factor")
        if st > 0:
            msg = self._lib.pbte_trace_last_error(self.handle).decode()
            raise np.linalg.LinAlgError(msg)
        self.factored = True

    def solve(self, rhs, out, precision: str = "fp64") -> None:
        """``out = K^-1 rhs`` (device float64 tensors of ``F * 3Q`` values)
        from the binary64 factors, or (``precision="fp32"``) from their
        binary32 copy, sums in binary64."""
        if rhs.numel() != self.n or out.numel() != self.n:
            raise ValueError("trace vectors do not match the system size")
        code = {"fp64": _native.PBTE_FP64, "fp32": _native.PBTE_FP32}[precision]
        self._check(self._lib.pbte_trace_solve(self.handle, code, ptr(rhs), ptr(out)), "trace_solve")
        self.solve_count += 1
