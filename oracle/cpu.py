"""ctypes wrapper of the C oracle (``pbte_oracle.c``).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``, always as the checker or as the
timed CPU restatement -- never by the product path.

``build()`` compiles ``libpbte_oracle.so`` next to this file with plain gcc
(OpenMP over directions, no FMA contraction, no fast-math).

The reference (pure Python + Numba) has no compiled sources, so there is no
``oracle/_ref`` build: the restatement is pinned by the reference's golden
vectors and frozen regression numbers instead (see the test suite).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "pbte_oracle.c"
LIB = HERE / "libpbte_oracle.so"

_f = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)


def build(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
           "-std=c11", str(SRC), "-o", str(tmp), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


class OrcProblem(C.Structure):
    _fields_ = [("ndir", C.c_int32), ("n_tris", C.c_int32), ("nd", C.c_int32),
                ("n_faces", C.c_int32),
                ("cxv", _f), ("cyv", _f), ("domega", _f),
                ("cv", C.c_double), ("vg", C.c_double), ("tau_r", C.c_double),
                ("tau_n", C.c_double), ("tau_c", C.c_double), ("pi", C.c_double),
                ("mass", _f), ("gx", _f), ("gy", _f), ("fcm", _f), ("ttfc", _f),
                ("nx", _f), ("ny", _f), ("tri_faces", _i), ("face_bc", _i), ("neighbour", _i),
                ("periodic_tri", _i), ("bc_type", _i), ("bc_temp", _f), ("flux_wall", _f),
                ("int_tri", _f), ("face_tri", _i), ("face_lfc", _i)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = C.CDLL(str(LIB))
        h.orc_lu_factor.restype = C.c_int
        h.orc_lu_factor.argtypes = [C.c_int, _f, _i]
        h.orc_lu_solve.argtypes = [C.c_int, _f, _i, _f]
        h.orc_build_operators.restype = C.c_int
        h.orc_build_operators.argtypes = [C.POINTER(OrcProblem), _f, _i]
        h.orc_sweep.argtypes = [C.POINTER(OrcProblem), _i, _f, _f, _f, _f, _i, _f, _f, C.c_int]
        h.orc_transport_residual.argtypes = [C.POINTER(OrcProblem), _f, _f, _f, _f, _f, _f]
        h.orc_moments.argtypes = [C.POINTER(OrcProblem), _f, _f, _f, _f, _f, _f, _f]
        h.orc_set_threads.argtypes = [C.c_int]
        h.orc_sum.restype = C.c_double
        h.orc_sum.argtypes = [_f, C.c_int]
        h.orc_residual.restype = C.c_double
        h.orc_residual.argtypes = [_f, _f, C.c_int]
        h.orc_wall_flux.argtypes = [C.POINTER(OrcProblem), _f, _f]
        h.orc_order_scan.restype = C.c_int
        h.orc_order_scan.argtypes = [C.c_int, C.c_int, _i, _f, _f, _f, _f, _i, _i, C.c_int]
        _lib = h
    return _lib


def _pf(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_f)


def _pi(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_i)


def lu_factor(a: np.ndarray):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    piv = np.zeros(a.shape[0], dtype=np.int32)
    info = lib().orc_lu_factor(a.shape[0], _pf(a), _pi(piv))
    return a, piv, info


def lu_solve(a, piv, b):
    b = np.ascontiguousarray(b, dtype=np.float64).copy()
    lib().orc_lu_solve(a.shape[0], _pf(np.ascontiguousarray(a)), _pi(np.ascontiguousarray(piv)), _pf(b))
    return b


def order_scan(mesh, vel, allow_break=False):
    cx, cy, _ = vel.flat()
    nd, nt = cx.size, mesh.n_tris
    order = np.zeros((nd, nt), dtype=np.int32)
    cyc = np.zeros(nd, dtype=np.int32)
    nb = np.ascontiguousarray(mesh.tri_neighbour, dtype=np.int32)
    nx = np.ascontiguousarray(mesh.tri_normal[:, :, 0])
    ny = np.ascontiguousarray(mesh.tri_normal[:, :, 1])
    st = lib().orc_order_scan(nd, nt, _pi(nb), _pf(nx), _pf(ny), _pf(np.ascontiguousarray(cx)),
                              _pf(np.ascontiguousarray(cy)), _pi(order), _pi(cyc),
                              1 if allow_break else 0)
    return st, order, cyc


class Oracle:
    """The restated reference solve on the host, for one case set-up."""

    def __init__(self, problem):
        pb = problem
        c = pb.case
        self.pb = pb
        cx, cy, dw = pb.vel.flat()
        self.keep = []

        def f(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self.keep.append(a)
            return _pf(a)

        def i(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            self.keep.append(a)
            return _pi(a)

        self.D, self.N, self.P, self.F = cx.size, pb.mesh.n_tris, c.ndof_tri, pb.mesh.n_faces
        self.flux_wall = np.zeros((self.F, self.P))
        p = OrcProblem()
        p.ndir, p.n_tris, p.nd, p.n_faces = self.D, self.N, self.P, self.F
        p.cxv, p.cyv, p.domega = f(cx), f(cy), f(dw)
        from math import pi
        p.cv, p.vg, p.tau_r, p.tau_n, p.tau_c, p.pi = (c.flow.cv, c.flow.vg, c.flow.tau_r,
                                                       c.flow.tau_n, c.tau_c, 4.0 * np.arctan(1.0))
        ig = pb.integrals
        p.mass, p.gx, p.gy = f(ig.int_tri_tri), f(ig.int_tri_tri_x), f(ig.int_tri_tri_y)
        p.fcm, p.ttfc = f(ig.int_tri_fc), f(ig.int_tri_tri_fc)
        p.nx, p.ny = f(pb.mesh.tri_normal[:, :, 0]), f(pb.mesh.tri_normal[:, :, 1])
        p.tri_faces, p.face_bc = i(pb.mesh.tri_faces), i(pb.mesh.face_bc)
        p.neighbour, p.periodic_tri = i(pb.mesh.tri_neighbour), i(pb.bcdata.periodic_tri)
        p.bc_type, p.bc_temp = i(pb.bcdata.bc_type), f(pb.bcdata.bc_temp)
        p.flux_wall = _pf(self.flux_wall)
        p.int_tri = f(ig.int_tri)
        p.face_tri, p.face_lfc = i(pb.mesh.face_tri), i(pb.mesh.face_lfc)
        self.p = p
        self.order = np.ascontiguousarray(pb.order.order, dtype=np.int32)
        self.lu = None
        self.vdf = np.zeros((self.D, self.N, self.P))
        self.ts = np.zeros((self.P, self.N))
        self.qxs = np.zeros((self.P, self.N))
        self.qys = np.zeros((self.P, self.N))
        self.temp = np.zeros(self.N)
        self.qx = np.zeros(self.N)
        self.qy = np.zeros(self.N)
        self.pbuf = np.zeros((self.D, self.F, self.P))

    def factorise(self):
        self.lu = np.zeros((self.D, self.N, self.P, self.P))
        self.piv = np.zeros((self.D, self.N, self.P), dtype=np.int32)
        return lib().orc_build_operators(C.byref(self.p), _pf(self.lu), _pi(self.piv))

    def sweep(self, use_pbuf=False):
        if self.lu is None:
            self.factorise()
        lib().orc_sweep(C.byref(self.p), _pi(self.order), _pf(self.ts), _pf(self.qxs),
                        _pf(self.qys), _pf(self.lu), _pi(self.piv), _pf(self.vdf),
                        _pf(self.pbuf), 1 if use_pbuf else 0)

    def moments(self):
        lib().orc_moments(C.byref(self.p), _pf(self.vdf), _pf(self.ts), _pf(self.qxs),
                          _pf(self.qys), _pf(self.temp), _pf(self.qx), _pf(self.qy))

    def wall_flux(self):
        lib().orc_wall_flux(C.byref(self.p), _pf(self.vdf), _pf(self.flux_wall))

    def transport_residual(self):
        num = np.zeros(self.D)
        den = np.zeros(self.D)
        lib().orc_transport_residual(C.byref(self.p), _pf(self.ts), _pf(self.qxs), _pf(self.qys),
                                     _pf(self.vdf), _pf(num), _pf(den))
        return num, den

    def true_residual(self):
        num, den = self.transport_residual()
        return float(np.sqrt(num.sum() / den.sum()))

    def run_cis(self, tol, tmax):
        """Source iteration as ``driver.py:170-215``; returns histories."""
        t_old = np.zeros(self.N)
        res_h, mass_h = [], []
        has_wall = bool(self.pb.bcdata.has_nonthermalising)
        step = 0
        while True:
            step += 1
            if has_wall:
                self.wall_flux()
            self.sweep()
            self.moments()
            res = lib().orc_residual(_pf(self.temp), _pf(t_old), self.N)
            t_old[:] = self.temp
            res_h.append(res)
            mass_h.append(lib().orc_sum(_pf(self.temp), self.N))
            if res < tol or step >= tmax:
                break
        return np.array(res_h), np.array(mass_h)
