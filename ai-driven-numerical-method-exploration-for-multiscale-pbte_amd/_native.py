"""ctypes binding of ``libpbte_hip.so`` (declared in ``include/pbte_hip.h``).

The library is loaded from this package directory only; if it is missing
the import fails loudly -- there is no CPU fallback for any kernel.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libpbte_hip.so"

PBTE_FP64 = 0
PBTE_FP32 = 1
PBTE_FP64_FP32 = 2          # pbte_factorise: both precisions from one pass

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class PbteDims(C.Structure):
    _fields_ = [("ndir", C.c_int32), ("n_tris", C.c_int32), ("nd", C.c_int32),
                ("nf", C.c_int32), ("n_faces", C.c_int32), ("nbc", C.c_int32)]


class PbteSetup(C.Structure):
    _fields_ = [("cxv", _f64p), ("cyv", _f64p), ("domega", _f64p),
                ("cv", C.c_double), ("vg", C.c_double), ("tau_r", C.c_double),
                ("tau_n", C.c_double), ("tau_c", C.c_double), ("pi", C.c_double),
                ("mass", _f64p), ("gx", _f64p), ("gy", _f64p), ("fcm", _f64p),
                ("ttfc", _f64p), ("nx", _f64p), ("ny", _f64p), ("int_tri", _f64p),
                ("tri_faces", _i32p), ("face_bc", _i32p), ("neighbour", _i32p),
                ("periodic_tri", _i32p), ("other_lfc", _i32p), ("bc_type", _i32p),
                ("bc_temp", _f64p), ("face_tri", _i32p), ("face_lfc", _i32p),
                ("order", _i32p), ("max_bundle", C.c_int32), ("block", C.c_int32),
                ("sweep_variant", C.c_int32), ("lds_slots_cap", C.c_int32),
                ("lds_value_bytes", C.c_int32), ("pairs_or_fail", C.c_int32)]


class PbteGsisData(C.Structure):
    _fields_ = [("cxv", _f64p), ("cyv", _f64p), ("domega", _f64p),
                ("inv_aa_sol", _f64p), ("aa_trace", _f64p), ("ba_sol", _f64p),     # device
                ("tri_fc_fc", _f64p), ("tri_hmin", _f64p), ("tri_faces", _i32p),
                ("t1", _i32p), ("l1", _i32p), ("t2", _i32p), ("l2", _i32p),
                ("flip", C.POINTER(C.c_uint8)), ("n_wall", C.c_int32), ("n_wall_th", C.c_int32),
                ("wall_face", _i32p), ("wall_tri", _i32p), ("wall_lfc", _i32p),
                ("wall_nx", _f64p), ("wall_ny", _f64p),
                ("tc5", C.c_double), ("tau_thr", C.c_double), ("drift", C.c_double),
                ("variant_b", C.c_int32), ("use_hmin", C.c_int32)]


class PbteTraceDesc(C.Structure):
    _fields_ = [("n_faces", C.c_int32), ("n_cells", C.c_int32), ("blk", C.c_int32),
                ("leaf_cells", C.c_int32), ("face_t1", _i32p), ("face_t2", _i32p),
                ("tri_faces", _i32p), ("cell_xy", _f64p), ("bsr_ptr", _i32p), ("bsr_col", _i32p)]


class PbteFaceIntegralsDesc(C.Structure):
    _fields_ = [("n_tris", C.c_int32), ("nd", C.c_int32), ("nf", C.c_int32), ("np_fc", C.c_int32),
                ("nodes", _f64p), ("tri_nodes", _i32p), ("tri_faces", _i32p), ("face_nodes", _i32p),
                ("face_len", _f64p), ("other", _i32p), ("off_x", _f64p), ("off_y", _f64p),
                ("C", _f64p), ("Cf", _f64p), ("mono", _i32p), ("yq", _f64p), ("lpq", _f64p)]


class PbteHdgDesc(C.Structure):
    _fields_ = [("n_tris", C.c_int32), ("nd", C.c_int32), ("nf", C.c_int32), ("n_faces", C.c_int32),
                ("variant_b", C.c_int32), ("cv", C.c_double), ("tau_r", C.c_double),
                ("tau_c", C.c_double), ("st", C.c_double * 3), ("drop_tol", C.c_double),
                ("int_tri_tri", C.c_void_p), ("int_tri_tri_x", C.c_void_p), ("int_tri_tri_y", C.c_void_p),
                ("int_tri_fc", C.c_void_p), ("int_tri_fc_fc", C.c_void_p), ("int_fc_fc", C.c_void_p),
                ("tri_normal", C.c_void_p), ("face_kind", C.c_void_p)]


_SIGS = {
    "pbte_build_order": (C.c_int, [C.c_int32, C.c_int32, _i32p, _f64p, _f64p, _f64p, _f64p,
                                   C.c_int32, _i32p, _i32p]),
    "pbte_ctx_create": (C.c_void_p, [C.POINTER(PbteDims), C.POINTER(PbteSetup), C.c_void_p,
                                     C.c_char_p, C.c_int32]),
    "pbte_ctx_destroy": (None, [C.c_void_p]),
    "pbte_batch_create": (C.c_void_p, [C.POINTER(C.c_void_p), C.c_int32, C.c_char_p, C.c_int32]),
    "pbte_batch_destroy": (None, [C.c_void_p]),
    "pbte_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "pbte_batch_sweep": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_void_p), C.c_int32]),
    "pbte_batch_sweep_async": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                         C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                         C.POINTER(C.c_void_p), C.c_int32]),
    "pbte_batch_join": (C.c_int, [C.c_void_p]),
    "pbte_last_error": (C.c_char_p, [C.c_void_p]),
    "pbte_set_cell_operators": (C.c_int, [C.c_void_p, _f64p, _f64p]),
    "pbte_set_fused": (C.c_int, [C.c_void_p, C.c_int32]),
    "pbte_sweep_path": (C.c_int, [C.c_void_p, C.c_int32]),
    "pbte_staging": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "pbte_schedule_info": (C.c_int, [C.c_void_p, _i64p]),
    "pbte_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "pbte_timing_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int32]),
    "pbte_factorise": (C.c_int, [C.c_void_p, C.c_int32]),
    "pbte_sweep": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]),
    "pbte_sweep_onthefly": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]),
    "pbte_moments": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbte_wall_flux": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "pbte_transport_residual": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, _f64p]),
    "pbte_residual_iterate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _f64p, _f64p]),
    "pbte_residual_iterate_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbte_periodic_gather": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "pbte_periodic_scatter": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "pbte_vec_multidot": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p,
                                    C.c_int64, _f64p]),
    "pbte_vec_update": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, _f64p,
                                  C.c_void_p, C.c_int64]),
    "pbte_vec_cgs2": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p,
                                C.c_int64, _f64p]),
    "pbte_vec_cgs2_begin": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p,
                                      C.c_int64, C.c_void_p]),
    "pbte_vec_cgs2_end": (C.c_int, [C.c_void_p, _f64p]),
    "pbte_vec_defect": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_int64]),
    "pbte_vec_relax": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_int64,
                                 C.POINTER(C.c_double)]),
    "pbte_vec_axpby": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_double, C.c_void_p,
                                 C.c_int64]),
    "pbte_gsis_setup": (C.c_int, [C.c_void_p, C.POINTER(PbteGsisData)]),
    "pbte_gsis_moments": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "pbte_gsis_source": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "pbte_gsis_rhs": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "pbte_gsis_local": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbte_gsis_uq": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbte_gsis_precond_setup": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbte_gsis_precond_src": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    "pbte_gsis_correct": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbte_gsis_correct_moments": (C.c_int, [C.c_void_p] + [C.c_void_p] * 11),
    "pbte_gsis_apply_correction": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_int32]),
    "pbte_hdg_local": (C.c_int, [C.POINTER(PbteHdgDesc), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_char_p, C.c_int32]),
    "pbte_hdg_assemble": (C.c_int, [C.POINTER(PbteHdgDesc), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                    C.c_void_p, C.c_void_p, _i64p, C.c_void_p, C.c_char_p, C.c_int32]),
    "pbte_hdg_assemble_static": (C.c_int, [C.POINTER(PbteHdgDesc), C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, _i64p, C.c_void_p,
                                           C.c_char_p, C.c_int32]),
    "pbte_trace_create": (C.c_void_p, [C.POINTER(PbteTraceDesc), C.c_void_p, C.c_char_p, C.c_int32]),
    "pbte_trace_destroy": (None, [C.c_void_p]),
    "pbte_trace_last_error": (C.c_char_p, [C.c_void_p]),
    "pbte_trace_info": (C.c_int, [C.c_void_p, _i64p]),
    "pbte_trace_factor": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pbte_trace_options": (C.c_int, [C.c_void_p, C.c_int32]),
    "pbte_trace_graphs": (C.c_int, [C.c_void_p, C.c_int32]),
    "pbte_trace_quality": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), _f64p]),
    "pbte_trace_solve": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "pbte_trace_level": (C.c_int, [C.c_void_p, C.c_int32, _i64p]),
    "pbte_gmsh_read": (C.c_void_p, [C.c_char_p, C.c_char_p, C.c_int32]),
    "pbte_batch_pipelined": (C.c_int, [C.c_void_p, C.c_int32]),
    "pbte_gmsh_sizes": (C.c_int, [C.c_void_p, _i64p]),
    "pbte_gmsh_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbte_gmsh_free": (None, [C.c_void_p]),
    "pbte_face_integrals": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with build_native.py "
                               "(there is no CPU fallback)")
        # the HIP runtime (PyTorch's) is brought up on the device before the
        # library's code objects register with it: registered first, the
        # library's later device allocations have failed with "no
        # ROCm-capable device".  Without a device (build and host-only
        # calls) there is nothing to bring up.
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
            torch.cuda.current_stream()
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def f64(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_f64p)


def i32(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_i32p)


class NativeError(RuntimeError):
    """A HIP or usage error reported by the native library."""


def check(ctx, status: int, what: str) -> int:
    if status < 0:
        msg = lib().pbte_last_error(ctx)
        raise NativeError(f"{what}: {msg.decode() if msg else 'unknown error'}")
    return status
