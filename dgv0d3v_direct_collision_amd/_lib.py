"""ctypes binding of libdgv_amd.so (the C ABI declared in include/dgv.h).

The library is the only compute path: if it is missing, importing this module
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# DGV_LIB_PATH selects another build of the same library (the sanitizer test's
# host-instrumented copy); there is no other compute path
LIB_PATH = os.environ.get("DGV_LIB_PATH") or os.path.join(HERE, "libdgv_amd.so")

DGV_OK = 0
ERRORS = {-1: "DGV_ERR_ARG", -2: "DGV_ERR_GRID", -3: "DGV_ERR_INDEX", -4: "DGV_ERR_NOMEM",
          -5: "DGV_ERR_HIP", -6: "DGV_ERR_STATE", -7: "DGV_ERR_IO", -8: "DGV_ERR_UNSUPPORTED"}
DGV_FP64, DGV_FP32_COMP, DGV_FP32_F64ACC, DGV_INT8_SPLIT = 0, 1, 2, 3
PRECISION_NAMES = {DGV_FP64: "FP64", DGV_FP32_COMP: "FP32_COMP", DGV_FP32_F64ACC: "FP32_F64ACC",
                   DGV_INT8_SPLIT: "INT8_SPLIT"}
DGV_KERNEL_AUTO, DGV_KERNEL_PAIRBLOCK, DGV_KERNEL_RECORD, DGV_KERNEL_MFMA = 0, 1, 2, 3
DGV_FRAMING_SUBRECORDS, DGV_FRAMING_UNSIGNED32 = 0, 1


class GridDesc(C.Structure):
    _fields_ = [("Mu", C.c_int32), ("Mv", C.c_int32), ("Mw", C.c_int32),
                ("su", C.c_int32), ("sv", C.c_int32), ("sw", C.c_int32),
                ("u_L", C.c_double), ("u_R", C.c_double), ("v_L", C.c_double),
                ("v_R", C.c_double), ("w_L", C.c_double), ("w_R", C.c_double)]


class AgenParams(C.Structure):
    _fields_ = [("Trad", C.c_double), ("min_sens", C.c_double), ("nquad", C.c_int32),
                ("moment_mode", C.c_int32)]


class DGVError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_i32 = C.c_int32
_i64 = C.c_int64
_dbl = C.c_double
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_fp = C.POINTER(C.c_float)

# name -> (restype, argtypes); every symbol of include/dgv.h
SIGNATURES = {
    "dgv_ctx_create": (C.c_int, [C.POINTER(GridDesc), C.c_int, _pp]),
    "dgv_ctx_destroy": (None, [_vp]),
    "dgv_last_error": (C.c_char_p, [_vp]),
    "dgv_ctx_num_nodes": (C.c_int, [_vp, _lp]),
    "dgv_ctx_canonical_cell": (C.c_int, [_vp, _ip]),
    "dgv_grid_geometry": (C.c_int, [C.POINTER(GridDesc), _dp, _dp, _dp, _dp, _ip, _lp]),
    "dgv_ctx_get_nodes": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "dgv_ctx_get_shifts": (C.c_int, [_vp, _ip, _ip, _ip, _ip]),
    "dgv_ctx_set_aarray": (C.c_int, [_vp, _ip, _i64, _dp, _ip, _ip, _i64, C.c_int]),
    "dgv_ctx_generate_aarray": (C.c_int, [_vp, C.POINTER(AgenParams)]),
    "dgv_agen_pair": (C.c_int, [C.POINTER(GridDesc), _i64, _i64, C.POINTER(AgenParams), _dp, _dp]),
    "dgv_ctx_aarray_size": (C.c_int, [_vp, _lp, _lp]),
    "dgv_ctx_workload": (C.c_int, [_vp, _lp]),
    "dgv_ctx_get_aarray_basis": (C.c_int, [_vp, _i32, _dp, _ip, _ip, _i64, _lp, C.c_int]),
    "dgv_ctx_get_aarray_bases": (C.c_int, [_vp, _ip, _i32, _dp, _ip, _ip, _i64, _lp, C.c_int]),
    "dgv_aarray_write_file": (C.c_int, [_vp, C.c_char_p]),
    "dgv_ctx_read_aarray_file": (C.c_int, [_vp, C.c_char_p]),
    "dgv_aarray_file_header": (C.c_int, [C.c_char_p, _lp, _lp]),
    "dgv_aarray_file_read": (C.c_int, [C.c_char_p, _ip, _i64, _dp, _ip, _ip, _ip, _i64]),
    "dgv_aarray_file_write": (C.c_int, [C.c_char_p, _ip, _i64, _dp, _ip, _ip, _ip, _i64, _i32]),
    "dgv_set_fortran_subrecord_limit": (C.c_int, [_i64]),
    "dgv_aarray_chunks_scan": (C.c_int, [C.POINTER(GridDesc), C.c_char_p, C.c_char_p, _i32, _i32, _i32, _lp]),
    "dgv_aarray_chunks_read": (C.c_int, [C.POINTER(GridDesc), C.c_char_p, C.c_char_p, _i32, _i32, _i32, _i64,
                                         _i64, _i32, _ip, _i64, _dp, _ip, _ip, _ip, _i64, _lp]),
    "dgv_ctx_write_aarray_chunks": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _i32, _i32, _i64, _i32, _ip]),
    "dgv_ctx_read_aarray_chunks": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _i32, _i32, _i32]),
    "dgv_ctx_trunc_aarray_threshold": (C.c_int, [_vp, _dbl]),
    "dgv_ctx_trunc_aarray_radius": (C.c_int, [_vp, _dbl]),
    "dgv_ctx_aarray_abs": (C.c_int, [_vp]),
    "dgv_aarray_chunk_name": (C.c_int, [C.POINTER(GridDesc), C.c_char_p, C.c_char_p, _i32, _i32,
                                        _i32, C.c_char_p, C.c_size_t]),
    "dgv_eval_q": (C.c_int, [_vp, _dp, _dp, _i64]),
    "dgv_eval_q_dev": (C.c_int, [_vp, _vp, _vp, _i64]),
    "dgv_eval_q_dev_noalloc": (C.c_int, [_vp, _vp, _vp, _i64]),
    "dgv_eval_q_a": (C.c_int, [_vp, _dp, _dp, _i64]),
    "dgv_eval_mixed": (C.c_int, [_vp, _dp, _dp, _dp, _i64]),
    "dgv_eval_mixed_dev": (C.c_int, [_vp, _vp, _vp, _vp, _i64]),
    "dgv_ctx_set_kernel": (C.c_int, [_vp, C.c_int]),
    "dgv_ctx_set_tile": (C.c_int, [_vp, _i32]),
    "dgv_ctx_set_projection": (C.c_int, [_vp, _i32]),
    "dgv_prepare_fma": (C.c_int, [_vp, _dp, _i64]),
    "dgv_prepare_fma_dev": (C.c_int, [_vp, _vp, _i64]),
    "dgv_eval_linear": (C.c_int, [_vp, _dp, _dp, _i64]),
    "dgv_eval_linear_dev": (C.c_int, [_vp, _vp, _i64, _vp, _i64, _i64, _i32]),
    "dgv_ctx_get_fma": (C.c_int, [_vp, _dp, _i64]),
    "dgv_ctx_get_fm": (C.c_int, [_vp, _dp, _i64]),
    "dgv_ctx_set_regime": (C.c_int, [_vp, _i32, _dbl, _dbl]),
    "dgv_ctx_get_regime": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int64)]),
    "dgv_check_solution_mode": (C.c_int, [_vp, _dp, _dp, _i64]),
    "dgv_ctx_autotune": (C.c_int, [_vp, _ip, _dp]),
    "dgv_spatial_oper": (C.c_int, [_vp, _dp, _dp, _i64, _dbl, _dbl]),
    "dgv_spatial_oper_dev": (C.c_int, [_vp, _vp, _vp, _i64, _dbl, _dbl]),
    "dgv_rk_step": (C.c_int, [_vp, _dp, _i64, _dbl, _dbl, _i32, _dp]),
    "dgv_rk_step_dev": (C.c_int, [_vp, _vp, _i64, _dbl, _dbl, _i32, _vp]),
    "dgv_rk_run_dev": (C.c_int, [_vp, _vp, _i64, _dbl, _dbl, _i32, _i64, C.c_int]),
    "dgv_ab_step": (C.c_int, [_vp, _dp, _i64, _dbl, _dbl, _i32, _dp]),
    "dgv_moments": (C.c_int, [_vp, _dp, _i64, _dp]),
    "dgv_moments_high": (C.c_int, [_vp, _dp, _i64, _dp]),
    "dgv_moments_high_dev": (C.c_int, [_vp, _vp, _i64, _dp]),
    "dgv_sol_file_name": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _i32, _i32, C.c_char_p, C.c_char_p, C.c_size_t]),
    "dgv_log_file_name": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _i32, _i32, C.c_char_p, C.c_char_p, C.c_size_t]),
    "dgv_sol_write_file": (C.c_int, [C.c_char_p, _dbl, _dbl, _i32, _i64, _dp, _dp]),
    "dgv_sol_write_init_file": (C.c_int, [C.c_char_p, _dbl, _dbl, _i32, _i64, _dp]),
    "dgv_sol_read_file": (C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int32), _i64, _dp, _dp]),
    "dgv_collision_moments": (C.c_int, [_vp, _dp, _i64, _dp]),
    "dgv_project_invariants": (C.c_int, [_vp, _dp, _i64]),
    "dgv_project_invariants_dev": (C.c_int, [_vp, _vp, _i64]),
    "dgv_batch_create": (C.c_int, [C.POINTER(_vp), _i32, _ip, _i32, _i32, _pp]),
    "dgv_batch_destroy": (None, [_vp]),
    "dgv_batch_layout": (C.c_int, [_vp, _i32, _lp, _lp, _lp]),
    "dgv_batch_set_f": (C.c_int, [_vp, _i32, _dp, _i64]),
    "dgv_batch_get_q": (C.c_int, [_vp, _i32, _dp, _i64]),
    "dgv_batch_set_f_rows_dev": (C.c_int, [_vp, _vp, _i64, _i32, _i32]),
    "dgv_batch_get_q_rows_dev": (C.c_int, [_vp, _vp, _i64, _i32, _i32]),
    "dgv_batch_get_f_rows_dev": (C.c_int, [_vp, _vp, _i64, _i32, _i32]),
    "dgv_batch_eval_q": (C.c_int, [_vp]),
    "dgv_batch_rk_run": (C.c_int, [_vp, _dbl, _dbl, _i32, _i64, C.c_int]),
    "dgv_batch_autotune": (C.c_int, [_vp, _ip, _dp]),
    "dgv_batch_set_tile": (C.c_int, [_vp, _i32, _i32]),
    "dgv_batch_set_projection": (C.c_int, [_vp, _i32]),
    "dgv_batch_device_buffers": (C.c_int, [_vp, _pp, _pp]),
    "dgv_batch_project_invariants": (C.c_int, [_vp]),
    "dgv_batch_set_probe_cells": (C.c_int, [_vp, _ip, _i32]),
    "dgv_batch_eval_q_probe": (C.c_int, [_vp]),
    "dgv_sur_create": (C.c_int, [_vp, _ip, _i32, _i32, C.c_uint64, _pp]),
    "dgv_sur_destroy": (None, [_vp]),
    "dgv_sur_set_scales": (C.c_int, [_vp, _dbl, _dbl]),
    "dgv_sur_layer_shape": (C.c_int, [_vp, _i32, _ip, _ip]),
    "dgv_sur_get_layer": (C.c_int, [_vp, _i32, _fp, _fp]),
    "dgv_sur_set_layer": (C.c_int, [_vp, _i32, _fp, _fp]),
    "dgv_sur_get_grad": (C.c_int, [_vp, _i32, _fp, _fp]),
    "dgv_sur_train_step": (C.c_int, [_vp, _vp, _vp, _i64, _i32, _dbl, _dp]),
    "dgv_sur_predict": (C.c_int, [_vp, _vp, _i64, _i64, _vp, _i32]),
    "dgv_sur_set_probe_cells": (C.c_int, [_vp, _ip, _i32]),
    "dgv_sur_eval": (C.c_int, [_vp, _vp, _i64, _i64, _vp, _dbl, _dp, _ip, _lp]),
    "dgv_sur_set_decomposition": (C.c_int, [_vp, _i32]),
    "dgv_iphi": (C.c_int, [_vp, _dp, _dp, _i64, _dbl]),
    "dgv_iphi_decomp": (C.c_int, [_vp, _dp, _dp, _i64, _dbl]),
    "dgv_sur_time_predict": (C.c_int, [_vp, _vp, _i64, _i64, _vp, _i32, _dp]),
    "dgv_ctx_synchronize": (C.c_int, [_vp]),
    "dgv_time_eval_q_dev": (C.c_int, [_vp, _vp, _vp, _i32, _dp, _dp]),
    "dgv_batch_time_eval": (C.c_int, [_vp, _i32, _dp, _dp]),
    "dgv_device_alloc": (C.c_int, [_vp, C.c_size_t, _pp]),
    "dgv_device_free": (C.c_int, [_vp, _vp]),
    "dgv_memcpy_htod": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "dgv_memcpy_dtoh": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "dgv_device_copy": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libdgv_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m "
                "dgv0d3v_direct_collision_amd.build` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, ctx=None) -> None:
    if rc != DGV_OK and rc is not None and rc < 0:
        msg = load().dgv_last_error(ctx)
        raise DGVError(rc, msg.decode() if msg else "")


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def fptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)
