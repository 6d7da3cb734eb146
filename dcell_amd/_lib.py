"""ctypes binding of libdcell_lsm.so, the C ABI declared in include/dcell_lsm.h.

The library is built in-tree (``dcell_amd/lib``) by ``__graft_entry__.build()``.
Loading fails loudly when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libdcell_lsm.so")

DCL_F64, DCL_F32 = 0, 1
DCL_LINEAR, DCL_CUBIC = 0, 1
DCL_EULER, DCL_RK2 = 0, 1
DCL_CURV_STD, DCL_CURV_LSS = 0, 1
DCL_PLS_REF_AZ = 1

# exported symbols of include/dcell_lsm.h, checked by tests/test_abi.py
EXPORTS = (
    "dcl_abi_version", "dcl_last_error", "dcl_device_count", "dcl_malloc", "dcl_free",
    "dcl_memcpy_htod", "dcl_memcpy_dtoh", "dcl_memcpy_dtod", "dcl_stream_synchronize",
    "dcl_host_malloc", "dcl_host_free", "dcl_stream_copy",
    "dcl_sl_advect", "dcl_curvature", "dcl_levelset_step",
    "dcl_pls_advect", "dcl_pls_ws_create", "dcl_pls_ws_destroy", "dcl_pls_correct",
    "dcl_pls_adjust_radii", "dcl_max_velocity", "dcl_band_compact", "dcl_band_compact_static", "dcl_irregular_nodes",
    "dcl_surface_quantities", "dcl_get_velocity", "dcl_get_velocity_planes", "dcl_reinitialize", "dcl_reinitialize_parallel",
    "dcl_real64_pack", "dcl_irregular_node_rows",
    "dcl_interface_metrics", "dcl_error_norms", "dcl_sl_adjoint",
    "dcl_sl_adjoint_record_bytes", "dcl_sl_adjoint_prepare", "dcl_sl_adjoint_step", "dcl_sl_adjoint_finish",
    "dcl_batch_create", "dcl_batch_destroy", "dcl_batch_layout", "dcl_batch_set_dt",
    "dcl_batch_set_config", "dcl_batch_get_config", "dcl_batch_set_class_zchunk", "dcl_batch_get_classes",
    "dcl_batch_set_claim", "dcl_batch_get_claim",
    "dcl_batch_step", "dcl_batch_band_step", "dcl_batch_autotune",
    "dcl_batch_graph_create", "dcl_batch_graph_launch", "dcl_batch_graph_destroy",
)


class ICoor(C.Structure):
    """iCoor (Common/Common.h:32-34)."""
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32)]


class Coor(C.Structure):
    """Coor (Common/Common.h:28-30)."""
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Window(C.Structure):
    _fields_ = [("p", ICoor), ("n", ICoor), ("d", Coor)]


class Velocity(C.Structure):
    _fields_ = [("data", C.c_void_p), ("win", Window), ("dof", C.c_int32), ("udof", C.c_int32)]


class ReinitCell(C.Structure):
    """dcl_reinit_cell (include/dcell_lsm.h)."""
    _fields_ = [("phi", C.c_void_p), ("win", Window), ("nodes", C.c_void_p), ("count", C.c_int64),
                ("band_width", C.c_double), ("phi_inf", C.c_double), ("band", C.c_void_p),
                ("band_capacity", C.c_int64), ("band_count", C.c_int64)]


class DclError(RuntimeError):
    def __init__(self, code, fn, msg):
        super().__init__(f"{fn} failed with code {code}: {msg}")
        self.code = code


_lib = None


def _declare(L):
    vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
    W, V = C.POINTER(Window), C.POINTER(Velocity)
    sig = {
        "dcl_abi_version": ([], i32),
        "dcl_last_error": ([], C.c_char_p),
        "dcl_device_count": ([C.POINTER(C.c_int)], i32),
        "dcl_malloc": ([C.POINTER(vp), C.c_size_t], i32),
        "dcl_free": ([vp], i32),
        "dcl_memcpy_htod": ([vp, vp, C.c_size_t, vp], i32),
        "dcl_memcpy_dtoh": ([vp, vp, C.c_size_t, vp], i32),
        "dcl_memcpy_dtod": ([vp, vp, C.c_size_t, vp], i32),
        "dcl_stream_synchronize": ([vp], i32),
        "dcl_host_malloc": ([C.POINTER(vp), C.c_size_t], i32),
        "dcl_host_free": ([vp], i32),
        "dcl_stream_copy": ([vp, vp, C.c_size_t, vp], i32),
        "dcl_sl_advect": ([vp, vp, W, V, vp, i64, vp, dbl, i32, i32, i32, vp], i32),
        "dcl_curvature": ([vp, vp, W, vp, i64, vp, i32, i32, vp], i32),
        "dcl_levelset_step": ([vp, vp, vp, W, V, dbl, i32, i32, i32, vp], i32),
        "dcl_pls_advect": ([vp, i64, V, Coor, dbl, i32, i32, vp], i32),
        "dcl_pls_ws_create": ([i64, i64, i32, vp, C.POINTER(vp)], i32),
        "dcl_pls_ws_destroy": ([vp], i32),
        "dcl_pls_correct": ([vp, i64, vp, W, vp, i64, vp, i32, i32, i32, vp, vp], i32),
        "dcl_pls_adjust_radii": ([vp, i64, vp, W, dbl, dbl, i32, i32, vp], i32),
        "dcl_max_velocity": ([V, vp, i64, i64, i32, i32, C.POINTER(dbl), vp], i32),
        "dcl_band_compact": ([vp, W, dbl, i32, vp, i64, C.POINTER(i64), vp], i32),
        "dcl_band_compact_static": ([vp, W, dbl, i32, vp, i64, C.POINTER(i64), vp], i32),
        "dcl_irregular_nodes": ([vp, W, vp, i64, i32, vp, i64, C.POINTER(i64), vp], i32),
        "dcl_surface_quantities": ([vp, W, vp, i64, i32, vp, vp], i32),
        "dcl_real64_pack": ([vp, i32, i64, vp, vp], i32),
        "dcl_irregular_node_rows": ([vp, vp, vp, i64, i32, vp, i64, C.POINTER(i64), vp], i32),
        "dcl_get_velocity": ([vp, C.POINTER(ICoor), W, i32, vp, vp], i32),
        "dcl_get_velocity_planes": ([vp, C.POINTER(ICoor), W, i32, vp, vp], i32),
        "dcl_reinitialize": ([C.POINTER(ReinitCell), i32, i64, vp], i32),
        "dcl_reinitialize_parallel": ([C.POINTER(ReinitCell), i32, i32, C.POINTER(C.c_int), vp], i32),
        "dcl_interface_metrics": ([vp, vp, W, dbl, i32, C.POINTER(dbl), vp], i32),
        "dcl_error_norms": ([vp, i32, vp, i32, i64, C.POINTER(dbl), vp], i32),
        "dcl_sl_adjoint": ([vp, W, V, vp, vp, vp, dbl, i32, i32, i32, vp], i32),
        "dcl_sl_adjoint_record_bytes": ([W, C.POINTER(i64)], i32),
        "dcl_sl_adjoint_prepare": ([W, V, dbl, i32, i32, i32, vp, vp], i32),
        "dcl_sl_adjoint_step": ([vp, W, vp, vp, vp, vp, dbl, i32, i32, vp], i32),
        "dcl_sl_adjoint_finish": ([W, V, vp, vp, dbl, i32, i32, vp], i32),
        "dcl_batch_create": ([i32, W, i32, i32, i32, C.POINTER(vp)], i32),
        "dcl_batch_destroy": ([vp], i32),
        "dcl_batch_layout": ([vp, vp, vp], i32),
        "dcl_batch_set_dt": ([vp, vp], i32),
        "dcl_batch_set_config": ([vp, i32, i32], i32),
        "dcl_batch_get_config": ([vp, vp, vp, vp, vp], i32),
        "dcl_batch_set_class_zchunk": ([vp, i32, i32], i32),
        "dcl_batch_get_classes": ([vp, vp, vp, vp], i32),
        "dcl_batch_step": ([vp, vp, vp, vp, vp, vp], i32),
        "dcl_batch_band_step": ([vp, vp, vp, vp, vp, vp, i64, vp], i32),
        "dcl_batch_autotune": ([vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, vp], i32),
        "dcl_batch_set_claim": ([vp, i32], i32),
        "dcl_batch_get_claim": ([vp, C.POINTER(i32)], i32),
        "dcl_batch_graph_create": ([vp, vp, vp, vp, vp, C.POINTER(vp)], i32),
        "dcl_batch_graph_launch": ([vp, vp], i32),
        "dcl_batch_graph_destroy": ([vp], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.argtypes = args
        f.restype = res


def lib():
    """Load libdcell_lsm.so (built in-tree); raise if it is not there."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(there is no CPU fallback for the level-set kernels)")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def check(rc, fn):
    if rc != 0:
        msg = lib().dcl_last_error()
        raise DclError(rc, fn, msg.decode() if msg else "")


def call(fn, *args):
    rc = getattr(lib(), fn)(*args)
    check(rc, fn)
    return rc
