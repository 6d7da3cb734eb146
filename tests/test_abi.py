"""The C ABI library loads and exports every symbol include/dcell_lsm.h declares;
argument checks answer without touching a GPU."""
import ctypes as C
import os
import re

import pytest

from dcell_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols(path):
    text = open(path).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(dcl_[a-z0-9_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    declared = header_symbols(os.path.join(ROOT, "include", "dcell_lsm.h"))
    assert declared, "no declarations parsed"
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(L.EXPORTS) == declared


def test_kernel_names_mention_no_unnamed_type():
    """A kernel template whose mangled name refers to an unnamed type (an
    enumerator of an anonymous enum inside a dependent parameter type) gets
    that type numbered differently by the host and the device compilations,
    and its launch then finds no device function.  The host stubs' kernel
    names are the mangled strings in the library's data."""
    blob = open(L.LIB_PATH, "rb").read()
    names = set(re.findall(rb"_ZN3dcl[0-9A-Za-z_$]+", blob))
    assert names, "no dcl kernel names found"
    bad = sorted(n.decode() for n in names if b"$_" in n)
    assert not bad, bad


def test_abi_version_and_argument_errors():
    lib = L.lib()
    assert lib.dcl_abi_version() == 1
    rc = lib.dcl_sl_advect(None, None, None, None, None, 0, None, 0.1, 0, 0, 0, None)
    assert rc == 85  # DCL_ERR_ARG_NULL
    assert b"phi0" in lib.dcl_last_error()
    w = L.Window(L.ICoor(0, 0, 0), L.ICoor(3, 3, 3), L.Coor(1, 1, 1))
    dummy = C.c_void_p(16)
    V = L.Velocity(dummy, w, 3, 0)
    rc = lib.dcl_sl_advect(dummy, C.c_void_p(32), C.byref(w), C.byref(V), None, 0, None, 0.1, 0, 0, 0, None)
    assert rc == 60  # dense windows need 5 nodes per axis
    rc = lib.dcl_sl_advect(dummy, C.c_void_p(32), C.byref(w), C.byref(V), None, 0, None, 0.1, 7, 0, 0, None)
    assert rc == 62  # unknown precision
    rc = lib.dcl_curvature(dummy, C.c_void_p(32), C.byref(w), None, 0, None, 0, 1, None)
    assert rc == 56  # LSS stencil is 2D only
    # component planes (dof 1): the SL entry points take them (these calls fail
    # later, on the window size), the others answer DCL_ERR_SUP
    w5 = L.Window(L.ICoor(0, 0, 0), L.ICoor(5, 5, 5), L.Coor(1, 1, 1))
    P = L.Velocity(dummy, w5, 1, 0)
    rc = lib.dcl_sl_advect(dummy, C.c_void_p(32), C.byref(w), C.byref(P), None, 0, None, 0.1, 0, 0, 0, None)
    assert rc == 60, lib.dcl_last_error()
    rc = lib.dcl_max_velocity(C.byref(P), dummy, 1, 12, 3, 0, C.byref(C.c_double()), None)
    assert rc == 56 and b"per node" in lib.dcl_last_error()
    rc = lib.dcl_sl_advect(dummy, C.c_void_p(32), C.byref(w5), C.byref(L.Velocity(dummy, w5, 1, 1)), None, 0,
                           None, 0.1, 0, 0, 0, None)
    assert rc == 62 and b"udof 0" in lib.dcl_last_error()


def test_buffer_load_paths_reject_windows_of_2_gib():
    """Kernels that read through buffer descriptors use signed 32-bit byte
    offsets: the adjoint step and the batch refuse windows of 2 GiB and more
    with DCL_ERR_ARG_SIZ before any device work."""
    lib = L.lib()
    dummy = C.c_void_p(16)
    big = L.Window(L.ICoor(0, 0, 0), L.ICoor(1024, 1024, 300), L.Coor(1, 1, 1))  # 2.5 GB of FP64
    rc = lib.dcl_sl_adjoint_step(dummy, C.byref(big), C.c_void_p(32), C.c_void_p(48), C.c_void_p(64),
                                 C.c_void_p(80), 0.1, L.DCL_F64, L.DCL_CUBIC, None)
    assert rc == 60 and b"2 GiB" in lib.dcl_last_error()
    cell = L.Window(L.ICoor(0, 0, 0), L.ICoor(600, 600, 600), L.Coor(1, 1, 1))   # velocity 2.6 GB in FP32
    out = C.c_void_p()
    rc = lib.dcl_batch_create(1, C.byref(cell), L.DCL_F32, L.DCL_CUBIC, L.DCL_RK2, C.byref(out))
    assert rc == 60 and b"2^31" in lib.dcl_last_error()


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", "/nonexistent/libdcell_lsm.so")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        L.lib()
