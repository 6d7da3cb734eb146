"""ctypes view of the C drop-in (libdcell_levelset.so over libdcell_compat.so).

Mirrors the reference's structs (LevelSetMethod.h:10-54, Grid.h:8-22,
LSM_private.h:45-68) so tests can build a LevelSet, fill its grids and band,
call the reference's entry points by name and read the results back.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib

LIB_LEVELSET = os.path.join(_lib.LIB_DIR, "libdcell_levelset.so")
LIB_COMPAT = os.path.join(_lib.LIB_DIR, "libdcell_compat.so")
PETSC_MAX_PATH_LEN = 4096


class ICoor(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class Coor(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class GridS(C.Structure):
    pass


GridP = C.POINTER(GridS)
GridS._fields_ = [
    ("v", C.c_void_p), ("name", C.c_char * PETSC_MAX_PATH_LEN), ("p", ICoor), ("n", ICoor),
    ("d", Coor), ("is2D", C.c_int), ("v1", C.POINTER(C.c_double)), ("grid", C.c_void_p),
    ("dof", C.c_int), ("SIZE", C.c_int), ("MAXSIZE", C.c_int), ("filePos", C.c_void_p),
    ("fileSize", C.c_void_p), ("Interpolate", C.c_void_p),
]


class LevelSetS(C.Structure):
    pass


LevelSetP = C.POINTER(LevelSetS)
ADVECT_FN = C.CFUNCTYPE(C.c_int, LevelSetP, GridP, C.c_double)


class ParticleLSS(C.Structure):
    pass


ParticleLSP = C.POINTER(ParticleLSS)
PADVECT_FN = C.CFUNCTYPE(C.c_int, ParticleLSP, Coor, GridP, C.c_double)
PCORR_FN = C.CFUNCTYPE(C.c_int, ParticleLSP, LevelSetP)

LevelSetS._fields_ = [
    ("phi", GridP), ("phi0", GridP), ("psi", LevelSetP), ("tmp", GridP), ("band", C.c_void_p),
    ("irregularNodes", C.c_void_p), ("bandWidth", C.c_double), ("PHI_INF", C.c_double),
    ("CFLthres", C.c_double), ("CFLcount", C.c_double), ("AdvectThres", C.c_int),
    ("AdvectCount", C.c_int), ("maxVel", C.c_double), ("pls", ParticleLSP), ("Advect", C.c_void_p),
]

ParticleLSS._fields_ = [
    ("R_MIN", C.c_double), ("R_MAX", C.c_double), ("D_INIT", C.c_double), ("S_INIT", C.c_int),
    ("L_MAX", C.c_int), ("G_TOL", C.c_double), ("E_DIST", C.c_double),
    ("AdvectParticles", C.c_void_p), ("ErrorCorrection", C.c_void_p), ("rnd", C.c_void_p),
    ("particles", C.c_void_p), ("particles0", C.c_void_p),
]


class IrregularNodeS(C.Structure):
    _fields_ = ([("pos", ICoor), ("axis", C.c_int), ("shift", C.c_int), ("signCenter", C.c_int),
                 ("signFace", C.c_int), ("GAP", C.c_int)]
                + [(f, C.c_double) for f in ("nx", "ny", "nz", "sx", "sy", "sz", "rx", "ry", "rz", "d",
                                             "k", "k_nn", "k_tt", "k_nt")]
                + [(f"{a}{i}{s}", C.c_double) for i in (1, 2, 3) for a, s in
                   (("F", ""), ("f", ""), ("f", "_n"), ("f", "_t"), ("f", "_nn"), ("f", "_tt"), ("f", "_nt"))]
                + [(f, C.c_double) for f in ("ftx", "fty", "ftz", "fa1", "fa2")]
                + [("X", Coor), ("op", Coor), ("numNei", C.c_int), ("uj", C.c_double)])


IrregularNodeP = C.POINTER(IrregularNodeS)
# InterfacialForce (IIM/ImmersedInterfaceMethod.h:15)
InterfacialForce = C.CFUNCTYPE(None, IrregularNodeP, C.c_void_p)


class IIMS(C.Structure):
    """struct _IIM (IIM/ImmersedInterfaceMethod.h:20-34)."""
    _fields_ = [("lc", C.c_void_p), ("lsq", C.c_void_p), ("Np", C.c_int), ("eps", C.c_double), ("mu", C.c_double),
                ("F", InterfacialForce), ("context", C.c_void_p), ("idx", C.c_void_p), ("coor", C.c_void_p),
                ("val", C.c_void_p), ("debug", C.c_void_p), ("dh", Coor), ("sidx", C.c_void_p)]


_L = None


class _Libs:
    """Looks a symbol up in the drop-in first, then in the containers."""

    def __init__(self, *libs):
        self._libs = libs

    def __getattr__(self, name):
        for L in self._libs:
            try:
                f = getattr(L, name)
            except AttributeError:
                continue
            setattr(self, name, f)
            return f
        raise AttributeError(name)


def lib():
    """Load the compat containers (globally, so the drop-in resolves them) and the drop-in."""
    global _L
    if _L is None:
        _lib.lib()
        for p in (LIB_COMPAT, LIB_LEVELSET):
            if not os.path.exists(p):
                raise RuntimeError(f"{p} is missing: build it with __graft_entry__.build()")
        compat = C.CDLL(LIB_COMPAT, mode=C.RTLD_GLOBAL)
        L = _Libs(C.CDLL(LIB_LEVELSET, mode=C.RTLD_GLOBAL), compat)
        vp, dbl, i32 = C.c_void_p, C.c_double, C.c_int
        L.LevelSetCreate.argtypes = [Coor, ICoor, ICoor, C.POINTER(LevelSetP)]
        L.LevelSetDestroy.argtypes = [LevelSetP]
        L.GridCreate.argtypes = [Coor, ICoor, ICoor, i32, C.POINTER(GridP)]
        L.GridDestroy.argtypes = [GridP]
        L.GridCopy.argtypes = [GridP, GridP]
        L.GridGet.argtypes = [GridP, vp]
        L.ArrayCreate.argtypes = [C.c_char_p, i32, C.POINTER(vp)]
        L.ArraySetSize.argtypes = [vp, i32]
        L.ArrayGetData.argtypes = [vp]
        L.ArrayGetData.restype = vp
        L.ArrayLength.argtypes = [vp]
        for f in ("IIMUpdateSurfaceQuantities_2D", "IIMUpdateSurfaceQuantities_3D"):
            getattr(L, f).argtypes = [C.POINTER(IIMS), LevelSetP]
        for f in ("LevelSetAdvectSL", "LevelSetAdvectSLRK2HalfStep", "LevelSetAdvectSLRK2FullStep",
                  "LevelSetAdvectAndReinit", "LevelSetAdvectPLS", "LevelSetAdvectPLSRK2HalfStep",
                  "LevelSetAdvectPLSRK2FullStep", "LevelSetCFLIncrement"):
            getattr(L, f).argtypes = [LevelSetP, GridP, dbl]
        L.LevelSet_MaxVelocity.argtypes = [LevelSetP, GridP]
        L.LevelSetCurvature.argtypes = [LevelSetP, GridP]
        L.LevelSetAdvectSLSetScheme.argtypes = [i32, i32]
        L.LevelSetPLSSetReferenceCornerZ.argtypes = [i32]
        L.ParticleLS_AdjustRadii.argtypes = [ParticleLSP, LevelSetP]
        L.LevelSetGPUFinalize.argtypes = []
        L.LevelSetGPUTransferStats.argtypes = [C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]
        _L = L
    return _L


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} returned {rc}: {_lib.lib().dcl_last_error().decode()}")


def grid_array(g: GridP) -> np.ndarray:
    """numpy view of a Grid's data shaped (nz, ny, nx[, dof]) (2D: (ny, nx[, dof]))."""
    s = g.contents
    shape = (s.n.y, s.n.x) if s.is2D else (s.n.z, s.n.y, s.n.x)
    if s.dof > 1:
        shape = shape + (s.dof,)
    return np.ctypeslib.as_array(s.v1, shape=(s.SIZE,)).reshape(shape)


def create_levelset(dh, p, n) -> LevelSetP:
    ls = LevelSetP()
    check(lib().LevelSetCreate(Coor(*dh), ICoor(*p), ICoor(*n), C.byref(ls)), "LevelSetCreate")
    return ls


def create_grid(dh, p, n, dof) -> GridP:
    g = GridP()
    check(lib().GridCreate(Coor(*dh), ICoor(*p), ICoor(*n), dof, C.byref(g)), "GridCreate")
    return g


def set_array(arr_handle, data: np.ndarray, elem_bytes: int):
    """Resize a reference Array to len(data) elements and copy raw data in."""
    data = np.ascontiguousarray(data)
    n = data.nbytes // elem_bytes
    check(lib().ArraySetSize(arr_handle, n), "ArraySetSize")
    if n:
        C.memmove(lib().ArrayGetData(arr_handle), data.ctypes.data, data.nbytes)


def get_array(arr_handle, dtype, width) -> np.ndarray:
    n = lib().ArrayLength(arr_handle)
    if n == 0:
        return np.zeros((0, width), dtype=dtype)
    nbytes = n * width * np.dtype(dtype).itemsize
    buf = (C.c_char * nbytes).from_address(lib().ArrayGetData(arr_handle))
    return np.frombuffer(buf, dtype=dtype).reshape(n, width).copy()


def new_array(name: bytes, elem_bytes: int):
    h = C.c_void_p()
    check(lib().ArrayCreate(name, elem_bytes, C.byref(h)), "ArrayCreate")
    return h


def attach_particles(ls: LevelSetP, particles: np.ndarray, dim: int, rmin=0.1, rmax=0.5):
    """Give ls a ParticleLS with these particles (the seeding itself,
    LevelSetInitializeParticles, is outside the drop-in) and the 2D/3D
    advection and correction entry points of the drop-in, as the reference
    sets them (ParticleLS.c:77-85)."""
    L = lib()
    pls = ParticleLSS()
    pls.R_MIN, pls.R_MAX, pls.D_INIT, pls.S_INIT = rmin, rmax, 3.0, 32 if dim == 3 else 16
    pls.L_MAX, pls.G_TOL, pls.E_DIST = 30, 1e-3, 0.5
    adv = L.ParticleLS_AdvectParticles3D if dim == 3 else L.ParticleLS_AdvectParticles2D
    cor = L.ParticleLS_ErrorCorrection3D if dim == 3 else L.ParticleLS_ErrorCorrection2D
    pls.AdvectParticles = C.cast(adv, C.c_void_p)
    pls.ErrorCorrection = C.cast(cor, C.c_void_p)
    pls.particles = new_array(b"particles", 32)
    pls.particles0 = None
    set_array(pls.particles, particles.astype(np.float64), 32)
    keep = pls
    ls.contents.pls = C.pointer(keep)
    ls.contents.Advect = C.cast(L.LevelSetAdvectPLS, C.c_void_p)
    return keep
