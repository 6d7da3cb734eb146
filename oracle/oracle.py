"""ctypes front end of the CPU oracle (oracle/lsm_oracle.c).

TEST INFRASTRUCTURE ONLY.  Arrays are numpy float64 in the reference's layout:
scalar windows shaped (nz, ny, nx) (or (ny, nx) in 2D), velocity windows
shaped (..., dof), particles shaped (count, 4) as {x, y, z, radius}.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblsm_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_tricubic.so")

LINEAR, CUBIC = 0, 1
EULER, RK2 = 0, 1
CURV_STD, CURV_LSS = 0, 1
PLS_REF_AZ = 1


class OrcWin(C.Structure):
    _fields_ = [("px", C.c_int32), ("py", C.c_int32), ("pz", C.c_int32),
                ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


@dataclass
class Win:
    """Grid window: global offset p, node counts n, spacing d (Common/Grid.h:11-13)."""
    p: tuple
    n: tuple
    d: tuple

    @property
    def is2d(self):
        return self.n[2] < 2

    @property
    def shape(self):
        return (self.n[1], self.n[0]) if self.is2d else (self.n[2], self.n[1], self.n[0])

    def c(self):
        return OrcWin(*self.p, *self.n, *self.d)


_lib = None
_ref = None
_P = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        W = C.POINTER(OrcWin)
        vp = C.c_void_p
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_interp_vel3d.argtypes = [vp, W, C.c_int, C.c_int, vp, vp]
        L.orc_interp_vel2d.argtypes = [vp, W, C.c_int, C.c_int, vp, vp]
        for f in ("orc_bilinear3d", "orc_bilinear2d", "orc_gridinterp3d", "orc_gridinterp2d",
                  "orc_tricubic", "orc_bicubic"):
            getattr(L, f).argtypes = [vp, W, vp]
            getattr(L, f).restype = C.c_double
        L.orc_cr_weights.argtypes = [C.c_double, vp]
        L.orc_cr_dweights.argtypes = [C.c_double, vp]
        L.orc_cr_stencil3.argtypes = [vp, C.c_double, C.c_double, C.c_double, vp]
        L.orc_sl_advect.argtypes = [vp, vp, vp, vp, C.c_int64, vp, W, W, C.c_int, C.c_int,
                                    C.c_double, C.c_int, C.c_int]
        L.orc_curv3d_at.argtypes = [vp, W, C.c_int, C.c_int, C.c_int]
        L.orc_curv3d_at.restype = C.c_double
        L.orc_curv2d_at.argtypes = [vp, W, C.c_int, C.c_int]
        L.orc_curv2d_at.restype = C.c_double
        L.orc_curv2d_lss_at.argtypes = [vp, W, C.c_int, C.c_int]
        L.orc_curv2d_lss_at.restype = C.c_double
        L.orc_curvature.argtypes = [vp, vp, vp, C.c_int64, vp, W, C.c_int]
        L.orc_pls_advect.argtypes = [vp, C.c_int64, vp, W, C.c_int, C.c_int, vp, C.c_double, C.c_int]
        L.orc_pls_correct.argtypes = [vp, C.c_int64, vp, vp, C.c_int64, W, C.c_int, C.c_int]
        L.orc_pls_adjust_radii.argtypes = [vp, C.c_int64, vp, W, C.c_double, C.c_double, C.c_int]
        L.orc_max_velocity.argtypes = [vp, W, C.c_int, C.c_int, vp, C.c_int64, C.c_int64, C.c_int]
        L.orc_max_velocity.restype = C.c_double
        L.orc_band_compact.argtypes = [vp, W, C.c_double, vp, C.c_int64]
        L.orc_band_compact.restype = C.c_int64
        L.orc_interface_metrics.argtypes = [vp, vp, W, C.c_double, vp]
        L.orc_sl_adjoint.argtypes = [vp, vp, vp, vp, vp, W, W, C.c_int, C.c_int, C.c_double,
                                     C.c_int, C.c_int]
        L.orc_error_norms.argtypes = [vp, vp, vp, C.c_int64, vp]
        L.orc_irregular_nodes.argtypes = [vp, W, vp, C.c_int64, vp, C.c_int64]
        L.orc_irregular_nodes.restype = C.c_int64
        L.orc_irr_tricubic.argtypes = [vp, C.c_double, C.c_double, C.c_double, vp]
        L.orc_irr_bicubic.argtypes = [vp, C.c_double, C.c_double, vp]
        L.orc_surface_quantities.argtypes = [vp, W, vp, C.c_int64, vp]
        L.orc_reinitialize.argtypes = [vp, W, vp, C.c_int64, C.c_double, C.c_double, vp, C.c_int64]
        L.orc_reinitialize.restype = C.c_int64
        L.orc_reinitialize_sweep.argtypes = [vp, W, vp, C.c_int64, C.c_double, C.c_double, vp, C.c_int64,
                                             C.c_int, C.POINTER(C.c_int)]
        L.orc_reinitialize_sweep.restype = C.c_int64
        _lib = L
    return _lib


def ref_tricubic():
    """The reference's TricubicInverse.h/TricubicInterpolation.h compiled from the
    reference tree, or None when that tree was not available at build time."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        R = C.CDLL(REF_PATH)
        R.ref_tricubic_eval.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_void_p]
        _ref = R
    return _ref


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _band(band):
    if band is None:
        return None
    return np.ascontiguousarray(band, dtype=np.int32).reshape(-1, 3)


def set_threads(n):
    lib().orc_set_threads(int(n))


def interp_vel(vel, vwin: Win, X, udof=0):
    vel = _f64(vel)
    X = _f64(X)
    dim = 2 if vwin.is2d else 3
    V = np.zeros(dim)
    f = lib().orc_interp_vel2d if dim == 2 else lib().orc_interp_vel3d
    f(_ptr(vel), C.byref(vwin.c()), vel.shape[-1], udof, _ptr(X), _ptr(V))
    return V


def bilinear(phi, win: Win, S):
    phi, S = _f64(phi), _f64(S)
    f = lib().orc_bilinear2d if win.is2d else lib().orc_bilinear3d
    return f(_ptr(phi), C.byref(win.c()), _ptr(S))


def gridinterp(phi, win: Win, X):
    phi, X = _f64(phi), _f64(X)
    f = lib().orc_gridinterp2d if win.is2d else lib().orc_gridinterp3d
    return f(_ptr(phi), C.byref(win.c()), _ptr(X))


def cubic(phi, win: Win, S):
    phi, S = _f64(phi), _f64(S)
    f = lib().orc_bicubic if win.is2d else lib().orc_tricubic
    return f(_ptr(phi), C.byref(win.c()), _ptr(S))


def cr_weights(t):
    w = np.zeros(4)
    lib().orc_cr_weights(float(t), _ptr(w))
    return w


def cr_dweights(t):
    w = np.zeros(4)
    lib().orc_cr_dweights(float(t), _ptr(w))
    return w


def cr_stencil3(st, x, y, z):
    st = _f64(st).reshape(64)
    out = np.zeros(4)
    lib().orc_cr_stencil3(_ptr(st), x, y, z, _ptr(out))
    return out


def ref_tricubic_eval(st, x, y, z):
    R = ref_tricubic()
    if R is None:
        return None
    st = _f64(st).reshape(64)
    out = np.zeros(4)
    R.ref_tricubic_eval(_ptr(st), x, y, z, _ptr(out))
    return out


def sl_advect(phi0, vel, win: Win, dt, interp=LINEAR, backtrace=EULER, band=None,
              vwin: Win | None = None, udof=0, phi=None):
    """Dense (band None) SL step returning phi, or band values (array of len(band))."""
    phi0 = _f64(phi0)
    vel = _f64(vel)
    vwin = vwin or win
    b = _band(band)
    if b is not None:
        out = np.zeros(len(b))
        lib().orc_sl_advect(_ptr(phi0), None, _ptr(vel), _ptr(b), len(b), _ptr(out),
                            C.byref(win.c()), C.byref(vwin.c()), vel.shape[-1], udof,
                            float(dt), interp, backtrace)
        return out
    out = np.empty_like(phi0) if phi is None else phi
    lib().orc_sl_advect(_ptr(phi0), _ptr(out), _ptr(vel), None, 0, None,
                        C.byref(win.c()), C.byref(vwin.c()), vel.shape[-1], udof,
                        float(dt), interp, backtrace)
    return out


def curvature(phi, win: Win, stencil=CURV_STD, band=None):
    phi = _f64(phi)
    b = _band(band)
    if b is not None:
        out = np.zeros(len(b))
        lib().orc_curvature(_ptr(phi), None, _ptr(b), len(b), _ptr(out), C.byref(win.c()), stencil)
        return out
    out = np.empty_like(phi)
    lib().orc_curvature(_ptr(phi), _ptr(out), None, 0, None, C.byref(win.c()), stencil)
    return out


def curv_at(phi, win: Win, x, y, z=0, stencil=CURV_STD):
    phi = _f64(phi)
    if win.is2d:
        f = lib().orc_curv2d_lss_at if stencil == CURV_LSS else lib().orc_curv2d_at
        return f(_ptr(phi), C.byref(win.c()), x, y)
    return lib().orc_curv3d_at(_ptr(phi), C.byref(win.c()), x, y, z)


def pls_advect(particles, vel, vwin: Win, dh, dt, dim=3, udof=0):
    P = _f64(particles).copy()
    vel = _f64(vel)
    dh = _f64(dh)
    lib().orc_pls_advect(_ptr(P), len(P), _ptr(vel), C.byref(vwin.c()), vel.shape[-1], udof,
                         _ptr(dh), float(dt), dim)
    return P


def pls_correct(particles, phi, win: Win, band=None, flags=0, dim=3):
    P = _f64(particles)
    out = _f64(phi).copy()
    b = _band(band)
    lib().orc_pls_correct(_ptr(P), len(P), _ptr(out), _ptr(b), 0 if b is None else len(b),
                          C.byref(win.c()), flags, dim)
    return out


def pls_adjust_radii(particles, phi, win: Win, rmin=0.1, rmax=0.5, dim=3):
    P = _f64(particles).copy()
    phi = _f64(phi)
    lib().orc_pls_adjust_radii(_ptr(P), len(P), _ptr(phi), C.byref(win.c()), rmin, rmax, dim)
    return P


def max_velocity(vel, vwin: Win, nodes, dim=3, udof=0, stride=12):
    vel = _f64(vel)
    nodes = np.ascontiguousarray(nodes)
    count = nodes.nbytes // stride
    return lib().orc_max_velocity(_ptr(vel), C.byref(vwin.c()), vel.shape[-1], udof,
                                  _ptr(nodes), count, stride, dim)


def band_compact(phi, win: Win, width):
    phi = _f64(phi)
    cap = phi.size
    out = np.zeros((cap, 3), dtype=np.int32)
    n = lib().orc_band_compact(_ptr(phi), C.byref(win.c()), float(width), _ptr(out), cap)
    return out[:n]


# one irregular-node record (orc_irregular_node, 88 bytes)
IRREG_DTYPE = np.dtype([("pos", np.int32, 3), ("axis", np.int32), ("shift", np.int32),
                        ("sign_center", np.int32), ("sign_face", np.int32), ("pad", np.int32),
                        ("d", np.float64), ("X", np.float64, 3), ("op", np.float64, 3)])
assert IRREG_DTYPE.itemsize == 88


def irregular_nodes(phi, win: Win, band):
    """LevelSetUpdateIrregularNodeList (LevelSetMethod/IrregularNodes.c:8-140,
    IrregularNodes3D.c:6-117): the records, in the reference's order."""
    phi = _f64(phi)
    band = np.ascontiguousarray(band, dtype=np.int32).reshape(-1, 3)
    cap = 10 * len(band)
    out = np.zeros(cap, dtype=IRREG_DTYPE)
    n = lib().orc_irregular_nodes(_ptr(phi), C.byref(win.c()), _ptr(band), len(band), _ptr(out), cap)
    return out[:n]


def irr_tricubic(st64, x, y, z):
    """p, px, py, pz of the projections' tricubic on a 4x4x4 stencil."""
    st = _f64(st64).reshape(-1)
    out = np.zeros(4)
    lib().orc_irr_tricubic(_ptr(st), float(x), float(y), float(z), _ptr(out))
    return out


def irr_bicubic(st16, x, y):
    st = _f64(st16).reshape(-1)
    out = np.zeros(3)
    lib().orc_irr_bicubic(_ptr(st), float(x), float(y), _ptr(out))
    return out


def interface_metrics(phi, win: Win, kappa=None, eps=1.5):
    phi = _f64(phi)
    k = None if kappa is None else _f64(kappa)
    out = np.zeros(4)
    lib().orc_interface_metrics(_ptr(phi), _ptr(k), C.byref(win.c()), float(eps), _ptr(out))
    return out


def sl_adjoint(phi0, vel, lam_next, win: Win, dt, interp=CUBIC, backtrace=RK2, udof=0):
    phi0, vel, lam_next = _f64(phi0), _f64(vel), _f64(lam_next)
    lam = np.zeros_like(phi0)
    gvel = np.zeros_like(vel)
    rc = lib().orc_sl_adjoint(_ptr(phi0), _ptr(vel), _ptr(lam_next), _ptr(lam), _ptr(gvel),
                              C.byref(win.c()), C.byref(win.c()), vel.shape[-1], udof,
                              float(dt), interp, backtrace)
    if rc:
        raise RuntimeError(f"orc_sl_adjoint failed: {rc}")
    return lam, gvel


def error_norms(a, b, mask=None):
    a, b = _f64(a).ravel(), _f64(b).ravel()
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    out = np.zeros(4)
    lib().orc_error_norms(_ptr(a), _ptr(b), _ptr(m), a.size, _ptr(out))
    return out


# surface quantities of one record (orc_surface, 104 bytes)
SURF_DTYPE = np.dtype([("n", np.float64, 3), ("s", np.float64, 3), ("r", np.float64, 3), ("k", np.float64),
                       ("k_nn", np.float64), ("k_tt", np.float64), ("k_nt", np.float64)])
assert SURF_DTYPE.itemsize == 104


def surface_quantities(phi, win: Win, records):
    """IIMUpdateSurfaceQuantities_3D / _2D (IIM/SurfaceQuantities.c:31-118) at
    the X of each irregular-node record (IRREG_DTYPE)."""
    phi = _f64(phi)
    recs = np.ascontiguousarray(records, dtype=IRREG_DTYPE)
    out = np.zeros(len(recs), dtype=SURF_DTYPE)
    lib().orc_surface_quantities(_ptr(phi), C.byref(win.c()), _ptr(recs), len(recs), _ptr(out))
    return out


def surface_at(phi, win: Win, X):
    """The surface quantities at arbitrary points X (n, 3) (records with only X set)."""
    X = np.asarray(X, dtype=np.float64).reshape(-1, 3)
    recs = np.zeros(len(X), dtype=IRREG_DTYPE)
    recs["X"] = X
    return surface_quantities(phi, win, recs)


def get_velocity(fluid, win: Win, vel):
    """LevelSetGetVelocity (LevelSetMethod/Advect.c:145-192): the velocity
    components (U_FACE..W_FACE) of the fluid's [z][y][x][p,u,v,w] array
    ([y][x][p,u,v] in 2D), node (0,0,0) at the global origin, copied into a
    copy of vel (win's shape + (dim,)) over the window's box clipped to the
    domain (:160-186); the rest of vel is kept."""
    fluid = np.asarray(fluid)
    out = np.array(vel, copy=True)
    if win.is2d:
        N = fluid.shape[:2]
        lo = [max(win.p[1], 0), max(win.p[0], 0)]
        hi = [min(win.p[1] + win.n[1], N[0]), min(win.p[0] + win.n[0], N[1])]
        if hi[0] > lo[0] and hi[1] > lo[1]:
            out[lo[0] - win.p[1]:hi[0] - win.p[1], lo[1] - win.p[0]:hi[1] - win.p[0], :] = \
                fluid[lo[0]:hi[0], lo[1]:hi[1], 1:3]
        return out
    N = fluid.shape[:3]
    p, n = (win.p[2], win.p[1], win.p[0]), (win.n[2], win.n[1], win.n[0])
    lo = [max(p[a], 0) for a in range(3)]
    hi = [min(p[a] + n[a], N[a]) for a in range(3)]
    if all(hi[a] > lo[a] for a in range(3)):
        out[tuple(slice(lo[a] - p[a], hi[a] - p[a]) for a in range(3))] = \
            fluid[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2], 1:4]
    return out


def reinitialize(phi, win: Win, records, band_width=6.0, phi_inf=7.5):
    """LevelSetReinitialize (LevelSetMethod/FastMarching.c:10-146): returns the
    new phi (a copy) and the new band, (n, 3) int32, in the reference's
    order, from the irregular-node records (IRREG_DTYPE) of phi."""
    out = np.array(_f64(phi), copy=True)
    recs = np.ascontiguousarray(records, dtype=IRREG_DTYPE)
    cap = out.size
    band = np.zeros((cap, 3), dtype=np.int32)
    n = lib().orc_reinitialize(_ptr(out), C.byref(win.c()), _ptr(recs), len(recs), float(band_width),
                               float(phi_inf), _ptr(band), cap)
    if n < 0:
        raise RuntimeError("reinitialisation: NaN update")
    return out, band[:n]


def reinitialize_sweep(phi, win: Win, records, band_width=6.0, phi_inf=7.5, max_sweeps=1000):
    """The parallel reinitialisation (dcl_reinitialize_parallel) restated:
    returns (phi copy, band, sweeps); the band is [boundary nodes in record
    order | - pass | + pass], each pass in window order."""
    out = np.array(_f64(phi), copy=True)
    recs = np.ascontiguousarray(records, dtype=IRREG_DTYPE)
    cap = out.size
    band = np.zeros((cap, 3), dtype=np.int32)
    sw = C.c_int(0)
    n = lib().orc_reinitialize_sweep(_ptr(out), C.byref(win.c()), _ptr(recs), len(recs), float(band_width),
                                     float(phi_inf), _ptr(band), cap, int(max_sweeps), C.byref(sw))
    if n == -3:
        raise RuntimeError(f"reinitialisation sweeps did not settle in {max_sweeps}")
    if n < 0:
        raise RuntimeError("reinitialisation: NaN update")
    return out, band[:n], sw.value


# ---------------------------------------------------------------- formats
REAL64_CLASSID = 1211214  # PETSc's VEC_FILE_CLASSID


def real64_bytes(values):
    """The .Real64 file GridWrite writes (Common/Serialize.c:23-55, VecView to
    a PETSc binary viewer): big-endian int32 class id and int32 count, then
    the values as big-endian doubles -- the layout
    Visualization/DCellVTK.py:24-32 reads (2 * 4 header bytes, '>f8')."""
    v = np.asarray(values, dtype=np.float64).ravel()
    return np.array([REAL64_CLASSID, v.size], dtype=">i4").tobytes() + v.astype(">f8").tobytes()


def irregular_rows(records, surf, extra=None, dim=3):
    """The rows of LevelSetWriteIrregularNodeList_3D / _2D
    (LevelSetMethod/IrregularNodes.c:178-219 / 145-172), as big-endian bytes:
    3D, every record, X, pos, n, k, f1, f1_n, f1_t, f1_nn, f1_tt, f1_nt, s,
    r, numNei (23 doubles); 2D, the records with axis != -1, X.x, X.y,
    pos.x, pos.y, n.x, n.y, f1, fa1, k, f1_n, f1_nn (11 doubles).  extra:
    the IIM fields per record (7 in 3D, 4 in 2D, in that order), zeros when
    None."""
    recs = np.ascontiguousarray(records, dtype=IRREG_DTYPE)
    sq = np.ascontiguousarray(surf, dtype=SURF_DTYPE)
    nx = 7 if dim == 3 else 4
    ex = np.zeros((len(recs), nx)) if extra is None else np.asarray(extra, dtype=np.float64).reshape(len(recs), nx)
    pos = recs["pos"].astype(np.float64)
    if dim == 3:
        rows = np.column_stack([recs["X"], pos, sq["n"], sq["k"], ex[:, :6], sq["s"], sq["r"], ex[:, 6]])
    else:
        keep = recs["axis"] != -1
        rows = np.column_stack([recs["X"][:, :2], pos[:, :2], sq["n"][:, :2], ex[:, 0], ex[:, 1], sq["k"],
                                ex[:, 2], ex[:, 3]])[keep]
    return np.ascontiguousarray(rows, dtype=np.float64).astype(">f8").tobytes()
