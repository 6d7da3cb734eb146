"""Host-side mirror of DCell's level-set update over the C ABI (torch tensors in HBM).

Each function enqueues the gfx950 kernels of ``libdcell_lsm.so`` on torch's
current HIP stream.  Tensors must be contiguous device tensors in the
reference's layout: scalar windows shaped (nz, ny, nx) (2D: (ny, nx)),
velocity windows shaped (..., dof), band lists (n, 3) int32 of global node
coordinates, particles (n, 4) as {x, y, z, radius}.  float64 tensors run the
FP64 path, float32 tensors the FP32 form (FP64 departure coordinates).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib as L

INTERP = {"linear": L.DCL_LINEAR, "cubic": L.DCL_CUBIC}
BACKTRACE = {"euler": L.DCL_EULER, "rk2": L.DCL_RK2}
STENCIL = {"std": L.DCL_CURV_STD, "lss": L.DCL_CURV_LSS}


@dataclass(frozen=True)
class GridWindow:
    """Geometry of a DCell Grid (Common/Grid.h:11-13): global offset p, node
    counts n (n[2] < 2 is 2D, Common/Grid.c:21) and spacing d."""
    p: tuple
    n: tuple
    d: tuple

    @property
    def is2d(self) -> bool:
        return self.n[2] < 2

    @property
    def dim(self) -> int:
        return 2 if self.is2d else 3

    @property
    def shape(self) -> tuple:
        return (self.n[1], self.n[0]) if self.is2d else (self.n[2], self.n[1], self.n[0])

    @property
    def nodes(self) -> int:
        return self.n[0] * self.n[1] * (1 if self.is2d else self.n[2])

    def c(self) -> L.Window:
        return L.Window(L.ICoor(*self.p), L.ICoor(*self.n), L.Coor(*self.d))

    @staticmethod
    def unit_box(n: int, dim: int = 3) -> "GridWindow":
        """Nodes (i h, j h, k h) on [0, 1]^dim with h = 1/(n-1)."""
        h = 1.0 / (n - 1)
        if dim == 2:
            return GridWindow((0, 0, 0), (n, n, 1), (h, h, h))
        return GridWindow((0, 0, 0), (n, n, n), (h, h, h))


def precision_of(t: torch.Tensor) -> int:
    if t.dtype == torch.float64:
        return L.DCL_F64
    if t.dtype == torch.float32:
        return L.DCL_F32
    raise TypeError(f"unsupported dtype {t.dtype}: float64 or float32 expected")


def stream_handle() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t: torch.Tensor, name: str, dtype=None) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} has dtype {t.dtype}, expected {dtype}")
    return t


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def velocity(vel: torch.Tensor, win: GridWindow, udof: int = 0, planes: bool = False) -> L.Velocity:
    """The C ABI's view of a velocity tensor: components per node, shaped
    (..., dof), or with ``planes`` one plane per component, shaped
    (dim, nz, ny, nx) (2D: (2, ny, nx)), passed as dof 1."""
    _dev(vel, "vel")
    if planes:
        dim = 3 if win.n[2] >= 2 else 2
        if udof != 0 or vel.shape[0] != dim or vel.numel() != win.nodes * dim:
            raise ValueError(f"component planes must be shaped ({dim}, window) with udof 0, "
                             f"got {tuple(vel.shape)}")
        return L.Velocity(C.c_void_p(vel.data_ptr()), win.c(), 1, 0)
    dof = vel.shape[-1]
    if vel.numel() != win.nodes * dof:
        raise ValueError(f"vel has {vel.numel()} elements, window {win.n} x dof {dof} needs "
                         f"{win.nodes * dof}")
    return L.Velocity(C.c_void_p(vel.data_ptr()), win.c(), dof, udof)


def _band(band):
    if band is None:
        return None, 0
    _dev(band, "band", torch.int32)
    if band.dim() != 2 or band.shape[1] != 3:
        raise ValueError("band must be shaped (n, 3)")
    return band, band.shape[0]


def advect_sl(phi0, vel, win: GridWindow, dt: float, interp="linear", backtrace="euler",
              band=None, vel_win: GridWindow | None = None, udof=0, out=None,
              band_values=False, planes=False):
    """One SL step (LevelSetAdvectSL, LevelSetMethod/AdvectSL.c:6-89).

    Dense (band None): returns the new phi window.  Band: writes the band
    nodes of ``out`` (a copy of phi0 if not given) and returns it, or with
    ``band_values=True`` returns just the band values.  ``planes``: vel holds
    one plane per component, shaped (dim, nz, ny, nx), instead of the
    components per node (..., dof)."""
    _dev(phi0, "phi0")
    prec = precision_of(phi0)
    if vel.dtype != phi0.dtype:
        raise TypeError("phi0 and vel must share a dtype")
    if phi0.numel() != win.nodes:
        raise ValueError(f"phi0 has {phi0.numel()} elements, window needs {win.nodes}")
    V = velocity(vel, vel_win or win, udof, planes)
    b, n = _band(band)
    if b is not None and n == 0:      # an empty tensor has a NULL data pointer: NULL means dense
        return phi0.new_empty(0) if band_values else (phi0.clone() if out is None else out)
    if b is not None and band_values:
        res = torch.empty(n, dtype=phi0.dtype, device=phi0.device)
        L.call("dcl_sl_advect", _p(phi0), None, C.byref(win.c()), C.byref(V), _p(b), n,
               _p(res), float(dt), prec, INTERP[interp], BACKTRACE[backtrace], stream_handle())
        return res
    if out is None:
        out = phi0.clone() if b is not None else torch.empty_like(phi0)
    _dev(out, "out", phi0.dtype)
    L.call("dcl_sl_advect", _p(phi0), _p(out), C.byref(win.c()), C.byref(V), _p(b), n, None,
           float(dt), prec, INTERP[interp], BACKTRACE[backtrace], stream_handle())
    return out


def curvature(phi, win: GridWindow, stencil="std", band=None, out=None, band_values=False):
    """Curvature at every node or at band nodes (GridFunction3D_Curv,
    Common/Grid3D.c:83-108; GridFunction2D_Curv, Common/Grid.c:231-254;
    GridFunction2D_CurvLSS, LevelSetMethod/Curvature.c:3-186)."""
    _dev(phi, "phi")
    prec = precision_of(phi)
    b, n = _band(band)
    if b is not None and n == 0:
        return phi.new_empty(0) if band_values else (torch.zeros_like(phi) if out is None else out)
    if b is not None and band_values:
        res = torch.empty(n, dtype=phi.dtype, device=phi.device)
        L.call("dcl_curvature", _p(phi), None, C.byref(win.c()), _p(b), n, _p(res), prec,
               STENCIL[stencil], stream_handle())
        return res
    if out is None:
        out = torch.zeros_like(phi)
    _dev(out, "out", phi.dtype)
    L.call("dcl_curvature", _p(phi), _p(out), C.byref(win.c()), _p(b), n, None, prec,
           STENCIL[stencil], stream_handle())
    return out


def levelset_step(phi0, vel, win: GridWindow, dt: float, interp="linear", backtrace="euler",
                  out_phi=None, out_kappa=None, udof=0, planes=False):
    """One dense update: SL step + curvature of the result.  Returns (phi,
    kappa).  ``planes`` as in advect_sl."""
    _dev(phi0, "phi0")
    prec = precision_of(phi0)
    if vel.dtype != phi0.dtype:
        raise TypeError("phi0 and vel must share a dtype")
    V = velocity(vel, win, udof, planes)
    out_phi = torch.empty_like(phi0) if out_phi is None else out_phi
    out_kappa = torch.empty_like(phi0) if out_kappa is None else out_kappa
    L.call("dcl_levelset_step", _p(phi0), _p(out_phi), _p(out_kappa), C.byref(win.c()),
           C.byref(V), float(dt), prec, INTERP[interp], BACKTRACE[backtrace], stream_handle())
    return out_phi, out_kappa


def get_velocity(fluid, win: GridWindow, out=None, planes=False):
    """LevelSetGetVelocity (LevelSetMethod/Advect.c:145-192) from a device
    fluid field: `fluid` is the fluid's (Nz, Ny, Nx, 4) array of (p, u, v, w)
    (2D: (Ny, Nx, 3) of (p, u, v)) over its domain with node (0, 0, 0) at the
    global origin.  Returns the velocity window (win's shape + (3,) or (2,)):
    (u, v, w) at the nodes of win inside the domain; the other nodes of `out`
    keep their values (zeros for a new tensor).  ``planes``: the window as
    component planes, shaped (3,) + win's shape ((2,) + in 2D), the layout
    advect_sl and levelset_step take with planes=True."""
    _dev(fluid, "fluid")
    d3 = fluid.dim() == 4
    if fluid.shape[-1] != (4 if d3 else 3):
        raise ValueError("fluid must be (Nz, Ny, Nx, 4) or (Ny, Nx, 3)")
    dim = 3 if d3 else 2
    shape = (dim,) + win.shape if planes else win.shape + (dim,)
    if out is None:
        out = torch.zeros(shape, dtype=fluid.dtype, device=fluid.device)
    _dev(out, "out", fluid.dtype)
    if tuple(out.shape) != shape:
        raise ValueError(f"out has shape {tuple(out.shape)}, expected {shape}")
    fn = L.ICoor(fluid.shape[-2], fluid.shape[-3], fluid.shape[0] if d3 else 1)
    L.call("dcl_get_velocity_planes" if planes else "dcl_get_velocity", _p(fluid), C.byref(fn), C.byref(win.c()), precision_of(fluid), _p(out),
           stream_handle())
    return out
