"""Particle level set over the C ABI: RK2 particle advection, escaped-particle
correction and radius adjustment (LevelSetMethod/ParticleLS3D.c,
ParticleLS2D.c, ParticleLS.c:158-185).  Particles are device tensors shaped
(n, 4) as {x, y, z, radius} in global index units, float64 or float32."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L
from .levelset import GridWindow, _band, _dev, _p, precision_of, stream_handle, velocity

REF_AZ = L.DCL_PLS_REF_AZ


def _particles(particles, dtype=None):
    """Particles must be a contiguous device tensor shaped (n, 4) of the given dtype."""
    _dev(particles, "particles", dtype)
    if particles.dim() != 2 or particles.shape[1] != 4:
        raise ValueError(f"particles must be shaped (n, 4), got {tuple(particles.shape)}")
    return particles


def _phi(phi, win: GridWindow, dtype):
    _dev(phi, "phi", dtype)
    if phi.numel() != win.nodes:
        raise ValueError(f"phi has {phi.numel()} elements, window needs {win.nodes}")
    return phi


def advect_particles(particles, vel, vel_win: GridWindow, dh, dt: float, dim: int = 3, udof=0):
    """In-place RK2 midpoint advection (ParticleLS_AdvectParticles3D/2D)."""
    _particles(particles)
    precision_of(particles)
    if particles.dtype != vel.dtype:
        raise TypeError("particles and vel must share a dtype")
    V = velocity(vel, vel_win, udof)
    L.call("dcl_pls_advect", _p(particles), particles.shape[0], C.byref(V), L.Coor(*dh),
           float(dt), dim, precision_of(particles), stream_handle())
    return particles


class CorrectionWorkspace:
    """phi+/phi- key grids and the touched-node list of the correction."""

    def __init__(self, max_nodes: int, max_events: int, dtype=torch.float64):
        self.dtype = dtype
        self.prec = L.DCL_F64 if dtype == torch.float64 else L.DCL_F32
        h = C.c_void_p()
        L.call("dcl_pls_ws_create", int(max_nodes), int(max_events), self.prec, stream_handle(),
               C.byref(h))
        self.handle = h

    def close(self):
        if self.handle:
            L.lib().dcl_pls_ws_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def correct(particles, phi, win: GridWindow, ws: CorrectionWorkspace, band=None, flags=0,
            band_values=False):
    """Escaped-particle correction (ParticleLS_ErrorCorrection3D/2D).  Updates phi
    in place (band nodes, or every node when band is None), or returns the
    merged band values when band_values=True."""
    _dev(phi, "phi")
    prec = precision_of(phi)
    _phi(phi, win, phi.dtype)
    _particles(particles, phi.dtype)
    b, n = _band(band)
    if b is not None and n == 0:
        b = torch.zeros((1, 3), dtype=torch.int32, device=phi.device)  # non-NULL, length 0
    res = None
    if b is not None and band_values:
        res = torch.empty(max(n, 1), dtype=phi.dtype, device=phi.device)
    L.call("dcl_pls_correct", _p(particles), particles.shape[0], _p(phi), C.byref(win.c()), _p(b), n,
           _p(res), flags, win.dim, prec, ws.handle, stream_handle())
    return res[:n] if res is not None else phi


def adjust_radii(particles, phi, win: GridWindow, rmin=0.1, rmax=0.5):
    """ParticleLS_AdjustRadii: r = s clamp(s phi(X), rmin, rmax), in place."""
    _particles(particles)
    _phi(phi, win, particles.dtype)
    L.call("dcl_pls_adjust_radii", _p(particles), particles.shape[0], _p(phi), C.byref(win.c()),
           float(rmin), float(rmax), win.dim, precision_of(phi), stream_handle())
    return particles
