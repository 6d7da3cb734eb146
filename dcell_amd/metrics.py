"""Reductions over the C ABI: max velocity (LevelSet_MaxVelocity,
LevelSetMethod/Advect.c:70-111), band compaction, interface metrics
(LevelSetDiracDelta3D/2D, NormalDirection.c:36-57) and error norms."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L
from .levelset import GridWindow, _dev, _p, precision_of, stream_handle, velocity


def max_velocity(vel, vel_win: GridWindow, nodes, stride_bytes: int = 12, udof=0) -> float:
    """max |V| at integer nodes (an int32 device tensor read with a byte stride)."""
    V = velocity(vel, vel_win, udof)
    _dev(nodes, "nodes")
    count = nodes.numel() * nodes.element_size() // stride_bytes
    out = C.c_double()
    L.call("dcl_max_velocity", C.byref(V), _p(nodes), count, stride_bytes, vel_win.dim,
           precision_of(vel), C.byref(out), stream_handle())
    return out.value


def band_compact(phi, win: GridWindow, width: float, capacity: int | None = None):
    """Nodes with |phi| < width in raster order, as an (n, 3) int32 device tensor."""
    _dev(phi, "phi")
    cap = win.nodes if capacity is None else capacity
    out = torch.empty((max(cap, 1), 3), dtype=torch.int32, device=phi.device)
    count = C.c_int64()
    L.call("dcl_band_compact", _p(phi), C.byref(win.c()), float(width), precision_of(phi), _p(out),
           cap, C.byref(count), stream_handle())
    return out[:min(count.value, cap)], count.value


def interface_metrics(phi, win: GridWindow, kappa=None, eps: float = 1.5):
    """(area, sum delta kappa dV, mean curvature, nodes in the delta band)."""
    _dev(phi, "phi")
    out = (C.c_double * 4)()
    L.call("dcl_interface_metrics", _p(phi), _p(kappa), C.byref(win.c()), float(eps),
           precision_of(phi), out, stream_handle())
    return tuple(out)


def error_norms(a, b):
    """(max |a-b|, sum (a-b)^2, max |b|, n)."""
    _dev(a, "a")
    _dev(b, "b")
    if a.numel() != b.numel():
        raise ValueError("a and b differ in size")
    out = (C.c_double * 4)()
    L.call("dcl_error_norms", _p(a), precision_of(a), _p(b), precision_of(b), a.numel(), out,
           stream_handle())
    return tuple(out)
