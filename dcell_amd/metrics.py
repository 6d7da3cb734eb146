"""Reductions over the C ABI: max velocity (LevelSet_MaxVelocity,
LevelSetMethod/Advect.c:70-111), band compaction, the irregular-node list
(LevelSetUpdateIrregularNodeList, LevelSetMethod/IrregularNodes.c:8-28) and
its surface quantities (IIMUpdateSurfaceQuantities_3D / _2D,
IIM/SurfaceQuantities.c:31-118), the reinitialisation that rebuilds phi
from it (LevelSetReinitialize, LevelSetMethod/FastMarching.c:10-146),
interface metrics (LevelSetDiracDelta3D/2D, NormalDirection.c:36-57) and
error norms."""
from __future__ import annotations

import ctypes as C

import numpy as np
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


def band_compact(phi, win: GridWindow, width: float, capacity: int | None = None, out=None):
    """Nodes with |phi| < width in raster order, as an (n, 3) int32 device
    tensor.  With ``out`` (an int32 device tensor shaped (capacity, 3)) the
    nodes are written there and the allocation-free entry point runs
    (dcl_band_compact_static: no device allocation, windows up to 2^28 nodes)."""
    _dev(phi, "phi")
    precision_of(phi)
    if phi.numel() != win.nodes:
        raise ValueError(f"phi has {phi.numel()} elements, window needs {win.nodes}")
    fn = "dcl_band_compact"
    if out is not None:
        _dev(out, "out", torch.int32)
        if out.dim() != 2 or out.shape[1] != 3:
            raise ValueError("out must be shaped (capacity, 3)")
        cap = out.shape[0]
        fn = "dcl_band_compact_static"
    else:
        cap = win.nodes if capacity is None else capacity
        out = torch.empty((max(cap, 1), 3), dtype=torch.int32, device=phi.device)
    count = C.c_int64()
    L.call(fn, _p(phi), C.byref(win.c()), float(width), precision_of(phi), _p(out),
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


# one dcl_irregular_node record (88 bytes)
IRREGULAR_DTYPE = np.dtype([("pos", np.int32, 3), ("axis", np.int32), ("shift", np.int32),
                            ("sign_center", np.int32), ("sign_face", np.int32), ("pad", np.int32),
                            ("d", np.float64), ("X", np.float64, 3), ("op", np.float64, 3)])
assert IRREGULAR_DTYPE.itemsize == 88


def irregular_nodes_into(phi, win: GridWindow, band, out) -> int:
    """Device-resident form: the records of the band nodes into `out` (a
    uint8 device tensor of capacity * 88 bytes), in the reference's order;
    returns the total count (records beyond the capacity are not written)."""
    _dev(phi, "phi")
    _dev(band, "band")
    _dev(out, "out")
    cap = out.numel() // IRREGULAR_DTYPE.itemsize
    count = C.c_int64()
    L.call("dcl_irregular_nodes", _p(phi), C.byref(win.c()), _p(band), band.shape[0], precision_of(phi),
           _p(out), cap, C.byref(count), stream_handle())
    return count.value


def irregular_nodes(phi, win: GridWindow, band, capacity: int | None = None):
    """The irregular-node records of the band nodes (an (n, 3) int32 device
    tensor), in the reference's order, as a numpy structured array
    (IRREGULAR_DTYPE) on the host, and the total count."""
    cap = 10 * band.shape[0] if capacity is None else capacity
    out = torch.empty(max(cap, 1) * IRREGULAR_DTYPE.itemsize, dtype=torch.uint8, device=phi.device)
    count = irregular_nodes_into(phi, win, band, out[:cap * IRREGULAR_DTYPE.itemsize])
    n = min(count, cap)
    host = out[:n * IRREGULAR_DTYPE.itemsize].cpu().numpy()
    return host.view(IRREGULAR_DTYPE), count


# the surface quantities of one record (dcl_surface, 104 bytes)
SURFACE_DTYPE = np.dtype([("n", np.float64, 3), ("s", np.float64, 3), ("r", np.float64, 3), ("k", np.float64),
                          ("k_nn", np.float64), ("k_tt", np.float64), ("k_nt", np.float64)])
assert SURFACE_DTYPE.itemsize == 104


def surface_quantities_into(phi, win: GridWindow, records, count: int, out) -> None:
    """Device-resident form: the surface quantities at the X of the first
    `count` records in `records` (a uint8 device tensor of 88-byte records, as
    irregular_nodes_into writes them) into `out` (a uint8 device tensor of at
    least count * 104 bytes), asynchronously on the current stream."""
    _dev(phi, "phi")
    _dev(records, "records")
    _dev(out, "out")
    if records.numel() < count * IRREGULAR_DTYPE.itemsize or out.numel() < count * SURFACE_DTYPE.itemsize:
        raise ValueError("records or out holds fewer than count entries")
    L.call("dcl_surface_quantities", _p(phi), C.byref(win.c()), _p(records), int(count), precision_of(phi),
           _p(out), stream_handle())


def surface_quantities(phi, win: GridWindow, records: np.ndarray) -> np.ndarray:
    """The surface quantities (SURFACE_DTYPE) at the X of host records
    (IRREGULAR_DTYPE), computed on phi's device."""
    recs = np.ascontiguousarray(records, dtype=IRREGULAR_DTYPE)
    n = len(recs)
    dev_recs = torch.from_numpy(recs.view(np.uint8).reshape(-1).copy()).to(phi.device)
    out = torch.empty(max(n, 1) * SURFACE_DTYPE.itemsize, dtype=torch.uint8, device=phi.device)
    surface_quantities_into(phi, win, dev_recs, n, out)
    return out[:n * SURFACE_DTYPE.itemsize].cpu().numpy().view(SURFACE_DTYPE)


def reinitialize(cells, band_width: float = 6.0, phi_inf: float = 7.5, heap_capacity: int = 0):
    """LevelSetReinitialize of a batch: `cells` is a list of (phi, win,
    records, count) with phi an FP64 device tensor (rewritten in place) and
    records a uint8 device tensor of irregular-node records (as
    irregular_nodes_into writes them).  Returns, per level set, the new band
    as an (n, 3) int32 device tensor in the reference's order."""
    arr, bands = _reinit_cells(cells, band_width, phi_inf)
    L.call("dcl_reinitialize", arr, len(cells), int(heap_capacity), stream_handle())
    return [b[:arr[i].band_count] for i, b in enumerate(bands)]


def _reinit_cells(cells, band_width, phi_inf):
    arr = (L.ReinitCell * len(cells))()
    bands = []
    for i, (phi, win, records, count) in enumerate(cells):
        _dev(phi, "phi", torch.float64)
        _dev(records, "records")
        if records.numel() < count * IRREGULAR_DTYPE.itemsize:
            raise ValueError("records holds fewer than count entries")
        band = torch.empty((max(win.nodes, 1), 3), dtype=torch.int32, device=phi.device)
        bands.append(band)
        arr[i].phi = phi.data_ptr()
        arr[i].win = win.c()
        arr[i].nodes = records.data_ptr() if count else None
        arr[i].count = int(count)
        arr[i].band_width = float(band_width)
        arr[i].phi_inf = float(phi_inf)
        arr[i].band = band.data_ptr()
        arr[i].band_capacity = win.nodes
    return arr, bands


def reinitialize_parallel(cells, band_width: float = 6.0, phi_inf: float = 7.5, max_sweeps: int = 0):
    """The parallel reinitialisation of a batch (dcl_reinitialize_parallel):
    same arguments as reinitialize.  Returns (bands, sweeps): per level set
    the new band ([projection nodes in record order | - side | + side], each
    side in window order) and the sweeps it took."""
    arr, bands = _reinit_cells(cells, band_width, phi_inf)
    sweeps = (C.c_int * len(cells))()
    L.call("dcl_reinitialize_parallel", arr, len(cells), int(max_sweeps), sweeps, stream_handle())
    return [b[:arr[i].band_count] for i, b in enumerate(bands)], list(sweeps)
