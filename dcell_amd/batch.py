"""Batched update of many independent level sets over the C ABI (dcl_batch_*).

The cells' windows (all 3D or all 2D) are packed back to back in one phi
buffer, one kappa buffer and one velocity buffer; ``Batch.step`` runs the SL
step and the curvature of every cell in two persistent launches,
``Batch.band_step`` the same at listed nodes only (band mode),
``Batch.autotune`` picks the tile shape, the occupancy and each size class's
z-chunk by timing them, and ``Batch.graph`` records a step as a hipGraph for
launch-bound loops.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .levelset import INTERP, BACKTRACE, GridWindow, _dev, _p, stream_handle

NSHAPES, NOCC = 12, 3  # 6 tile shapes x 2 SL kernels (tile gather, LDS ring), x 3 occupancies
NCONFIG = NSHAPES * NOCC
NCLASS, NZC = 3, 4     # cell-size classes x ring z-chunks tried (16, 32, 64, 128)
ZCHUNKS = (16, 32, 64, 128)
CLAIMS = (0, 1, 4, 16)  # tiles per claim tried (0: static stride)


class Batch:
    def __init__(self, windows: list[GridWindow], dtype=torch.float32, interp="cubic", backtrace="rk2"):
        self.windows = list(windows)
        self.dtype = dtype
        self.prec = L.DCL_F64 if dtype == torch.float64 else L.DCL_F32
        arr = (L.Window * len(windows))(*[w.c() for w in windows])
        h = C.c_void_p()
        L.call("dcl_batch_create", len(windows), arr, self.prec, INTERP[interp], BACKTRACE[backtrace],
               C.byref(h))
        self.handle = h
        po = np.zeros(len(windows) + 1, dtype=np.int64)
        vo = np.zeros(len(windows) + 1, dtype=np.int64)
        L.call("dcl_batch_layout", h, po.ctypes.data_as(C.c_void_p), vo.ctypes.data_as(C.c_void_p))
        self.phi_offsets, self.vel_offsets = po, vo
        self.nodes = int(po[-1])
        self.graphs = []

    @property
    def dim(self) -> int:
        return self.windows[0].dim

    def new_field(self, components: int = 1) -> torch.Tensor:
        return torch.zeros(self.nodes * components, dtype=self.dtype, device="cuda:0")

    def cell(self, t: torch.Tensor, c: int, components: int = 1) -> torch.Tensor:
        """View of cell c in a packed buffer, shaped like its window."""
        w = self.windows[c]
        if components == 1:
            return t[self.phi_offsets[c]:self.phi_offsets[c + 1]].view(w.shape)
        return t[self.vel_offsets[c]:self.vel_offsets[c + 1]].view(w.shape + (components,))

    def set_dt(self, dt):
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        L.call("dcl_batch_set_dt", self.handle, dt.ctypes.data_as(C.c_void_p))

    def set_config(self, shape: int, occupancy: int):
        L.call("dcl_batch_set_config", self.handle, shape, occupancy)

    def config(self):
        s, o, b, k = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        tile = (C.c_int32 * 3)()
        L.call("dcl_batch_get_config", self.handle, C.byref(s), C.byref(o), tile, C.byref(b))
        L.call("dcl_batch_get_claim", self.handle, C.byref(k))
        return dict(shape=s.value, occupancy=o.value, tile=list(tile), blocks=b.value, tiles_per_claim=k.value)

    def set_claim(self, tiles_per_claim: int):
        L.call("dcl_batch_set_claim", self.handle, tiles_per_claim)

    def classes(self):
        """Per size class: the largest cell axis it holds (None: unbounded), its
        cell count and its ring z-chunk."""
        e, n, z = (C.c_int32 * NCLASS)(), (C.c_int32 * NCLASS)(), (C.c_int32 * NCLASS)()
        L.call("dcl_batch_get_classes", self.handle, e, n, z)
        return [dict(max_edge=e[k] if e[k] >= 0 else None, cells=n[k], zchunk=z[k]) for k in range(NCLASS)]

    def set_class_zchunk(self, cls: int, zchunk: int):
        L.call("dcl_batch_set_class_zchunk", self.handle, cls, zchunk)

    def _check(self, *fields):
        for t, comps in zip(fields, (1, 1, 1, self.dim)):
            _dev(t, "field", self.dtype)
            if t.numel() != self.nodes * comps:
                raise ValueError(f"a packed field has {t.numel()} elements, the batch needs {self.nodes * comps}")

    def step(self, phi0, phi, kappa, vel):
        self._check(phi0, phi, kappa, vel)
        L.call("dcl_batch_step", self.handle, _p(phi0), _p(phi), _p(kappa), _p(vel), stream_handle())

    def band_step(self, phi0, phi, kappa, vel, band):
        """SL and curvature at the listed nodes only: band is an (n, 4) int32
        device tensor of {x, y, z, cell} (global coordinates of that cell)."""
        self._check(phi0, phi, kappa, vel)
        _dev(band, "band", torch.int32)
        if band.dim() != 2 or band.shape[1] != 4:
            raise ValueError("band must be shaped (n, 4)")
        if band.shape[0]:
            cells = band[:, 3]
            if int(cells.min()) < 0 or int(cells.max()) >= len(self.windows):
                raise ValueError("band cell index out of range")
        L.call("dcl_batch_band_step", self.handle, _p(phi0), _p(phi), _p(kappa), _p(vel),
               _p(band) if band.shape[0] else None, band.shape[0], stream_handle())

    def autotune(self, phi0, phi, kappa, vel, reps: int = 3):
        """Returns (best ms per step, the shape x occupancy table, the size
        class x z-chunk table, the tiles-per-claim table); -1 where a
        configuration was not timed."""
        self._check(phi0, phi, kappa, vel)
        table = np.zeros(NCONFIG)
        ztable = np.zeros(NCLASS * NZC)
        ctable = np.zeros(len(CLAIMS))
        best = C.c_double()
        L.call("dcl_batch_autotune", self.handle, _p(phi0), _p(phi), _p(kappa), _p(vel), reps,
               table.ctypes.data_as(C.c_void_p), ztable.ctypes.data_as(C.c_void_p),
               ctable.ctypes.data_as(C.c_void_p), C.byref(best), stream_handle())
        return best.value, table.reshape(NSHAPES, NOCC), ztable.reshape(NCLASS, NZC), ctable

    def graph(self, phi0, phi, kappa, vel):
        self._check(phi0, phi, kappa, vel)
        g = C.c_void_p()
        L.call("dcl_batch_graph_create", self.handle, _p(phi0), _p(phi), _p(kappa), _p(vel), C.byref(g))
        self.graphs.append(g)
        return g

    @staticmethod
    def launch(graph):
        L.call("dcl_batch_graph_launch", graph, stream_handle())

    def close(self):
        for g in self.graphs:
            L.lib().dcl_batch_graph_destroy(g)
        self.graphs = []
        if self.handle:
            L.lib().dcl_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
