"""A level set kept in HBM through LevelSetAdvectAndReinit.

``DeviceLevelSet`` holds the state DCell's LevelSet carries on this path
(LevelSetMethod.h:56-94: phi, phi0, the band, the irregular-node list,
maxVel, CFLcount / CFLthres, AdvectCount / AdvectThres, bandWidth, PHI_INF)
as device tensors, and ``advect_and_reinit`` runs the reference's step
(LevelSetAdvectAndReinit, LevelSetMethod/Advect.c:31-50) with every stage on
the GPU and nothing copied to the host but two scalars:

1. GridCopy(phi, phi0), then the SL step over the band
   (LevelSetAdvectSL, AdvectSL.c:6-89);
2. LevelSetCFLIncrement (Advect.c:54-66): max |V| at the positions of the
   irregular-node list left by the previous step (LevelSet_MaxVelocity,
   Advect.c:70-111), CFLcount += dt * maxVel / h.x, AdvectCount += 1;
3. LevelSetUpdateIrregularNodeList (IrregularNodes.c:8-140);
4. past either threshold, both counts to 0 and LevelSetReinitialize
   (FastMarching.c:10-146): the exact fast marching, or the parallel sweeps
   (``reinit="parallel"``), which give the new band.

The thresholds default to the reference's (LevelSet.c:21-24: CFLthres 3,
AdvectThres 1000).
"""
from __future__ import annotations

import torch

from . import metrics as M
from .levelset import GridWindow, _dev, advect_sl


class DeviceLevelSet:
    """A level set's state in HBM (FP64, as DCell's)."""

    def __init__(self, phi: torch.Tensor, win: GridWindow, band_width: float = 6.0, phi_inf: float = 7.5,
                 cfl_thres: float = 3.0, advect_thres: int = 1000, interp: str = "linear",
                 backtrace: str = "euler", reinit: str = "exact"):
        _dev(phi, "phi", torch.float64)
        if reinit not in ("exact", "parallel"):
            raise ValueError("reinit must be 'exact' or 'parallel'")
        self.win = win
        self.phi = phi
        self.phi0 = torch.empty_like(phi)
        self.band_width, self.phi_inf = float(band_width), float(phi_inf)
        self.cfl_thres, self.advect_thres = float(cfl_thres), int(advect_thres)
        self.cfl_count, self.advect_count, self.max_vel = 0.0, 0, 0.0
        self.interp, self.backtrace, self.reinit = interp, backtrace, reinit
        self.band, _ = M.band_compact(phi, win, self.band_width)
        self.band = self.band.contiguous()
        self.records = torch.empty(0, dtype=torch.uint8, device=phi.device)
        self.count = 0
        self.reinits = 0
        self.last_sweeps = 0

    def update_irregular_nodes(self) -> int:
        """LevelSetUpdateIrregularNodeList of the current phi and band."""
        need = max(10 * self.band.shape[0], 1) * M.IRREGULAR_DTYPE.itemsize
        if self.records.numel() < need:
            self.records = torch.empty(need, dtype=torch.uint8, device=self.phi.device)
        count = M.irregular_nodes_into(self.phi, self.win, self.band, self.records)
        if count * M.IRREGULAR_DTYPE.itemsize > self.records.numel():
            self.records = torch.empty(count * M.IRREGULAR_DTYPE.itemsize, dtype=torch.uint8, device=self.phi.device)
            count = M.irregular_nodes_into(self.phi, self.win, self.band, self.records)
        self.count = count
        return count

    def reinitialize(self) -> None:
        """LevelSetReinitialize: phi rebuilt from the irregular-node list, the new band."""
        cells = [(self.phi, self.win, self.records, self.count)]
        if self.reinit == "parallel":
            (band,), (self.last_sweeps,) = M.reinitialize_parallel(cells, self.band_width, self.phi_inf)
        else:
            (band,) = M.reinitialize(cells, self.band_width, self.phi_inf)
        self.band = band.contiguous()
        self.reinits += 1

    def advect_and_reinit(self, vel: torch.Tensor, dt: float) -> bool:
        """One LevelSetAdvectAndReinit step; returns whether it reinitialised."""
        self.phi0.copy_(self.phi)
        advect_sl(self.phi0, vel, self.win, dt, self.interp, self.backtrace, band=self.band, out=self.phi)
        nodes = self.records[:self.count * M.IRREGULAR_DTYPE.itemsize]
        self.max_vel = M.max_velocity(vel, self.win, nodes, stride_bytes=M.IRREGULAR_DTYPE.itemsize) \
            if self.count else 0.0
        self.cfl_count += dt * self.max_vel / self.win.d[0]
        self.advect_count += 1
        self.update_irregular_nodes()
        if self.advect_count > self.advect_thres or self.cfl_count > self.cfl_thres:
            self.cfl_count = 0.0
            self.advect_count = 0
            self.reinitialize()
            return True
        return False
