"""The batched update (dcl_batch_*): every cell of a packed batch against the
oracle's dense step of that cell alone; every tile configuration gives the
same bits; the hipGraph of a step reproduces the eager step over repeated
replays; the autotuner picks one of the timed configurations.

FP64: <= 1e-12 of max |phi| (kappa over |phi| < 6 against its max);
FP32: oracle on float32-rounded inputs, 2e-6 (phi) and 5e-5 (kappa).
"""
import numpy as np
import pytest
import torch

import oracle as O
import dcell_amd as D
from dcell_amd import configs
from dcell_amd.batch import NOCC, NSHAPES, Batch
from helpers import owin, rel_max, rounded, to_np

pytestmark = pytest.mark.gpu

EDGES = [9, 12, 16, 21, 26, 33, 40]


def _build(dtype, interp, bt, seed=5):
    g = configs.philox(seed)
    wins, phis, vels, dts = [], [], [], []
    for n in EDGES:
        phi, vel, dt = configs.cell_fields(n, g, "cpu", torch.float64)
        wins.append(D.GridWindow.unit_box(n))
        phis.append(rounded(phi.numpy(), dtype))
        vels.append(rounded(vel.numpy(), dtype))
        dts.append(4 * dt)                         # departures of ~2 cells
    b = Batch(wins, dtype, interp, bt)
    phi0, phi1, kap, vel = b.new_field(), b.new_field(), b.new_field(), b.new_field(3)
    for c in range(len(EDGES)):
        b.cell(phi0, c).copy_(torch.from_numpy(phis[c]))
        b.cell(vel, c, 3).copy_(torch.from_numpy(vels[c]))
    b.set_dt(dts)
    return b, (phi0, phi1, kap, vel), wins, phis, vels, dts


@pytest.mark.parametrize("dtype,tp,tk", [(torch.float64, 1e-12, 1e-12), (torch.float32, 2e-6, 5e-5)])
@pytest.mark.parametrize("interp,bt", [("cubic", "rk2"), ("linear", "euler")])
@pytest.mark.parametrize("shape", [1, 7])   # SL by tile gathers / by LDS-ring columns
def test_batch_matches_per_cell_oracle(dev, dtype, tp, tk, interp, bt, shape):
    b, (phi0, phi1, kap, vel), wins, phis, vels, dts = _build(dtype, interp, bt)
    b.set_config(shape, 1)
    b.step(phi0, phi1, kap, vel)
    for c, w in enumerate(wins):
        ref = O.sl_advect(phis[c], vels[c], owin(w), dts[c], O.CUBIC if interp == "cubic" else O.LINEAR,
                          O.RK2 if bt == "rk2" else O.EULER)
        assert rel_max(to_np(b.cell(phi1, c)), ref) <= tp, c
        kr = O.curvature(rounded(to_np(b.cell(phi1, c)), dtype), owin(w))
        sel = np.abs(ref) < 6
        assert rel_max(to_np(b.cell(kap, c))[sel], kr[sel], np.max(np.abs(kr[sel]))) <= tk, c
    b.close()


def test_batch_configs_identical_graph_replays_autotune(dev):
    b, (phi0, phi1, kap, vel), wins, *_ = _build(torch.float32, "cubic", "rk2")
    b.step(phi0, phi1, kap, vel)
    ref_phi, ref_kap = phi1.clone(), kap.clone()
    for shape in range(NSHAPES):
        for occ in range(NOCC):
            b.set_config(shape, occ)
            phi1.zero_()
            kap.zero_()
            b.step(phi0, phi1, kap, vel)
            assert torch.equal(phi1, ref_phi) and torch.equal(kap, ref_kap), (shape, occ)
    best, table = b.autotune(phi0, phi1, kap, vel, reps=2)
    assert np.all(table > 0) and best == pytest.approx(table.min())
    cfg = b.config()
    assert table[cfg["shape"], cfg["occupancy"]] == pytest.approx(best)
    # a graph of one step, replayed with eager work between the replays
    g = b.graph(phi0, phi1, kap, vel)
    for _ in range(4):
        phi1.fill_(7.0)
        kap.fill_(7.0)
        torch.cuda.synchronize()
        Batch.launch(g)
        torch.cuda.synchronize()
        assert torch.equal(phi1, ref_phi) and torch.equal(kap, ref_kap)
    b.close()


def test_batch_argument_errors(dev):
    with pytest.raises(D._lib.DclError):
        Batch([D.GridWindow((0, 0, 0), (9, 9, 1), (0.1, 0.1, 0.1))])        # 2D cell
    b = Batch([D.GridWindow.unit_box(9)])
    f = b.new_field()
    with pytest.raises(D._lib.DclError):
        b.step(f, b.new_field(), b.new_field(), b.new_field(3))             # dt not set
    b.close()
