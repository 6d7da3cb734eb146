"""The batched update (dcl_batch_*): every cell of a packed batch against the
oracle's dense step of that cell alone, for 3D and 2D cells; the band-mode
step against the oracle's band step of every cell, nodes outside the band
untouched; every tile configuration and z-chunk gives the same bits; the
hipGraph of a step reproduces the eager step over repeated replays; the
autotuner picks one of the timed configurations.

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
from helpers import owin, random_velocity, rel_max, rounded, smooth_field, to_np

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
    # every ring z-chunk of every size class: the same bits
    b.set_config(7, 1)
    for cls in range(3):
        for zc in (8, 16, 64, 128):
            b.set_class_zchunk(cls, zc)
            phi1.zero_()
            kap.zero_()
            b.step(phi0, phi1, kap, vel)
            assert torch.equal(phi1, ref_phi) and torch.equal(kap, ref_kap), (cls, zc)
        b.set_class_zchunk(cls, 32)
    # every work distribution: the same bits
    for k in (0, 1, 3, 16):
        b.set_claim(k)
        for _ in range(2):                                 # the second launch starts from reset counters
            phi1.zero_()
            kap.zero_()
            b.step(phi0, phi1, kap, vel)
            assert torch.equal(phi1, ref_phi) and torch.equal(kap, ref_kap), k
    best, table, ztable, ctable = b.autotune(phi0, phi1, kap, vel, reps=2)
    assert np.all(table > 0) and np.all(ctable > 0)
    cfg = b.config()
    assert best <= table.min() * (1 + 1e-9) and best <= ctable.min() * (1 + 1e-9)
    assert best == pytest.approx(min(table.min(), ctable.min(), ztable[ztable > 0].min(initial=1e300)))
    assert cfg["tiles_per_claim"] in (0, 1, 4, 16)
    if cfg["shape"] >= 6:
        assert (ztable > 0).any()
        for c in b.classes():
            assert c["zchunk"] in (16, 32, 64, 128)
    else:
        assert not (ztable > 0).any()
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


EDGES2D = [9, 24, 37, 64, 90, 131]


def _build2d(dtype, interp, bt, seed=8):
    rng = np.random.default_rng(seed)
    wins, phis, vels, dts = [], [], [], []
    for q, n in enumerate(EDGES2D):
        h = 1.0 / (n - 1)
        wins.append(D.GridWindow((q - 2, 3 - q, 0), (n, n + q % 2, 1), (h, h, h)))
        phi = smooth_field(max(n, n + q % 2), 2, rng)[: n + q % 2, :n]
        vel = random_velocity(max(n, n + q % 2), 2, rng, 1.0)[: n + q % 2, :n]
        phis.append(rounded(np.ascontiguousarray(phi), dtype))
        vels.append(rounded(np.ascontiguousarray(vel), dtype))
        dts.append(2.5 * h)
    b = Batch(wins, dtype, interp, bt)
    phi0, phi1, kap, vel = b.new_field(), b.new_field(), b.new_field(), b.new_field(2)
    for c in range(len(wins)):
        b.cell(phi0, c).copy_(torch.from_numpy(phis[c]))
        b.cell(vel, c, 2).copy_(torch.from_numpy(vels[c]))
    b.set_dt(dts)
    return b, (phi0, phi1, kap, vel), wins, phis, vels, dts


@pytest.mark.parametrize("dtype,tp,tk", [(torch.float64, 1e-12, 1e-12), (torch.float32, 2e-6, 5e-5)])
@pytest.mark.parametrize("interp,bt", [("cubic", "rk2"), ("linear", "euler")])
def test_batch_2d_matches_per_cell_oracle(dev, dtype, tp, tk, interp, bt):
    b, (phi0, phi1, kap, vel), wins, phis, vels, dts = _build2d(dtype, interp, bt)
    best, table, ztable, ctable = b.autotune(phi0, phi1, kap, vel, reps=1)
    assert table[0].min() > 0 and np.all(table[1:] == -1) and np.all(ztable == -1) and np.all(ctable > 0)
    g = b.graph(phi0, phi1, kap, vel)
    phi1.fill_(7.0)
    kap.fill_(7.0)
    Batch.launch(g)
    for c, w in enumerate(wins):
        ref = O.sl_advect(phis[c], vels[c], owin(w), dts[c], O.CUBIC if interp == "cubic" else O.LINEAR,
                          O.RK2 if bt == "rk2" else O.EULER)
        got = to_np(b.cell(phi1, c))
        assert rel_max(got, ref) <= tp, c
        kr = O.curvature(rounded(got, dtype), owin(w))
        sel = np.abs(ref) < 6
        assert rel_max(to_np(b.cell(kap, c))[sel], kr[sel], np.max(np.abs(kr[sel]))) <= tk, c
    b.close()


def _band_list(wins, phis, width, rng):
    """Nodes with |phi| < width of every cell as {x, y, z, cell} in a random
    order, plus each cell's (n, 3) band."""
    rows, per = [], []
    for c, (w, phi) in enumerate(zip(wins, phis)):
        idx = np.argwhere(np.abs(phi) < width)[:, ::-1]           # x, y(, z) local
        q = np.zeros((len(idx), 3), dtype=np.int32)
        q[:, :w.dim] = idx + np.array(w.p[:w.dim])
        q[:, 2] = q[:, 2] if w.dim == 3 else w.p[2]
        per.append(q)
        rows.append(np.c_[q, np.full(len(q), c, dtype=np.int32)])
    allrows = np.concatenate(rows)
    return allrows[rng.permutation(len(allrows))].astype(np.int32), per


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("dtype,tp,tk", [(torch.float64, 1e-12, 1e-12), (torch.float32, 2e-6, 5e-5)])
@pytest.mark.parametrize("interp,bt", [("cubic", "rk2"), ("linear", "euler")])
def test_batch_band_step(dev, dim, dtype, tp, tk, interp, bt):
    if dim == 3:
        b, (phi0, phi1, kap, vel), wins, phis, vels, dts = _build(dtype, interp, bt)
    else:
        b, (phi0, phi1, kap, vel), wins, phis, vels, dts = _build2d(dtype, interp, bt)
    rng = np.random.default_rng(11)
    rows, per = _band_list(wins, phis, 4.0, rng)
    band = torch.from_numpy(rows).to("cuda:0")
    phi1.fill_(7.0)
    kap.fill_(7.0)
    b.band_step(phi0, phi1, kap, vel, band)
    for c, w in enumerate(wins):
        q = per[c]
        ref = O.sl_advect(phis[c], vels[c], owin(w), dts[c], O.CUBIC if interp == "cubic" else O.LINEAR,
                          O.RK2 if bt == "rk2" else O.EULER, band=q)
        got = to_np(b.cell(phi1, c))
        loc = tuple(q[:, a] - w.p[a] for a in range(w.dim))[::-1]
        assert rel_max(got[loc], ref, np.max(np.abs(phis[c]))) <= tp, c
        mask = np.zeros(got.shape, dtype=bool)
        mask[loc] = True
        assert np.all(got[~mask] == 7.0) and np.all(to_np(b.cell(kap, c))[~mask] == 7.0), c
        # curvature at the band nodes from the stepped phi as the GPU left it
        kr = O.curvature(rounded(got, dtype), owin(w), band=q)
        gk = to_np(b.cell(kap, c))[loc]
        sel = np.abs(ref) < 2
        assert rel_max(gk[sel], kr[sel], np.max(np.abs(kr[sel]))) <= tk, c
    b.close()


def test_batch_argument_errors(dev):
    with pytest.raises(D._lib.DclError):
        Batch([D.GridWindow.unit_box(9), D.GridWindow((0, 0, 0), (9, 9, 1), (0.1, 0.1, 0.1))])  # 3D and 2D
    b = Batch([D.GridWindow.unit_box(9)])
    f = b.new_field()
    with pytest.raises(D._lib.DclError):
        b.step(f, b.new_field(), b.new_field(), b.new_field(3))             # dt not set
    b.close()
