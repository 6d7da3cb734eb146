"""Parity of the gfx950 reductions (max velocity, band compaction, interface
metrics, error norms) with the oracle."""
import math

import numpy as np
import pytest
import torch

import oracle as O
import dcell_amd as D
from dcell_amd import configs
from dcell_amd import metrics as M
from helpers import owin, random_velocity, smooth_field, to_dev, to_np

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [2, 3])
def test_max_velocity_with_irregular_node_stride(dev, dim):
    rng = np.random.default_rng(70 + dim)
    n = 33
    win = D.GridWindow((1, 2, 3), (n, n, n) if dim == 3 else (n, n, 1), (0.1, 0.1, 0.1))
    vel = random_velocity(n, dim, rng, 1.0)
    nodes = rng.integers(0, n - 1, size=(5000, 3)).astype(np.int32) + np.array(win.p, dtype=np.int32)
    if dim == 2:
        nodes[:, 2] = win.p[2]
    # the same nodes embedded in 40-byte records, as IrregularNode arrays are read
    rec = np.zeros((5000, 10), dtype=np.int32)
    rec[:, :3] = nodes
    ref = O.max_velocity(vel, owin(win), nodes, dim=dim)
    got = M.max_velocity(to_dev(vel), win, to_dev(rec, torch.int32), stride_bytes=40)
    assert got == pytest.approx(ref, rel=1e-15)
    got32 = M.max_velocity(to_dev(vel, torch.float32), win, to_dev(nodes, torch.int32))
    assert got32 == pytest.approx(ref, rel=1e-6)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_band_compact_raster_order(dev, dtype):
    rng = np.random.default_rng(8)
    n = 70
    win = D.GridWindow((-5, 3, 9), (n, n + 3, n - 2), (1, 1, 1))
    phi = smooth_field(n, 3, rng)[: n - 2, :, :]
    phi = np.concatenate([phi, phi[:, :3, :]], axis=1)
    b, count = M.band_compact(to_dev(phi, dtype), win, 3.0)
    ref = O.band_compact(np.asarray(phi, np.float32).astype(np.float64) if dtype == torch.float32 else phi,
                         owin(win), 3.0)
    assert count == len(ref)
    assert np.array_equal(to_np(b).astype(np.int32), ref)
    # capacity smaller than the band: the first entries, the full count
    b2, count2 = M.band_compact(to_dev(phi, dtype), win, 3.0, capacity=100)
    assert count2 == len(ref) and np.array_equal(to_np(b2).astype(np.int32), ref[:100])


def test_band_compact_many_scan_tiles(dev):
    """A window of 288M nodes: 1.1M count blocks, so the scan of the block
    counts runs over more than 1024 tiles of 1024 (the carry across chunks of
    the tile sums).  Checked against torch.nonzero, whose indices come in the
    same raster order."""
    nx, ny, nz = 1024, 1024, 275
    win = D.GridWindow((3, -7, 11), (nx, ny, nz), (1, 1, 1))
    z = torch.arange(nz, device="cuda:0", dtype=torch.float32).view(nz, 1, 1)
    x = torch.arange(nx, device="cuda:0", dtype=torch.float32).view(1, 1, nx)
    phi = (z - 137.3) + 0.8 * torch.sin(0.05 * x) * torch.ones((1, ny, 1), device="cuda:0")
    b, count = M.band_compact(phi.contiguous(), win, 1.5, capacity=8_000_000)
    ref = torch.nonzero(phi.abs() < 1.5)  # (z, y, x)
    assert count == ref.shape[0]
    ref_xyz = ref.flip(1).to(torch.int32) + torch.tensor([3, -7, 11], device="cuda:0", dtype=torch.int32)
    assert torch.equal(b, ref_xyz)
    del phi, b, ref, ref_xyz
    torch.cuda.empty_cache()


def test_interface_metrics_and_norms(dev):
    m = 129
    win = D.GridWindow.unit_box(m)
    phi = configs.sphere_sdf(m, (0.5, 0.5, 0.5), 0.3)
    kap = O.curvature(phi, owin(win))
    ref = O.interface_metrics(phi, owin(win), kap)
    got = M.interface_metrics(to_dev(phi), win, to_dev(kap))
    assert np.allclose(got, ref, rtol=1e-12)
    assert got[0] == pytest.approx(4 * math.pi * 0.09, rel=2e-3)
    got32 = M.interface_metrics(to_dev(phi, torch.float32), win, to_dev(kap, torch.float32))
    assert got32[0] == pytest.approx(ref[0], rel=1e-5) and got32[2] == pytest.approx(ref[2], rel=1e-5)
    w2 = D.GridWindow.unit_box(257, 2)
    c = configs.circle_sdf(257)
    assert M.interface_metrics(to_dev(c), w2)[0] == pytest.approx(O.interface_metrics(c, owin(w2))[0], rel=1e-12)
    rng = np.random.default_rng(2)
    a, b = rng.normal(size=100003), rng.normal(size=100003)
    got = M.error_norms(to_dev(a), to_dev(b))
    ref = O.error_norms(a, b)
    assert np.allclose(got, ref, rtol=1e-12)
    got = M.error_norms(to_dev(a, torch.float32), to_dev(b))
    assert got[0] == pytest.approx(np.max(np.abs(a.astype(np.float32).astype(np.float64) - b)), rel=1e-12)
