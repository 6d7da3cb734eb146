"""Parity of the gfx950 parallel reinitialisation (dcl_reinitialize_parallel)
with the oracle's restatement of its sweeps (reinitialize_sweep, itself
pinned against the fast marching in test_oracle_reinit.py): phi bit for bit
over the whole window, the band in the same order and the same number of
sweeps.  Cases: a scaled sphere in an offset window, a noisy sphere (the
values-only-go-down phase), a torus (cycles), an axis plane reaching the
window's faces, a 2D circle, a batch of level sets of different sizes that
settle after different numbers of sweeps, a 2D batch holding a level set
without records, an empty record list, a band capacity below the band's
length (the count still whole, the prefix written), too few sweeps allowed
and a batch mixing 2D and 3D (errors), and a 294M-node window (the
large-window load path) checked against the closed-form distance to a
plane."""
import numpy as np
import pytest
import torch

import oracle as O
import dcell_amd as D
from dcell_amd import metrics as M
from helpers import owin, to_dev, to_np

pytestmark = pytest.mark.gpu


def _case(phi, win, width=8.0):
    recs = O.irregular_nodes(phi, owin(win), O.band_compact(phi, owin(win), width))
    dev_recs = torch.from_numpy(recs.view(np.uint8).reshape(-1).copy()).to("cuda:0")
    return recs, dev_recs


def _check(phi, win, recs, got_phi, got_band, got_sweeps):
    ref_phi, ref_band, ref_sweeps = O.reinitialize_sweep(phi, owin(win), recs)
    assert got_sweeps == ref_sweeps
    assert to_np(got_phi).tobytes() == ref_phi.tobytes()
    np.testing.assert_array_equal(to_np(got_band), ref_band)


def _sphere(n, c, r, scale, noise=0.0, seed=0):
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    sdf = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r
    return scale * sdf + noise * np.random.default_rng(seed).normal(size=sdf.shape)


def _run_one(phi, win):
    recs, drecs = _case(phi, win)
    dphi = to_dev(phi)
    print(f"reinitialize_parallel window {win.n} records {len(recs)}", flush=True)
    (band,), (sweeps,) = M.reinitialize_parallel([(dphi, win, drecs, len(recs))])
    _check(phi, win, recs, dphi, band, sweeps)
    return sweeps


@pytest.mark.parametrize("noise", [0.0, 0.3])
def test_parallel_sphere_bit_exact(dev, noise):
    n = 48
    win = D.GridWindow((2, -3, 5), (n, n, n), (1.0, 1.0, 1.0))
    _run_one(_sphere(n, (23.3, 22.6, 24.2), 13.4, 2.0, noise, seed=4), win)


def test_parallel_torus(dev):
    n = 64
    win = D.GridWindow((0, 0, 0), (n, n, n), (1.0, 1.0, 1.0))
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    phi = 1.7 * (np.hypot(np.hypot(x - 31.7, y - 32.4) - 18.0, z - 31.1) - 7.3)
    assert _run_one(phi, win) > 20


def test_parallel_plane_and_edge(dev):
    n = 40
    win = D.GridWindow((0, 0, 0), (n, n, 24), (1.0, 1.0, 1.0))
    z, y, x = np.meshgrid(np.arange(24.0), np.arange(float(n)), np.arange(float(n)), indexing="ij")
    _run_one(3.0 * (y - 20.3), win)


def test_parallel_circle_2d(dev):
    n = 96
    win = D.GridWindow((0, 0, 0), (n, n, 1), (1.0, 1.0, 1.0))
    y, x = np.meshgrid(np.arange(float(n)), np.arange(float(n)), indexing="ij")
    _run_one(2.0 * (np.hypot(x - 47.3, y - 49.1) - 30.2), win)


def test_parallel_batch(dev):
    cells, host = [], []
    for i, n in enumerate((24, 40, 33, 70)):
        win = D.GridWindow((i, -i, 2 * i), (n, n + 1, n - 1), (1.0, 1.0, 1.0))
        phi = np.ascontiguousarray(_sphere(n, (n / 2 + 0.3, n / 2 - 0.2, n / 2), n / 4 + 0.1, 1.5, 0.1, seed=i)
                                   [: n - 1, :, :])
        phi = np.concatenate([phi, phi[:, :1, :]], axis=1)  # (n-1, n+1, n)
        recs, drecs = _case(phi, win)
        dphi = to_dev(phi)
        cells.append((dphi, win, drecs, len(recs)))
        host.append((phi, win, recs))
    print("reinitialize_parallel batch of 4", flush=True)
    bands, sweeps = M.reinitialize_parallel(cells)
    assert len(set(sweeps)) > 1  # the level sets settle after different sweeps
    for (phi, win, recs), (dphi, _, _, _), band, sw in zip(host, cells, bands, sweeps):
        _check(phi, win, recs, dphi, band, sw)


def test_parallel_empty_and_sweep_limit(dev):
    n = 20
    win = D.GridWindow((0, 0, 0), (n, n, n), (1.0, 1.0, 1.0))
    phi = _sphere(n, (9.5, 9.5, 9.5), 5.2, 1.0)
    dphi = to_dev(phi)
    empty = torch.zeros(0, dtype=torch.uint8, device="cuda:0")
    (band,), (sweeps,) = M.reinitialize_parallel([(dphi, win, empty, 0)])
    assert band.shape[0] == 0 and sweeps == 1
    assert np.array_equal(to_np(dphi), np.where(phi > 0, 7.5, -7.5))
    recs, drecs = _case(phi, win)
    with pytest.raises(RuntimeError, match="did not settle"):
        M.reinitialize_parallel([(to_dev(phi), win, drecs, len(recs))], max_sweeps=3)


def test_parallel_batch_2d_with_empty_cell(dev):
    """A 2D batch: circles of different sizes and a level set without
    records (it settles after one sweep, all +-PHI_INF)."""
    cells, host = [], []
    for i, n in enumerate((40, 72, 30, 57)):
        win = D.GridWindow((i - 3, 2 * i, 0), (n, n + 3, 1), (1.0, 1.0, 1.0))
        y, x = np.meshgrid(np.arange(float(n + 3)), np.arange(float(n)), indexing="ij")
        phi = 1.5 * (np.hypot(x - n / 2 + 0.3, y - n / 2 - 0.4) - n / 4 - 0.2)
        if i == 2:
            drecs = torch.zeros(0, dtype=torch.uint8, device="cuda:0")
            recs = np.zeros(0, dtype=O.IRREG_DTYPE)
        else:
            recs, drecs = _case(phi, win)
        dphi = to_dev(phi)
        cells.append((dphi, win, drecs, len(recs)))
        host.append((phi, win, recs))
    bands, sweeps = M.reinitialize_parallel(cells)
    assert sweeps[2] == 1 and bands[2].shape[0] == 0
    for (phi, win, recs), (dphi, _, _, _), band, sw in zip(host, cells, bands, sweeps):
        _check(phi, win, recs, dphi, band, sw)


def test_parallel_band_capacity_and_mixed_dims(dev):
    import ctypes as C
    from dcell_amd import _lib as L
    n = 36
    win = D.GridWindow((0, 0, 0), (n, n, n), (1.0, 1.0, 1.0))
    phi = _sphere(n, (17.5, 18.2, 17.1), 9.3, 2.0)
    recs, drecs = _case(phi, win)
    ref_phi, ref_band, ref_sw = O.reinitialize_sweep(phi, owin(win), recs)
    cap = len(ref_band) // 3
    dphi = to_dev(phi)
    band = torch.full((cap, 3), -7, dtype=torch.int32, device="cuda:0")
    cell = L.ReinitCell()
    cell.phi, cell.win, cell.nodes, cell.count = dphi.data_ptr(), win.c(), drecs.data_ptr(), len(recs)
    cell.band_width, cell.phi_inf, cell.band, cell.band_capacity = 6.0, 7.5, band.data_ptr(), cap
    sw = (C.c_int * 1)()
    L.call("dcl_reinitialize_parallel", C.byref(cell), 1, 0, sw, None)
    assert cell.band_count == len(ref_band) and sw[0] == ref_sw
    np.testing.assert_array_equal(to_np(band), ref_band[:cap])
    assert to_np(dphi).tobytes() == ref_phi.tobytes()
    win2 = D.GridWindow((0, 0, 0), (20, 20, 1), (1.0, 1.0, 1.0))
    phi2 = torch.zeros((20, 20), dtype=torch.float64, device="cuda:0")
    empty = torch.zeros(0, dtype=torch.uint8, device="cuda:0")
    with pytest.raises(RuntimeError, match="mixes 2D and 3D"):
        M.reinitialize_parallel([(to_dev(phi), win, drecs, len(recs)), (phi2, win2, empty, 0)])


def test_parallel_large_window_flat_path(dev):
    """A window of 2^28 nodes or more (700 x 700 x 600) takes the large-window
    path (64-bit pointer loads instead of 32-bit buffer offsets).  phi = 3 x
    the distance to the plane z = 300.3, so the answer is closed-form: the
    exact distance on every node within band_width (1e-12, each z-plane
    uniform), +-PHI_INF elsewhere, and the band is exactly those nodes."""
    nx, ny, nz = 700, 700, 600
    assert nx * ny * nz >= 1 << 28
    win = D.GridWindow((0, 0, 0), (nx, ny, nz), (1.0, 1.0, 1.0))
    sdf = torch.arange(nz, dtype=torch.float64, device="cuda:0") - 300.3
    phi = (3.0 * sdf).view(nz, 1, 1).expand(nz, ny, nx).contiguous()
    band, nb = M.band_compact(phi, win, 8.0)
    recs = torch.empty(10 * nb * M.IRREGULAR_DTYPE.itemsize, dtype=torch.uint8, device="cuda:0")
    count = M.irregular_nodes_into(phi, win, band.contiguous(), recs)
    del band
    print(f"reinitialize_parallel window {win.n} records {count}", flush=True)
    (band2,), (sweeps,) = M.reinitialize_parallel([(phi, win, recs, count)])
    inside = sdf.abs() <= 6.0
    planes = phi.view(nz, -1)
    assert torch.equal(planes.amax(1), planes.amin(1))
    vals = planes[:, 0]
    assert (vals[inside] - sdf[inside]).abs().max().item() < 1e-12
    assert torch.all(vals[~inside].abs() == 7.5)
    assert band2.shape[0] == int(inside.sum()) * nx * ny
    assert sweeps <= 10
