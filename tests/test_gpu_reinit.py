"""Parity of the gfx950 reinitialisation (dcl_reinitialize, LevelSetReinitialize)
with the oracle: phi bit for bit over the whole window and the band in the
same order (both run the reference's heap with its comparisons and the same
second-order updates in the same operand order).  Cases: a scaled sphere, a
perturbed sphere (many second-order / first-order switches and near-ties),
an axis plane (exact ties along the plane), a 2D circle, a band that
touches the window's edge, a batch of level sets of different sizes run
side by side, an empty record list, and a heap too small (an error)."""
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


def _check(phi, win, recs, got_phi, got_band):
    ref_phi, ref_band = O.reinitialize(phi, owin(win), recs)
    assert to_np(got_phi).tobytes() == ref_phi.tobytes()
    np.testing.assert_array_equal(to_np(got_band), ref_band)


def _sphere(n, c, r, scale, noise=0.0, seed=0):
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    sdf = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r
    return scale * sdf + noise * np.random.default_rng(seed).normal(size=sdf.shape)


@pytest.mark.parametrize("noise", [0.0, 0.3])
def test_reinit_sphere_bit_exact(dev, noise):
    n = 48
    win = D.GridWindow((2, -3, 5), (n, n, n), (1.0, 1.0, 1.0))
    phi = _sphere(n, (23.3, 22.6, 24.2), 13.4, 2.0, noise, seed=4)
    recs, drecs = _case(phi, win)
    dphi = to_dev(phi)
    (band,) = M.reinitialize([(dphi, win, drecs, len(recs))])
    _check(phi, win, recs, dphi, band)


def test_reinit_plane_and_edge(dev):
    n = 40
    win = D.GridWindow((0, 0, 0), (n, n, 24), (1.0, 1.0, 1.0))
    z, y, x = np.meshgrid(np.arange(24.0), np.arange(float(n)), np.arange(float(n)), indexing="ij")
    phi = 3.0 * (y - 20.3)  # a plane; the band reaches the x and z faces of the window
    recs, drecs = _case(phi, win)
    dphi = to_dev(phi)
    (band,) = M.reinitialize([(dphi, win, drecs, len(recs))])
    _check(phi, win, recs, dphi, band)


def test_reinit_circle_2d(dev):
    n = 96
    win = D.GridWindow((0, 0, 0), (n, n, 1), (1.0, 1.0, 1.0))
    y, x = np.meshgrid(np.arange(float(n)), np.arange(float(n)), indexing="ij")
    phi = 2.0 * (np.hypot(x - 47.3, y - 49.1) - 30.2)
    recs, drecs = _case(phi, win)
    dphi = to_dev(phi)
    (band,) = M.reinitialize([(dphi, win, drecs, len(recs))])
    _check(phi, win, recs, dphi, band)


def test_reinit_batch(dev):
    cells, host = [], []
    for i, n in enumerate((24, 40, 33, 56)):
        win = D.GridWindow((i, -i, 2 * i), (n, n + 1, n - 1), (1.0, 1.0, 1.0))
        phi = np.ascontiguousarray(_sphere(n, (n / 2 + 0.3, n / 2 - 0.2, n / 2), n / 4 + 0.1, 1.5, 0.1, seed=i)
                                   [: n - 1, :, :])
        phi = np.concatenate([phi, phi[:, :1, :]], axis=1)  # (n-1, n+1, n)
        recs, drecs = _case(phi, win)
        dphi = to_dev(phi)
        cells.append((dphi, win, drecs, len(recs)))
        host.append((phi, win, recs))
    bands = M.reinitialize(cells)
    for (phi, win, recs), (dphi, _, _, _), band in zip(host, cells, bands):
        _check(phi, win, recs, dphi, band)


def test_reinit_empty_and_heap_error(dev):
    n = 20
    win = D.GridWindow((0, 0, 0), (n, n, n), (1.0, 1.0, 1.0))
    phi = _sphere(n, (9.5, 9.5, 9.5), 5.2, 1.0)
    dphi = to_dev(phi)
    empty = torch.zeros(0, dtype=torch.uint8, device="cuda:0")
    (band,) = M.reinitialize([(dphi, win, empty, 0)])
    assert band.shape[0] == 0
    assert np.array_equal(to_np(dphi), np.where(phi > 0, 7.5, -7.5))
    recs, drecs = _case(phi, win)
    with pytest.raises(RuntimeError, match="heap full"):
        M.reinitialize([(to_dev(phi), win, drecs, len(recs))], heap_capacity=8)
