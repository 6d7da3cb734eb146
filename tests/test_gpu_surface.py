"""Parity of the gfx950 surface quantities (dcl_surface_quantities,
IIMUpdateSurfaceQuantities_3D / _2D) with the oracle: every field of every
record bit for bit, NaNs at the same places (both evaluate the functionals, the interpolation, the
tangent frame and the rotation one IEEE double operation at a time in the
reference's operand order).  Cases: the records of a perturbed sphere in FP64
and FP32, a 2D Zalesak disk, points near and beyond the window's edges
(clamped reads), an empty list, and a sample of the 256^3 sphere."""
import numpy as np
import pytest
import torch

import oracle as O
import dcell_amd as D
from dcell_amd import configs
from dcell_amd import metrics as M
from helpers import owin, rounded, to_dev

pytestmark = pytest.mark.gpu


def _same(got, ref):
    """Bit for bit, except that a NaN's sign and payload are free (the GPU's
    default NaN is positive, x86's negative): NaNs must sit at the same
    places.  Constant fields (clamped reads beyond the window) give 0/0."""
    assert len(got) == len(ref)
    for f in got.dtype.names:
        a, b = np.array(got[f]), np.array(ref[f])
        na, nb = np.isnan(a), np.isnan(b)
        assert np.array_equal(na, nb), f
        a[na] = 0.0
        b[nb] = 0.0
        assert a.tobytes() == b.tobytes(), f


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_surface_3d_bit_exact(dev, dtype):
    rng = np.random.default_rng(31)
    n = 40
    h = 1.0 / (n - 1)
    win = D.GridWindow((3, -2, 7), (n, n, n), (h, 1.1 * h, 0.9 * h))
    phi = (configs.sphere_sdf(n, (0.5, 0.45, 0.52), 0.28) + 0.1 * rng.normal(size=(n, n, n))) * h
    phi = rounded(phi, dtype)
    band = O.band_compact(phi, owin(win), 3.0 * h)
    recs = O.irregular_nodes(phi, owin(win), band)
    assert len(recs) > 1000
    ref = O.surface_quantities(phi, owin(win), recs)
    got = M.surface_quantities(to_dev(phi, dtype), win, recs)
    _same(got, ref)


def test_surface_2d_bit_exact(dev):
    n = 128
    h = 1.0 / (n - 1)
    win = D.GridWindow((0, 0, 0), (n, n, 1), (h, h, h))
    phi = configs.zalesak_sdf(n) * h
    recs = O.irregular_nodes(phi, owin(win), O.band_compact(phi, owin(win), 3.0 * h))
    ref = O.surface_quantities(phi, owin(win), recs)
    got = M.surface_quantities(to_dev(phi), win, recs)
    _same(got, ref)


def test_surface_edges_and_empty(dev):
    rng = np.random.default_rng(5)
    n = 16
    win = D.GridWindow((-4, 2, 1), (n, n + 2, n - 1), (0.5, 0.5, 0.5))
    phi = rng.normal(size=(n - 1, n + 2, n))
    lo = np.array([-4, 2, 1], float)
    hi = lo + np.array([n, n + 2, n - 1]) - 1
    X = rng.uniform(lo - 3, hi + 3, size=(2000, 3))  # inside, on and beyond every face
    recs = np.zeros(len(X), dtype=O.IRREG_DTYPE)
    recs["X"] = X
    ref = O.surface_quantities(phi, owin(win), recs)
    got = M.surface_quantities(to_dev(phi), win, recs)
    _same(got, ref)
    none = M.surface_quantities(to_dev(phi), win, recs[:0])
    assert len(none) == 0


def test_surface_c2_size_sample(dev):
    n = 256
    win = D.GridWindow.unit_box(n)
    phi = configs.sphere_sdf(n) / (n - 1)
    rng = np.random.default_rng(17)
    band = O.band_compact(phi, owin(win), 2.0 / (n - 1))
    sample = band[np.sort(rng.choice(len(band), 4000, replace=False))]
    recs = O.irregular_nodes(phi, owin(win), sample)
    ref = O.surface_quantities(phi, owin(win), recs)
    got = M.surface_quantities(to_dev(phi), win, recs)
    _same(got, ref)
