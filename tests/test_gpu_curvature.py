"""Parity of the gfx950 curvature stencils (dcl_curvature) with the oracle.

FP64: max |kappa_gpu - kappa_oracle| <= 1e-12 max |kappa| over nodes with
|phi| < 6 (the interface band; away from it a random field's gradient can
vanish and kappa is ill-conditioned).  FP32: <= 5e-5 max |kappa| against the
oracle on float32-rounded phi.
"""
import numpy as np
import pytest
import torch

import oracle as O
import dcell_amd as D
from dcell_amd import configs
from helpers import owin, rel_max, rounded, smooth_field, to_dev, to_np

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 5e-5)])
@pytest.mark.parametrize("dim,stencil,kernel", [(3, "std", "tiled"), (3, "std", "gather"), (2, "std", "tiled"),
                                                (2, "lss", "tiled")])
def test_curvature_dense_and_band(dev, monkeypatch, dtype, tol, dim, stencil, kernel):
    # kernel: the 3D plane-ring kernel, or (DCL_DENSE_KERNEL=gather) the per-node
    # kernel that windows of 2 GiB and more take
    monkeypatch.setenv("DCL_DENSE_KERNEL", kernel)
    rng = np.random.default_rng(31 + dim)
    n = 36 if dim == 3 else 80
    h = 1.0 / (n - 1)
    win = D.GridWindow((4, -7, 2), (n, n, n) if dim == 3 else (n, n, 1), (h, 0.8 * h, 1.1 * h))
    phi = rounded(smooth_field(n, dim, rng), dtype)
    st = O.CURV_LSS if stencil == "lss" else O.CURV_STD
    k = to_np(D.curvature(to_dev(phi, dtype), win, stencil))
    ref = O.curvature(phi, owin(win), st)
    sel = np.abs(phi) < 6
    assert rel_max(k[sel], ref[sel], np.max(np.abs(ref[sel]))) <= tol
    assert np.array_equal(np.isfinite(k), np.isfinite(ref))
    # band form at random nodes, edges included (kappa = 0 where the stencil leaves the window)
    nodes = rng.integers(0, n, size=(500, dim))
    band = np.zeros((500, 3), dtype=np.int32)
    band[:, :dim] = nodes + np.array(win.p[:dim])
    band[:, 2] = band[:, 2] if dim == 3 else win.p[2]
    kb = to_np(D.curvature(to_dev(phi, dtype), win, stencil, band=to_dev(band, torch.int32),
                           band_values=True))
    rb = O.curvature(phi, owin(win), st, band=band)
    assert rel_max(kb, rb, np.max(np.abs(ref[sel]))) <= tol


def test_curvature_sphere_full_size(dev):
    """256^3 sphere (the C2 initial state): kappa -> 2/r near the interface,
    and GPU == oracle on 20000 sampled nodes."""
    n = 256
    h = 1.0 / (n - 1)
    win = D.GridWindow.unit_box(n)
    phi = configs.sphere_sdf(n, (0.5, 0.5, 0.5), 0.3)
    k = to_np(D.curvature(to_dev(phi), win))
    sel = np.abs(phi) < 0.5
    assert np.max(np.abs(k[sel] - 2 / (0.3 + phi[sel] * h))) < 1e-3 * 2 / 0.3
    rng = np.random.default_rng(3)
    band = rng.integers(0, n, size=(20000, 3)).astype(np.int32)
    ref = O.curvature(phi, owin(win), band=band)
    assert rel_max(k[band[:, 2], band[:, 1], band[:, 0]], ref, 2 / 0.3) <= 1e-12


def test_curvature_zero_gradient_is_zero(dev):
    n = 12
    win = D.GridWindow.unit_box(n)
    flat = np.full((n, n, n), 1.5)
    assert np.all(to_np(D.curvature(to_dev(flat), win)) == 0)
    win2 = D.GridWindow((0, 0, 0), (n, n, 1), (0.1, 0.1, 0.1))
    assert np.all(to_np(D.curvature(to_dev(flat[0]), win2)) == 0)
