"""Parity of the gfx950 particle level set (dcl_pls_*) with the oracle.

Advection and radii: FP64 <= 1e-12 (relative to the window size for
positions); FP32 against the oracle on float32-rounded inputs.  Correction:
FP64 bit-exact (max/min and a single sqrt per corner); FP32 <= 1e-6 of
max |phi|.  Collisions (many particles on one corner) and the reference's
A.z = floor(X.y) variant are covered.
"""
import numpy as np
import pytest
import torch

import oracle as O
import dcell_amd as D
from dcell_amd import configs
from dcell_amd import pls as PL
from helpers import owin, random_velocity, rel_max, rounded, smooth_field, to_dev, to_np

pytestmark = pytest.mark.gpu


def _setup(dim, n, seed, per_cell=4):
    rng = np.random.default_rng(seed)
    h = 1.0 / (n - 1)
    p = (2, -3, 4)
    win = D.GridWindow(p, (n, n, n) if dim == 3 else (n, n, 1), (h, h, h))
    phi = smooth_field(n, dim, rng)
    vel = random_velocity(n, dim, rng, 1.0)
    nodes = np.argwhere(np.abs(phi) < 3.0)[:, ::-1]
    P = np.zeros((len(nodes) * per_cell, 4))
    P[:, :dim] = np.repeat(nodes, per_cell, 0) + rng.uniform(-0.5, 0.5, (len(nodes) * per_cell, dim))
    P[:, :dim] = np.clip(P[:, :dim], 0, n - 1) + np.array(p[:dim])
    if dim == 2:
        P[:, 2] = p[2]
    sgn = rng.choice([-1.0, 1.0], size=len(P))
    P[:, 3] = sgn * rng.uniform(0.1, 0.5, size=len(P))
    return win, phi, vel, P, rng


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-6)])
@pytest.mark.parametrize("dim", [2, 3])
def test_advect_particles(dev, dtype, tol, dim):
    win, phi, vel, P, _ = _setup(dim, 40 if dim == 3 else 96, 40 + dim)
    P, vel = rounded(P, dtype), rounded(vel, dtype)
    dt = 1.5 * win.d[0]
    Pd = to_dev(P, dtype)
    PL.advect_particles(Pd, to_dev(vel, dtype), win, win.d, dt, dim)
    ref = O.pls_advect(P, vel, owin(win), win.d, dt, dim)
    got = to_np(Pd)
    assert rel_max(got[:, :dim], ref[:, :dim], float(max(win.n))) <= tol
    assert np.array_equal(got[:, 3], ref[:, 3])


def _interp(phi, win, P, dim):
    """Plain (bi/tri)linear phi at particle positions (numpy)."""
    pos = P[:, :3] - np.array(win.p, dtype=np.float64)
    if dim == 3:
        return configs.trilinear_numpy(phi, pos)
    return configs.trilinear_numpy(np.stack([phi, phi]), np.c_[pos[:, :2], np.zeros(len(pos))])


def _off_threshold(P, phi, win, dim, margin=1e-4):
    """Drop particles whose escape test sits within margin of a threshold, where
    float32 and float64 arithmetic may decide differently."""
    d = _interp(phi, win, P, dim)
    keep = (np.abs(d) > margin) & (np.abs(np.abs(d) - np.abs(P[:, 3])) > margin)
    return P[keep]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("flags", [0, PL.REF_AZ])
def test_correction_dense_and_band(dev, dtype, dim, flags):
    if dim == 2 and flags:
        pytest.skip("the A.z variant is 3D only")
    win, phi, _, P, rng = _setup(dim, 32 if dim == 3 else 80, 50 + dim, per_cell=8)
    phi, P = rounded(phi, dtype), rounded(P, dtype)
    # many particles sit on the wrong side: flip signs of a third of them
    flip = rng.random(len(P)) < 0.33
    P[flip, 3] *= -1
    if dtype == torch.float32:
        P = _off_threshold(P, phi, win, dim)
    ws = PL.CorrectionWorkspace(win.nodes, 8 * len(P), dtype)
    # dense: every node merges
    pd = to_dev(phi, dtype)
    PL.correct(to_dev(P, dtype), pd, win, ws, flags=flags)
    ref = O.pls_correct(P, phi, owin(win), None, flags, dim)
    got = to_np(pd)
    assert (got != phi).sum() > 0
    if dtype == torch.float64:
        assert np.array_equal(got, ref)
    else:
        assert rel_max(got, ref, np.max(np.abs(phi))) <= 1e-6
    # band form (a random half of the nodes), compact values and scatter, same workspace again
    nodes = np.argwhere(np.ones(phi.shape))[:, ::-1]
    sel = rng.random(len(nodes)) < 0.5
    band = np.zeros((sel.sum(), 3), dtype=np.int32)
    band[:, :dim] = nodes[sel] + np.array(win.p[:dim])
    band[:, 2] = band[:, 2] if dim == 3 else win.p[2]
    ref = O.pls_correct(P, phi, owin(win), band, flags, dim)
    vals = to_np(PL.correct(to_dev(P, dtype), to_dev(phi, dtype), win, ws, band=to_dev(band, torch.int32),
                            flags=flags, band_values=True))
    idx = tuple(band[:, a] - win.p[a] for a in range(dim))[::-1]
    if dtype == torch.float64:
        assert np.array_equal(vals, ref[idx])
    else:
        assert rel_max(vals, ref[idx], np.max(np.abs(phi))) <= 1e-6
    pd = to_dev(phi, dtype)
    PL.correct(to_dev(P, dtype), pd, win, ws, band=to_dev(band, torch.int32), flags=flags)
    assert rel_max(to_np(pd), ref, np.max(np.abs(phi))) <= (0 if dtype == torch.float64 else 1e-6)
    ws.close()


def test_correction_collisions_and_reuse(dev):
    """4000 escaped particles on the corners of one cell: the scatter is a pure
    max/min, so the result equals the serial oracle exactly; the workspace is
    back to neutral afterwards (a second call with no escapes changes nothing)."""
    n = 12
    win = D.GridWindow((0, 0, 0), (n, n, n), (0.1, 0.1, 0.1))
    phi = np.full((n, n, n), 2.0)
    rng = np.random.default_rng(7)
    P = np.zeros((4000, 4))
    P[:, :3] = 5 + rng.random((4000, 3))
    P[:, 3] = -rng.uniform(0.1, 0.5, 4000)
    ws = PL.CorrectionWorkspace(win.nodes, 8 * 4000)
    pd = to_dev(phi)
    PL.correct(to_dev(P), pd, win, ws)
    ref = O.pls_correct(P, phi, owin(win), None, 0, 3)
    assert np.array_equal(to_np(pd), ref)
    assert (ref != phi).sum() == 8
    calm = P.copy()
    calm[:, 3] *= -1                                   # same side: nobody escapes
    pd = to_dev(phi)
    PL.correct(to_dev(calm), pd, win, ws)
    assert np.array_equal(to_np(pd), phi)
    ws.close()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("dim", [2, 3])
def test_adjust_radii(dev, dtype, dim):
    win, phi, _, P, _ = _setup(dim, 30 if dim == 3 else 70, 60 + dim)
    phi, P = rounded(phi, dtype), rounded(P, dtype)
    Pd = to_dev(P, dtype)
    PL.adjust_radii(Pd, to_dev(phi, dtype), win)
    ref = O.pls_adjust_radii(P, phi, owin(win), 0.1, 0.5, dim)
    got = to_np(Pd)
    assert np.array_equal(got[:, :3], P[:, :3])
    assert rel_max(got[:, 3], ref[:, 3], 0.5) <= (1e-12 if dtype == torch.float64 else 1e-6)


def test_workspace_limits_are_errors(dev):
    win = D.GridWindow.unit_box(16)
    ws = PL.CorrectionWorkspace(100, 80)
    P = torch.zeros((20, 4), dtype=torch.float64, device="cuda:0")
    phi = torch.zeros(win.shape, dtype=torch.float64, device="cuda:0")
    with pytest.raises(D._lib.DclError):
        PL.correct(P, phi, win, ws)                    # window larger than the workspace
    ws.close()
