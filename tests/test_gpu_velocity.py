"""dcl_get_velocity (LevelSetGetVelocity, LevelSetMethod/Advect.c:145-192)
against the oracle, exactly (a copy): windows inside the fluid domain,
across its low and high corners, and outside it, in FP64 and FP32, 3D and
2D, per node and as component planes; nodes outside the domain keep the
output's previous values; a fluid array that is not aligned to its records
is an error."""
import numpy as np
import pytest
import torch

import oracle as O
import dcell_amd as D
from helpers import owin, rounded, to_dev, to_np

pytestmark = pytest.mark.gpu

WINDOWS_3D = [((4, 5, 6), (20, 18, 16)), ((-3, -5, -2), (17, 19, 11)), ((30, 25, 20), (12, 15, 9)),
              ((50, 0, 0), (4, 4, 4))]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_get_velocity_3d(dev, dtype):
    rng = np.random.default_rng(2)
    fluid = rounded(rng.normal(size=(27, 31, 36, 4)), dtype)  # (Nz, Ny, Nx, p u v w)
    fd = to_dev(fluid, dtype)
    for p, n in WINDOWS_3D:
        win = D.GridWindow(p, n, (0.1, 0.1, 0.1))
        init = rounded(rng.normal(size=(n[2], n[1], n[0], 3)), dtype)
        out = to_dev(init, dtype)
        D.get_velocity(fd, win, out=out)
        ref = O.get_velocity(fluid, owin(win), init)
        assert np.array_equal(to_np(out).astype(np.float64), ref), (p, n)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_get_velocity_planes_3d(dev, dtype):
    """The component-plane form writes the same values, plane c holding
    component c of every node; nodes outside the domain keep their values."""
    rng = np.random.default_rng(4)
    fluid = rounded(rng.normal(size=(27, 31, 36, 4)), dtype)
    fd = to_dev(fluid, dtype)
    for p, n in WINDOWS_3D:
        win = D.GridWindow(p, n, (0.1, 0.1, 0.1))
        init = rounded(rng.normal(size=(n[2], n[1], n[0], 3)), dtype)
        out = to_dev(np.ascontiguousarray(np.moveaxis(init, -1, 0)), dtype)
        D.get_velocity(fd, win, out=out, planes=True)
        ref = O.get_velocity(fluid, owin(win), init)
        assert np.array_equal(np.moveaxis(to_np(out).astype(np.float64), 0, -1), ref), (p, n)


def test_get_velocity_planes_2d_feeds_advect(dev):
    """2D planes from the fluid, then one SL step: the same values as the
    per-node velocity."""
    rng = np.random.default_rng(5)
    fluid = rng.normal(size=(40, 50, 3))
    fd = to_dev(fluid)
    win = D.GridWindow((5, 7, 0), (20, 22, 1), (0.05, 0.05, 1.0))
    node = D.get_velocity(fd, win)
    planes = D.get_velocity(fd, win, planes=True)
    assert tuple(planes.shape) == (2, 22, 20)
    assert np.array_equal(np.moveaxis(to_np(planes), 0, -1), to_np(node))
    phi0 = to_dev(rng.normal(size=(22, 20)))
    a = D.advect_sl(phi0, node, win, 0.02, "cubic", "rk2")
    b = D.advect_sl(phi0, planes, win, 0.02, "cubic", "rk2", planes=True)
    assert torch.equal(a, b)


def test_get_velocity_2d(dev):
    rng = np.random.default_rng(3)
    fluid = rng.normal(size=(40, 50, 3))  # (Ny, Nx, p u v)
    for p, n in (((5, 7, 0), (20, 22, 1)), ((-4, 30, 0), (16, 15, 1))):
        win = D.GridWindow(p, n, (0.1, 0.1, 1.0))
        got = D.get_velocity(to_dev(fluid), win)
        ref = O.get_velocity(fluid, owin(win), np.zeros((n[1], n[0], 2)))
        assert np.array_equal(to_np(got), ref)


def test_get_velocity_alignment_error(dev):
    fd = torch.zeros(8 * 8 * 8 * 4 + 1, dtype=torch.float64, device="cuda:0")
    bad = fd[1:].view(8, 8, 8, 4)
    with pytest.raises(RuntimeError, match="aligned"):
        D.get_velocity(bad, D.GridWindow((0, 0, 0), (4, 4, 4), (1, 1, 1)))
