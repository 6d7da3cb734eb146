"""Parity of the gfx950 irregular-node list (dcl_irregular_nodes) with the
oracle: the same records, in the same order, bit for bit (both evaluate the
projections' Newton iterations one IEEE double operation at a time in the
same order).  Cases: a perturbed sphere (many Newton failures exercise the
quadratic / linear fallbacks), FP32 input read as double, a 2D disk, band
nodes on the window's edge (clamped reads), a capacity below the count, and
an empty band.  Blocks whose records fit in LDS are staged there, others
write directly: the sphere cases take the first path, the random field the
second."""
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
    assert len(got) == len(ref)
    assert got.tobytes() == ref.tobytes()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_irregular_nodes_3d_bit_exact(dev, dtype):
    rng = np.random.default_rng(3)
    n = 40
    h = 1.0 / (n - 1)
    win = D.GridWindow((3, -2, 7), (n, n, n), (h, h, h))
    phi = configs.sphere_sdf(n, (0.5, 0.45, 0.52), 0.28) + 0.2 * rng.normal(size=(n, n, n))
    phi = rounded(phi, dtype)
    band = O.band_compact(phi, owin(win), 3.0)
    band = band[rng.permutation(len(band))]
    ref = O.irregular_nodes(phi, owin(win), band)
    got, count = M.irregular_nodes(to_dev(phi, dtype), win, to_dev(band, torch.int32))
    assert count == len(ref)
    _same(got, ref)
    # a capacity that ends inside a block's records
    cap = len(ref) // 2 + 37
    part, count = M.irregular_nodes(to_dev(phi, dtype), win, to_dev(band, torch.int32), capacity=cap)
    assert count == len(ref)
    _same(part, ref[:cap])


def test_irregular_nodes_2d_bit_exact(dev):
    n = 128
    h = 1.0 / (n - 1)
    win = D.GridWindow((0, 0, 0), (n, n, 1), (h, h, h))
    phi = configs.zalesak_sdf(n)
    band = O.band_compact(phi, owin(win), 3.0)
    ref = O.irregular_nodes(phi, owin(win), band)
    got, count = M.irregular_nodes(to_dev(phi), win, to_dev(band, torch.int32))
    assert count == len(ref) and len(ref) > 500
    _same(got, ref)


def test_irregular_nodes_edges_capacity_empty(dev):
    rng = np.random.default_rng(9)
    n = 16
    win = D.GridWindow((0, 0, 0), (n, n, n), (1.0, 1.0, 1.0))
    phi = rng.normal(size=(n, n, n))                     # sign changes everywhere
    nodes = np.argwhere(np.ones((n, n, n), bool))[:, ::-1].astype(np.int32)
    edge = nodes[np.any((nodes == 0) | (nodes == n - 1), axis=1)]
    band = np.concatenate([edge[:300], nodes[rng.choice(len(nodes), 300, replace=False)]])
    ref = O.irregular_nodes(phi, owin(win), band)
    got, count = M.irregular_nodes(to_dev(phi), win, to_dev(band, torch.int32))
    assert count == len(ref)
    _same(got, ref)
    cap = len(ref) // 3
    part, count = M.irregular_nodes(to_dev(phi), win, to_dev(band, torch.int32), capacity=cap)
    assert count == len(ref)
    _same(part, ref[:cap])
    empty = torch.zeros((0, 3), dtype=torch.int32, device="cuda:0")
    none, count = M.irregular_nodes(to_dev(phi), win, empty)
    assert count == 0 and len(none) == 0


def test_irregular_nodes_c2_size_sample(dev):
    """The 256^3 sphere of the headline workload: every band node within 2
    cells, GPU against the oracle on a seeded sample of 4000 of them."""
    n = 256
    win = D.GridWindow.unit_box(n)
    phi = configs.sphere_sdf(n)
    band = O.band_compact(phi, owin(win), 2.0)
    rng = np.random.default_rng(17)
    sample = band[np.sort(rng.choice(len(band), 4000, replace=False))]
    ref = O.irregular_nodes(phi, owin(win), sample)
    got, count = M.irregular_nodes(to_dev(phi), win, to_dev(sample, torch.int32))
    _same(got, ref)
