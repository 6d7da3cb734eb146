"""Pins of the oracle's reinitialisation (LevelSetReinitialize,
LevelSetMethod/FastMarching.c:10-146, second-order updates of
FastMarching3D.c / FastMarching2D.c, heap of Common/Heap.c).

- Axis-aligned planes: the projections are exact (linear field), the
  second-order one-sided differences are exact on a linear distance, and the
  neighbours along the plane tie, so every band value is the exact signed
  distance (1e-12); the band is exactly the nodes within band_width.
- Spheres and circles given as a scaled distance (2x): the result is the
  distance to the surface to a few hundredths of a cell, signs kept.
- Structure: the band starts with the projection records' nodes in record
  order, holds no node twice, and is exactly the set of nodes whose new phi
  is not +-PHI_INF; every value stays within band_width (or PHI_INF).
- The - and + passes are independent: flipping the sign of phi (and of the
  records' sign_center) flips the result.
"""
import numpy as np
import pytest

import oracle as O
from oracle import Win

BW, INF = 6.0, 7.5


def _records(phi, w, width=8.0):
    return O.irregular_nodes(phi, w, O.band_compact(phi, w, width))


@pytest.mark.parametrize("axis,offset", [(0, 0.3), (1, -0.45), (2, 0.0)])
def test_axis_plane_exact(axis, offset):
    n = 32
    w = Win((0, 0, 0), (n, n, n), (1, 1, 1))
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    coord = (x, y, z)[axis]
    sdf = coord - (n / 2 + offset)
    phi = 3.0 * sdf  # not a distance
    out, band = O.reinitialize(phi, w, _records(phi, w))
    inside = np.abs(sdf) <= BW
    assert np.max(np.abs(out[inside] - sdf[inside])) < 1e-12
    assert np.all(np.abs(out[~inside]) == INF)
    assert len(band) == inside.sum()


def test_sphere_scaled_distance():
    n = 64
    w = Win((0, 0, 0), (n, n, n), (1, 1, 1))
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    c = np.array([31.3, 30.6, 32.2])
    sdf = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - 17.4
    phi = 2.0 * sdf
    recs = _records(phi, w)
    out, band = O.reinitialize(phi, w, recs)
    m = np.abs(out) < INF
    assert np.abs(out[m] - sdf[m]).max() < 0.06
    assert np.all(np.sign(out[m]) == np.sign(sdf[m]))
    # structure
    proj = recs[recs["axis"] == -1]["pos"]
    np.testing.assert_array_equal(band[:len(proj)], proj)
    assert len({tuple(b) for b in band}) == len(band)
    got = {tuple(b) for b in band}
    want = {(int(i), int(j), int(k)) for k, j, i in zip(*np.nonzero(m))}
    assert got == want
    assert np.all(np.abs(out[m]) <= BW)
    # the sign passes are mirror images
    recs2 = recs.copy()
    recs2["sign_center"] *= -1
    out2, band2 = O.reinitialize(-phi, w, recs2)
    nz = out != 0
    np.testing.assert_array_equal(out2[nz], -out[nz])


def test_circle_2d():
    n = 96
    w = Win((0, 0, 0), (n, n, 1), (1, 1, 1))
    y, x = np.meshgrid(np.arange(n, dtype=float), np.arange(n, dtype=float), indexing="ij")
    sdf = np.hypot(x - 47.3, y - 49.1) - 30.2
    phi = 2.0 * sdf
    out, band = O.reinitialize(phi, w, _records(phi, w))
    m = np.abs(out) < INF
    assert m.sum() == len(band)
    assert np.abs(out[m] - sdf[m]).max() < 0.06
