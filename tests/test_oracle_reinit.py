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

The parallel reinitialisation's restatement (reinitialize_sweep) is pinned
against the fast marching above and against exact distances:
- planes: the exact distance (1e-12), the same band;
- a sphere, a torus, two nearby spheres and a 2D circle: the same band set
  as the fast marching; on the sphere and circle at least 99% of the band
  values bit for bit and all within 2e-3; everywhere within 1.5e-2, while
  the distance error of both against the exact distance is ten times that
  and the same to 1e-3 (the sweeps are as accurate as the fast marching);
- a sweep that took nodes in the wrong order (a neighbour counted as
  accepted whatever its value) would miss these by far: a plant of that
  mistake is shown to fail them;
- the sweeps settle, and the sign flip flips the result.
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


# ---------------------------------------------------------------- sweeps


@pytest.mark.parametrize("axis,offset", [(0, 0.3), (2, -0.45)])
def test_sweep_plane_exact(axis, offset):
    n = 32
    w = Win((0, 0, 0), (n, n, n), (1, 1, 1))
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    sdf = (x, y, z)[axis] - (n / 2 + offset)
    phi = 3.0 * sdf
    recs = _records(phi, w)
    out, band, sweeps = O.reinitialize_sweep(phi, w, recs)
    inside = np.abs(sdf) <= BW
    assert np.max(np.abs(out[inside] - sdf[inside])) < 1e-12
    assert np.all(np.abs(out[~inside]) == INF)
    _, fband = O.reinitialize(phi, w, recs)
    assert {tuple(b) for b in band} == {tuple(b) for b in fband}
    assert sweeps <= 10


def _torus(n):
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    return np.hypot(np.hypot(x - 31.7, y - 32.4) - 18.0, z - 31.1) - 7.3


def _two_spheres(n):
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    s1 = np.sqrt((x - 20.2) ** 2 + (y - 31.5) ** 2 + (z - 32.3) ** 2) - 11.1
    s2 = np.sqrt((x - 44.6) ** 2 + (y - 32.1) ** 2 + (z - 31.7) ** 2) - 11.4
    return np.minimum(s1, s2)


def _sphere_sdf(n):
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    return np.sqrt((x - 31.3) ** 2 + (y - 30.6) ** 2 + (z - 32.2) ** 2) - 17.4


def _sweep_vs_fmm(phi, sdf, w, max_d=0.0):
    recs = _records(phi, w)
    a, fband = O.reinitialize(phi, w, recs)
    b, band, sweeps = O.reinitialize_sweep(phi, w, recs)
    ma, mb = np.abs(a) < INF, np.abs(b) < INF
    # the band sets agree but at the band's edge, where the two values
    # straddle band_width
    assert np.all(np.abs(a[ma & ~mb]) > BW - max_d) and np.all(np.abs(b[mb & ~ma]) > BW - max_d)
    assert (ma ^ mb).sum() <= 1e-4 * ma.sum()
    # the band: projection nodes in record order, then - side, + side in window order
    proj = recs[(recs["axis"] == -1)]["pos"]
    np.testing.assert_array_equal(band[:len(proj)], proj)
    assert len({tuple(q) for q in band} ^ {tuple(q) for q in fband}) == (ma ^ mb).sum()
    ma = ma & mb
    rest = band[len(proj):]
    idx = (rest[:, 2], rest[:, 1], rest[:, 0]) if a.ndim == 3 else (rest[:, 1], rest[:, 0])
    flat = np.ravel_multi_index(idx, a.shape)
    side = b.ravel()[flat] > 0
    assert not np.any(side[:-1] & ~side[1:])                      # all - before all +
    assert np.all(np.diff(flat[~side]) > 0) and np.all(np.diff(flat[side]) > 0)
    d = np.abs(a[ma] - b[ma])
    ea, eb = np.abs(a[ma] - sdf[ma]), np.abs(b[ma] - sdf[ma])
    return d, (a[ma] == b[ma]).mean(), ea, eb, sweeps


@pytest.mark.parametrize("shape,exact_frac,max_d", [
    ("sphere", 0.99, 2e-3), ("torus", 0.85, 1.5e-2), ("two_spheres", 0.8, 1.5e-2)])
def test_sweep_matches_fast_marching(shape, exact_frac, max_d):
    n = 64
    w = Win((0, 0, 0), (n, n, n), (1, 1, 1))
    sdf = {"sphere": _sphere_sdf, "torus": _torus, "two_spheres": _two_spheres}[shape](n)
    d, frac, ea, eb, sweeps = _sweep_vs_fmm(1.7 * sdf, sdf, w, max_d)
    assert frac >= exact_frac
    assert d.max() <= max_d
    # as accurate as the fast marching against the exact distance
    assert ea.max() > 10 * d.max() or shape == "sphere"
    assert abs(ea.max() - eb.max()) < 1e-3 and abs(ea.mean() - eb.mean()) < 1e-3
    assert sweeps < 64


def test_sweep_circle_2d_and_sign_flip():
    n = 96
    w = Win((0, 0, 0), (n, n, 1), (1, 1, 1))
    y, x = np.meshgrid(np.arange(n, dtype=float), np.arange(n, dtype=float), indexing="ij")
    sdf = np.hypot(x - 47.3, y - 49.1) - 30.2
    d, frac, ea, eb, _ = _sweep_vs_fmm(2.0 * sdf, sdf, w)
    assert frac >= 0.99 and d.max() < 1e-6
    recs = _records(2.0 * sdf, w)
    out, _, _ = O.reinitialize_sweep(2.0 * sdf, w, recs)
    recs2 = recs.copy()
    recs2["sign_center"] *= -1
    out2, _, _ = O.reinitialize_sweep(-2.0 * sdf, w, recs2)
    nz = out != 0
    np.testing.assert_array_equal(out2[nz], -out[nz])


def test_sweep_noisy_settles():
    """A bumpy, noisy field (many near-ties): the sweeps settle (the
    values-only-go-down phase ends the cycles) and keep the band set."""
    n = 64
    w = Win((0, 0, 0), (n, n, n), (1, 1, 1))
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    s = _sphere_sdf(n)
    phi = s * (1 + 0.3 * np.sin(x / 3.1) * np.cos(y / 2.3)) + 0.2 * np.random.default_rng(3).normal(size=s.shape)
    recs = _records(phi, w)
    a, _ = O.reinitialize(phi, w, recs)
    b, band, sweeps = O.reinitialize_sweep(phi, w, recs)
    assert sweeps < 64
    ma, mb = np.abs(a) < INF, np.abs(b) < INF
    assert (ma ^ mb).sum() <= 0.001 * ma.sum()
    m = ma & mb
    assert np.median(np.abs(a[m] - b[m])) < 1e-6
    with pytest.raises(RuntimeError, match="did not settle"):
        O.reinitialize_sweep(phi, w, recs, max_sweeps=5)


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_sweep_pins_catch_a_wrong_order(plant):
    """Plant a mistake in the oracle's sweep: count every usable neighbour as
    accepted before the node (1), replay pushes after the node's own value
    came up (2), or keep the last push instead of the least (3).  Each fails
    a pin of test_sweep_matches_fast_marching (or breaks the update)."""
    import ctypes as C
    from oracle import oracle as OO
    flag = C.c_int.in_dll(OO.lib(), "orc_sweep_plant")
    n = 64
    w = Win((0, 0, 0), (n, n, n), (1, 1, 1))
    caught = False
    for make, exact_frac, max_d in ((_sphere_sdf, 0.99, 2e-3), (_two_spheres, 0.8, 1.5e-2)):
        phi = 1.7 * make(n)
        recs = _records(phi, w)
        a, _ = O.reinitialize(phi, w, recs)
        flag.value = plant
        try:
            b, _, _ = O.reinitialize_sweep(phi, w, recs)
        except RuntimeError as e:
            caught = "NaN" in str(e)
            break
        finally:
            flag.value = 0
        m = (np.abs(a) < INF) & (np.abs(b) < INF)
        if (a[m] == b[m]).mean() < exact_frac or np.abs(a[m] - b[m]).max() > max_d:
            caught = True
            break
    assert caught
