"""Pins of the oracle's irregular-node list (LevelSetUpdateIrregularNodeList,
LevelSetMethod/IrregularNodes.c:8-140, IrregularNodes3D.c:6-117, with the
orthogonal projections of OrthogonalProjection3D.c / OrthogonalProjection.c).

- The projections' tricubic / bicubic evaluator against the reference's own
  TricubicInverse.h + TricubicInterpolation.h compiled from the reference
  tree, and against the independent Catmull-Rom evaluator.
- Planes (lines): the tricubic reproduces a linear field, so Newton lands on
  the exact foot point of the perpendicular, and every crossing record lies
  on the plane.
- Spheres (circles): projections land on the analytic surface up to the
  interpolation error, at the analytic distance from their node.
- Record counts and order against an independent count of the sign changes.
"""
import numpy as np
import pytest

import oracle as O
from oracle import Win
from dcell_amd import configs


def _band(phi, w, width=3.0):
    return O.band_compact(phi, w, width)


# --------------------------------------------------------------- evaluators
def test_projection_tricubic_matches_reference_header_and_cr():
    rng = np.random.default_rng(21)
    worst_ref = worst_cr = 0.0
    for _ in range(300):
        st = rng.normal(size=64) * rng.uniform(0.1, 10)
        x, y, z = rng.random(3)
        a = O.irr_tricubic(st, x, y, z)
        worst_cr = max(worst_cr, np.max(np.abs(a - O.cr_stencil3(st, x, y, z))) / max(1.0, np.abs(st).max()))
        if O.ref_tricubic() is not None:
            b = O.ref_tricubic_eval(st, x, y, z)
            worst_ref = max(worst_ref, np.max(np.abs(a - b)) / max(1.0, np.abs(st).max()))
    assert worst_cr < 1e-12, worst_cr
    assert worst_ref < 1e-12, worst_ref


def test_projection_bicubic_is_z_constant_tricubic():
    rng = np.random.default_rng(22)
    for _ in range(100):
        st2 = rng.normal(size=16)
        x, y = rng.random(2)
        st3 = np.tile(st2, 4)  # the same 4x4 layer at every z
        a = O.irr_bicubic(st2, x, y)
        b = O.irr_tricubic(st3, x, y, rng.random())
        np.testing.assert_allclose(a, b[:3], rtol=0, atol=1e-12 * max(1.0, np.abs(st2).max()))
        assert abs(b[3]) < 1e-12 * max(1.0, np.abs(st2).max())


# --------------------------------------------------------------- planes
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_plane_projection_exact_and_crossings_on_plane(seed):
    rng = np.random.default_rng(seed)
    n = 20
    w = Win((2, -3, 5), (n, n, n), (1.0, 1.0, 1.0))
    nrm = rng.normal(size=3)
    nrm /= np.linalg.norm(nrm)
    z, y, x = np.meshgrid(np.arange(n) + 5, np.arange(n) - 3, np.arange(n) + 2, indexing="ij")
    off = -(nrm @ (np.array([2, -3, 5]) + n / 2)) + rng.uniform(-0.5, 0.5)
    phi = nrm[0] * x + nrm[1] * y + nrm[2] * z + off
    band = _band(phi, w)
    inner = np.all((band - np.array([2, -3, 5]) >= 2) & (band - np.array([2, -3, 5]) <= n - 3), axis=1)
    recs = O.irregular_nodes(phi, w, band[inner])
    assert len(recs) > 0
    X = recs["X"]
    on = X @ nrm + off
    assert np.max(np.abs(on)) < 1e-11
    pr = recs[recs["shift"] == -1]
    pos = pr["pos"].astype(float)
    foot = pos - (pos @ nrm + off)[:, None] * nrm[None, :]
    assert np.max(np.abs(pr["X"] - foot)) < 1e-11
    np.testing.assert_allclose(pr["op"], pr["X"] - pos, atol=1e-12)
    # crossing records: d in [0, 1], sign_face flips at d = 1/2
    cr = recs[recs["shift"] >= 0]
    assert np.all((cr["d"] >= 0) & (cr["d"] <= 1))
    assert np.all(cr["sign_face"] == np.where(cr["d"] < 0.5, -cr["sign_center"], cr["sign_center"]))


def test_line_2d_projection_exact():
    n = 40
    w = Win((0, 0, 0), (n, n, 1), (1.0, 1.0, 1.0))
    a = np.array([0.6, 0.8])
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    off = -(a @ np.array([n / 2 + 0.3, n / 2 - 0.1]))
    phi = a[0] * x + a[1] * y + off
    band = _band(phi, w)
    keep = np.all((band[:, :2] >= 2) & (band[:, :2] <= n - 3), axis=1)
    recs = O.irregular_nodes(phi, w, band[keep])
    X = recs["X"][:, :2]
    assert np.max(np.abs(X @ a + off)) < 1e-11
    pr = recs[recs["shift"] == -1]
    pos = pr["pos"][:, :2].astype(float)
    # the nearer of the bicubic foot point (distance D) and the linear axis
    # crossing (distance D / |a_axis| > D for this slanted line): the foot point
    foot = pos - (pos @ a + off)[:, None] * a[None, :]
    assert np.max(np.abs(pr["X"][:, :2] - foot)) < 1e-11
    assert np.all(pr["X"][:, 2] == 0)


# --------------------------------------------------------------- spheres
def test_sphere_projection_on_surface():
    n = 48
    h = 1.0 / (n - 1)
    w = Win((0, 0, 0), (n, n, n), (h, h, h))
    c = np.array([0.52, 0.47, 0.5])
    r = 0.27
    phi = configs.sphere_sdf(n, tuple(c), r)
    recs = O.irregular_nodes(phi, w, _band(phi, w))
    pr = recs[recs["shift"] == -1]
    assert len(pr) > 1000
    rad = np.linalg.norm(pr["X"] - c / h, axis=1)
    assert np.max(np.abs(rad - r / h)) < 2e-3          # on the surface (cells)
    pos = pr["pos"].astype(float)
    dist = np.linalg.norm(pr["X"] - pos, axis=1)
    exact = np.abs(np.linalg.norm(pos - c / h, axis=1) - r / h)
    assert np.max(np.abs(dist - exact)) < 2e-3         # at the distance of the node


def test_circle_projection_on_curve():
    n = 96
    h = 1.0 / (n - 1)
    w = Win((0, 0, 0), (n, n, 1), (h, h, h))
    phi = configs.circle_sdf(n, (0.48, 0.53), 0.31)
    recs = O.irregular_nodes(phi, w, _band(phi, w))
    pr = recs[recs["shift"] == -1]
    rad = np.hypot(pr["X"][:, 0] - 0.48 / h, pr["X"][:, 1] - 0.53 / h)
    assert np.max(np.abs(rad - 0.31 / h)) < 2e-3


# --------------------------------------------------------------- structure
def _brute_counts3(phi, band, p):
    """Records per band node from the sign-change conditions, written
    independently (numpy, whole-array)."""
    b = band - np.asarray(p)
    P = lambda dx, dy, dz: phi[b[:, 2] + dz, b[:, 1] + dy, b[:, 0] + dx]  # noqa: E731
    c = P(0, 0, 0)
    star = np.zeros(len(b), bool)
    for d in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
        star |= c * P(*d) <= 0
    cnt = star.astype(int)
    faces = [(-1, 0, 0), (0, -1, 0), (0, 0, -1)]
    for g in faces:
        cnt += (c * P(*g) <= 0).astype(int)
    for g in faces:
        hi = (P(*g) + c) / 2
        for lo in faces:
            if lo == g:
                continue
            low = (P(g[0] + lo[0], g[1] + lo[1], g[2] + lo[2]) + P(*lo)) / 2
            cnt += (hi * low <= 0).astype(int)
    return cnt


def test_record_counts_and_order():
    n = 32
    rng = np.random.default_rng(5)
    w = Win((0, 0, 0), (n, n, n), (1.0, 1.0, 1.0))
    phi = configs.sphere_sdf(n, (0.5, 0.5, 0.5), 0.3) + 0.3 * rng.normal(size=(n, n, n))
    band = _band(phi, w)
    band = band[np.all((band >= 2) & (band <= n - 3), axis=1)]
    band = band[rng.permutation(len(band))]
    recs = O.irregular_nodes(phi, w, band)
    cnt = _brute_counts3(phi, band, (0, 0, 0))
    assert len(recs) == cnt.sum()
    # records come node by node in band order, projection first
    owner = np.repeat(np.arange(len(band)), cnt)
    np.testing.assert_array_equal(recs["pos"], band[owner])
    first = np.concatenate([[True], owner[1:] != owner[:-1]])
    assert np.all(recs["shift"][~first] != -1)
    # a prefix of the band gives the prefix of the records
    some = O.irregular_nodes(phi, w, band[:50])
    assert len(some) == cnt[:50].sum()
    assert some.tobytes() == recs[:len(some)].tobytes()
