"""Pins of the oracle's surface quantities at irregular nodes
(IIMUpdateSurfaceQuantities_3D / _2D, IIM/SurfaceQuantities.c:31-118).

- Quadratic fields: central differences are exact on quadratics and the
  trilinear (bilinear) interpolation of a linear function is exact, so the
  normal at any X is the exact unit gradient there, and the rotated Hessian
  gives k_nn = -s.A.s / |g|, k_tt = -r.A.r / |g|, k_nt = -r.A.s / |g| with
  the constant Hessian A; k_nn + k_tt = -(tr A - n.A.n) / |g| holds for any
  tangent frame.
- The tangent frame: orthonormal, r = n x s, and s in the plane the
  reference's branch picks (s.z = 0 when nx^2 + ny^2 >= nx^2 + nz^2, else
  s.y = 0).
- At a node, k equals the node's 19-point (9-point) curvature, which the
  main oracle tests pin; on spheres (circles) the normals are radial, k is
  2/r (1/r) and k_nn = k_tt = -1/r, k_nt = 0, to the discretisation error.
"""
import numpy as np
import pytest

import oracle as O
from oracle import Win
from dcell_amd import configs


def _quadratic(n, rng, h):
    A = rng.normal(size=(3, 3))
    A = A + A.T
    g = rng.normal(size=3) * 3
    c0 = rng.normal()
    z, y, x = np.meshgrid(*(np.arange(n) * h,) * 3, indexing="ij")
    P = np.stack([x, y, z], -1)
    phi = c0 + P @ g + 0.5 * np.einsum("...i,ij,...j->...", P, A, P)
    return phi, A, g


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_quadratic_3d_exact_normal_and_hessian(seed):
    rng = np.random.default_rng(seed)
    n = 12
    h = (0.1, 0.13, 0.08)
    w = Win((0, 0, 0), (n, n, n), h)
    A = rng.normal(size=(3, 3))
    A = A + A.T
    g = rng.normal(size=3) * 3
    z, y, x = np.meshgrid(*(np.arange(n) * h[a] for a in (2, 1, 0)), indexing="ij")
    phi = 0.3 + g[0] * x + g[1] * y + g[2] * z + 0.5 * (
        A[0, 0] * x * x + A[1, 1] * y * y + A[2, 2] * z * z + 2 * (A[0, 1] * x * y + A[0, 2] * x * z + A[1, 2] * y * z))
    X = rng.uniform(2, n - 3, size=(200, 3))
    sq = O.surface_at(phi, w, X)
    P = X * np.array(h)
    grad = g + P @ A
    gn = np.linalg.norm(grad, axis=1)
    nrm = grad / gn[:, None]
    scale = np.abs(A).max() / gn.min()
    np.testing.assert_allclose(sq["n"], nrm, atol=1e-12)
    s, r = sq["s"], sq["r"]
    knn = -np.einsum("ni,ij,nj->n", s, A, s) / gn
    ktt = -np.einsum("ni,ij,nj->n", r, A, r) / gn
    knt = -np.einsum("ni,ij,nj->n", r, A, s) / gn
    np.testing.assert_allclose(sq["k_nn"], knn, atol=1e-10 * scale)
    np.testing.assert_allclose(sq["k_tt"], ktt, atol=1e-10 * scale)
    np.testing.assert_allclose(sq["k_nt"], knt, atol=1e-10 * scale)
    # frame-free: the tangential trace of the Hessian
    tr = -(np.trace(A) - np.einsum("ni,ij,nj->n", nrm, A, nrm)) / gn
    np.testing.assert_allclose(sq["k_nn"] + sq["k_tt"], tr, atol=1e-10 * scale)


def test_tangent_frame():
    rng = np.random.default_rng(7)
    n = 24
    h = 1.0 / (n - 1)
    w = Win((0, 0, 0), (n, n, n), (h, h, h))
    phi = configs.sphere_sdf(n, (0.5, 0.5, 0.5), 0.3) + 0.02 * rng.normal(size=(n, n, n))
    X = rng.uniform(3, n - 4, size=(500, 3))
    sq = O.surface_at(phi, w, X)
    N, S, R = sq["n"], sq["s"], sq["r"]
    for a, b in ((N, S), (N, R), (S, R)):
        assert np.abs((a * b).sum(1)).max() < 1e-14
    for a in (N, S, R):
        assert np.abs(np.linalg.norm(a, axis=1) - 1).max() < 1e-14
    np.testing.assert_allclose(R, np.cross(N, S), atol=1e-14)
    first = N[:, 0] ** 2 + N[:, 1] ** 2 >= N[:, 0] ** 2 + N[:, 2] ** 2
    assert np.all(S[first, 2] == 0) and np.all(S[~first, 1] == 0)
    assert first.any() and (~first).any()


def test_k_at_nodes_is_node_curvature():
    rng = np.random.default_rng(8)
    n = 20
    h = 1.0 / (n - 1)
    w = Win((0, 0, 0), (n, n, n), (h, h, h))
    phi = configs.sphere_sdf(n, (0.5, 0.45, 0.5), 0.3) + 0.05 * rng.normal(size=(n, n, n))
    nodes = rng.integers(2, n - 2, size=(50, 3))
    sq = O.surface_at(phi, w, nodes.astype(float))
    ref = np.array([O.curv_at(phi, w, *q) for q in nodes])
    np.testing.assert_allclose(sq["k"], ref, rtol=1e-13, atol=1e-13)


def test_sphere():
    n = 48
    h = 1.0 / (n - 1)
    w = Win((0, 0, 0), (n, n, n), (h, h, h))
    c = np.array([0.5, 0.48, 0.52])
    r = 0.3
    phi = configs.sphere_sdf(n, tuple(c), r) * h
    recs = O.irregular_nodes(phi, w, O.band_compact(phi, w, 3 * h))
    sq = O.surface_quantities(phi, w, recs)
    X = recs["X"] * h
    radial = (X - c) / np.linalg.norm(X - c, axis=1)[:, None]
    assert np.abs(sq["n"] - radial).max() < 2e-3
    assert np.abs(sq["k"] * r / 2 - 1).max() < 1e-2
    assert np.abs(sq["k_nn"] * r + 1).max() < 1e-2
    assert np.abs(sq["k_tt"] * r + 1).max() < 1e-2
    assert np.abs(sq["k_nt"] * r).max() < 1e-2


def test_quadratic_2d_and_circle():
    rng = np.random.default_rng(4)
    n = 30
    h = (0.07, 0.09)
    w = Win((0, 0, 0), (n, n, 1), (h[0], h[1], 1.0))
    A = rng.normal(size=(2, 2))
    A = A + A.T
    g = rng.normal(size=2) * 3
    y, x = np.meshgrid(np.arange(n) * h[1], np.arange(n) * h[0], indexing="ij")
    phi = 0.1 + g[0] * x + g[1] * y + 0.5 * (A[0, 0] * x * x + 2 * A[0, 1] * x * y + A[1, 1] * y * y)
    X = np.zeros((100, 3))
    X[:, :2] = rng.uniform(2, n - 3, size=(100, 2))
    sq = O.surface_at(phi, w, X)
    grad = g + (X[:, :2] * np.array(h)) @ A
    nrm = grad / np.linalg.norm(grad, axis=1)[:, None]
    np.testing.assert_allclose(sq["n"][:, :2], nrm, atol=1e-12)
    np.testing.assert_allclose(sq["s"][:, 0], nrm[:, 1], atol=1e-12)
    np.testing.assert_allclose(sq["s"][:, 1], -nrm[:, 0], atol=1e-12)
    assert np.all(sq["n"][:, 2] == 0) and np.all(sq["r"] == 0)
    assert np.all(sq["k_nn"] == sq["k"]) and np.all(sq["k_tt"] == 0) and np.all(sq["k_nt"] == 0)
    # a circle: k = 1/r
    m = 96
    hh = 1.0 / (m - 1)
    w2 = Win((0, 0, 0), (m, m, 1), (hh, hh, hh))
    phi2 = configs.circle_sdf(m, (0.48, 0.53), 0.31) * hh
    recs = O.irregular_nodes(phi2, w2, O.band_compact(phi2, w2, 3 * hh))
    sq2 = O.surface_quantities(phi2, w2, recs)
    assert np.abs(sq2["k"] * 0.31 - 1).max() < 1e-2
