"""Synthetic inputs of the level-set workloads (numpy, float64, reference layout).

Conventions: node-centred grids on the unit box, node (i, j, k) at
(i h, j h, k h) with h = 1/(N-1) (the reference's tests use dx = 1/(d1-1),
LevelSetMethod/tests/resize.c:15); phi in grid units (phi_phys / h, as after
the reference's reinitialisation, FastMarching.c:68-70); MAC velocity with
component c sampled half a cell below the node along axis c
(FluidField/tests/testSL.c:157-159).  Random draws use numpy's Philox
bit generator with the stated seeds.
"""
from __future__ import annotations

import math

import numpy as np


def philox(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def node_coords(n: int, dim: int = 3):
    h = 1.0 / (n - 1)
    x = np.arange(n) * h
    if dim == 2:
        Y, X = np.meshgrid(x, x, indexing="ij")
        return X, Y
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return X, Y, Z


def sphere_sdf(n: int, center=(0.35, 0.35, 0.35), radius=0.15) -> np.ndarray:
    """Signed distance to a sphere in grid units, negative inside, shaped (n, n, n)."""
    h = 1.0 / (n - 1)
    X, Y, Z = node_coords(n)
    d = np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2) - radius
    return d / h


def circle_sdf(n: int, center=(0.5, 0.5), radius=0.25) -> np.ndarray:
    h = 1.0 / (n - 1)
    X, Y = node_coords(n, 2)
    return (np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2) - radius) / h


def _seg_dist(px, py, ax, ay, bx, by):
    dx, dy = bx - ax, by - ay
    t = np.clip(((px - ax) * dx + (py - ay) * dy) / (dx * dx + dy * dy), 0.0, 1.0)
    return np.hypot(px - (ax + t * dx), py - (ay + t * dy))


def zalesak_sdf(n: int = 128, center=(0.5, 0.75), radius=0.15, slot_w=0.05,
                slot_top=0.85) -> np.ndarray:
    """Exact signed distance (grid units) to Zalesak's slotted disk: the disk
    minus the slot |x - cx| <= slot_w/2 from the disk's bottom up to slot_top."""
    h = 1.0 / (n - 1)
    X, Y = node_coords(n, 2)
    cx, cy = center
    hw = slot_w / 2
    yc = cy - math.sqrt(radius ** 2 - hw ** 2)  # where the slot walls meet the circle
    rx, ry = X - cx, Y - cy
    rr = np.hypot(rx, ry)
    # nearest point on the full circle; inside the removed arc -> use the arc ends
    safe = np.where(rr > 0, rr, 1.0)
    qx, qy = cx + radius * rx / safe, cy + radius * ry / safe
    removed = (np.abs(qx - cx) < hw) & (qy < cy)
    d_arc = np.where(removed,
                     np.minimum(np.hypot(X - (cx - hw), Y - yc), np.hypot(X - (cx + hw), Y - yc)),
                     np.abs(rr - radius))
    d_l = _seg_dist(X, Y, cx - hw, yc, cx - hw, slot_top)
    d_r = _seg_dist(X, Y, cx + hw, yc, cx + hw, slot_top)
    d_t = _seg_dist(X, Y, cx - hw, slot_top, cx + hw, slot_top)
    d = np.minimum(np.minimum(d_arc, d_l), np.minimum(d_r, d_t))
    inside = (rr < radius) & ~((np.abs(X - cx) <= hw) & (Y <= slot_top))
    return np.where(inside, -d, d) / h


def rotation_velocity(n: int = 128) -> np.ndarray:
    """Rigid rotation u = 2 pi (0.5 - y), v = 2 pi (x - 0.5), MAC-sampled, (n, n, 2)."""
    h = 1.0 / (n - 1)
    i = np.arange(n)
    J, I = np.meshgrid(i, i, indexing="ij")
    vel = np.empty((n, n, 2))
    vel[..., 0] = 2 * math.pi * (0.5 - J * h)          # u at ((i - 1/2) h, j h)
    vel[..., 1] = 2 * math.pi * (I * h - 0.5)          # v at (i h, (j - 1/2) h)
    return vel


def deformation_velocity(n: int) -> np.ndarray:
    """LeVeque's divergence-free deformation field (time factor excluded),
    MAC-sampled, shaped (n, n, n, 3):
      u =  2 sin^2(pi x) sin(2 pi y) sin(2 pi z)
      v = -sin(2 pi x) sin^2(pi y) sin(2 pi z)
      w = -sin(2 pi x) sin(2 pi y) sin^2(pi z)"""
    h = 1.0 / (n - 1)
    c = np.arange(n) * h
    f = c - 0.5 * h                                  # face coordinates
    s2 = lambda a: np.sin(2 * math.pi * a)           # noqa: E731
    sq = lambda a: np.sin(math.pi * a) ** 2          # noqa: E731
    vel = np.empty((n, n, n, 3))
    z, y, x = c[:, None, None], c[None, :, None], c[None, None, :]
    vel[..., 0] = 2 * sq(f[None, None, :]) * s2(y) * s2(z)
    vel[..., 1] = -s2(x) * sq(f[None, :, None]) * s2(z)
    vel[..., 2] = -s2(x) * s2(y) * sq(f[:, None, None])
    return vel


def deformation_dt(n: int, step: int, dt: float | None = None) -> float:
    """dt of step ``step`` with the cos(pi t / 3) time factor folded in: the SL
    step uses the velocity only through V dt, so V0 cos(pi t/3) with dt equals
    V0 with dt cos(pi t/3), t evaluated at t_n + dt/2."""
    h = 1.0 / (n - 1)
    dt = h if dt is None else dt
    t = step * dt + 0.5 * dt
    return dt * math.cos(math.pi * t / 3.0)


def random_fourier_velocity(n: int, modes: int, kmax: int, amp: float, seed: int,
                            kmin: int = 1) -> np.ndarray:
    """Sum of divergence-free modes A sin(2 pi k.x + theta) with A perpendicular to
    k, |k|_inf in [kmin, kmax], A ~ N(0, amp); MAC-sampled, shaped (n, n, n, 3)."""
    g = philox(seed)
    h = 1.0 / (n - 1)
    c = np.arange(n) * h
    vel = np.zeros((n, n, n, 3))
    for _ in range(modes):
        while True:
            k = g.integers(-kmax, kmax + 1, size=3)
            if kmin <= np.abs(k).max() <= kmax:
                break
        a = g.normal(0.0, amp, size=3)
        a = a - k * (a @ k) / (k @ k)           # project out the k component
        theta = g.uniform(0, 2 * math.pi)
        for comp in range(3):
            shift = [0.0, 0.0, 0.0]
            shift[comp] = -0.5 * h
            x = (c + shift[0])[None, None, :]
            y = (c + shift[1])[None, :, None]
            z = (c + shift[2])[:, None, None]
            vel[..., comp] += a[comp] * np.sin(2 * math.pi * (k[0] * x + k[1] * y + k[2] * z) + theta)
    return vel


def seed_particles(phi: np.ndarray, per_cell: int, band: float, seed: int,
                   rmin: float = 0.1, rmax: float = 0.5) -> np.ndarray:
    """Particles {x, y, z, radius} (global index units): ``per_cell`` uniform
    positions in the cell [i-1/2, i+1/2]^3 of each node with |phi| < band
    (the cell convention of LevelSetInitializeParticles, ParticleLS.c:56-62),
    sign from trilinear phi at the particle, radius s * clamp(|phi|, rmin, rmax)."""
    g = philox(seed)
    nodes = np.argwhere(np.abs(phi) < band)               # (m, 3) as (z, y, x)
    m = len(nodes)
    pos = np.repeat(nodes[:, ::-1].astype(np.float64), per_cell, axis=0)
    pos += g.uniform(-0.5, 0.5, size=pos.shape)
    n = np.array(phi.shape[::-1])
    pos = np.clip(pos, 0.0, n - 1.000001)
    ph = trilinear_numpy(phi, pos)
    s = np.where(ph >= 0, 1.0, -1.0)
    r = s * np.clip(np.abs(ph), rmin, rmax)
    P = np.empty((m * per_cell, 4))
    P[:, :3] = pos
    P[:, 3] = r
    return P


def trilinear_numpy(phi: np.ndarray, pos: np.ndarray) -> np.ndarray:
    """Plain trilinear interpolation of phi (z, y, x) at index positions (m, 3)."""
    nz, ny, nx = phi.shape
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    i = np.clip(np.floor(x).astype(int), 0, nx - 2)
    j = np.clip(np.floor(y).astype(int), 0, ny - 2)
    k = np.clip(np.floor(z).astype(int), 0, nz - 2)
    fx, fy, fz = x - i, y - j, z - k
    out = np.zeros(len(pos))
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                w = (fx if di else 1 - fx) * (fy if dj else 1 - fy) * (fz if dk else 1 - fz)
                out += w * phi[k + dk, j + dj, i + di]
    return out


# ----------------------------------------------------------------------------
# batch of cells (torch, so the full-size batch is generated on the device)
def _harmonics(x, y, z):
    """The 24 real solid harmonics of degrees 1..4 (unnormalised) of a unit
    direction (x, y, z)."""
    x2, y2, z2 = x * x, y * y, z * z
    return [
        x, y, z,
        x * y, y * z, 3 * z2 - 1, x * z, x2 - y2,
        (3 * x2 - y2) * y, x * y * z, y * (5 * z2 - 1), z * (5 * z2 - 3), x * (5 * z2 - 1),
        z * (x2 - y2), x * (x2 - 3 * y2),
        x * y * (x2 - y2), y * z * (3 * x2 - y2), x * y * (7 * z2 - 1), y * z * (7 * z2 - 3),
        35 * z2 * z2 - 30 * z2 + 3, x * z * (7 * z2 - 3), (x2 - y2) * (7 * z2 - 1),
        x * z * (x2 - 3 * y2), x2 * (x2 - 3 * y2) - y2 * (3 * x2 - y2),
    ]


def cell_fields(n: int, g: np.random.Generator, device, dtype, radius=0.25, modes=8, amp=0.05,
                vmodes=6, kmax=4, vamp=0.5):
    """phi (grid units) and MAC velocity of one cell on [0, 1]^3 with n nodes
    per axis: a sphere of `radius` perturbed by `modes` random real spherical
    harmonics of degree <= 4 (peak amplitude amp * radius), in a sum of
    `vmodes` random divergence-free Fourier modes (|k|_inf 1..kmax, A ~ N(0, vamp))."""
    import torch
    h = 1.0 / (n - 1)
    c = torch.arange(n, device=device, dtype=torch.float64) * h
    Z, Y, X = torch.meshgrid(c, c, c, indexing="ij")
    dx, dy, dz = X - 0.5, Y - 0.5, Z - 0.5
    r = torch.sqrt(dx * dx + dy * dy + dz * dz).clamp_min(1e-12)
    basis = _harmonics(dx / r, dy / r, dz / r)
    pick = g.choice(len(basis), size=modes, replace=False)
    coef = g.normal(size=modes)
    pert = sum(float(a) * basis[i] for a, i in zip(coef, pick))
    pert = pert * (amp / float(pert.abs().max()))
    phi = (r - radius * (1.0 + pert)) / h
    vel = torch.zeros((n, n, n, 3), device=device, dtype=torch.float64)
    for _ in range(vmodes):
        while True:
            k = g.integers(-kmax, kmax + 1, size=3)
            if 1 <= np.abs(k).max() <= kmax:
                break
        a = g.normal(0.0, vamp, size=3)
        a = a - k * (a @ k) / (k @ k)
        theta = g.uniform(0, 2 * math.pi)
        for comp in range(3):
            sh = [0.0, 0.0, 0.0]
            sh[comp] = -0.5 * h
            arg = 2 * math.pi * (k[0] * (X + sh[0]) + k[1] * (Y + sh[1]) + k[2] * (Z + sh[2])) + theta
            vel[..., comp] += float(a[comp]) * torch.sin(arg)
    vmax = float(vel.abs().max())
    dt = 0.5 * h / vmax if vmax > 0 else h
    return phi.to(dtype), vel.to(dtype), dt


def batch_edges(ncells: int = 512, seed: int = 4, lo: int = 32, hi: int = 128) -> np.ndarray:
    """Cell edges log-uniform in [lo, hi] (Philox, seed)."""
    g = philox(seed)
    return np.rint(np.exp(g.uniform(math.log(lo), math.log(hi), size=ncells))).astype(int)


def cell_fields_2d(n: int, g: np.random.Generator, device, dtype, radius=0.25, modes=4, amp=0.05,
                   vmodes=6, kmax=4, vamp=0.5):
    """phi (grid units) and MAC velocity of one 2D cell on [0, 1]^2 with n nodes
    per axis, the 2D counterpart of cell_fields: a circle of `radius`
    perturbed by `modes` random angular modes cos/sin(m theta), m 1..4 (peak
    amplitude amp * radius), in a sum of `vmodes` random divergence-free
    Fourier modes A sin(2 pi k.x + theta) with A perpendicular to k
    (|k|_inf 1..kmax, A ~ N(0, vamp))."""
    import torch
    h = 1.0 / (n - 1)
    c = torch.arange(n, device=device, dtype=torch.float64) * h
    Y, X = torch.meshgrid(c, c, indexing="ij")
    dx, dy = X - 0.5, Y - 0.5
    r = torch.sqrt(dx * dx + dy * dy).clamp_min(1e-12)
    th = torch.atan2(dy, dx)
    pert = torch.zeros_like(r)
    for _ in range(modes):
        m = int(g.integers(1, 5))
        pert = pert + float(g.normal()) * (torch.cos(m * th) if g.random() < 0.5 else torch.sin(m * th))
    pert = pert * (amp / max(float(pert.abs().max()), 1e-12))
    phi = (r - radius * (1.0 + pert)) / h
    vel = torch.zeros((n, n, 2), device=device, dtype=torch.float64)
    for _ in range(vmodes):
        while True:
            k = g.integers(-kmax, kmax + 1, size=2)
            if 1 <= np.abs(k).max() <= kmax:
                break
        a = g.normal(0.0, vamp, size=2)
        a = a - k * (a @ k) / (k @ k)
        theta = g.uniform(0, 2 * math.pi)
        for comp in range(2):
            sh = [0.0, 0.0]
            sh[comp] = -0.5 * h
            arg = 2 * math.pi * (k[0] * (X + sh[0]) + k[1] * (Y + sh[1])) + theta
            vel[..., comp] += float(a[comp]) * torch.sin(arg)
    vmax = float(vel.abs().max())
    dt = 0.5 * h / vmax if vmax > 0 else h
    return phi.to(dtype), vel.to(dtype), dt
