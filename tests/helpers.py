"""Shared helpers of the GPU parity tests."""
import numpy as np
import torch

import oracle as O
from dcell_amd import GridWindow


def owin(w: GridWindow) -> O.Win:
    return O.Win(tuple(w.p), tuple(w.n), tuple(w.d))


def to_dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda:0", dtype=dtype)


def to_np(t):
    torch.cuda.synchronize()
    return t.detach().double().cpu().numpy()


def rounded(a, dtype):
    """The values a float32 run actually sees, back in float64."""
    return np.asarray(a, dtype=np.float32).astype(np.float64) if dtype == torch.float32 else a


def rel_max(a, b, scale=None):
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b)) if scale is None else scale
    return float(np.max(np.abs(a - b)) / max(s, 1e-300)) if a.size else 0.0


def smooth_field(n, dim, rng, offset=(0, 0, 0)):
    """A smooth signed-distance-like field in grid units with nonzero gradient
    almost everywhere, plus a random low-frequency perturbation."""
    shape = (n, n) if dim == 2 else (n, n, n)
    idx = np.indices(shape, dtype=np.float64)[::-1]      # x, y, (z)
    c = rng.uniform(0.4, 0.6, size=dim) * (n - 1)
    r = np.sqrt(sum((idx[a] - c[a]) ** 2 for a in range(dim)))
    pert = sum(rng.normal() * np.sin(2 * np.pi * (idx[a] / (n - 1)) * rng.integers(1, 3) + rng.uniform(0, 6))
               for a in range(dim))
    return r - 0.3 * (n - 1) + 0.5 * pert


def random_velocity(n, dim, rng, amp):
    shape = (n, n, 2) if dim == 2 else (n, n, n, 3)
    idx = np.indices(shape[:-1], dtype=np.float64)[::-1]
    vel = np.zeros(shape)
    for c in range(dim):
        for _ in range(3):
            k = rng.integers(1, 3, size=dim)
            ph = rng.uniform(0, 2 * np.pi)
            vel[..., c] += amp * rng.normal() * np.sin(sum(2 * np.pi * k[a] * idx[a] / (n - 1) for a in range(dim)) + ph)
    return vel
