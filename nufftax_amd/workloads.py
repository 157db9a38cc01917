"""Seeded synthetic inputs for the benchmark workloads (numpy.random.default_rng, fixed draw order).

Each generator returns float64/complex128 numpy arrays; callers cast to the
working precision.  Draw order is part of the definition (same seed -> same data).
"""

from __future__ import annotations

import numpy as np


def uniform_1d_type1(seed: int = 0, M: int = 100_000):
    """1D type 1: x ~ U[-pi, pi), c = N(0,1) + i N(0,1); n_modes 1024, eps 1e-6."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-np.pi, np.pi, M)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    return x, c


def radial_2d(seed: int = 0, spokes: int = 512, samples: int = 7812, jitter: float = 1e-3, n_modes=(512, 512)):
    """2D radial trajectory: spokes x samples points, angles U[0, pi), radius U[-pi, pi),
    Gaussian jitter sigma=1e-3, wrapped to [-pi, pi).  Returns x, y, c (type-1 strengths)
    and f (type-2 modes, shape (n2, n1)).  Draw order: theta, r, jitter x, jitter y, c, f."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.0, np.pi, spokes)
    r = rng.uniform(-np.pi, np.pi, (spokes, samples))
    jx = rng.normal(0.0, jitter, (spokes, samples))
    jy = rng.normal(0.0, jitter, (spokes, samples))
    x = (r * np.cos(theta)[:, None] + jx).ravel()
    y = (r * np.sin(theta)[:, None] + jy).ravel()
    x = np.mod(x + np.pi, 2 * np.pi) - np.pi
    y = np.mod(y + np.pi, 2 * np.pi) - np.pi
    M = x.size
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    f = rng.standard_normal((n_modes[1], n_modes[0])) + 1j * rng.standard_normal((n_modes[1], n_modes[0]))
    return x, y, c, f


def uniform_nd(seed: int, M: int, dim: int):
    """Points U[-pi, pi)^dim and strengths N(0,1) + i N(0,1).  Draw order: coords (x, y, z), c."""
    rng = np.random.default_rng(seed)
    coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(dim)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    return coords, c


def gaussian_mixture_3d(seed: int = 0, M: int = 8_000_000, n_clusters: int = 16, sigma: float = 0.3):
    """3D clustered points: mixture of n_clusters Gaussians (sigma 0.3), centres U[-pi, pi)^3,
    wrapped.  Draw order: centres, cluster labels, offsets, c."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-np.pi, np.pi, (n_clusters, 3))
    lab = rng.integers(0, n_clusters, M)
    pts = centres[lab] + rng.normal(0.0, sigma, (M, 3))
    pts = np.mod(pts + np.pi, 2 * np.pi) - np.pi
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    return [pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy()], c


def type3_3d(seed: int = 0, M: int = 200_000, N: int = 200_000, s_max: float = 64.0):
    """3D type 3: sources U[-pi, pi)^3, targets U[-s_max, s_max]^3, c complex N(0,1)."""
    rng = np.random.default_rng(seed)
    xs = [rng.uniform(-np.pi, np.pi, M) for _ in range(3)]
    ss = [rng.uniform(-s_max, s_max, N) for _ in range(3)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    return xs, ss, c
