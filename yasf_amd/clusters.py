"""Seeded synthetic particle clusters for the benchmark and test workloads.

The reference has no generator (it reads cluster files); these follow the
workload descriptions: radii log-normal (median, sigma_ln) or uniform, centres
by random sequential addition (RSA) uniformly inside a sphere of radius
``R = (sum r^3 / volume_fraction)^(1/3)`` with no overlaps.
"""

from __future__ import annotations

import numpy as np


def lognormal_radii(rng: np.random.Generator, n: int, median: float, sigma_ln: float) -> np.ndarray:
    return median * np.exp(sigma_ln * rng.standard_normal(n))


def rsa_in_sphere(rng: np.random.Generator, radii: np.ndarray, volume_fraction: float,
                  max_attempts: int = 200000) -> np.ndarray:
    """Non-overlapping centres (|r_i - r_j| >= r_i + r_j), largest spheres first.

    Each sphere draws uniform candidates in the ball of radius R - r until one
    overlaps no placed sphere.  Overlap candidates come from a k-d tree of the
    placed spheres (rebuilt every 128 placements) plus a direct check of the
    ones placed since the last rebuild."""
    from scipy.spatial import cKDTree

    n = radii.size
    R = (np.sum(radii ** 3) / volume_fraction) ** (1.0 / 3.0)
    order = np.argsort(-radii)
    pos = np.zeros((n, 3))
    placed = np.zeros((n, 3))
    placed_r = np.zeros(n)
    rmax = radii.max()
    count = 0
    tree, in_tree = None, 0
    for idx in order:
        r = radii[idx]
        lim = R - r
        if lim <= 0:
            raise ValueError("volume fraction too high for the largest sphere")
        for _ in range(max_attempts):
            c = rng.uniform(-lim, lim, 3)
            if c @ c > lim * lim:
                continue
            ok = True
            if tree is not None:
                near = tree.query_ball_point(c, r + rmax)
                if near:
                    d = placed[near] - c
                    ok = not np.any(np.einsum("ij,ij->i", d, d) < (r + placed_r[near]) ** 2)
            if ok and count > in_tree:
                d = placed[in_tree:count] - c
                ok = not np.any(np.einsum("ij,ij->i", d, d) < (r + placed_r[in_tree:count]) ** 2)
            if ok:
                pos[idx] = c
                placed[count] = c
                placed_r[count] = r
                count += 1
                if count - in_tree >= 128:
                    tree, in_tree = cKDTree(placed[:count]), count
                break
        else:
            raise RuntimeError("RSA failed to place a sphere; lower the volume fraction")
    return pos


def cluster(seed: int, n: int, volume_fraction: float, median: float = 100.0, sigma_ln: float = 0.2,
            uniform: tuple[float, float] | None = None):
    """(positions [n,3], radii [n]) in nm."""
    rng = np.random.default_rng(seed)
    if uniform is not None:
        radii = rng.uniform(uniform[0], uniform[1], n)
    else:
        radii = lognormal_radii(rng, n, median, sigma_ln)
    return rsa_in_sphere(rng, radii, volume_fraction), radii
