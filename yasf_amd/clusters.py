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
    """Non-overlapping centres (|r_i - r_j| >= r_i + r_j), largest spheres first."""
    n = radii.size
    R = (np.sum(radii ** 3) / volume_fraction) ** (1.0 / 3.0)
    order = np.argsort(-radii)
    pos = np.zeros((n, 3))
    placed = np.zeros(n, dtype=bool)
    cell = 2.0 * radii.max()
    grid: dict = {}

    def key(p):
        return tuple(np.floor(p / cell).astype(int))

    for idx in order:
        r = radii[idx]
        lim = R - r
        if lim <= 0:
            raise ValueError("volume fraction too high for the largest sphere")
        for _ in range(max_attempts):
            c = rng.uniform(-lim, lim, 3)
            if c @ c > lim * lim:
                continue
            k = key(c)
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for j in grid.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                            d = c - pos[j]
                            if d @ d < (r + radii[j]) ** 2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pos[idx] = c
                placed[idx] = True
                grid.setdefault(k, []).append(idx)
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
