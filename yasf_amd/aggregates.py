"""Cluster geometry: seeded fractal aggregates and the reference's cluster files.

``fractal_aggregate`` builds a monodisperse aggregate by sequential
cluster-cluster aggregation with a prescribed fractal law
``N = kf (Rg / a)^Df``: clusters are merged pairwise, smallest first, and the
centre-of-mass distance Gamma of each merge is the one that gives the merged
cluster the law's radius of gyration,

    Gamma^2 = N^2/(N1 N2) Rg^2 - N/N2 Rg1^2 - N/N1 Rg2^2,   N = N1 + N2

(the tunable cluster-cluster construction of Filippov, Zurita & Rosner,
J. Colloid Interface Sci. 229 (2000) 261).  The second cluster is placed, in a
random orientation, so that one of its monomers touches one of the first
cluster's and no two monomers overlap; where no contact is possible at exactly
Gamma (small clusters), the nearest feasible distance is used, and where every
tried orientation overlaps, Gamma is raised by 10% per 100 tries.  The reference
reads such aggregates from PyFracVAL files (``read_cluster``); it has no
generator of its own.
"""

from __future__ import annotations

from io import StringIO
from pathlib import Path

import numpy as np


def _rg2(p: np.ndarray) -> float:
    c = p.mean(axis=0)
    return float(np.mean(np.sum((p - c) ** 2, axis=1)))


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _rotation_to(u: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Rotation matrix taking unit vector u to unit vector w."""
    v = np.cross(u, w)
    c = float(u @ w)
    if c < -1.0 + 1e-12:                       # opposite: half turn about any normal
        n = np.cross(u, [1.0, 0.0, 0.0])
        if np.linalg.norm(n) < 1e-6:
            n = np.cross(u, [0.0, 1.0, 0.0])
        n /= np.linalg.norm(n)
        return 2.0 * np.outer(n, n) - np.eye(3)
    K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + K + K @ K / (1.0 + c)


def _merge(rng, A: np.ndarray, B: np.ndarray, a: float, Df: float, kf: float, tol: float,
           attempts: int) -> np.ndarray:
    from scipy.spatial import cKDTree

    N1, N2 = len(A), len(B)
    N = N1 + N2
    rg = a * (N / kf) ** (1.0 / Df)
    gamma2 = N * N / (N1 * N2) * rg * rg - N / N2 * _rg2(A) - N / N1 * _rg2(B)
    cA = A.mean(axis=0)
    Ac = A - cA
    Bc0 = B - B.mean(axis=0)
    rA = np.linalg.norm(Ac, axis=1)
    rB = np.linalg.norm(Bc0, axis=1)
    tree = cKDTree(A)
    for attempt in range(attempts):
        if attempt and attempt % 100 == 0:
            gamma2 *= 1.1                       # every tried contact overlaps: back off
        gamma = float(np.sqrt(max(gamma2, 0.0)))
        # contact pair (i in A, j in B) from which Gamma is reachable: with B
        # turned so that its monomer j points back along A_i's radius, the
        # centre distance over the contact sphere spans |r_i + r_j -+ 2a|
        j = int(rng.integers(N2))
        ok = np.flatnonzero(np.abs(rA + rB[j] - gamma) <= 2 * a)
        i = int(rng.choice(ok)) if ok.size else int(np.argmin(np.abs(rA + rB[j] - gamma)))
        u = Ac[i] / rA[i] if rA[i] > 0 else _random_rotation(rng)[:, 0]
        R = np.eye(3)
        if rB[j] > 0:
            R = _rotation_to(Bc0[j] / rB[j], -u)
        ang = rng.uniform(0, 2 * np.pi)  # random twist about u
        K = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
        twist = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)
        Bc = Bc0 @ (twist @ R).T
        # B_j = A_i + 2a v, |v| = 1; centre distance |d + 2a v| = Gamma with d = A_i - cA - Bc_j
        d = Ac[i] - Bc[j]
        dn = float(np.linalg.norm(d))
        if dn < 1e-12 * a:
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
        else:
            c = (gamma * gamma - dn * dn - 4 * a * a) / (4 * a * dn)
            c = min(1.0, max(-1.0, c))          # nearest feasible distance
            dh = d / dn
            e = rng.normal(size=3)
            e -= (e @ dh) * dh
            e /= np.linalg.norm(e)
            v = c * dh + np.sqrt(max(0.0, 1.0 - c * c)) * e
        Bp = Bc + (A[i] + 2 * a * v - Bc[j])
        dist, _ = tree.query(Bp, k=1)
        if np.all(dist >= 2 * a * (1.0 - tol)):
            return np.vstack([A, Bp])
    raise RuntimeError("cluster-cluster merge found no overlap-free contact; raise attempts")


def fractal_aggregate(seed: int, N: int, Df: float = 1.8, kf: float = 1.3, monomer_radius: float = 40.0,
                      tol: float = 1e-6, attempts: int = 2000):
    """(positions [N, 3], radii [N]) of a seeded fractal aggregate of N monomers
    (nm), centred on its centre of mass."""
    if N < 1:
        raise ValueError("N must be >= 1")
    rng = np.random.default_rng(seed)
    a = float(monomer_radius)
    clusters = [np.zeros((1, 3)) for _ in range(N)]
    while len(clusters) > 1:
        clusters.sort(key=len)
        merged = []
        k = 0
        while k + 1 < len(clusters):
            merged.append(_merge(rng, clusters[k], clusters[k + 1], a, Df, kf, tol, attempts))
            k += 2
        if k < len(clusters):
            merged.append(clusters[k])
        clusters = merged
    pos = clusters[0] - clusters[0].mean(axis=0)
    return pos, np.full(N, a)


def radius_of_gyration(positions: np.ndarray) -> float:
    return float(np.sqrt(_rg2(positions)))


def read_cluster(path, delimiter: str | None = None):
    """Sphere table (N x >=4: x, y, z, r[, refractive index]) and metadata of a
    cluster file, as the reference's Config reads it (yasfpy/config.py:200-240):
    ``.csv`` with a header row, ``.dat`` whitespace columns with a YAML header in
    ``#`` comment lines (PyFracVAL), or any file with an explicit delimiter
    ("whitespace" for runs of blanks)."""
    import yaml

    path = Path(path)
    metadata = None
    if delimiter is not None:
        sep = None if delimiter == "whitespace" else delimiter
        spheres = np.loadtxt(path, delimiter=sep, ndmin=2, comments="#")
    elif path.suffix == ".csv":
        spheres = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)
    elif path.suffix == ".dat":
        spheres = np.loadtxt(path, ndmin=2)
        header = [ln[1:] for ln in path.read_text().splitlines(keepends=True) if ln.startswith("#")]
        metadata = yaml.safe_load(StringIO("".join(header))) if header else None
    else:
        raise NotImplementedError(f"cluster file type {path.suffix!r}: give a delimiter")
    if spheres.shape[1] < 4:
        raise ValueError("the cluster file needs at least 4 columns (x, y, z, r)")
    return spheres, metadata
