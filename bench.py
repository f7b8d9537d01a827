"""Benchmark: dense FP64 GMRES solves of a 2,000-sphere cluster on one MI355X.

Workload (seeded): 2,000 spheres, radii log-normal (median 100 nm, sigma_ln
0.2), random sequential addition in a sphere at volume fraction 0.2, n = 1.5 +
0.01i in vacuum, lambda = 550 nm, plane wave along +z (TE), lmax = 6 (96
coefficients per sphere, 192,000 unknowns), complex128 matrix-free coupling,
block-diagonal preconditioner with about 8 spheres per box, GMRES (restart
100) to 1e-6 relative residual.

One step = one solve: preconditioner build + LU on the device, then device
GMRES from x0 = b.  Inputs (geometry, T-matrix, right-hand side) are resident
on the device before timing.  Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)


def workload(seed=2, N=2000, lmax=6, lam=550.0, vf=0.2):
    from yasf_amd.clusters import cluster
    from yasf_amd.coefficients import mie_coefficients, planewave_coefficients
    from yasf_amd.model import Particles

    pos, radii = cluster(seed, N, vf)
    p = Particles(pos, radii, np.full(N, 1.5 + 0.01j))
    omega = np.array([2 * np.pi / lam])
    k = omega * 1.0
    mie = mie_coefficients(lmax, p.unique_radius_index_pairs, omega, k)
    t = mie[p.single_unique_array_idx, :, 0].ravel()
    a = planewave_coefficients(lmax, pos, k, 0.0, 0.0, 1)[:, :, 0].ravel()
    return pos, radii, float(k[0]), t, t * a


def factorised_flops_per_pair(lmax: int) -> int:
    """8 x the number of distinct (p, m1, n2) source terms a pair contributes."""
    from yasf_amd import _native

    ab5 = _native.translation_table(lmax)
    ms = []
    for tau in (1, 2):
        for l in range(1, lmax + 1):
            ms.extend(range(-l, l + 1))
    ms = np.array(ms)
    n2, n1, p = np.nonzero(ab5 != 0)  # ab5[n2][n1][p] multiplies Q[p][m(n1)][n2]
    triples = {(int(pp), int(ms[a]), int(b)) for b, a, pp in zip(n2, n1, p)}
    return 8 * len(triples)


def direct_flops_per_pair(lmax: int) -> int:
    """8 T_nz + 8 nmax^2: the direct per-pair count of the reference loop."""
    from yasf_amd import _native

    ab5 = _native.translation_table(lmax)
    return 8 * int(np.count_nonzero(ab5)) + 8 * ab5.shape[0] ** 2


def cpu_baseline(pos, k, lmax, x, sample_particles, threads):
    """Oracle restatement of particle_interaction_scalar (C/OpenMP) on a sample of
    target particles against all sources; returns seconds per full matvec."""
    import oracle
    from oracle import restate as R

    used = oracle.load_native().set_threads(threads)
    N = pos.shape[0]
    R.coupling_matvec(lmax, pos, k, x, s1_lo=0, n_s1=1)  # load + warm the library
    t0 = time.perf_counter()
    R.coupling_matvec(lmax, pos, k, x, s1_lo=0, n_s1=sample_particles)
    dt = time.perf_counter() - t0
    return dt * N / sample_particles, used


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-precond", action="store_true")
    ap.add_argument("--per-box", type=float, default=8.0)
    ap.add_argument("--box-cap", type=int, default=12, help="bisect boxes above this many spheres")
    ap.add_argument("--restart", type=int, default=100)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=48, help="target particles in the CPU sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every run here uses a single MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not run).", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)

    from yasf_amd import _native
    from yasf_amd.preconditioner import (MAX_BLOCK_ROWS, edge_size_for, make_particle_partition,
                                         split_oversized)

    lmax, N = 6, 2000
    pos, radii, k, t, b = workload()
    s = _native.Session(pos, lmax, [k])
    s.set_tmatrix(t)
    use_pc = not args.no_precond
    boxes = None
    if use_pc:
        part = split_oversized(make_particle_partition(pos, edge_size_for(pos, args.per_box)), pos,
                               min(MAX_BLOCK_ROWS // 96, args.box_cap))
        boxes = np.empty(N, dtype=np.int32)
        for i, members in enumerate(part):
            boxes[members] = i

    def step():
        if use_pc:
            s.precond_build(boxes)
        return s.gmres(b, b, rtol=1e-6, restart=args.restart, maxiter=1000, use_precond=use_pc)

    for _ in range(args.warmup):
        step()
        print("warmup step done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    lat = []
    results = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        results.append(step())
        lat.append(time.perf_counter() - t0)
        print(f"step {i}: {lat[-1]:.3f} s, {results[-1][2]} iterations, info {results[-1][1]}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    x, info, iters, res, _ = results[-1]

    # Matvec kernel timing with HIP events on the session stream.
    pair_ms, mv_ms = s.time_matvec(reps=5)
    pairs = N * (N - 1)
    f_fact = factorised_flops_per_pair(lmax) * pairs
    f_direct = direct_flops_per_pair(lmax) * pairs
    macs_issued, _ = _native.plan_stats(lmax)
    ntiles = (N + 15) // 16
    f_issued = macs_issued * 8 * (ntiles * 16) * N  # MFMA flops incl. tile padding
    achieved = f_fact / (pair_ms * 1e-3) / 1e12

    line = {
        "metric": "solves/s + p50/p99 latency at 1e-6 rel. residual; matvec TFLOP/s vs gfx950 peak",
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (fp64)",
        "data": "synthetic (seeded RSA cluster, seed 2)",
        "config": {"workload": "dense FP64 solve: 2000 spheres, lmax 6, vf 0.2, lambda 550 nm, "
                               "GMRES(100) to 1e-6" + (", block-diag precond ~8/box" if use_pc else ""),
                   "particles": N, "lmax": lmax, "unknowns": N * 96,
                   "preconditioner": use_pc, "restart": args.restart,
                   "box_cap": args.box_cap if use_pc else None},
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "gmres_iterations": iters,
        "gmres_info": info,
        "relative_residual": res,
        "matvec_ms": mv_ms,
        "pair_kernel_ms": pair_ms,
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": None,
            "kernel": "pair_kernel<double> (FP64 MFMA 16x16x4)",
            "flops_per_launch": f_fact,
            "issued_mfma_tflops": f_issued / (pair_ms * 1e-3) / 1e12,
            "direct_count_tflops_per_matvec": f_direct / (mv_ms * 1e-3) / 1e12,
        },
    }
    if not args.skip_cpu:
        threads = min(16, os.cpu_count() or 1)
        xr = np.random.default_rng(0).normal(size=t.size) + 0j
        sec, threads = cpu_baseline(pos, k, lmax, xr, args.cpu_sample, threads)
        matvecs = iters + max(1, iters // args.restart + 1) + 1
        line["cpu_baseline"] = {
            "value": 1.0 / (sec * matvecs),
            "unit": "solves/s (extrapolated: matvec time x GPU matvec count, no preconditioner)",
            "cores": threads,
            "kind": "port",
            "sample": f"oracle particle_interaction_scalar, {args.cpu_sample} target spheres x "
                      f"{N} sources, scaled to the full matvec ({sec:.2f} s/matvec)",
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
