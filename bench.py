"""Benchmarks of the multiple-sphere T-matrix solve on one MI355X.

Default workload (``--workload solve``, seeded): 2,000 spheres, radii
log-normal (median 100 nm, sigma_ln 0.2), random sequential addition in a
sphere at volume fraction 0.2, n = 1.5 + 0.01i in vacuum, lambda = 550 nm,
plane wave along +z (TE), lmax = 6 (96 coefficients per sphere, 192,000
unknowns), complex128 matrix-free coupling over a resident FP64 pair-factor
table, block-diagonal preconditioner (about 8 spheres per box, boxes bisected
above 8), GMRES (restart 100) to 1e-6 relative residual.

One step = one solve: preconditioner build + LU on the device, then device
GMRES from x0 = b.  Geometry, T-matrix and pair table are resident on the
device before timing (the pair-table build, about 2 ms, is not in the timed
step); b and x are torch tensors in HBM, so no vector crosses PCIe inside the
timed region.  The same system is then solved with the mixed-precision path
(FP64 iterative refinement around GMRES with the complex64 pair GEMM) and
reported beside it.  Prints one JSON line.

``--workload precision``: 10,000 spheres (radii uniform 50-150 nm, n =
1.33 + 0.001i, volume fraction 0.15), lmax 5, lambda 600 nm: the
mixed-precision solve (complex64 pair GEMM in GMRES(50), FP64 refinement to
1e-6) against the all-FP64 solve to 1e-6, with the error of the coefficients
and cross sections measured against an FP64 solve to 1e-10 (block-diagonal
preconditioner in all three).

``--workload serving``: 512 seeded solve requests of uneven size (N
log-uniform in [20, 3000], lmax 3..8, lambda 300-1000 nm, random index and
incidence) through the batched mixed-precision server, all queued at the start
(or Poisson arrivals with ``--rate``), against the all-FP64 single-solve path
one request at a time; prints one JSON line with solves/s and latency
percentiles of both.  Requests are submitted as spheres + incidence and set
up (Mie T-matrix, plane-wave coefficients) on the device at admission;
``--host-setup`` submits host-computed t and a_inc instead.

``--workload aggregates``: serving at scale on fractal aggregates (seeded
cluster-cluster aggregation, Df 1.8, kf 1.3, monomer radius 40 nm, n = 1.7 +
0.1i, N uniform in [1000, 20000] from seed 5), lmax 4; per aggregate a sample
of the 64 wavelengths of linspace(350, 900) nm (``--wavelengths``, evenly
spread) x ``--directions`` random incidence directions, each a request; the
requests of one (aggregate, wavelength) share one operator on the device.
Solves/s and latency percentiles of the batch, all queued at the start; a few
requests are re-solved on the all-FP64 single-solve path for the error and a
per-request FP64 time.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 matrix (= vector)
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix
METRIC = "solves/s + p50/p99 latency at 1e-6 rel. residual; matvec TFLOP/s vs gfx950 peak"
TRAFFIC_FILE = ROOT / "profiles" / "pair_table_gemm_fp64_2000spheres_lmax6_traffic.json"


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


def trimmed_table_macs_per_pair(lmax: int) -> int:
    """Complex MACs per particle pair that the FP64 table GEMM issues when
    each order runs over only the column tiles it feeds (the one-order-at-a-
    time shape): per (dm, order p), 16 columns per tile up to the last tile
    holding a source column with a nonzero ab5 term at p (columns of a dm
    ordered by l descending, as the coupling plan orders them)."""
    from yasf_amd import _native

    ab5 = _native.translation_table(lmax)
    ls, ms = [], []
    for _tau in (1, 2):
        for l in range(1, lmax + 1):
            for m in range(-l, l + 1):
                ls.append(l)
                ms.append(m)
    ls, ms = np.array(ls), np.array(ms)
    nz = ab5 != 0  # [n2][n1][p]
    total = 0
    for dm in range(-2 * lmax, 2 * lmax + 1):
        cols = [n2 for n2 in range(len(ms)) if abs(ms[n2] - dm) <= lmax]
        cols.sort(key=lambda n2: -ls[n2])  # stable: ties keep mode order
        need = np.zeros(2 * lmax + 1, dtype=int)
        for c, n2 in enumerate(cols):
            n1 = np.nonzero(ms[n2] - ms == dm)[0]
            ps = np.nonzero(nz[n2, n1, :].any(axis=0))[0]
            need[ps] = np.maximum(need[ps], c // 16 + 1)
        total += 16 * int(need.sum())
    return total


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


def host_cpu_model() -> str:
    """CPU model of this host (as lscpu reports it)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def host_threads() -> tuple[int, int]:
    """(threads used, CPUs this process may run on): every CPU of the
    process's affinity set, within the thread allotment OMP_NUM_THREADS
    states where the host sets one."""
    affinity = len(os.sched_getaffinity(0))
    try:
        allotment = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        allotment = 0
    return (min(affinity, allotment) if allotment > 0 else affinity), affinity


def small_solve_comparison(threads: int, reps: int = 20):
    """The 50-sphere CPU-runnable configuration (seed 1, lmax 4, volume
    fraction 0.1, n 1.5+0.01i, 550 nm, plane wave +z TE) solved to 1e-6 with
    the block-diagonal preconditioner (boxes of at most 8 spheres), in full on
    the host by the oracle (its particle_interaction_scalar restatement under
    SciPy's GMRES, blocks assembled and LU-factored on the host) with
    `threads` and with one thread, and on the GPU (preconditioner build + LU +
    device GMRES per solve, p50 of `reps`)."""
    import oracle
    from oracle import restate as R
    from yasf_amd import _native
    from yasf_amd.preconditioner import edge_size_for, make_particle_partition, split_oversized

    lmax = 4
    pos, radii, k, t, b = workload(seed=1, N=50, lmax=lmax, lam=550.0, vf=0.1)
    part = split_oversized(make_particle_partition(pos, edge_size_for(pos, 8.0)), pos, 8)
    boxes = np.empty(pos.shape[0], dtype=np.int32)
    for i, members in enumerate(part):
        boxes[members] = i
    s = _native.Session(pos, lmax, [k])
    s.set_tmatrix(t)
    lat = []
    for _ in range(reps + 2):
        t0 = time.perf_counter()
        s.precond_build(boxes)
        x_gpu, info, iters, res, _ = s.gmres(b, b, rtol=1e-6, restart=100, maxiter=1000, use_precond=True)
        lat.append(time.perf_counter() - t0)
    s.close()
    gpu_s = float(np.percentile(lat[2:], 50))
    out = {"workload": "50 spheres, lmax 4, vf 0.1, 550 nm, block-diag precond (<= 8/box), GMRES(100) to 1e-6",
           "gpu_seconds_p50": gpu_s, "gpu_iterations": iters, "gpu_info": info}
    lib = oracle.load_native()
    for label, th in (("cpu", threads), ("cpu_1_thread", 1)):
        used = lib.set_threads(th)
        t0 = time.perf_counter()
        M = R.block_preconditioner(lmax, pos, k, t, part)
        x_cpu, info_c, it_c = R.gmres_solve(lambda v: R.master_matvec(lmax, pos, k, t, v), b, rtol=1e-6,
                                            restart=100, psolve=M)
        sec = time.perf_counter() - t0
        out[label] = {"seconds": sec, "threads": used, "iterations": it_c, "info": int(info_c)}
    out["cpu_over_gpu_time"] = out["cpu"]["seconds"] / gpu_s
    out["solution_rel_diff_gpu_vs_cpu"] = float(np.linalg.norm(x_gpu - x_cpu) / np.linalg.norm(x_cpu))
    return out


def pmc_calibration():
    """Known-size kernels for the HBM-traffic counters: a 1 GiB fill (writes
    1 GiB) and a 1 GiB device copy (reads and writes 1 GiB)."""
    import torch

    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def traffic_record(kernel: str):
    """HBM bytes per launch of one FP64 table-GEMM instance on this workload,
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (null when
    absent)."""
    try:
        rec = json.loads(TRAFFIC_FILE.read_text())
        return float(rec["kernels"][kernel]["hbm_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def run_solve(args):
    import torch

    from yasf_amd import _native
    from yasf_amd.preconditioner import (MAX_BLOCK_ROWS, edge_size_for, make_particle_partition,
                                         split_oversized)

    if args.pmc_calibration:
        pmc_calibration()
    lmax, N = 6, 2000
    pos, radii, k, t, b = workload()
    s = _native.Session(pos, lmax, [k])
    s.set_tmatrix(t)
    s.set_matvec_mode(_native.MATVEC_TABLE)
    tuned = {name: s.tune_matvec(precision=prec, reps=10)[0]
             for name, prec in (("fp64", _native.FP64), ("fp32", _native.FP32))}
    use_pc = not args.no_precond
    boxes = None
    if use_pc:
        part = split_oversized(make_particle_partition(pos, edge_size_for(pos, args.per_box)), pos,
                               min(MAX_BLOCK_ROWS // 96, args.box_cap))
        boxes = np.empty(N, dtype=np.int32)
        for i, members in enumerate(part):
            boxes[members] = i

    # b and x live in HBM (device-pointer GMRES), so the timed solves move no vectors over PCIe
    b_dev = torch.from_numpy(np.ascontiguousarray(b, dtype=np.complex128)).to("cuda")
    x_dev = torch.empty_like(b_dev)

    def step():
        if use_pc:
            s.precond_build(boxes)
        x_dev.copy_(b_dev)  # x0 = b
        return s.gmres_device(b_dev, x_dev, rtol=1e-6, restart=args.restart, maxiter=1000, use_precond=use_pc)

    def step_mixed():
        if use_pc:
            s.precond_build(boxes)
        return s.gmres_refined(b, b, rtol=1e-6, inner_rtol=mixed_inner_rtol(args), restart=args.restart,
                               inner_maxiter=1000, max_refinements=10, use_precond=use_pc)

    def timed_mixed(inner):
        s.set_inner_precision(inner)
        step_mixed()  # warm: builds the inner operator's table
        lat_i = []
        res_i = None
        for _ in range(args.steps):
            t0 = time.perf_counter()
            res_i = step_mixed()
            lat_i.append(time.perf_counter() - t0)
        s.set_inner_precision(_native.FP32)
        return lat_i, res_i

    for _ in range(args.warmup):
        step()
        step_mixed()
        print("warmup step done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    lat = []
    results = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        results.append(step())
        lat.append(time.perf_counter() - t0)
        print(f"step {i}: {lat[-1]:.3f} s, {results[-1][1]} iterations, info {results[-1][0]}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    info, iters, res, _ = results[-1]
    x = x_dev.cpu().numpy()

    # mixed precision on the same system, same preconditioner
    lat_m = []
    mixed = None
    for i in range(args.steps):
        t0 = time.perf_counter()
        mixed = step_mixed()
        lat_m.append(time.perf_counter() - t0)
    torch.cuda.synchronize()
    xm, info_m, steps_m, iters_m, res_m, _ = mixed
    # the same refinement with the inner operator on the fp16 pair GEMM
    lat_h, mixed_h = timed_mixed(_native.FP16X2)
    xh, info_h, steps_h, iters_h, res_h, _ = mixed_h
    # the same solve with the preconditioner's factors kept in complex128
    fp64_factors = None
    if use_pc:
        s.set_precond_factor_precision(_native.FP64)
        lat64f = []
        for i in range(max(2, args.steps // 4)):
            t0 = time.perf_counter()
            r64f = step()
            lat64f.append(time.perf_counter() - t0)
        torch.cuda.synchronize()
        fp64_factors = {"solves_per_s": len(lat64f) / sum(lat64f), "p50_latency_s": float(np.percentile(lat64f, 50)),
                        "gmres_iterations": r64f[1], "relative_residual": r64f[2], "info": r64f[0]}
        s.set_precond_factor_precision(_native.FP32)
    a_inc = b / t
    c64 = s.cross_sections(a_inc, x)
    cm = s.cross_sections(a_inc, xm)
    chh = s.cross_sections(a_inc, xh)

    # Pair-kernel timing with HIP events on the session stream.
    pair_ms, mv_ms = s.time_matvec(reps=5)
    pair32_ms, mv32_ms = s.time_matvec(precision=_native.FP32, reps=5)
    pair16x2_ms, _ = s.time_matvec(precision=_native.FP16X2, reps=5)
    pair16_ms, _ = s.time_matvec(precision=_native.FP16, reps=5)
    pairs = N * (N - 1)
    f_fact = factorised_flops_per_pair(lmax) * pairs
    f_direct = direct_flops_per_pair(lmax) * pairs
    macs_issued, _ = _native.plan_stats(lmax)
    if tuned["fp64"] == 2:  # one order at a time over the tiles it feeds
        macs_issued = trimmed_table_macs_per_pair(lmax)
    ntiles = (N + 15) // 16
    # real MFMA flops issued, tile padding included: 3 real products (6 flop)
    # per complex MAC in the FP64 GEMM (Gauss's form)
    f_issued = macs_issued * 6 * (ntiles * 16) * N
    achieved = f_fact / (pair_ms * 1e-3) / 1e12
    waves64, tiles64 = _native.Session.TABLE_SHAPES[tuned["fp64"]]
    # instance name as rocprofv3 reports it: <T, waves, orders per wave (lmax <= 7), column tiles,
    # one order at a time>; FP64 shape 2 runs jobs of up to 6 column tiles one order at a time
    if tuned["fp64"] == 2:
        kernel64 = "pair_table_gemm_kernel<double, 8, 1, 6, true>"
    else:
        kernel64 = f"pair_table_gemm_kernel<double, {waves64}, {-(-15 // waves64)}, {tiles64}, false>"
    achieved32 = f_fact / (pair32_ms * 1e-3) / 1e12

    line = {
        "metric": METRIC,
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
                   "preconditioner": use_pc, "precond_factors": "complex64 storage, FP64 LU and substitutions",
                   "restart": args.restart,
                   "box_cap": args.box_cap if use_pc else None,
                   "matvec": "resident FP64 pair-factor table + ab5 contraction",
                   "table_gemm_shape": {k: {"waves": _native.Session.TABLE_SHAPES[v][0],
                                            "column_tiles": 6 if (k == "fp64" and v == 2)
                                            else _native.Session.TABLE_SHAPES[v][1],
                                            "orders_at_once": 1 if (k == "fp64" and v == 2) else "all"}
                                        for k, v in tuned.items()}},
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
            "traffic": traffic_record(kernel64),
            "kernel": kernel64 + " (FP64 MFMA 16x16x4 over the resident pair table)",
            "flops_per_launch": f_fact,
            "issued_mfma_tflops": f_issued / (pair_ms * 1e-3) / 1e12,
            "direct_count_tflops_per_matvec": f_direct / (mv_ms * 1e-3) / 1e12,
        },
        "fp64_precond_factors": fp64_factors,
        "mixed_precision": {
            "solves_per_s": args.steps / sum(lat_m),
            "p50_latency_s": float(np.percentile(lat_m, 50)),
            "speedup_vs_fp64": float(np.percentile(lat, 50) / np.percentile(lat_m, 50)),
            "refinements": steps_m,
            "inner_iterations": iters_m,
            "relative_residual_fp64": res_m,
            "info": info_m,
            "c_ext_rel_diff_vs_fp64": abs(cm[0] - c64[0]) / abs(c64[0]),
            "c_sca_rel_diff_vs_fp64": abs(cm[1] - c64[1]) / abs(c64[1]),
            "pair_kernel_ms_fp32": pair32_ms,
            "pair_kernel_tflops_fp32": achieved32,
            "pair_kernel_frac_fp32_peak": achieved32 / PEAK_FP32_TFLOPS,
        },
        "mixed_precision_fp16x2_inner": {
            "solves_per_s": len(lat_h) / sum(lat_h),
            "p50_latency_s": float(np.percentile(lat_h, 50)),
            "speedup_vs_fp64": float(np.percentile(lat, 50) / np.percentile(lat_h, 50)),
            "refinements": steps_h,
            "inner_iterations": iters_h,
            "relative_residual_fp64": res_h,
            "info": info_h,
            "c_ext_rel_diff_vs_fp64": abs(chh[0] - c64[0]) / abs(c64[0]),
            "c_sca_rel_diff_vs_fp64": abs(chh[1] - c64[1]) / abs(c64[1]),
            "pair_kernel_ms_fp16x2": pair16x2_ms,
            "pair_kernel_ms_fp16": pair16_ms,
        },
    }
    if not args.skip_cpu:
        threads, affinity = host_threads()
        xr = np.random.default_rng(0).normal(size=t.size) + 0j
        sec, threads = cpu_baseline(pos, k, lmax, xr, args.cpu_sample, threads)
        sample1 = max(1, args.cpu_sample // 6)
        sec1, _ = cpu_baseline(pos, k, lmax, xr, sample1, 1)
        matvecs = iters + max(1, iters // args.restart + 1) + 1
        line["cpu_baseline"] = {
            "value": 1.0 / (sec * matvecs),
            "unit": "solves/s (extrapolated: CPU matvec time x the GPU solve's matvec count; "
                    "the preconditioner's build and applies not counted)",
            "cores": threads,
            "kind": "port",
            "sample": f"oracle particle_interaction_scalar (full p-range loop), {args.cpu_sample} target spheres x "
                      f"{N} sources, scaled to the full matvec ({sec:.2f} s/matvec on {threads} threads)",
            "cpu_model": host_cpu_model(),
            "cpus_in_affinity": affinity,
            "one_thread": {"value": 1.0 / (sec1 * matvecs), "seconds_per_matvec": sec1,
                           "sample": f"{sample1} target spheres x {N} sources, scaled"},
            "small_solve_in_full": small_solve_comparison(threads),
        }
    if not args.no_serving:
        line["serving"] = serving_report(args, lambda msg: print(msg, file=sys.stderr, flush=True))
    print(json.dumps(line), flush=True)


def run_precision(args):
    import torch

    from yasf_amd import _native
    from yasf_amd.clusters import cluster
    from yasf_amd.coefficients import mie_coefficients, planewave_coefficients
    from yasf_amd.model import Particles
    from yasf_amd.preconditioner import (MAX_BLOCK_ROWS, edge_size_for, make_particle_partition,
                                         split_oversized)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    N, lmax, lam = 10000, 5, 600.0
    nmax = 2 * lmax * (lmax + 2)
    pos, radii = cluster(3, N, 0.15, uniform=(50.0, 150.0))
    p = Particles(pos, radii, np.full(N, 1.33 + 0.001j))
    omega = np.array([2 * np.pi / lam])
    k = float(omega[0])
    t = mie_coefficients(lmax, p.unique_radius_index_pairs, omega, omega)[p.single_unique_array_idx, :, 0].ravel()
    a = planewave_coefficients(lmax, pos, omega, 0.0, 0.0, 1)[:, :, 0].ravel()
    b = t * a
    part = split_oversized(make_particle_partition(pos, edge_size_for(pos, args.per_box)), pos,
                           min(MAX_BLOCK_ROWS // nmax, args.box_cap))
    boxes = np.empty(N, dtype=np.int32)
    for i, members in enumerate(part):
        boxes[members] = i

    def session():
        s = _native.Session(pos, lmax, [k])
        s.set_tmatrix(t)
        s.precond_build(boxes)
        return s

    out = {}
    # mixed precision first, in its own session, so its complex64 table is resident;
    # pair tables (and the FP64 slab) are allocated and built before timing
    s = session()
    s.tune_matvec(precision=_native.FP32, reps=3)
    s.tune_matvec(precision=_native.FP64, reps=3)  # on the first FP64 slab of the residuals
    s.time_matvec(precision=_native.FP64, reps=1)
    lat = []
    for i in range(max(1, args.steps)):
        t0 = time.perf_counter()
        xm, info_m, steps_m, iters_m, res_m, hist = s.gmres_refined(
            b, b, rtol=1e-6, inner_rtol=mixed_inner_rtol(args), restart=50, inner_maxiter=1000, max_refinements=10,
            use_precond=True)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        log(f"mixed {i}: {lat[-1]:.2f} s, {steps_m} refinements, {iters_m} inner iterations, residual {res_m:.2e}")
    pair32_ms, _ = s.time_matvec(precision=_native.FP32, reps=3)
    cm = s.cross_sections(a, xm)
    s.close()
    # the same refinement with the inner GMRES on the two-term fp16 pair GEMM
    s = session()
    s.set_inner_precision(_native.FP16X2)
    s.time_matvec(precision=_native.FP16X2, reps=1)  # the fp16-term table first, as the complex64 one above
    s.tune_matvec(precision=_native.FP64, reps=3)  # the FP64 residuals' slabs
    lat_h = []
    for i in range(max(1, args.steps)):
        t0 = time.perf_counter()
        xh, info_h, steps_h, iters_h, res_h, _ = s.gmres_refined(
            b, b, rtol=1e-6, inner_rtol=mixed_inner_rtol(args), restart=50, inner_maxiter=1000, max_refinements=10,
            use_precond=True)
        torch.cuda.synchronize()
        lat_h.append(time.perf_counter() - t0)
        log(f"mixed fp16x2 {i}: {lat_h[-1]:.2f} s, {steps_h} refinements, {iters_h} inner iterations, "
            f"residual {res_h:.2e}")
    pair16x2_ms, _ = s.time_matvec(precision=_native.FP16X2, reps=3)
    ch = s.cross_sections(a, xh)
    s.close()
    s = session()
    s.tune_matvec(precision=_native.FP64, reps=3)
    lat64 = []
    for i in range(max(1, args.steps)):
        t0 = time.perf_counter()
        x64, info64, it64, res64, _ = s.gmres(b, b, rtol=1e-6, restart=50, maxiter=1000, use_precond=True)
        torch.cuda.synchronize()
        lat64.append(time.perf_counter() - t0)
        log(f"fp64 {i}: {lat64[-1]:.2f} s, {it64} iterations, residual {res64:.2e}")
    t64 = float(np.percentile(lat64, 50))
    pair64_ms, _ = s.time_matvec(reps=3)
    c64 = s.cross_sections(a, x64)
    xr, info_r, it_r, res_r, _ = s.gmres(b, x64, rtol=1e-10, restart=50, maxiter=1000, use_precond=True)
    cr = s.cross_sections(a, xr)
    s.close()
    log(f"fp64 reference to 1e-10: {it_r} more iterations, residual {res_r:.2e}")
    rel = lambda u, v: float(np.linalg.norm(u - v) / np.linalg.norm(v))
    line = {
        "metric": METRIC,
        "value": 1.0 / float(np.percentile(lat, 50)),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(lat),
        "warmup": 0,
        "ms_per_step": 1e3 * float(np.percentile(lat, 50)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 pair GEMM inside FP64 refinement (fp32/fp64)",
        "data": "synthetic (seeded RSA cluster, seed 3)",
        "config": {"workload": "precision: 10000 spheres, lmax 5, radii U(50,150) nm, n 1.33+0.001i, vf 0.15, "
                               "lambda 600 nm, GMRES(50) inner, FP64 refinement to 1e-6, block-diag precond",
                   "particles": N, "lmax": lmax, "unknowns": N * nmax},
        "mixed": {"seconds": float(np.percentile(lat, 50)), "refinements": steps_m, "inner_iterations": iters_m,
                  "relative_residual_fp64": res_m, "info": info_m, "pair_kernel_ms_fp32": pair32_ms},
        "fp64": {"seconds": t64, "iterations": it64, "relative_residual": res64, "info": info64,
                 "pair_kernel_ms": pair64_ms},
        "mixed_fp16x2_inner": {"seconds": float(np.percentile(lat_h, 50)), "refinements": steps_h,
                               "inner_iterations": iters_h, "relative_residual_fp64": res_h, "info": info_h,
                               "pair_kernel_ms_fp16x2": pair16x2_ms,
                               "speedup_vs_fp64": t64 / float(np.percentile(lat_h, 50))},
        "speedup_vs_fp64": t64 / float(np.percentile(lat, 50)),
        "error_vs_fp64_1e-10": {
            "mixed_coefficients": rel(xm, xr), "fp64_coefficients": rel(x64, xr),
            "mixed_c_ext": abs(cm[0] - cr[0]) / abs(cr[0]), "mixed_c_sca": abs(cm[1] - cr[1]) / abs(cr[1]),
            "mixed_fp16x2_coefficients": rel(xh, xr),
            "mixed_fp16x2_c_ext": abs(ch[0] - cr[0]) / abs(cr[0]),
            "mixed_fp16x2_c_sca": abs(ch[1] - cr[1]) / abs(cr[1]),
            "fp64_c_ext": abs(c64[0] - cr[0]) / abs(cr[0]), "fp64_c_sca": abs(c64[1] - cr[1]) / abs(cr[1]),
            "reference_residual": res_r},
    }
    print(json.dumps(line), flush=True)


def mixed_inner_rtol(args) -> float:
    """Inner GMRES reduction of the mixed-precision solves (default 1e-4)."""
    return args.inner_rtol if args.inner_rtol is not None else 1e-4


def serving_report(args, log):
    """The 512-request serving stream (seed 4) through the batched server:
    closed loop (all queued at the start) with the inner iterations on the
    two-term fp16 pair GEMM (complex64 accuracy; the headline), on the
    one-term fp16 GEMM and on the complex64 GEMM; Poisson arrivals at
    args.rate per second (two-term fp16); against the all-FP64 single-solve
    path one request at a time.  Returns a dict."""
    import torch

    from yasf_amd import _native
    from yasf_amd.serving import make_workload, serve, solve_sequential

    t0 = time.perf_counter()
    reqs = make_workload(args.requests, seed=4)
    log(f"serving workload: {len(reqs)} requests prepared on the host in {time.perf_counter() - t0:.1f} s")
    srv = _native.Server(restart=50)
    tiles = {}
    for L in range(3, 9):
        tiles[L] = {}
        for cls, particles in (("small", 400), ("large", 2000)):
            (waves, ctiles), ms = srv.autotune(L, particles=particles, reps=3)
            tiles[L][cls] = {"waves": waves, "column_tiles": ctiles}
    log(f"autotuned complex64 tile shapes: {tiles}")
    device_setup = not args.host_setup

    def summary(rep):
        return {**rep.as_dict(),
                "converged": int(sum(1 for r in rep.results.values() if r["info"] == 0)),
                "mean_inner_iterations": float(np.mean([r["iterations"] for r in rep.results.values()])),
                "mean_cycles": float(np.mean([r["cycles"] for r in rep.results.values()]))}

    # reduction asked of each restart cycle: 1e-6 for operators of complex64
    # accuracy (most requests then converge after one cycle and one FP64
    # residual), 1e-4 for the one-term fp16 operator (relative error ~1e-4)
    inner_rtol = {"fp16x2": 1e-6, "fp16": 1e-4, "fp32": 1e-6}
    if args.inner_rtol is not None:
        inner_rtol = {k: args.inner_rtol for k in inner_rtol}
    runs = {}
    for name, prec in (("fp16x2", _native.FP16X2), ("fp16", _native.FP16), ("fp32", _native.FP32)):
        srv.set_inner_precision(prec)
        srv.set_inner_rtol(inner_rtol[name])
        if args.warmup:
            serve(srv, make_workload(8, seed=99, n_range=(20, 200)))
        torch.cuda.synchronize()
        runs[name] = serve(srv, reqs, progress=log, device_setup=device_setup)
        log(f"served closed-loop, inner {name}: {runs[name].as_dict()}")
    closed = runs["fp16x2"]
    out = {"requests": len(reqs), "restart": 50, "inner_rtol": inner_rtol, "tile_shapes": tiles,
           "setup": "device (Mie + plane wave at admission)" if device_setup else "host",
           "inner_precision": "fp16x2: pair factors and x as two fp16 terms each, fp16 MFMA, FP32 accumulation "
                              "(complex64 accuracy); FP64 residuals decide convergence",
           "closed_loop": summary(closed),
           "closed_loop_inner_fp16": summary(runs["fp16"]),
           "closed_loop_inner_fp32": summary(runs["fp32"])}
    if math.isfinite(args.rate) and args.rate > 0:
        srv.set_inner_precision(_native.FP16X2)
        srv.set_inner_rtol(inner_rtol["fp16x2"])
        timed = make_workload(args.requests, seed=4, rate=args.rate)
        poisson = serve(srv, timed, progress=log, device_setup=device_setup)
        log(f"served Poisson {args.rate}/s: {poisson.as_dict()}")
        out["poisson"] = {"arrival_rate_per_s": args.rate, **summary(poisson)}
    srv.close()
    if not args.no_baseline:
        base = solve_sequential(reqs, progress=log)
        log(f"sequential FP64: {base.as_dict()}")
        out["fp64_single_solve"] = {**base.as_dict(),
                                    "converged": int(sum(1 for r in base.results.values() if r["info"] == 0))}
        out["speedup_vs_fp64_single_solve"] = closed.solves_per_s / base.solves_per_s
        out["speedup_vs_fp64_single_solve_inner_fp16"] = runs["fp16"].solves_per_s / base.solves_per_s
        out["speedup_vs_fp64_single_solve_inner_fp32"] = runs["fp32"].solves_per_s / base.solves_per_s
        for name, rep in runs.items():
            out[f"max_rel_diff_c_ext_{name}"] = max(abs(rep.results[i]["c_ext"] - base.results[i]["c_ext"])
                                                    / abs(base.results[i]["c_ext"]) for i in rep.results)
            out[f"max_rel_diff_c_sca_{name}"] = max(abs(rep.results[i]["c_sca"] - base.results[i]["c_sca"])
                                                    / abs(base.results[i]["c_sca"]) for i in rep.results)
    return out


def run_serving(args):
    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    rep = serving_report(args, log)
    closed = rep["closed_loop"]
    line = {
        "metric": METRIC,
        "value": closed["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": rep["requests"],
        "warmup": args.warmup,
        "ms_per_step": 1e3 * closed["makespan_s"] / rep["requests"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "two-term fp16 pair GEMM (complex64 accuracy) inside FP64 refinement",
        "data": "synthetic (seeded request stream, seed 4)",
        "config": {"workload": "serving: 512 requests, N log-uniform [20, 3000], lmax 3..8, "
                               "lambda 300-1000 nm, random index and incidence, rtol 1e-6",
                   "requests": rep["requests"], "arrivals": "all queued at start (closed loop)"},
        "p50_latency_s": closed["p50_latency_s"],
        "p99_latency_s": closed["p99_latency_s"],
        "serving": rep,
    }
    print(json.dumps(line), flush=True)


def run_aggregates(args):
    import torch

    from yasf_amd import _native
    from yasf_amd.serving import make_aggregate_workload, serve, solve_sequential

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    rng = np.random.default_rng(5)
    counts = [int(rng.uniform(1000, 20000)) for _ in range(args.aggregates)]
    if args.particles:
        counts = [args.particles] * args.aggregates
    grid = np.linspace(350.0, 900.0, 64)
    pick = np.unique(np.round(np.linspace(0, 63, args.wavelengths)).astype(int))
    lams = [float(grid[i]) for i in pick]
    t0 = time.perf_counter()
    reqs = make_aggregate_workload(counts, lams, args.directions, seed=5)
    log(f"workload: aggregates N={counts}, wavelengths {[round(v, 1) for v in lams]} nm, "
        f"{args.directions} directions: {len(reqs)} requests prepared in {time.perf_counter() - t0:.1f} s")
    # inner operator: complex64, or the fp16 pair GEMM in two terms (complex64
    # accuracy) or one; reduction per restart cycle 1e-4, 1e-6 for fp16x2
    prec = {"fp32": _native.FP32, "fp16x2": _native.FP16X2, "fp16": _native.FP16}[args.inner_precision]
    inner_rtol = args.inner_rtol if args.inner_rtol is not None else (
        1e-6 if args.inner_precision == "fp16x2" else 1e-4)
    srv = _native.Server(restart=50, inner_rtol=inner_rtol)
    (waves, ctiles), ms = srv.autotune(4, particles=3000, reps=3)
    log(f"autotuned lmax-4 tile shape (large requests): {waves} waves, {ctiles} column tiles ({ms})")
    srv.set_inner_precision(prec)
    if args.warmup:
        serve(srv, make_aggregate_workload([200], [500.0], 2, seed=99))
    torch.cuda.synchronize()
    rep = serve(srv, reqs, progress=log, device_setup=not args.host_setup)
    srv.close()
    log(f"served: {rep.as_dict()}")
    infos = [r["info"] for r in rep.results.values()]
    line = {
        "metric": METRIC,
        "value": rep.solves_per_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": args.warmup,
        "ms_per_step": 1e3 * rep.makespan_s / len(reqs),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "c64 pair GEMM inside FP64 refinement (fp32/fp64)",
                  "fp16x2": "two-term fp16 pair GEMM (FP32 accumulation) inside FP64 refinement",
                  "fp16": "one-term fp16 pair GEMM (FP32 accumulation) inside FP64 refinement"}[args.inner_precision],
        "data": "synthetic (seeded fractal aggregates, seed 5)",
        "config": {"workload": "serving at scale: fractal aggregates Df 1.8, kf 1.3, monomer 40 nm, "
                               "n 1.7+0.1i, lmax 4; sampled wavelengths x incidence directions, rtol 1e-6",
                   "particles": counts, "wavelengths_nm": lams, "directions": args.directions,
                   "requests": len(reqs), "operators": len(counts) * len(lams),
                   "restart": 50, "inner_precision": args.inner_precision, "inner_rtol": inner_rtol,
                   "tile_shape": [waves, ctiles],
                   "setup": "host" if args.host_setup else "device (Mie + plane wave at admission)"},
        "p50_latency_s": rep.p50_s,
        "p99_latency_s": rep.p99_s,
        "makespan_s": rep.makespan_s,
        "converged": int(sum(1 for i in infos if i == 0)),
        "mean_inner_iterations": float(np.mean([r["iterations"] for r in rep.results.values()])),
        "mean_cycles": float(np.mean([r["cycles"] for r in rep.results.values()])),
    }
    if args.fp64_check:
        # first request of each operator on the all-FP64 single-solve path
        firsts = {}
        for r in reqs:
            firsts.setdefault(r.operator_key, r)
        sample = list(firsts.values())[: args.fp64_check]
        base = solve_sequential(sample, progress=log)
        line["fp64_single_solve"] = {
            "requests": len(sample), "seconds_per_solve": base.makespan_s / len(sample),
            "max_rel_diff_c_ext": max(abs(rep.results[r.index]["c_ext"] - base.results[r.index]["c_ext"])
                                      / abs(base.results[r.index]["c_ext"]) for r in sample),
            "max_rel_diff_c_sca": max(abs(rep.results[r.index]["c_sca"] - base.results[r.index]["c_sca"])
                                      / abs(base.results[r.index]["c_sca"]) for r in sample)}
        line["speedup_vs_fp64_single_solve"] = rep.solves_per_s * base.makespan_s / len(sample)
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["solve", "precision", "serving", "aggregates"], default="solve")
    ap.add_argument("--no-precond", action="store_true")
    ap.add_argument("--per-box", type=float, default=8.0)
    ap.add_argument("--box-cap", type=int, default=8, help="bisect boxes above this many spheres")
    ap.add_argument("--restart", type=int, default=100)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=48, help="target particles in the CPU sample")
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="run a 1 GiB fill and copy first (reference kernels for rocprofv3 --pmc)")
    ap.add_argument("--requests", type=int, default=512, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=24.0,
                    help="serving: Poisson arrivals per second for the timed-arrival run (inf: closed loop only)")
    ap.add_argument("--no-serving", action="store_true",
                    help="solve: leave out the serving sub-object (the 512-request stream)")
    ap.add_argument("--no-baseline", action="store_true", help="serving: skip the sequential FP64 run")
    ap.add_argument("--inner-rtol", type=float, default=None,
                    help="mixed precision: least relative residual reduction of each inner GMRES cycle "
                         "(default 1e-4; serving: 1e-6 with the fp16x2 and complex64 inner operators)")
    ap.add_argument("--host-setup", action="store_true",
                    help="serving/aggregates: submit host-computed t and a_inc instead of spheres + incidence "
                         "(default: the server sets them up on the device)")
    ap.add_argument("--aggregates", type=int, default=1, help="aggregates: number of aggregates")
    ap.add_argument("--particles", type=int, default=0, help="aggregates: fixed N instead of the seeded draw")
    ap.add_argument("--wavelengths", type=int, default=2, help="aggregates: wavelengths sampled of 64")
    ap.add_argument("--directions", type=int, default=16, help="aggregates: incidence directions")
    ap.add_argument("--inner-precision", choices=["fp32", "fp16x2", "fp16"], default="fp16x2",
                    help="aggregates: pair GEMM of the inner GMRES iterations: two fp16 terms (complex64 "
                         "accuracy, the default), complex64, or one fp16 term")
    ap.add_argument("--fp64-check", type=int, default=1,
                    help="aggregates: operators whose first request is re-solved in FP64 (0: none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: runs on more than one GPU are not supported "
              "(the sharding plan is documented, not implemented).", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    if args.workload == "serving":
        run_serving(args)
    elif args.workload == "aggregates":
        run_aggregates(args)
    elif args.workload == "precision":
        run_precision(args)
    else:
        run_solve(args)


if __name__ == "__main__":
    main()
