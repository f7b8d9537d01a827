"""Serving workload and driver for the batched engine (``_native.Server``).

The workload is a seeded stream of independent scattering requests of uneven
size: per request a random cluster (N log-uniform in [20, 3000] spheres,
radii log-normal with median 100 nm and sigma_ln 0.2, random sequential
addition at volume fraction 0.1), lmax uniform in {3..8}, wavelength uniform in
300-1000 nm, refractive index real part uniform in 1.3-1.8 and imaginary part
in 0-0.05, and a plane wave (TE) from a uniformly random direction; vacuum
medium.  Each request is solved to 1e-6 relative residual and reports C_ext and
C_sca.  The host prepares each request's T-matrix diagonal (Mie coefficients,
mie_coefficients.py:29-161) and incident coefficients
(initial_field_coefficients.py:79-142) before it is submitted.

``serve`` drives a server under Poisson arrivals at a given rate (or with all
requests queued at once, rate = inf) and reports solves/s and latency
percentiles; ``solve_sequential`` is the one-request-at-a-time reference: a
fresh FP64 session and device GMRES per request.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from yasf_amd.clusters import cluster
from yasf_amd.coefficients import mie_coefficients, planewave_coefficients
from yasf_amd.model import Particles


@dataclass
class ServeRequest:
    index: int
    positions: np.ndarray
    lmax: int
    wavelength: float
    k: float
    t: np.ndarray        # T-matrix diagonal, N * nmax complex128
    a_inc: np.ndarray    # incident coefficients, N * nmax complex128
    refractive_index: complex
    polar: float
    azimuthal: float
    arrival_s: float = 0.0
    operator_key: int = 0   # requests with the same nonzero key share one operator
    radii: np.ndarray | None = None  # sphere radii (for the device set-up of t)

    @property
    def particles(self) -> int:
        return self.positions.shape[0]

    @property
    def unknowns(self) -> int:
        return self.particles * 2 * self.lmax * (self.lmax + 2)


def make_request(rng: np.random.Generator, index: int, n_range=(20, 3000), lmax_range=(3, 8),
                 lam_range=(300.0, 1000.0), volume_fraction: float = 0.1) -> ServeRequest:
    N = int(math.floor(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
    N = max(n_range[0], min(n_range[1], N))
    lmax = int(rng.integers(lmax_range[0], lmax_range[1] + 1))
    lam = float(rng.uniform(*lam_range))
    m = complex(rng.uniform(1.3, 1.8), rng.uniform(0.0, 0.05))
    beta = float(math.acos(rng.uniform(-1.0, 1.0)))
    alpha = float(rng.uniform(0.0, 2 * math.pi))
    seed = int(rng.integers(0, 2 ** 31 - 1))
    pos, radii = cluster(seed, N, volume_fraction)
    p = Particles(pos, radii, np.full(N, m))
    omega = np.array([2 * math.pi / lam])
    k = omega * 1.0
    mie = mie_coefficients(lmax, p.unique_radius_index_pairs, omega, k)
    t = np.ascontiguousarray(mie[p.single_unique_array_idx, :, 0]).ravel()
    a = np.ascontiguousarray(planewave_coefficients(lmax, pos, k, beta, alpha, 1)[:, :, 0]).ravel()
    return ServeRequest(index, pos, lmax, lam, float(k[0]), t, a, m, beta, alpha, radii=radii)


def make_workload(count: int = 512, seed: int = 4, rate: float = math.inf, **kw) -> list[ServeRequest]:
    """`count` requests; arrival times of a Poisson process of `rate` per second
    (all zero for rate = inf)."""
    rng = np.random.default_rng(seed)
    reqs = [make_request(rng, i, **kw) for i in range(count)]
    arr = np.random.default_rng(seed + 1000)
    t = 0.0
    for r in reqs:
        if math.isfinite(rate):
            t += float(arr.exponential(1.0 / rate))
        r.arrival_s = t
    return reqs


def make_aggregate_workload(particle_counts, wavelengths, directions: int, seed: int = 5,
                            lmax: int = 4, m: complex = 1.7 + 0.1j, monomer_radius: float = 40.0,
                            Df: float = 1.8, kf: float = 1.3) -> list[ServeRequest]:
    """Fractal-aggregate requests: for every aggregate (one per entry of
    particle_counts) and wavelength, `directions` plane waves (TE) from seeded
    random directions; the requests of one (aggregate, wavelength) share their
    operator key.  All arrive at the start."""
    from yasf_amd.aggregates import fractal_aggregate

    rng = np.random.default_rng(seed)
    reqs = []
    for ai, N in enumerate(particle_counts):
        pos, radii = fractal_aggregate(seed * 1000 + ai, int(N), Df, kf, monomer_radius)
        p = Particles(pos, radii, np.full(int(N), m))
        for wi, lam in enumerate(wavelengths):
            omega = np.array([2 * math.pi / lam])
            k = omega * 1.0
            mie = mie_coefficients(lmax, p.unique_radius_index_pairs, omega, k)
            t = np.ascontiguousarray(mie[p.single_unique_array_idx, :, 0]).ravel()
            key = (ai + 1) * 100000 + wi + 1
            for _ in range(directions):
                beta = float(math.acos(rng.uniform(-1.0, 1.0)))
                alpha = float(rng.uniform(0.0, 2 * math.pi))
                a = np.ascontiguousarray(planewave_coefficients(lmax, pos, k, beta, alpha, 1)[:, :, 0]).ravel()
                reqs.append(ServeRequest(len(reqs), pos, lmax, float(lam), float(k[0]), t, a, m, beta, alpha,
                                         operator_key=key, radii=radii))
    return reqs


@dataclass
class ServeReport:
    completed: int
    makespan_s: float
    solves_per_s: float
    p50_s: float
    p99_s: float
    results: dict = field(default_factory=dict)

    def as_dict(self):
        return {"completed": self.completed, "makespan_s": self.makespan_s,
                "solves_per_s": self.solves_per_s, "p50_latency_s": self.p50_s,
                "p99_latency_s": self.p99_s}


def serve(server, requests: list[ServeRequest], rtol: float = 1e-6, keep_solutions: bool = False,
          progress=None, device_setup: bool = False) -> ServeReport:
    """Submits each request at its arrival time (relative to the start) and
    steps the server until all have completed.  Latency = completion -
    arrival; solves/s = requests / (last completion - start).  device_setup:
    submit radii, refractive index and incidence, and let the server set up t
    and a_inc on the device (TE plane wave in vacuum, as the workloads here)."""
    if len({r.index for r in requests}) != len(requests):
        raise ValueError("request indices must be distinct (results are keyed by index)")
    order = sorted(requests, key=lambda r: r.arrival_s)
    ids = {}
    results = {}
    t0 = time.perf_counter()
    nxt = 0
    total = len(order)
    last_report = t0
    while len(results) < total:
        now = time.perf_counter() - t0
        while nxt < total and order[nxt].arrival_s <= now:
            r = order[nxt]
            if device_setup:
                rid = server.submit_spheres(r.positions, r.radii, r.refractive_index, r.lmax, r.k, r.polar,
                                            r.azimuthal, polarization=1, rtol=rtol, operator_key=r.operator_key)
            else:
                rid = server.submit(r.positions, r.lmax, r.k, r.t, r.a_inc, rtol=rtol,
                                    operator_key=r.operator_key)
            ids[rid] = r
            nxt += 1
        active, queued, _ = server.step()
        done_t = time.perf_counter() - t0
        while True:
            res = server.poll(with_solution=keep_solutions)
            if res is None:
                break
            r = ids.pop(res["id"])
            res["latency_s"] = done_t - r.arrival_s
            res["request"] = r.index
            results[r.index] = res
        if active == 0 and queued == 0 and nxt < total:
            wait = order[nxt].arrival_s - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        if progress and time.perf_counter() - last_report > 10.0:
            last_report = time.perf_counter()
            progress(f"serving: {len(results)}/{total} done, {active} active, {queued} queued")
    makespan = max(res["latency_s"] + requests_by_index(requests)[i].arrival_s
                   for i, res in results.items())
    lat = np.array([res["latency_s"] for res in results.values()])
    return ServeReport(len(results), makespan, len(results) / makespan, float(np.percentile(lat, 50)),
                       float(np.percentile(lat, 99)), results)


def requests_by_index(requests):
    return {r.index: r for r in requests}


def solve_sequential(requests: list[ServeRequest], rtol: float = 1e-6, restart: int = 50,
                     progress=None) -> ServeReport:
    """One request at a time on the all-FP64 single-solve path: a session per
    request (resident FP64 pair table when it fits), device GMRES from x0 = b,
    then C_ext / C_sca.  All requests are available at the start."""
    from yasf_amd import _native

    results = {}
    t0 = time.perf_counter()
    last_report = t0
    for r in requests:
        s = _native.Session(r.positions, r.lmax, [r.k])
        s.set_tmatrix(r.t[np.newaxis, :])
        b = r.t * r.a_inc
        x, info, iters, res, _ = s.gmres(b, b, rtol=rtol, restart=restart, maxiter=1000, history_cap=1)
        c_ext, c_sca = s.cross_sections(r.a_inc, x)
        s.close()
        results[r.index] = {"info": info, "iterations": iters, "residual": res, "c_ext": c_ext,
                            "c_sca": c_sca, "latency_s": time.perf_counter() - t0}
        if progress and time.perf_counter() - last_report > 10.0:
            last_report = time.perf_counter()
            progress(f"sequential: {len(results)}/{len(requests)} done")
    makespan = time.perf_counter() - t0
    lat = np.array([v["latency_s"] for v in results.values()])
    return ServeReport(len(results), makespan, len(results) / makespan, float(np.percentile(lat, 50)),
                       float(np.percentile(lat, 99)), results)


def main(argv=None):
    """Command-line driver: serves the seeded request stream on one server and
    prints the report as one JSON line.  Example:
    ``python3 -m yasf_amd.serving --requests 512 --precision fp16x2 --eager``."""
    import argparse
    import json

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--rate", type=float, default=math.inf, help="Poisson arrivals per second (inf: all at once)")
    ap.add_argument("--precision", choices=["fp16x2", "fp16", "fp32"], default="fp16x2",
                    help="inner operator precision")
    ap.add_argument("--restart", type=int, default=50)
    ap.add_argument("--inner-rtol", type=float, default=1e-4)
    ap.add_argument("--eager", action="store_true",
                    help="launch each iteration's kernels directly instead of replaying a captured graph")
    ap.add_argument("--iterations-per-step", type=int, default=None,
                    help="batched Arnoldi iterations between host synchronisations (server default if omitted)")
    args = ap.parse_args(argv)

    import torch

    torch.cuda.set_device(0)
    from yasf_amd import _native

    prec = {"fp16x2": _native.FP16X2, "fp16": _native.FP16, "fp32": _native.FP32}[args.precision]
    reqs = make_workload(args.requests, seed=args.seed, rate=args.rate)
    srv = _native.Server(restart=args.restart, inner_rtol=args.inner_rtol, use_graph=not args.eager)
    try:
        if not args.eager:
            for lmax in range(3, 9):
                for particles in (400, 2000):
                    srv.autotune(lmax, particles=particles, reps=3)
        srv.set_inner_precision(prec)
        if args.iterations_per_step is not None:
            srv.set_iterations_per_step(args.iterations_per_step)
        rep = serve(srv, reqs, device_setup=True, progress=lambda m: print(m, flush=True))
    finally:
        srv.close()
    out = rep.as_dict()
    res = list(rep.results.values())
    out.update(precision=args.precision, rate=args.rate, graph=not args.eager, inner_rtol=args.inner_rtol,
               converged=sum(1 for r in res if r.get("info") == 0),
               mean_cycles=float(np.mean([r["cycles"] for r in res])),
               mean_iterations=float(np.mean([r["iterations"] for r in res])),
               max_residual=float(max(r["residual"] for r in res)))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
