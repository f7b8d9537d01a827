"""Batched serving engine (yasfhip_server_*) on the GPU.

Each served solution is checked independently through the single-session
FP64 operator (itself pinned to the oracle at 1e-12): FP64 relative residual
<= 1e-6, the reported residual and cross sections consistent with that
operator, and C_ext / C_sca within 1e-6 of a tight (1e-10) all-FP64 solve.
Batch composition must not change any result: runs with and without the HIP
graph agree bit for bit, and a small arena (forcing queueing and FP64
residuals from slabs or on the fly) agrees to rounding, as does the residual
pool's slab.  Requests that share an operator by key give the same solutions
bit for bit as unshared ones; without pair tables (on-the-fly complex64 GEMMs)
the solutions still meet the FP64 criterion."""

import numpy as np
import pytest

from yasf_amd import _native
from yasf_amd.serving import make_aggregate_workload, make_request, serve, solve_sequential

pytestmark = pytest.mark.gpu


def _requests(count, seed, n_range=(20, 160)):
    rng = np.random.default_rng(seed)
    return [make_request(rng, i, n_range=n_range) for i in range(count)]


@pytest.fixture(scope="module")
def mixed_requests():
    return _requests(10, 7)


@pytest.fixture(scope="module")
def served(mixed_requests):
    srv = _native.Server(memory_bytes=4 << 30, restart=50, inner_rtol=1e-4)
    report = serve(srv, mixed_requests, keep_solutions=True)
    srv.close()
    return report


def test_served_solutions_meet_fp64_criterion(mixed_requests, served):
    assert served.completed == len(mixed_requests)
    assert {r.lmax for r in mixed_requests} >= {3, 8} or len({r.lmax for r in mixed_requests}) >= 4
    for r in mixed_requests:
        res = served.results[r.index]
        assert res["info"] == 0, (r.index, r.particles, r.lmax)
        s = _native.Session(r.positions, r.lmax, [r.k])
        s.set_tmatrix(r.t[np.newaxis, :])
        b = r.t * r.a_inc
        x = res["x"]
        rel = np.linalg.norm(b - s.master_matvec(x)) / np.linalg.norm(b)
        assert rel <= 1e-6
        assert res["residual"] == pytest.approx(rel, rel=1e-6, abs=1e-12)
        ce, cs = s.cross_sections(r.a_inc, x)
        assert res["c_ext"] == pytest.approx(ce, rel=1e-10)
        assert res["c_sca"] == pytest.approx(cs, rel=1e-10)
        xt, info, _, _, _ = s.gmres(b, b, rtol=1e-10, restart=100, maxiter=100)
        assert info == 0
        ce64, cs64 = s.cross_sections(r.a_inc, xt)
        assert abs(res["c_ext"] - ce64) <= 1e-6 * abs(ce64)
        assert abs(res["c_sca"] - cs64) <= 1e-6 * abs(cs64)
        s.close()


def test_graph_eager_and_small_arena_agree(mixed_requests, served):
    eager = _native.Server(memory_bytes=4 << 30, restart=50, inner_rtol=1e-4, use_graph=False)
    rep_eager = serve(eager, mixed_requests, keep_solutions=True)
    eager.close()
    # an arena that holds about two of these requests at a time
    small = _native.Server(memory_bytes=160 << 20, restart=50, inner_rtol=1e-4)
    rep_small = serve(small, mixed_requests, keep_solutions=True)
    forms = small.residual_forms()
    assert forms["slab"] + forms["on_the_fly"] > 0
    small.close()
    for r in mixed_requests:
        a = served.results[r.index]
        e = rep_eager.results[r.index]
        assert e["info"] == 0 and np.array_equal(a["x"], e["x"])
        assert a["iterations"] == e["iterations"] and a["cycles"] == e["cycles"]
        assert a["c_ext"] == e["c_ext"] and a["c_sca"] == e["c_sca"]
        # in the small arena some FP64 residuals are computed without the
        # resident FP64 table (from FP64 slabs, or on the fly): same sums,
        # agreement to rounding
        sm = rep_small.results[r.index]
        assert sm["info"] == 0
        assert np.linalg.norm(sm["x"] - a["x"]) <= 1e-12 * np.linalg.norm(a["x"])
        assert abs(sm["c_ext"] - a["c_ext"]) <= 1e-12 * abs(a["c_ext"])


def test_iterations_per_step_agree(mixed_requests, served):
    # several batched Arnoldi iterations per step (one graph replay or one
    # eager sequence, one host synchronisation): each request runs the same
    # iterations, only its refreshes wait for the step's end, so the results
    # agree bit for bit
    for k, use_graph in ((4, True), (3, False)):
        srv = _native.Server(memory_bytes=4 << 30, restart=50, inner_rtol=1e-4, use_graph=use_graph)
        srv.set_iterations_per_step(k)
        rep = serve(srv, mixed_requests, keep_solutions=True)
        srv.close()
        for r in mixed_requests:
            a, b = served.results[r.index], rep.results[r.index]
            assert b["info"] == 0 and np.array_equal(a["x"], b["x"])
            assert a["iterations"] == b["iterations"] and a["cycles"] == b["cycles"]
            assert a["c_ext"] == b["c_ext"] and a["c_sca"] == b["c_sca"]
    with pytest.raises(ValueError):
        _native.Server(memory_bytes=1 << 30).set_iterations_per_step(17)


def test_residual_pool_agrees(mixed_requests):
    # no per-operator tables or slabs (on-the-fly complex64 GEMMs) with and
    # without a residual pool of 128 MB (10 FP64 tiles of the largest request,
    # 160 spheres at lmax 8): with it every FP64 residual builds the pool's
    # slab in turn, without it generates the pair factors on the fly; same
    # sums, agreement to rounding
    reps, forms = [], []
    for pool in (0, 128 << 20):
        srv = _native.Server(memory_bytes=2 << 30, restart=50, inner_rtol=1e-4)
        srv.set_table_policy(_native.Server.TABLES_ON_THE_FLY)
        if pool:
            with pytest.raises(ValueError):
                srv.set_residual_pool(4 << 30)  # beyond the arena
            srv.set_residual_pool(pool)
        reps.append(serve(srv, mixed_requests, keep_solutions=True))
        forms.append(srv.residual_forms())
        if pool:
            r0 = mixed_requests[0]
            srv.submit(r0.positions, r0.lmax, r0.k, r0.t, r0.a_inc)
            with pytest.raises(ValueError):
                srv.set_residual_pool(0)  # needs an idle server
        srv.close()
    assert forms[0]["on_the_fly"] > 0 and forms[0]["on_the_fly"] == sum(forms[0].values())
    assert forms[1]["pool"] == forms[0]["on_the_fly"] and forms[1]["on_the_fly"] == 0
    for r in mixed_requests:
        a, p = reps[0].results[r.index], reps[1].results[r.index]
        assert a["info"] == 0 and p["info"] == 0
        assert np.linalg.norm(p["x"] - a["x"]) <= 1e-12 * np.linalg.norm(a["x"])
        assert abs(p["c_ext"] - a["c_ext"]) <= 1e-12 * abs(a["c_ext"])


def test_sequential_reference_agrees(mixed_requests, served):
    seq = solve_sequential(mixed_requests[:4])
    for r in mixed_requests[:4]:
        a, b = served.results[r.index], seq.results[r.index]
        assert b["info"] == 0 and b["residual"] <= 1e-6
        assert abs(a["c_ext"] - b["c_ext"]) <= 2e-6 * abs(b["c_ext"])
        assert abs(a["c_sca"] - b["c_sca"]) <= 2e-6 * abs(b["c_sca"])


def test_autotune_and_tile_shapes(mixed_requests):
    srv = _native.Server(memory_bytes=2 << 30)
    shapes = _native.Server.TILE_SHAPES
    for L in (3, 5, 8):
        # blocks of 4 and 8 waves take any lmax; 2-wave blocks up to 5 orders per wave
        usable = [(w, ct) for w, ct in shapes if w >= 4 or -(-(2 * L + 1) // w) <= 5]
        best, ms = srv.autotune(L, particles=96, reps=2)
        assert best in usable
        for i, shape in enumerate(shapes):
            assert (ms[i] > 0) == (shape in usable)
        assert ms[shapes.index(best)] == pytest.approx(ms[ms > 0].min())
    with pytest.raises(ValueError):
        srv.set_tile_shape(8, 2, 2)   # 17 orders over 2 waves exceeds the kernel's range
    with pytest.raises(ValueError):
        srv.set_tile_shape(4, 2, 8)   # not an instantiated shape
    # every usable fixed shape gives the same solution to rounding (blocks cut
    # leftover orders along K by their job's order count, which the shape's
    # column-tile grouping sets)
    reqs = mixed_requests[:3]
    outs = []
    for waves, ct in ((4, 2), (8, 2), (4, 4), (8, 4), (8, 8)):
        for L in range(3, 9):
            srv.set_tile_shape(L, waves, ct)
        outs.append(serve(srv, reqs, keep_solutions=True))
    for r in reqs:
        x0 = outs[0].results[r.index]["x"]
        for o in outs[1:]:
            x = o.results[r.index]["x"]
            assert np.linalg.norm(x - x0) <= 1e-5 * np.linalg.norm(x0)
            assert o.results[r.index]["c_ext"] == pytest.approx(outs[0].results[r.index]["c_ext"], rel=1e-6)
    srv.close()


def test_limits_and_errors():
    req = _requests(1, 11, n_range=(40, 40))[0]
    srv = _native.Server(memory_bytes=1 << 20)   # far too small for any request
    rid = srv.submit(req.positions, req.lmax, req.k, req.t, req.a_inc)
    srv.step()
    res = srv.poll()
    assert res["id"] == rid and res["info"] == -1
    srv.close()
    srv = _native.Server(memory_bytes=1 << 30, restart=5)
    srv.submit(req.positions, req.lmax, req.k, req.t, req.a_inc, rtol=1e-15, max_cycles=2)
    while True:
        srv.step()
        res = srv.poll()
        if res is not None:
            break
    assert res["info"] == 1 and res["cycles"] == 2 and res["iterations"] <= 10
    with pytest.raises(ValueError):
        srv.submit(req.positions, 9, req.k, req.t, req.a_inc)
    with pytest.raises(ValueError):
        _native.Server(restart=0)
    srv.close()


@pytest.fixture(scope="module")
def aggregate_requests():
    # two aggregates x two wavelengths x three incidence directions
    return make_aggregate_workload([48, 90], [450.0, 700.0], 3, seed=6)


def _fp64_check(r, res, tol=1e-6):
    s = _native.Session(r.positions, r.lmax, [r.k])
    s.set_tmatrix(r.t[np.newaxis, :])
    b = r.t * r.a_inc
    rel = np.linalg.norm(b - s.master_matvec(res["x"])) / np.linalg.norm(b)
    ce, cs = s.cross_sections(r.a_inc, res["x"])
    s.close()
    assert res["info"] == 0 and rel <= tol
    assert res["c_ext"] == pytest.approx(ce, rel=1e-10)
    assert res["c_sca"] == pytest.approx(cs, rel=1e-10)


def test_shared_operators_match_unshared(aggregate_requests):
    reqs = aggregate_requests
    assert len({r.operator_key for r in reqs}) == 4 and len(reqs) == 12
    srv = _native.Server(memory_bytes=2 << 30, restart=50, inner_rtol=1e-4)
    shared = serve(srv, reqs, keep_solutions=True)
    srv.close()
    import copy
    plain = [copy.copy(r) for r in reqs]
    for r in plain:
        r.operator_key = 0
    srv = _native.Server(memory_bytes=2 << 30, restart=50, inner_rtol=1e-4)
    unshared = serve(srv, plain, keep_solutions=True)
    srv.close()
    for r in reqs:
        a, b = shared.results[r.index], unshared.results[r.index]
        assert np.array_equal(a["x"], b["x"])
        assert a["iterations"] == b["iterations"] and a["c_ext"] == b["c_ext"]
        _fp64_check(r, a)
    # directions of one (aggregate, wavelength) differ in solution, not operator
    k0 = [r for r in reqs if r.operator_key == reqs[0].operator_key]
    assert not np.array_equal(shared.results[k0[0].index]["x"], shared.results[k0[1].index]["x"])


def test_operator_key_rules(aggregate_requests):
    r0 = aggregate_requests[0]
    other = next(r for r in aggregate_requests if r.particles != r0.particles)
    srv = _native.Server(memory_bytes=1 << 30)
    srv.submit(r0.positions, r0.lmax, r0.k, r0.t, r0.a_inc, operator_key=77)
    with pytest.raises(ValueError):   # same key, different size
        srv.submit(other.positions, other.lmax, other.k, other.t, other.a_inc, operator_key=77)
    with pytest.raises(ValueError):   # same key, different wave number
        srv.submit(r0.positions, r0.lmax, 1.01 * r0.k, r0.t, r0.a_inc, operator_key=77)
    while srv.poll() is None:
        srv.step()
    # the key is free again once its requests have completed
    rid = srv.submit(other.positions, other.lmax, other.k, other.t, other.a_inc, operator_key=77)
    while True:
        srv.step()
        res = srv.poll()
        if res is not None:
            break
    assert res["id"] == rid and res["info"] == 0
    srv.close()


def test_on_the_fly_pair_gemm(aggregate_requests, mixed_requests, served):
    # no resident pair tables: every complex64 GEMM and FP64 residual
    # generates its pair factors on the fly; keyed and unkeyed requests mixed
    import copy
    extra = [copy.copy(r) for r in mixed_requests[:3]]
    for i, r in enumerate(extra):
        r.index = 1000 + i
    reqs = aggregate_requests[:6] + extra
    srv = _native.Server(memory_bytes=2 << 30, restart=50, inner_rtol=1e-4)
    srv.set_table_policy(_native.Server.TABLES_ON_THE_FLY)
    rep = serve(srv, reqs, keep_solutions=True)
    srv.close()
    for r in reqs:
        _fp64_check(r, rep.results[r.index])
    for r, e in zip(mixed_requests[:3], extra):
        a, b = rep.results[e.index], served.results[r.index]
        assert abs(a["c_ext"] - b["c_ext"]) <= 2e-6 * abs(b["c_ext"])
        assert abs(a["c_sca"] - b["c_sca"]) <= 2e-6 * abs(b["c_sca"])


def test_slab_operators_match_tables(aggregate_requests, mixed_requests, served):
    # slabs of two FP64 (four complex64) target tiles: several slabs per GEMM,
    # one build per slab for all requests of an operator; same kernels per
    # tile as the resident tables, so the same solutions bit for bit
    import copy
    extra = [copy.copy(r) for r in mixed_requests[:4]]
    for i, r in enumerate(extra):
        r.index = 2000 + i
    reqs = aggregate_requests + extra
    srv = _native.Server(memory_bytes=2 << 30, restart=50, inner_rtol=1e-4)
    tab = serve(srv, reqs, keep_solutions=True)
    srv.close()
    srv = _native.Server(memory_bytes=2 << 30, restart=50, inner_rtol=1e-4)
    srv.set_table_policy(_native.Server.TABLES_SLABS, slab_tiles=2)
    slab = serve(srv, reqs, keep_solutions=True)
    srv.close()
    # a small arena: the 90-sphere operators keep a complex64 table with an
    # FP64 slab of one tile for the residual and C_sca (FP64 table too large)
    srv = _native.Server(memory_bytes=48 << 20, restart=50, inner_rtol=1e-4)
    mixed = serve(srv, aggregate_requests, keep_solutions=True)
    srv.close()
    for r in reqs:
        a, b = tab.results[r.index], slab.results[r.index]
        assert b["info"] == 0
        assert np.array_equal(a["x"], b["x"]) and a["iterations"] == b["iterations"]
        assert a["c_ext"] == b["c_ext"]
        # C_sca: regular pair factors from a slab instead of on the fly
        assert b["c_sca"] == pytest.approx(a["c_sca"], rel=1e-12)
        if r in aggregate_requests:
            m = mixed.results[r.index]
            assert np.array_equal(a["x"], m["x"]) and m["c_sca"] == pytest.approx(a["c_sca"], rel=1e-12)
    for r, e in zip(mixed_requests[:4], extra):
        assert np.array_equal(slab.results[e.index]["x"], served.results[r.index]["x"])
    with pytest.raises(ValueError):
        _native.Server(memory_bytes=1 << 30).set_table_policy(3)


def test_scattering_methods_agree(mixed_requests, served):
    # completions' C_sca by far-field quadrature (default) and by the
    # regular pair sum agree to rounding; solutions are unaffected
    srv = _native.Server(memory_bytes=4 << 30, restart=50, inner_rtol=1e-4)
    srv.set_scattering_method(_native.CSCA_PAIR_SUM)
    pair = serve(srv, mixed_requests[:5], keep_solutions=True)
    srv.close()
    for r in mixed_requests[:5]:
        a, b = served.results[r.index], pair.results[r.index]
        assert np.array_equal(a["x"], b["x"]) and a["c_ext"] == b["c_ext"]
        assert b["c_sca"] == pytest.approx(a["c_sca"], rel=1e-10)
    with pytest.raises(ValueError):
        _native.Server(memory_bytes=1 << 30).set_scattering_method(2)


@pytest.mark.parametrize("precision", [_native.FP16X2, _native.FP16], ids=["fp16x2", "fp16"])
def test_half_inner_precision(mixed_requests, served, aggregate_requests, precision):
    # the inner iterations on the fp16 pair GEMM (fp16-term tables); FP64
    # residuals still decide convergence, so every solution meets the FP64
    # criterion and its cross sections agree with the complex64 server's
    import copy
    extra = [copy.copy(r) for r in aggregate_requests[:6]]
    for i, r in enumerate(extra):
        r.index = 3000 + i  # serve() keys results by request index
    reqs = mixed_requests + extra
    runs = []
    for use_graph in (True, False):
        srv = _native.Server(memory_bytes=4 << 30, restart=50, inner_rtol=1e-4, use_graph=use_graph)
        srv.set_inner_precision(precision)
        runs.append(serve(srv, reqs, keep_solutions=True))
        srv.close()
    half, eager = runs
    for r in reqs:
        a = half.results[r.index]
        _fp64_check(r, a)
        e = eager.results[r.index]
        assert np.array_equal(a["x"], e["x"]) and a["iterations"] == e["iterations"]
    for r in mixed_requests:
        a, b = half.results[r.index], served.results[r.index]
        assert abs(a["c_ext"] - b["c_ext"]) <= 2e-6 * abs(b["c_ext"])
        assert abs(a["c_sca"] - b["c_sca"]) <= 2e-6 * abs(b["c_sca"])
    srv = _native.Server(memory_bytes=1 << 30)
    r0 = mixed_requests[0]
    srv.submit(r0.positions, r0.lmax, r0.k, r0.t, r0.a_inc)
    with pytest.raises(ValueError):  # not on a busy server
        srv.set_inner_precision(precision)
    srv.close()


@pytest.mark.parametrize("precision", [_native.FP16X2, _native.FP16], ids=["fp16x2", "fp16"])
def test_half_shared_and_slab_operators(aggregate_requests, precision):
    # operators shared by key run their requests' fp16 GEMMs as one launch
    # over the resident fp16 table, or over fp16 slabs rebuilt range by range
    # (a slab of one FP64 tile holds 1 two-term or 3 one-term fp16 target
    # tiles of the 90-sphere operators, so their GEMMs span several slabs);
    # each block computes what the per-request GEMM computes, so every form
    # gives the same solutions bit for bit
    import copy
    reqs = aggregate_requests
    plain = [copy.copy(r) for r in reqs]
    for r in plain:
        r.operator_key = 0
    runs = {}
    for name, batch, slab_tiles in (("shared", reqs, 0), ("unshared", plain, 0), ("slabs", reqs, 1),
                                    ("unshared_slabs", plain, 1)):
        srv = _native.Server(memory_bytes=2 << 30, restart=50, inner_rtol=1e-4)
        srv.set_inner_precision(precision)
        if slab_tiles:
            srv.set_table_policy(_native.Server.TABLES_SLABS, slab_tiles=slab_tiles)
        runs[name] = serve(srv, batch, keep_solutions=True)
        srv.close()
    for r in reqs:
        a = runs["shared"].results[r.index]
        _fp64_check(r, a)
        for name in ("unshared", "slabs", "unshared_slabs"):
            b = runs[name].results[r.index]
            assert b["info"] == 0
            assert np.array_equal(a["x"], b["x"]) and a["iterations"] == b["iterations"], name
            assert a["c_ext"] == b["c_ext"]
            assert b["c_sca"] == pytest.approx(a["c_sca"], rel=1e-12)


@pytest.mark.parametrize("precision", [_native.FP16X2, _native.FP16], ids=["fp16x2", "fp16"])
def test_half_operator_groups_across_orders(precision):
    # five requests per operator at lmax 3 (two column tiles per job: four
    # requests per fp16 GEMM block, then a group of one) and at lmax 6 (six
    # tiles: one request per block), each with its own incident field; the
    # groups give the unshared solutions bit for bit
    import copy
    from yasf_amd.coefficients import planewave_coefficients
    rng = np.random.default_rng(21)
    reqs = []
    for key, lmax in ((31, 3), (32, 6)):
        base = make_request(rng, 0, n_range=(60, 70), lmax_range=(lmax, lmax))
        for d in range(5):
            r = copy.copy(base)
            r.index = 100 * key + d
            r.operator_key = key
            beta, alpha = float(np.arccos(rng.uniform(-1, 1))), float(rng.uniform(0, 2 * np.pi))
            k = np.array([r.k])
            r.a_inc = np.ascontiguousarray(
                planewave_coefficients(lmax, r.positions, k, beta, alpha, 1)[:, :, 0]).ravel()
            reqs.append(r)
    plain = [copy.copy(r) for r in reqs]
    for r in plain:
        r.operator_key = 0
    runs = []
    for batch in (reqs, plain):
        srv = _native.Server(memory_bytes=2 << 30, restart=50, inner_rtol=1e-4)
        srv.set_inner_precision(precision)
        runs.append(serve(srv, batch, keep_solutions=True))
        srv.close()
    shared, unshared = runs
    for r in reqs:
        a, b = shared.results[r.index], unshared.results[r.index]
        assert a["info"] == 0 and np.array_equal(a["x"], b["x"]) and a["iterations"] == b["iterations"]
        # C_sca: a key's completions share their quadrature launches, the
        # unkeyed ones run one each; the same bits
        assert a["c_ext"] == b["c_ext"] and a["c_sca"] == b["c_sca"]
    for key in (31, 32):
        _fp64_check(reqs[0] if key == 31 else reqs[5], shared.results[100 * key])


def test_inner_rtol_setter(mixed_requests, served):
    # a smaller reduction per restart cycle with the fp16x2 inner operator:
    # fewer cycles (FP64 residuals) in total, and every solution still meets
    # the FP64 criterion
    srv = _native.Server(memory_bytes=4 << 30, restart=50, inner_rtol=1e-4)
    srv.set_inner_precision(_native.FP16X2)
    srv.set_inner_rtol(1e-6)
    with pytest.raises(ValueError):
        srv.set_inner_rtol(0.0)
    with pytest.raises(ValueError):
        srv.set_inner_rtol(1.0)
    rep = serve(srv, mixed_requests, keep_solutions=True)
    srv.close()
    for r in mixed_requests:
        _fp64_check(r, rep.results[r.index])
    cycles = sum(rep.results[r.index]["cycles"] for r in mixed_requests)
    assert cycles <= sum(served.results[r.index]["cycles"] for r in mixed_requests)
