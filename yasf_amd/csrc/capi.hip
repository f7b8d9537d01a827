// C ABI (include/yasfhip.h): session state, device buffers and the GMRES
// driver.  The GMRES control flow restates SciPy's gmres
// (scipy/sparse/linalg/_isolve/iterative.py:692-840, the solver the reference
// calls from yasfpy/solver.py:108-121); each restart cycle's Arnoldi steps,
// Givens rotations and back substitution run on the device, and the host
// takes the per-cycle decisions (true residual, SciPy's ptol update).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "../../include/yasfhip.h"
#include "internal.hpp"

namespace yasfhip {

namespace {

thread_local std::string g_last_error;

template <typename T>
T *dmalloc(size_t count) {
    void *p = nullptr;
    if (count == 0)
        count = 1;
    YH_CHECK(hipMalloc(&p, count * sizeof(T)));
    return static_cast<T *>(p);
}

template <typename T>
T *upload(const std::vector<T> &v, hipStream_t s) {
    T *d = dmalloc<T>(v.size());
    if (!v.empty())
        YH_CHECK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return d;
}

}  // namespace

}  // namespace yasfhip

using namespace yasfhip;

struct yasfhip_session {
    int N = 0, L = 0, nmax = 0, W = 0, Npad = 0, ntiles1 = 0;
    long n = 0;
    std::vector<double> k;
    std::vector<double> pos_host;
    hipStream_t stream = nullptr;
    CouplingPlan plan;
    DevicePlan dp{};
    double4 *pos = nullptr;
    void *Q = nullptr;  // pair-GEMM output, sized for complex128
    double2 *t = nullptr;
    bool t_set = false;
    double2 *xin = nullptr, *yout = nullptr;
    float2 *x32 = nullptr;  // complex64 copy of the pair GEMM's x
    // GMRES workspace
    int vcap = 0;
    double2 *V = nullptr, *bvec = nullptr, *xvec = nullptr, *rvec = nullptr, *avec = nullptr;
    double2 *rb = nullptr, *rx = nullptr, *rr = nullptr;  // iterative refinement: b, x, residual
    double *scal = nullptr, *partials = nullptr;
    unsigned *counter = nullptr;
    double *scal_host = nullptr;
    // Arnoldi state of the device GMRES (ArnoldiBuffers, sized for restart
    // vcap - 1): Hessenberg, Givens, coefficients, partial sums, state word
    void *arn_mem = nullptr;
    ArnoldiState *arn_snap = nullptr;  // pinned: two polling slots and the cycle end
    double *arn_hist = nullptr;        // pinned: a cycle's residual estimates
    hipEvent_t poll_ev[2] = {nullptr, nullptr};
    BlockPrecond *precond = nullptr;
    int precond_channel = -1;
    bool precond_fp32 = true;  // LU factors stored in complex64 for the apply (FP64 arithmetic)
    int csca_method = YASFHIP_CSCA_FAR_FIELD;  // how yasfhip_cross_sections computes C_sca
    // Resident pair-factor tables, one per precision (0 = FP64, 1 = FP32).
    int matvec_mode = YASFHIP_MATVEC_AUTO;
    void *table[2] = {nullptr, nullptr};
    int table_channel[2] = {-1, -1};
    bool table_refused[2] = {false, false};
    // table GEMM shape per precision until the tuner picks one: shape 2 in
    // both (FP64: 8 waves, one order at a time over jobs of up to 6 tiles;
    // FP32: 8 waves, jobs of 8 tiles), the fastest at 600 spheres lmax 4,
    // 2,000 spheres lmax 6 and 2,771 spheres lmax 8
    int table_grouping[2] = {2, 2};
    double near2 = kAllPairs;          // squared near-field radius of W (kAllPairs: every pair)
    // Slab of the pair table (a few target tiles), rebuilt tile range by tile
    // range inside every matvec when the full table does not fit.
    void *slab = nullptr;
    size_t slab_bytes = 0;
    int slab_tiles_req = 0;  // tiles per slab fixed by the caller (0: what fits)
    bool slab_refused = false;
    // fp16-term pair table (hi term, then lo term when split) and the
    // working memory of its GEMM
    void *htable = nullptr;
    int htable_channel = -1;
    bool htable_split = false;
    void *xh = nullptr;
    unsigned *hcolmax = nullptr;
    int *hxexp = nullptr, *hpexp = nullptr;
    unsigned long long *hdmin = nullptr;
    int inner_precision = YASFHIP_FP32;  // operator of the refined solve's inner GMRES
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};

    ~yasfhip_session() {
        if (precond)
            precond_destroy(precond);
        void *bufs[] = {pos,  Q,    t,    xin,   yout,    x32,     V,    bvec,     xvec, rvec,
                        avec, arn_mem, scal, partials, counter, rb, rx, rr,
                        htable, xh, hcolmax, hxexp, hpexp, hdmin};
        for (void *b : bufs)
            if (b)
                (void)hipFree(b);
        free_device_plan(dp);
        for (void *b : table)
            if (b)
                (void)hipFree(b);
        if (slab)
            (void)hipFree(slab);
        if (scal_host)
            (void)hipHostFree(scal_host);
        for (void *h : {static_cast<void *>(arn_snap), static_cast<void *>(arn_hist)})
            if (h)
                (void)hipHostFree(h);
        for (auto &e : ev)
            if (e)
                (void)hipEventDestroy(e);
        for (auto &e : poll_ev)
            if (e)
                (void)hipEventDestroy(e);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    Reducer reducer() const { return Reducer{partials, counter}; }

    // Makes the resident pair table of this precision hold `channel`; false
    // when the matvec should generate the pair factors on the fly instead.
    bool table_for(int channel, bool fp32) {
        const int ix = fp32 ? 1 : 0;
        if (matvec_mode == YASFHIP_MATVEC_ON_THE_FLY)
            return false;
        if (table[ix] && table_channel[ix] == channel)
            return true;
        const size_t bytes = pair_table_bytes(dp, N, fp32);
        if (!table[ix]) {
            if (matvec_mode == YASFHIP_MATVEC_AUTO) {
                if (table_refused[ix])
                    return false;
                size_t free_b = 0, total_b = 0;
                YH_CHECK(hipMemGetInfo(&free_b, &total_b));
                if (bytes > free_b / 4 * 3) {  // leave room for solver state and other sessions
                    table_refused[ix] = true;
                    return false;
                }
            }
            YH_CHECK(hipMalloc(&table[ix], bytes));
        }
        launch_pair_table_build(dp, pos, N, k[channel], fp32, table[ix], stream, 0, -1, nullptr, false, near2);
        table_channel[ix] = channel;
        return true;
    }

    // Makes the fp16-term table hold `channel` (with its lo term when split).
    // It is a resident table only: the matvec mode must allow one (TABLE, or
    // AUTO with room for it).
    void half_table_for(int channel, bool split) {
        YH_REQUIRE(matvec_mode == YASFHIP_MATVEC_TABLE || matvec_mode == YASFHIP_MATVEC_AUTO,
                   "the fp16 pair GEMM needs a resident table (matvec mode TABLE or AUTO)");
        if (htable && htable_channel == channel && (htable_split || !split))
            return;
        const size_t term = half_table_term_bytes(dp, N);
        if (htable && split && !htable_split) {
            YH_CHECK(hipFree(htable));
            htable = nullptr;
        }
        if (!htable) {
            const size_t bytes = split ? 2 * term : term;
            if (matvec_mode == YASFHIP_MATVEC_AUTO) {
                size_t free_b = 0, total_b = 0;
                YH_CHECK(hipMemGetInfo(&free_b, &total_b));
                YH_REQUIRE(bytes <= free_b / 4 * 3, "the fp16 pair table (" + std::to_string(bytes >> 20) +
                                                        " MiB) does not fit in three quarters of the free device "
                                                        "memory (" + std::to_string(free_b >> 20) + " MiB)");
            }
            YH_CHECK(hipMalloc(&htable, bytes));
            htable_split = split;
        }
        if (!xh) {
            xh = dmalloc<unsigned char>(half_x_bytes(dp, N));
            hcolmax = dmalloc<unsigned>(nmax);
            hxexp = dmalloc<int>(nmax + 1);
            hpexp = dmalloc<int>(2 * L + 1);
            hdmin = dmalloc<unsigned long long>(1);
        }
        launch_half_table_build(dp, pos, N, k[channel], htable, htable_split, hpexp, hdmin, stream, near2);
        htable_channel = channel;
    }

    // Target tiles per slab for this precision once the slab buffer exists;
    // 0 when the matvec should generate the pair factors on the fly.  AUTO
    // and SLAB size the buffer to a quarter of the free memory (capped at the
    // full FP64 table); a caller-fixed tile count sizes it for FP64 tiles.
    int slab_tiles(bool fp32) {
        if (matvec_mode == YASFHIP_MATVEC_ON_THE_FLY || matvec_mode == YASFHIP_MATVEC_TABLE)
            return 0;
        const int ntiles1 = (N + 15) / 16;
        if (!slab) {
            if (slab_refused)
                return 0;
            const size_t tile64 = pair_table_tile_bytes(dp, N, false);
            size_t bytes = 0;
            if (slab_tiles_req > 0) {
                bytes = tile64 * std::min(slab_tiles_req, ntiles1);
            } else {
                size_t free_b = 0, total_b = 0;
                YH_CHECK(hipMemGetInfo(&free_b, &total_b));
                bytes = std::min(free_b / 4, tile64 * ntiles1);
                bytes = bytes / tile64 * tile64;
            }
            if (bytes < tile64) {
                slab_refused = true;
                return 0;
            }
            YH_CHECK(hipMalloc(&slab, bytes));
            slab_bytes = bytes;
        }
        int tiles = static_cast<int>(slab_bytes / pair_table_tile_bytes(dp, N, fp32));
        if (slab_tiles_req > 0)
            tiles = std::min(tiles, slab_tiles_req);
        return std::min(tiles, ntiles1);
    }

    // gate (device int, may be null): every kernel does nothing unless
    // *gate == 1 when it starts.
    void pair_gemm(int channel, const void *xv, bool fp32, const int *gate = nullptr) {
        const int ix = fp32 ? 1 : 0;
        if (matvec_mode != YASFHIP_MATVEC_SLAB && table_for(channel, fp32)) {
            launch_pair_table_gemm(dp, N, table[ix], xv, fp32, table_grouping[ix], Q, stream, 0, -1, gate);
            return;
        }
        const int tiles = slab_tiles(fp32);
        if (tiles == 0) {
            launch_pair(dp, pos, N, k[channel], xv, fp32, Q, stream, gate, near2);
            return;
        }
        const int ntiles1 = (N + 15) / 16;
        for (int t0 = 0; t0 < ntiles1; t0 += tiles) {
            const int ntl = std::min(tiles, ntiles1 - t0);
            launch_pair_table_build(dp, pos, N, k[channel], fp32, slab, stream, t0, ntl, gate, false, near2);
            launch_pair_table_gemm(dp, N, slab, xv, fp32, table_grouping[ix], Q, stream, t0, ntl, gate);
        }
    }

    // Q = the regular (J) pair GEMM of complex128 x for C_sca: over slabs of
    // the FP64 regular table (build, multiply, next slab) where the session
    // streams slabs (SLAB mode, or AUTO once its matvecs allocated the slab
    // buffer), else with the factors generated on the fly.
    void pair_gemm_regular(int channel, const void *xv) {
        const int tiles = (slab || matvec_mode == YASFHIP_MATVEC_SLAB) ? slab_tiles(false) : 0;
        if (tiles == 0) {
            launch_pair_regular(dp, pos, N, k[channel], xv, Q, stream);
            return;
        }
        const int ntiles1 = (N + 15) / 16;
        for (int t0 = 0; t0 < ntiles1; t0 += tiles) {
            const int ntl = std::min(tiles, ntiles1 - t0);
            launch_pair_table_build(dp, pos, N, k[channel], false, slab, stream, t0, ntl, nullptr, true);
            launch_pair_table_gemm(dp, N, slab, xv, false, table_grouping[0], Q, stream, t0, ntl);
        }
    }

    // y = W x (mode 0) or x - t W x (mode 1) for device vectors of complex128.
    // YASFHIP_FP32: the pair GEMM runs in complex64 on a complex64 copy of x;
    // YASFHIP_FP16X2 / YASFHIP_FP16: on the fp16 matrix cores over the
    // fp16-term table and an fp16-term copy of x (complex64 partial sums);
    // the ab5 contraction and the epilogue stay FP64.  gate: as pair_gemm.
    void matvec_dev(int channel, const double2 *x, double2 *y, int mode, int precision,
                    const int *gate = nullptr) {
        const double2 *tt = t + static_cast<size_t>(channel) * n;
        if (precision == YASFHIP_FP16X2 || precision == YASFHIP_FP16) {
            const bool split = precision == YASFHIP_FP16X2;
            half_table_for(channel, split);
            launch_half_x(dp, N, x, xh, split, hcolmax, hxexp, stream, gate);
            launch_half_gemm(dp, N, htable, xh, split, hpexp, hxexp, Q, stream, gate);
            launch_contract(dp, N, Q, x, tt, mode, y, true, false, stream, 0, -1, gate);
        } else if (precision == YASFHIP_FP64) {
            pair_gemm(channel, x, false, gate);
            launch_contract(dp, N, Q, x, tt, mode, y, false, false, stream, 0, -1, gate);
        } else {
            launch_to_fp32(reinterpret_cast<float *>(x32), reinterpret_cast<const double *>(x), 2 * n,
                           stream, gate);
            pair_gemm(channel, x32, true, gate);
            launch_contract(dp, N, Q, x, tt, mode, y, true, false, stream, 0, -1, gate);
        }
    }

    // Arnoldi buffers of restart capacity cap = vcap - 1, carved from arn_mem.
    static size_t arnoldi_bytes(int cap, int nblk) {
        const size_t c = cap, R1 = c + 1, R2 = c + 2;
        return sizeof(double2) * (c * R1 + c + R1 + 2 * R1 + 2 * R2 * nblk + c) + sizeof(double) * (2 * c) +
               sizeof(ArnoldiState) + 16 * 256;
    }

    ArnoldiBuffers arnoldi(int restart) const {
        const int cap = vcap - 1;
        const size_t c = cap, R1 = c + 1, R2 = c + 2;
        ArnoldiBuffers a{};
        a.V = V;
        a.n = static_cast<long>(n);
        a.restart = restart;
        a.nblk = arnoldi_blocks(a.n);
        char *p = static_cast<char *>(arn_mem);
        auto take = [&](size_t bytes) {
            char *q = p;
            p += (bytes + 255) & ~size_t(255);
            return q;
        };
        a.H = reinterpret_cast<double2 *>(take(sizeof(double2) * c * R1));
        a.gs = reinterpret_cast<double2 *>(take(sizeof(double2) * c));
        a.g = reinterpret_cast<double2 *>(take(sizeof(double2) * R1));
        a.h1 = reinterpret_cast<double2 *>(take(sizeof(double2) * R1));
        a.h2 = reinterpret_cast<double2 *>(take(sizeof(double2) * R1));
        a.partA = reinterpret_cast<double2 *>(take(sizeof(double2) * R2 * a.nblk));
        a.partB = reinterpret_cast<double2 *>(take(sizeof(double2) * R2 * a.nblk));
        a.y = reinterpret_cast<double2 *>(take(sizeof(double2) * c));
        a.gc = reinterpret_cast<double *>(take(sizeof(double) * c));
        a.hist = reinterpret_cast<double *>(take(sizeof(double) * c));
        a.st = reinterpret_cast<ArnoldiState *>(take(sizeof(ArnoldiState)));
        return a;
    }

    void ensure_gmres(int restart) {
        if (vcap >= restart + 1)
            return;
        for (void *b : {static_cast<void *>(V), arn_mem})
            if (b)
                YH_CHECK(hipFree(b));
        V = nullptr;
        arn_mem = nullptr;
        if (arn_hist)
            YH_CHECK(hipHostFree(arn_hist));
        arn_hist = nullptr;
        vcap = 0;
        V = dmalloc<double2>(static_cast<size_t>(restart + 1) * n);
        arn_mem = dmalloc<char>(arnoldi_bytes(restart, arnoldi_blocks(static_cast<long>(n))));
        YH_CHECK(hipHostMalloc(reinterpret_cast<void **>(&arn_hist), restart * sizeof(double),
                               hipHostMallocDefault));
        if (!arn_snap)
            YH_CHECK(hipHostMalloc(reinterpret_cast<void **>(&arn_snap), 3 * sizeof(ArnoldiState),
                                   hipHostMallocDefault));
        for (auto &e : poll_ev)
            if (!e)
                YH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        vcap = restart + 1;
    }

    // Host-visible reduction results after a stream sync.
    void fetch_scalars(int count) {
        YH_CHECK(hipMemcpyAsync(scal_host, scal, count * sizeof(double), hipMemcpyDeviceToHost, stream));
        YH_CHECK(hipStreamSynchronize(stream));
    }

    double norm_dev(const double2 *x) {
        launch_norm2(x, n, reducer(), scal, stream);
        fetch_scalars(1);
        return std::sqrt(scal_host[0]);
    }

    void psolve(const double2 *r, double2 *z, bool use_precond, const int *gate = nullptr) {
        if (use_precond)
            precond_apply(precond, r, z, stream, gate);
        else if (z != r)
            YH_CHECK(hipMemcpyAsync(z, r, n * sizeof(double2), hipMemcpyDeviceToDevice, stream));
    }

    struct GmresOutcome {
        int info = 0;         // 0 converged, else maxiter (SciPy)
        int iterations = 0;   // inner (Arnoldi) iterations
        double rnorm = 0.0;   // final ||b - A x|| with the operator used
        double bnorm = 0.0;
    };

    // SciPy's gmres(A, b, x0, M, restart, rtol, atol=0, maxiter) on device
    // vectors (scipy/sparse/linalg/_isolve/iterative.py:692-840): bb is the
    // right-hand side, x holds x0 on entry and the solution on exit; neither
    // may alias rvec, avec or V.  precision selects the pair GEMM of every
    // operator application (the Arnoldi process stays FP64).
    // history: presid/||b|| per inner iteration, up to history_cap values.
    //
    // A restart cycle runs on the device without host round trips: the host
    // queues the cycle's Arnoldi steps (operator, preconditioner,
    // launch_arnoldi_step) in groups of kStepsPerPoll, every kernel gated by
    // the cycle's state word, and reads that word back one group behind, so
    // the device does not wait for the host inside a cycle; steps queued past
    // the device's stop decision return at once.  The host syncs once per
    // cycle, for the true residual that drives SciPy's outer loop.
    static constexpr int kStepsPerPoll = 4;

    GmresOutcome gmres_run(int channel, const double2 *bb, double2 *x, double rtol, int restart,
                           int maxiter, bool pc, int precision, double *history, int history_cap) {
        YH_REQUIRE(restart >= 1 && restart <= arnoldi_max_restart(),
                   "restart must be in [1, " + std::to_string(arnoldi_max_restart()) + "]");
        GmresOutcome out;
        hipStream_t st = stream;
        double2 *r = rvec, *av = avec;
        int nhist = 0;
        double rnorm = 0.0;
        const double bnrm2 = norm_dev(bb);
        out.bnorm = bnrm2;
        if (bnrm2 == 0.0) {  // SciPy returns b
            YH_CHECK(hipMemcpyAsync(x, bb, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
            return out;
        }
        const double atol = rtol * bnrm2;
        const double eps = DBL_EPSILON;
        psolve(bb, av, pc);
        const double Mb_nrm2 = norm_dev(av);
        double ptol_max_factor = 1.0;
        double ptol = Mb_nrm2 * std::min(ptol_max_factor, atol / bnrm2);
        double presid = 0.0;
        bool breakdown = false;
        ensure_gmres(restart);
        const ArnoldiBuffers a = arnoldi(restart);
        const int *gate = &a.st->active;

        for (int iteration = 0; iteration < maxiter; ++iteration) {
            if (iteration == 0) {
                // r = b - A x if x.any() else b
                launch_norm2(x, n, reducer(), scal, st);
                fetch_scalars(1);
                if (scal_host[0] != 0.0) {
                    matvec_dev(channel, x, av, 1, precision);
                    launch_residual(r, bb, av, n, reducer(), scal, st);
                    fetch_scalars(1);
                } else {
                    YH_CHECK(hipMemcpyAsync(r, bb, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
                    launch_norm2(r, n, reducer(), scal, st);
                    fetch_scalars(1);
                }
                rnorm = std::sqrt(scal_host[0]);
                if (rnorm < atol) {
                    out.rnorm = rnorm;
                    return out;
                }
            }
            // v[0] = M r / ||M r||, g = [||M r||, 0, ...]
            psolve(r, V, pc);
            launch_norm2(V, n, reducer(), scal, st);
            launch_arnoldi_start(a, scal, ptol, st);
            int queued = 0;
            for (int group = 0; queued < restart; ++group) {
                const int end = std::min(restart, queued + kStepsPerPoll);
                for (; queued < end; ++queued) {
                    const double2 *vc = V + static_cast<size_t>(queued) * n;
                    double2 *w = V + static_cast<size_t>(queued + 1) * n;
                    if (pc) {
                        matvec_dev(channel, vc, av, 1, precision, gate);
                        psolve(av, w, true, gate);
                    } else {
                        matvec_dev(channel, vc, w, 1, precision, gate);
                    }
                    launch_arnoldi_step(a, queued, st);
                }
                YH_CHECK(hipMemcpyAsync(arn_snap + (group & 1), a.st, sizeof(ArnoldiState), hipMemcpyDeviceToHost,
                                        st));
                YH_CHECK(hipEventRecord(poll_ev[group & 1], st));
                if (group > 0) {
                    YH_CHECK(hipEventSynchronize(poll_ev[(group - 1) & 1]));
                    if (arn_snap[(group - 1) & 1].active != 1)
                        break;
                }
            }
            // x += V y, y = H^-1 g; then the true residual
            launch_arnoldi_end(a, x, st);
            matvec_dev(channel, x, av, 1, precision);
            launch_residual(r, bb, av, n, reducer(), scal, st);
            YH_CHECK(hipMemcpyAsync(arn_snap + 2, a.st, sizeof(ArnoldiState), hipMemcpyDeviceToHost, st));
            YH_CHECK(hipMemcpyAsync(arn_hist, a.hist, restart * sizeof(double), hipMemcpyDeviceToHost, st));
            fetch_scalars(1);
            const ArnoldiState cycle = arn_snap[2];
            out.iterations += cycle.j;
            for (int k = 0; k < cycle.j && history && nhist < history_cap; ++k)
                history[nhist++] = arn_hist[k] / bnrm2;
            presid = cycle.presid;
            breakdown = cycle.breakdown != 0;
            rnorm = std::sqrt(scal_host[0]);
            if (rnorm <= atol)
                break;
            else if (breakdown)
                break;
            else if (presid <= ptol)
                ptol_max_factor = std::max(eps, 0.25 * ptol_max_factor);
            else
                ptol_max_factor = std::min(1.0, 1.5 * ptol_max_factor);
            ptol = presid * std::min(ptol_max_factor, atol / rnorm);
        }
        out.rnorm = rnorm;
        out.info = (rnorm <= atol) ? 0 : maxiter;
        return out;
    }
};

namespace {

int fail(const yasfhip::Error &e) {
    g_last_error = e.what();
    return e.code;
}

}  // namespace

int yasfhip::set_last_error(const std::string &msg, int code) {
    g_last_error = msg;
    return code;
}

namespace {

template <typename F>
int guarded(F &&f) {
    try {
        f();
        g_last_error.clear();
        return kOk;
    } catch (const yasfhip::Error &e) {
        return fail(e);
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return kInternal;
    }
}

}  // namespace

extern "C" {

const char *yasfhip_last_error(void) { return g_last_error.c_str(); }

int yasfhip_translation_table(int lmax, double *out) {
    return guarded([&] {
        YH_REQUIRE(out, "out is NULL");
        YH_REQUIRE(lmax >= 1 && lmax <= kMaxLmax, "lmax must be in [1, 8]");
        const auto &ab5 = translation_table_cached(lmax);
        std::memcpy(out, ab5.data(), ab5.size() * sizeof(cplx));
    });
}

int yasfhip_plan_stats(int lmax, long long *pair_macs_per_pair, long long *nnz_ab5) {
    return guarded([&] {
        const auto &plan = coupling_plan_cached(lmax);
        if (pair_macs_per_pair)
            *pair_macs_per_pair = plan.pair_macs_per_pair;
        if (nnz_ab5)
            *nnz_ab5 = plan.nnz_ab5;
    });
}

int yasfhip_create(const double *positions, int particle_number, int lmax, const double *k_medium,
                   int channels, yasfhip_session **out) {
    return guarded([&] {
        YH_REQUIRE(out, "out is NULL");
        YH_REQUIRE(positions && k_medium, "positions/k_medium is NULL");
        YH_REQUIRE(particle_number >= 1, "particle_number must be >= 1");
        YH_REQUIRE(channels >= 1, "channels must be >= 1");
        for (int w = 0; w < channels; ++w)
            YH_REQUIRE(std::isfinite(k_medium[w]) && k_medium[w] > 0.0,
                       "k_medium must be real, finite and positive");
        auto s = std::make_unique<yasfhip_session>();
        s->N = particle_number;
        s->L = lmax;
        s->W = channels;
        s->plan = coupling_plan_cached(lmax);
        s->nmax = s->plan.nmax;
        s->n = static_cast<long>(s->N) * s->nmax;
        s->ntiles1 = (s->N + 15) / 16;
        s->Npad = s->ntiles1 * 16;
        s->k.assign(k_medium, k_medium + channels);
        s->pos_host.assign(positions, positions + 3 * static_cast<size_t>(particle_number));
        YH_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        for (auto &e : s->ev)
            YH_CHECK(hipEventCreate(&e));
        hipStream_t st = s->stream;

        std::vector<double4> pos(s->Npad, make_double4(0, 0, 0, 0));
        for (int i = 0; i < s->N; ++i)
            pos[i] = make_double4(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], 0.0);
        s->pos = upload(pos, st);

        const auto &pl = s->plan;
        s->dp = upload_device_plan(pl, st);

        s->Q = dmalloc<double2>(static_cast<size_t>(s->ntiles1) * pl.nq * 256);
        s->t = dmalloc<double2>(static_cast<size_t>(s->W) * s->n);
        s->xin = dmalloc<double2>(s->n);
        s->yout = dmalloc<double2>(s->n);
        s->x32 = dmalloc<float2>(s->n);
        s->bvec = dmalloc<double2>(s->n);
        s->xvec = dmalloc<double2>(s->n);
        s->rvec = dmalloc<double2>(s->n);
        s->avec = dmalloc<double2>(s->n);
        s->scal = dmalloc<double>(8);
        s->partials = dmalloc<double>(kRedPartials);
        s->counter = dmalloc<unsigned>(1);
        YH_CHECK(hipMemsetAsync(s->counter, 0, sizeof(unsigned), st));
        YH_CHECK(hipHostMalloc(reinterpret_cast<void **>(&s->scal_host), 8 * sizeof(double),
                               hipHostMallocDefault));
        YH_CHECK(hipStreamSynchronize(st));
        *out = s.release();
    });
}

int yasfhip_destroy(yasfhip_session *s) {
    return guarded([&] { delete s; });
}

int yasfhip_set_tmatrix(yasfhip_session *s, const double *t) {
    return guarded([&] {
        YH_REQUIRE(s && t, "NULL argument");
        YH_CHECK(hipMemcpyAsync(s->t, t, static_cast<size_t>(s->W) * s->n * sizeof(double2),
                                hipMemcpyHostToDevice, s->stream));
        YH_CHECK(hipStreamSynchronize(s->stream));
        s->t_set = true;
    });
}

static void host_matvec(yasfhip_session *s, int channel, const double *x, double *y, int precision,
                        int mode) {
    YH_REQUIRE(s && x && y, "NULL argument");
    YH_REQUIRE(channel >= 0 && channel < s->W, "channel out of range");
    YH_REQUIRE(precision >= YASFHIP_FP64 && precision <= YASFHIP_FP16, "unknown precision");
    if (mode == 1)
        YH_REQUIRE(s->t_set, "T-matrix not set (yasfhip_set_tmatrix)");
    YH_CHECK(hipMemcpyAsync(s->xin, x, s->n * sizeof(double2), hipMemcpyHostToDevice, s->stream));
    s->matvec_dev(channel, s->xin, s->yout, mode, precision);
    YH_CHECK(hipMemcpyAsync(y, s->yout, s->n * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
    YH_CHECK(hipStreamSynchronize(s->stream));
}

int yasfhip_coupling_matvec(yasfhip_session *s, int channel, const double *x, double *y,
                            int precision) {
    return guarded([&] { host_matvec(s, channel, x, y, precision, 0); });
}

int yasfhip_master_matvec(yasfhip_session *s, int channel, const double *x, double *y,
                          int precision) {
    return guarded([&] { host_matvec(s, channel, x, y, precision, 1); });
}

namespace {

// device: b and x_inout are device pointers on the session's device, copied
// within the device instead of over PCIe.
int gmres_solve(yasfhip_session *s, int channel, const double *b, double *x_inout, double rtol, int restart,
                int maxiter, int use_precond, int *info, int *iterations, double *residual, double *history,
                int history_cap, bool device) {
    return guarded([&] {
        YH_REQUIRE(s && b && x_inout, "NULL argument");
        YH_REQUIRE(channel >= 0 && channel < s->W, "channel out of range");
        YH_REQUIRE(s->t_set, "T-matrix not set (yasfhip_set_tmatrix)");
        YH_REQUIRE(rtol >= 0.0, "rtol must be >= 0");
        YH_REQUIRE(maxiter >= 1, "maxiter must be >= 1");
        if (use_precond)
            YH_REQUIRE(s->precond && s->precond_channel == channel,
                       "preconditioner not prepared for this channel");
        if (device) {
            for (const void *p : {static_cast<const void *>(b), static_cast<const void *>(x_inout)}) {
                hipPointerAttribute_t attr{};
                const hipError_t e = hipPointerGetAttributes(&attr, p);
                if (e != hipSuccess)
                    (void)hipGetLastError();  // unknown (host) memory: clear the error, refuse below
                YH_REQUIRE(e == hipSuccess && attr.type == hipMemoryTypeDevice, "b and x_inout must be device memory");
                // the n complex128 values must lie inside one allocation
                hipDeviceptr_t base = nullptr;
                size_t bytes = 0;
                const hipError_t r = hipMemGetAddressRange(&base, &bytes, const_cast<void *>(p));
                if (r != hipSuccess)
                    (void)hipGetLastError();
                const char *lo = static_cast<const char *>(base);
                const char *q = static_cast<const char *>(p);
                YH_REQUIRE(r == hipSuccess && q >= lo &&
                               static_cast<size_t>(q - lo) + static_cast<size_t>(s->n) * sizeof(double2) <= bytes,
                           "b and x_inout must each hold n complex128 values inside one device allocation");
            }
        }
        const hipMemcpyKind in = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        const hipMemcpyKind out = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        const long n = s->n;
        if (restart <= 0)
            restart = 20;
        restart = static_cast<int>(std::min<long>(restart, n));
        s->ensure_gmres(restart);
        hipStream_t st = s->stream;
        YH_CHECK(hipMemcpyAsync(s->bvec, b, n * sizeof(double2), in, st));
        YH_CHECK(hipMemcpyAsync(s->xvec, x_inout, n * sizeof(double2), in, st));
        const auto o = s->gmres_run(channel, s->bvec, s->xvec, rtol, restart, maxiter, use_precond != 0,
                                    YASFHIP_FP64, history, history ? history_cap : 0);
        YH_CHECK(hipMemcpyAsync(x_inout, s->xvec, n * sizeof(double2), out, st));
        YH_CHECK(hipStreamSynchronize(st));
        if (info)
            *info = o.info;
        if (iterations)
            *iterations = o.iterations;
        if (residual)
            *residual = o.bnorm > 0 ? o.rnorm / o.bnorm : 0.0;
    });
}

}  // namespace

int yasfhip_gmres_solve(yasfhip_session *s, int channel, const double *b, double *x_inout,
                        double rtol, int restart, int maxiter, int use_precond, int *info,
                        int *iterations, double *residual, double *history, int history_cap) {
    return gmres_solve(s, channel, b, x_inout, rtol, restart, maxiter, use_precond, info, iterations, residual,
                       history, history_cap, false);
}

int yasfhip_gmres_solve_device(yasfhip_session *s, int channel, const double *b, double *x_inout,
                               double rtol, int restart, int maxiter, int use_precond, int *info,
                               int *iterations, double *residual, double *history, int history_cap) {
    return gmres_solve(s, channel, b, x_inout, rtol, restart, maxiter, use_precond, info, iterations, residual,
                       history, history_cap, true);
}

int yasfhip_gmres_solve_refined(yasfhip_session *s, int channel, const double *b, double *x_inout,
                                double rtol, double inner_rtol, int restart, int inner_maxiter,
                                int max_refinements, int use_precond, int *info, int *refinements,
                                int *iterations, double *residual, double *history,
                                int history_cap) {
    return guarded([&] {
        YH_REQUIRE(s && b && x_inout, "NULL argument");
        YH_REQUIRE(channel >= 0 && channel < s->W, "channel out of range");
        YH_REQUIRE(s->t_set, "T-matrix not set (yasfhip_set_tmatrix)");
        YH_REQUIRE(rtol >= 0.0, "rtol must be >= 0");
        YH_REQUIRE(inner_rtol > 0.0 && inner_rtol < 1.0, "inner_rtol must be in (0, 1)");
        YH_REQUIRE(inner_maxiter >= 1, "inner_maxiter must be >= 1");
        YH_REQUIRE(max_refinements >= 1, "max_refinements must be >= 1");
        if (use_precond)
            YH_REQUIRE(s->precond && s->precond_channel == channel,
                       "preconditioner not prepared for this channel");
        const long n = s->n;
        if (restart <= 0)
            restart = 20;
        restart = static_cast<int>(std::min<long>(restart, n));
        s->ensure_gmres(restart);
        if (!s->rb) {
            s->rb = dmalloc<double2>(n);
            s->rx = dmalloc<double2>(n);
            s->rr = dmalloc<double2>(n);
        }
        hipStream_t st = s->stream;
        if (s->inner_precision == YASFHIP_FP32)
            s->table_for(channel, true);  // the inner solves' table first: they make most matvecs
        else
            s->half_table_for(channel, s->inner_precision == YASFHIP_FP16X2);
        YH_CHECK(hipMemcpyAsync(s->rb, b, n * sizeof(double2), hipMemcpyHostToDevice, st));
        YH_CHECK(hipMemcpyAsync(s->rx, x_inout, n * sizeof(double2), hipMemcpyHostToDevice, st));
        const double bnorm = s->norm_dev(s->rb);
        int steps = 0, inner = 0, nhist = 0;
        double rnorm = 0.0;
        // FP64 residual r = b - A x of the current iterate
        auto residual64 = [&] {
            s->matvec_dev(channel, s->rx, s->avec, 1, YASFHIP_FP64);
            launch_residual(s->rr, s->rb, s->avec, n, s->reducer(), s->scal, st);
            s->fetch_scalars(1);
            rnorm = std::sqrt(s->scal_host[0]);
            if (history && nhist < history_cap)
                history[nhist++] = bnorm > 0 ? rnorm / bnorm : 0.0;
        };
        if (bnorm == 0.0) {
            YH_CHECK(hipMemsetAsync(s->rx, 0, n * sizeof(double2), st));
        } else {
            residual64();
            while (rnorm > rtol * bnorm && steps < max_refinements) {
                // correction A d = r by GMRES with the complex64 pair GEMM, asked
                // for the reduction still missing, but not below inner_rtol
                const double target = std::max(inner_rtol, 0.5 * rtol * bnorm / rnorm);
                YH_CHECK(hipMemsetAsync(s->xvec, 0, n * sizeof(double2), st));
                const auto o = s->gmres_run(channel, s->rr, s->xvec, target, restart, inner_maxiter,
                                            use_precond != 0, s->inner_precision, nullptr, 0);
                inner += o.iterations;
                ++steps;
                launch_add(s->rx, s->xvec, n, st);
                const double before = rnorm;
                residual64();
                if (o.iterations == 0 || rnorm >= before)
                    break;  // no progress: the complex64 operator cannot reduce r further
            }
        }
        YH_CHECK(hipMemcpyAsync(x_inout, s->rx, n * sizeof(double2), hipMemcpyDeviceToHost, st));
        YH_CHECK(hipStreamSynchronize(st));
        if (info)
            *info = (rnorm <= rtol * bnorm) ? 0 : 1;
        if (refinements)
            *refinements = steps;
        if (iterations)
            *iterations = inner;
        if (residual)
            *residual = bnorm > 0 ? rnorm / bnorm : 0.0;
    });
}

int yasfhip_cross_sections(yasfhip_session *s, int channel, const double *a_inc, const double *f,
                           double *c_ext, double *c_sca) {
    return guarded([&] {
        YH_REQUIRE(s && a_inc && f && c_ext && c_sca, "NULL argument");
        YH_REQUIRE(channel >= 0 && channel < s->W, "channel out of range");
        const long n = s->n;
        hipStream_t st = s->stream;
        YH_CHECK(hipMemcpyAsync(s->xin, f, n * sizeof(double2), hipMemcpyHostToDevice, st));
        YH_CHECK(hipMemcpyAsync(s->avec, a_inc, n * sizeof(double2), hipMemcpyHostToDevice, st));
        if (s->csca_method == YASFHIP_CSCA_PAIR_SUM) {
            // y = W_J f (regular translation, coincident pairs included), then
            // <f, y> and <a, f> (conj on the left operand)
            s->pair_gemm_regular(channel, s->xin);
            launch_contract(s->dp, s->N, s->Q, s->xin, s->t, 0, s->yout, false, false, st);
            launch_mgs_step(s->yout, nullptr, nullptr, s->xin, n, s->reducer(), s->scal, st);
            launch_mgs_step(s->xin, nullptr, nullptr, s->avec, n, s->reducer(), s->scal + 4, st);
            s->fetch_scalars(8);
        } else {
            // C_sca by quadrature of the far field, C_ext from <a, f>
            double r2 = 0.0;
            for (int i = 0; i < s->N; ++i) {
                const double *p = &s->pos_host[3 * static_cast<size_t>(i)];
                r2 = std::max(r2, p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            }
            int nth = 0, nph = 0;
            csca_quadrature_grid(s->k[channel] * std::sqrt(r2), s->L, &nth, &nph);
            DeviceBuffer<double> part(csca_quadrature_partials(nth, nph));
            launch_csca_quadrature(s->pos, s->xin, s->N, s->L, s->k[channel], nth, nph, part.get(), s->scal, st);
            launch_mgs_step(s->xin, nullptr, nullptr, s->avec, n, s->reducer(), s->scal + 4, st);
            s->fetch_scalars(8);
        }
        const double k2 = s->k[channel] * s->k[channel];
        *c_ext = -s->scal_host[4] / k2 * M_PI;
        *c_sca = std::hypot(s->scal_host[0], s->scal_host[1]) / k2 * M_PI;
    });
}

int yasfhip_set_scattering_method(yasfhip_session *s, int method) {
    return guarded([&] {
        YH_REQUIRE(s, "session is NULL");
        YH_REQUIRE(method == YASFHIP_CSCA_FAR_FIELD || method == YASFHIP_CSCA_PAIR_SUM,
                   "method must be YASFHIP_CSCA_FAR_FIELD or YASFHIP_CSCA_PAIR_SUM");
        s->csca_method = method;
    });
}

int yasfhip_set_inner_precision(yasfhip_session *s, int precision) {
    return guarded([&] {
        YH_REQUIRE(s, "session is NULL");
        YH_REQUIRE(precision == YASFHIP_FP32 || precision == YASFHIP_FP16X2 || precision == YASFHIP_FP16,
                   "inner precision must be YASFHIP_FP32, YASFHIP_FP16X2 or YASFHIP_FP16");
        s->inner_precision = precision;
    });
}

int yasfhip_far_field(yasfhip_session *s, int channel, const double *f, int num_directions,
                      const double *polar, const double *azimuthal, double *e_theta, double *e_phi) {
    return guarded([&] {
        YH_REQUIRE(s && f && polar && azimuthal && e_theta && e_phi, "NULL argument");
        YH_REQUIRE(channel >= 0 && channel < s->W, "channel out of range");
        YH_REQUIRE(num_directions >= 1, "num_directions must be >= 1");
        const int A = num_directions;
        const int Apad = (A + 15) / 16 * 16;
        std::vector<double4> dir(Apad, make_double4(0, 0, 1, 0));
        for (int a = 0; a < A; ++a) {
            YH_REQUIRE(std::isfinite(polar[a]) && std::isfinite(azimuthal[a]), "angles must be finite");
            const double sth = std::sin(polar[a]);
            dir[a] = make_double4(sth * std::cos(azimuthal[a]), sth * std::sin(azimuthal[a]),
                                  std::cos(polar[a]), 0.0);
        }
        hipStream_t st = s->stream;
        // per-call buffers: the far field is evaluated once per solution
        double4 *ddir = upload(dir, st);
        std::vector<double> ang(2 * static_cast<size_t>(A));
        std::copy(polar, polar + A, ang.begin());
        std::copy(azimuthal, azimuthal + A, ang.begin() + A);
        double *dang = upload(ang, st);
        double2 *F = dmalloc<double2>(static_cast<size_t>(A) * s->nmax);
        double2 *eout = dmalloc<double2>(2 * static_cast<size_t>(A));
        auto release = [&] {
            (void)hipStreamSynchronize(st);
            for (void *p : {static_cast<void *>(ddir), static_cast<void *>(dang), static_cast<void *>(F),
                            static_cast<void *>(eout)})
                (void)hipFree(p);
        };
        try {
            YH_CHECK(hipMemcpyAsync(s->xin, f, s->n * sizeof(double2), hipMemcpyHostToDevice, st));
            launch_far_field(s->pos, s->xin, s->N, s->L, s->k[channel], ddir, dang, dang + A, A, F, eout,
                             eout + A, st);
            YH_CHECK(hipMemcpyAsync(e_theta, eout, A * sizeof(double2), hipMemcpyDeviceToHost, st));
            YH_CHECK(hipMemcpyAsync(e_phi, eout + A, A * sizeof(double2), hipMemcpyDeviceToHost, st));
            YH_CHECK(hipStreamSynchronize(st));
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}

int yasfhip_precond_build(yasfhip_session *s, int channel, const int *box_of_particle,
                          int num_boxes) {
    return guarded([&] {
        YH_REQUIRE(s && box_of_particle, "NULL argument");
        YH_REQUIRE(channel >= 0 && channel < s->W, "channel out of range");
        YH_REQUIRE(s->t_set, "T-matrix not set (yasfhip_set_tmatrix)");
        YH_REQUIRE(num_boxes >= 1 && num_boxes <= s->N, "num_boxes out of range");
        for (int i = 0; i < s->N; ++i)
            YH_REQUIRE(box_of_particle[i] >= 0 && box_of_particle[i] < num_boxes,
                       "box index out of range");
        if (s->precond) {
            precond_destroy(s->precond);
            s->precond = nullptr;
            s->precond_channel = -1;
        }
        s->precond = precond_create(s->plan, translation_table_cached(s->L), s->pos_host.data(), s->N,
                                    box_of_particle, num_boxes, s->k[channel],
                                    s->t + static_cast<size_t>(channel) * s->n, s->precond_fp32, s->stream);
        s->precond_channel = channel;
    });
}

int yasfhip_set_precond_factor_precision(yasfhip_session *s, int precision) {
    return guarded([&] {
        YH_REQUIRE(s, "NULL argument");
        YH_REQUIRE(precision == YASFHIP_FP64 || precision == YASFHIP_FP32, "unknown precision");
        s->precond_fp32 = precision == YASFHIP_FP32;
    });
}

int yasfhip_precond_apply(yasfhip_session *s, const double *r, double *z) {
    return guarded([&] {
        YH_REQUIRE(s && r && z, "NULL argument");
        YH_REQUIRE(s->precond, "preconditioner not prepared (yasfhip_precond_build)");
        YH_CHECK(hipMemcpyAsync(s->xin, r, s->n * sizeof(double2), hipMemcpyHostToDevice, s->stream));
        precond_apply(s->precond, s->xin, s->yout, s->stream);
        YH_CHECK(hipMemcpyAsync(z, s->yout, s->n * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
        YH_CHECK(hipStreamSynchronize(s->stream));
    });
}

int yasfhip_set_matvec_mode(yasfhip_session *s, int mode) {
    return guarded([&] {
        YH_REQUIRE(s, "NULL argument");
        YH_REQUIRE(mode == YASFHIP_MATVEC_AUTO || mode == YASFHIP_MATVEC_ON_THE_FLY ||
                       mode == YASFHIP_MATVEC_TABLE || mode == YASFHIP_MATVEC_SLAB,
                   "unknown matvec mode");
        s->matvec_mode = mode;
        if (mode == YASFHIP_MATVEC_ON_THE_FLY || mode == YASFHIP_MATVEC_SLAB)
            for (int ix = 0; ix < 2; ++ix)
                if (s->table[ix]) {
                    YH_CHECK(hipFree(s->table[ix]));
                    s->table[ix] = nullptr;
                    s->table_channel[ix] = -1;
                }
        if ((mode == YASFHIP_MATVEC_ON_THE_FLY || mode == YASFHIP_MATVEC_TABLE) && s->slab) {
            YH_CHECK(hipFree(s->slab));
            s->slab = nullptr;
            s->slab_bytes = 0;
        }
        s->table_refused[0] = s->table_refused[1] = false;
        s->slab_refused = false;
    });
}

int yasfhip_mie_coefficients(int lmax, int particle_number, const double *radius, const double *relative_index,
                             double k_medium, double *t) {
    return guarded([&] {
        YH_REQUIRE(radius && relative_index && t, "NULL argument");
        YH_REQUIRE(lmax >= 1 && lmax <= kMaxLmax, "lmax must be in [1, 8]");
        YH_REQUIRE(particle_number >= 1, "particle_number must be >= 1");
        YH_REQUIRE(std::isfinite(k_medium) && k_medium > 0.0, "k_medium must be real, finite and positive");
        const int N = particle_number;
        const size_t n = static_cast<size_t>(N) * 2 * lmax * (lmax + 2);
        std::vector<double4> sph(N);
        for (int i = 0; i < N; ++i) {
            YH_REQUIRE(radius[i] > 0.0 && std::isfinite(radius[i]), "radii must be positive and finite");
            sph[i] = make_double4(radius[i], relative_index[2 * i], relative_index[2 * i + 1], 0.0);
        }
        DeviceBuffer<double4> d_sph(N);
        DeviceBuffer<double2> d_t(n);
        YH_CHECK(hipMemcpy(d_sph.get(), sph.data(), sizeof(double4) * N, hipMemcpyHostToDevice));
        launch_mie(d_sph.get(), N, lmax, k_medium, d_t.get(), nullptr);
        YH_CHECK(hipMemcpy(t, d_t.get(), sizeof(double2) * n, hipMemcpyDeviceToHost));
    });
}

int yasfhip_planewave_coefficients(int lmax, int particle_number, const double *positions, double k_medium,
                                   double polar, double azimuthal, int polarization, double amplitude,
                                   const double *focal_point, double *a) {
    return guarded([&] {
        YH_REQUIRE(positions && a, "NULL argument");
        YH_REQUIRE(lmax >= 1 && lmax <= kMaxLmax, "lmax must be in [1, 8]");
        YH_REQUIRE(particle_number >= 1, "particle_number must be >= 1");
        YH_REQUIRE(std::isfinite(k_medium) && k_medium > 0.0, "k_medium must be real, finite and positive");
        const int N = particle_number, nmax = 2 * lmax * (lmax + 2);
        const size_t n = static_cast<size_t>(N) * nmax;
        std::vector<double2> c(nmax);
        planewave_angular(lmax, polar, azimuthal, polarization, amplitude, c.data());
        std::vector<double4> pos(N);
        for (int i = 0; i < N; ++i)
            pos[i] = make_double4(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], 0.0);
        const double zero[3] = {0.0, 0.0, 0.0};
        DeviceBuffer<double4> d_pos(N);
        DeviceBuffer<double2> d_c(nmax), d_a(n);
        YH_CHECK(hipMemcpy(d_pos.get(), pos.data(), sizeof(double4) * N, hipMemcpyHostToDevice));
        YH_CHECK(hipMemcpy(d_c.get(), c.data(), sizeof(double2) * nmax, hipMemcpyHostToDevice));
        launch_planewave(d_pos.get(), N, lmax, k_medium, polar, azimuthal, focal_point ? focal_point : zero,
                         d_c.get(), d_a.get(), nullptr);
        YH_CHECK(hipMemcpy(a, d_a.get(), sizeof(double2) * n, hipMemcpyDeviceToHost));
    });
}

int yasfhip_set_near_field_radius(yasfhip_session *s, double radius) {
    return guarded([&] {
        YH_REQUIRE(s, "NULL argument");
        YH_REQUIRE(!std::isnan(radius), "radius must not be NaN");
        YH_CHECK(hipStreamSynchronize(s->stream));
        s->near2 = (radius >= 0.0 && std::isfinite(radius)) ? radius * radius : kAllPairs;
        s->table_channel[0] = s->table_channel[1] = -1;  // tables are rebuilt on their next use
    });
}

int yasfhip_set_slab_tiles(yasfhip_session *s, int tiles) {
    return guarded([&] {
        YH_REQUIRE(s, "NULL argument");
        YH_REQUIRE(tiles >= 0, "tiles must be >= 0");
        YH_CHECK(hipStreamSynchronize(s->stream));
        if (s->slab)
            YH_CHECK(hipFree(s->slab));
        s->slab = nullptr;
        s->slab_bytes = 0;
        s->slab_refused = false;
        s->slab_tiles_req = tiles;
    });
}

int yasfhip_tune_matvec(yasfhip_session *s, int channel, int precision, int reps, int *grouping,
                        double *ms) {
    return guarded([&] {
        YH_REQUIRE(s && grouping && ms, "NULL argument");
        YH_REQUIRE(channel >= 0 && channel < s->W, "channel out of range");
        YH_REQUIRE(precision == YASFHIP_FP64 || precision == YASFHIP_FP32, "unknown precision");
        YH_REQUIRE(reps >= 1, "reps must be >= 1");
        const bool fp32 = precision == YASFHIP_FP32;
        const int ix = fp32 ? 1 : 0;
        for (int g = 0; g < kTableShapes; ++g)
            ms[g] = -1.0;
        *grouping = s->table_grouping[ix];
        hipStream_t st = s->stream;
        // the resident table, else the first slab of one (tile range [0, ntl))
        const void *tab = nullptr;
        int ntl = -1;
        if (s->matvec_mode != YASFHIP_MATVEC_SLAB && s->table_for(channel, fp32)) {
            tab = s->table[ix];
        } else if (const int tiles = s->slab_tiles(fp32)) {
            ntl = tiles;
            launch_pair_table_build(s->dp, s->pos, s->N, s->k[channel], fp32, s->slab, st, 0, ntl, nullptr, false,
                                    s->near2);
            tab = s->slab;
        } else {
            return;  // pair factors on the fly: nothing to tune
        }
        // realistic operands: the GEMM's speed depends on its data
        launch_fill_test(reinterpret_cast<double *>(s->xin), 2 * s->n, st);
        const void *xv = s->xin;
        if (fp32) {
            launch_to_fp32(reinterpret_cast<float *>(s->x32), reinterpret_cast<const double *>(s->xin),
                           2 * s->n, st);
            xv = s->x32;
        }
        double best = 1e300;
        for (int g = 0; g < kTableShapes; ++g) {
            if (!table_shape_offered(g, fp32))
                continue;  // wider FP64 jobs exceed a wave's registers
            launch_pair_table_gemm(s->dp, s->N, tab, xv, fp32, g, s->Q, st, 0, ntl);  // warm-up
            YH_CHECK(hipEventRecord(s->ev[0], st));
            for (int r = 0; r < reps; ++r)
                launch_pair_table_gemm(s->dp, s->N, tab, xv, fp32, g, s->Q, st, 0, ntl);
            YH_CHECK(hipEventRecord(s->ev[1], st));
            YH_CHECK(hipEventSynchronize(s->ev[1]));
            float t = 0.f;
            YH_CHECK(hipEventElapsedTime(&t, s->ev[0], s->ev[1]));
            ms[g] = t / reps;
            if (ms[g] < best) {
                best = ms[g];
                *grouping = g;
            }
        }
        s->table_grouping[ix] = *grouping;
    });
}

int yasfhip_set_table_shape(yasfhip_session *s, int precision, int shape) {
    return guarded([&] {
        YH_REQUIRE(s, "NULL argument");
        YH_REQUIRE(precision == YASFHIP_FP64 || precision == YASFHIP_FP32, "unknown precision");
        YH_REQUIRE(table_shape_offered(shape, precision == YASFHIP_FP32),
                   "table GEMM shape not offered in this precision");
        s->table_grouping[precision == YASFHIP_FP32 ? 1 : 0] = shape;
    });
}

int yasfhip_time_matvec(yasfhip_session *s, int channel, int precision, int reps, double *ms_pair,
                        double *ms_total) {
    return guarded([&] {
        YH_REQUIRE(s && ms_pair && ms_total, "NULL argument");
        YH_REQUIRE(channel >= 0 && channel < s->W, "channel out of range");
        YH_REQUIRE(reps >= 1, "reps must be >= 1");
        YH_REQUIRE(precision >= YASFHIP_FP64 && precision <= YASFHIP_FP16, "unknown precision");
        const bool fp32 = precision == YASFHIP_FP32;
        const bool half = precision == YASFHIP_FP16X2 || precision == YASFHIP_FP16;
        const bool split = precision == YASFHIP_FP16X2;
        hipStream_t st = s->stream;
        double pair_ms = 0.0, total_ms = 0.0;
        launch_fill_test(reinterpret_cast<double *>(s->xin), 2 * s->n, st);  // realistic operands
        if (fp32)
            launch_to_fp32(reinterpret_cast<float *>(s->x32), reinterpret_cast<const double *>(s->xin),
                           2 * s->n, st);
        if (half) {
            s->half_table_for(channel, split);
            launch_half_x(s->dp, s->N, s->xin, s->xh, split, s->hcolmax, s->hxexp, st);
        }
        // Back-to-back launches between two events, as the solver issues
        // them: reps pair GEMMs, then reps whole matvecs (pair GEMM +
        // contraction).  A (re)build of the table is not part of either.
        const void *xv = fp32 ? static_cast<const void *>(s->x32) : static_cast<const void *>(s->xin);
        if (half)
            s->half_table_for(channel, split);
        else
            s->table_for(channel, fp32);
        auto pair = [&] {
            if (half)
                launch_half_gemm(s->dp, s->N, s->htable, s->xh, split, s->hpexp, s->hxexp, s->Q, st);
            else
                s->pair_gemm(channel, xv, fp32);
        };
        pair();  // warm-up
        YH_CHECK(hipEventRecord(s->ev[0], st));
        for (int r = 0; r < reps; ++r)
            pair();
        YH_CHECK(hipEventRecord(s->ev[1], st));
        for (int r = 0; r < reps; ++r) {
            pair();
            launch_contract(s->dp, s->N, s->Q, s->xin, s->t + static_cast<size_t>(channel) * s->n, 0,
                            s->yout, fp32 || half, false, st);
        }
        YH_CHECK(hipEventRecord(s->ev[2], st));
        YH_CHECK(hipEventSynchronize(s->ev[2]));
        float a = 0.f, b = 0.f;
        YH_CHECK(hipEventElapsedTime(&a, s->ev[0], s->ev[1]));
        YH_CHECK(hipEventElapsedTime(&b, s->ev[1], s->ev[2]));
        pair_ms = a;
        total_ms = b;
        *ms_pair = pair_ms / reps;
        *ms_total = total_ms / reps;
    });
}

}  // extern "C"
