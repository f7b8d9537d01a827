// Batched serving of many independent solves on one device.
//
// Requests (geometry, lmax, wave number, T-matrix diagonal, incident
// coefficients) are admitted first-in first-out while they fit in a device
// memory arena (one allocation made at start-up, carved per request; nothing
// is allocated per request).  An admitted request keeps its complex64
// pair-factor table, Krylov basis, Hessenberg/Givens state and vectors
// resident until it completes.
//
// Each request runs GMRES-IR: restart cycles of GMRES whose operator applies
// the complex64 pair GEMM (contraction, epilogue, Arnoldi and Givens in FP64),
// and after every cycle the FP64 residual b - A x (FP64 matvec: over a
// resident FP64 table when the request is at most a quarter of the arena,
// otherwise with the pair factors generated on the fly)
// decides convergence (||r|| <= rtol ||b||, the reference's criterion,
// yasfpy/solver.py:108-121) and seeds the next cycle, which is asked for the
// reduction max(inner_rtol, rtol ||b|| / (2 ||r||)).  x0 = b as in
// yasfpy/scattered_field_coefficients.py:95-156.
//
// One engine step advances every request that is inside a cycle by one
// Arnoldi iteration with a fixed sequence of batched launches (one grid over
// all such requests each): normalise + complex64 copy, pair GEMM over the
// resident tables (one launch per tile shape), contraction + master epilogue,
// classical Gram-Schmidt with one reorthogonalisation (CGS2: dots, reduce,
// update + dots, reduce, update + norm), and the Givens update.  The
// sequence is captured once per set of active requests in a HIP graph and
// replayed; requests outside a cycle skip every kernel through their device
// state word.  Cycle ends, FP64 residuals, admission and completion run
// between replays.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>

#include "../../include/yasfhip.h"
#include "givens.hpp"
#include "half_gemm.hpp"
#include "table_gemm.hpp"

namespace yasfhip {

namespace {

constexpr int kChunk = 2048;        // vector elements per block of the Arnoldi kernels
constexpr int kVecThreads = 256;
constexpr int kPerThread = kChunk / kVecThreads;
constexpr int kMaxRestart = 128;
// Tile shapes of the batched complex64 pair GEMM: waves per (job, target
// tile) block and the column-tile grouping of the job list (index into
// kTableGrouping).  Orders per wave, ceil((2L+1)/waves), must not exceed
// kShapeNipmax (5 with 2 or 4 waves, 3 with 8).
struct TileShape {
    int waves, grouping;
};
constexpr int kShapes = 6;
constexpr TileShape kShape[kShapes] = {{2, 0}, {4, 0}, {8, 0}, {4, 1}, {8, 1}, {8, 2}};
constexpr int kMaxGemmGroups = kShapes;
constexpr int kLargeRequest = 1000;  // particles from which a request uses the large-size tile shape

enum : int { kIdle = 0, kArnoldi = 1, kCycleDone = 2 };

// Static description of an active request (device copy rebuilt when the set
// of active requests changes).
struct BatchReq {
    const double4 *pos;
    const float2 *table;  // complex64 pair factors in MFMA A layout
    const double2 *t;     // T-matrix diagonal
    double2 *V;           // [restart+1][n] Krylov basis (FP64)
    float2 *x32;          // complex64 copy of the current basis vector
    float *Q;             // pair-GEMM output [ntiles1][nq][16][16] complex64
    double2 *partial;     // [chunks][restart+1] dot-product partials
    double *npartial;     // [2][chunks] squared-norm partials (before, after)
    double2 *H;           // [restart][restart+1] Hessenberg, column-major
    double2 *h1, *h2;     // [restart+1] Gram-Schmidt coefficients of the two passes
    double2 *g;           // [restart+1] rotated residual vector
    double *gc;           // [restart] Givens cosines
    double2 *gs;          // [restart] Givens sines
    const PairJob *jobs;
    const int *colmap, *crow, *cidx, *dmoff, *order_tiles;
    const double2 *cval;
    // operators with an fp16-term table (the fp16 pair GEMM): the table, its
    // order exponents, the bins of up to 8 column tiles it runs, and the
    // request's fp16 copy of the basis vector with its column maxima and
    // exponents
    const _Float16 *htab;
    const int *pexp;
    const HalfBin *hbins;
    _Float16 *xh;
    unsigned *colmax;
    int *xexp;
    long n;
    int N, lmax, nmax, ntiles1, N4, nq, maxcols, chunks, restart, N32;
};

// Dynamic state, written by the device inside a cycle and by the host between.
struct ReqState {
    int state;
    int j;          // columns done in this cycle; V[j] is the next basis vector
    int breakdown;
    int pad;
    double inv_norm;  // pending scale of V[j] (the norm found by the Givens step)
    double target;    // absolute presid target of the cycle
    double presid;
};

// Largest r with start[r] <= b (start[0] = 0).
__device__ __forceinline__ int find_owner(const int *__restrict__ start, int count, int b) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return uniform(lo);
}

__device__ __forceinline__ double2 wave_sum2(double2 v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_down(v.x, off, 64);
        v.y += __shfl_down(v.y, off, 64);
    }
    return v;
}

// V[j] *= inv_norm (normalising the vector the previous Givens step sized)
// and its complex64 copy for the pair GEMM.  Grid: chunks of every request.
__global__ __launch_bounds__(kVecThreads) void serve_prep_kernel(const BatchReq *__restrict__ reqs,
                                                                 const ReqState *__restrict__ st,
                                                                 const int *__restrict__ cstart, int nreq) {
    const int r = find_owner(cstart, nreq, blockIdx.x);
    if (st[r].state != kArnoldi)
        return;
    const BatchReq &q = reqs[r];
    const long c = blockIdx.x - cstart[r];
    const double s = st[r].inv_norm;
    double2 *v = q.V + static_cast<long>(st[r].j) * q.n;
    const long e0 = c * kChunk, e1 = min(e0 + kChunk, q.n);
    if (q.colmax && c == 0)  // the column maxima of the fp16 copy start from zero
        for (int i = threadIdx.x; i < q.nmax; i += kVecThreads)
            q.colmax[i] = 0u;
    for (long e = e0 + threadIdx.x; e < e1; e += kVecThreads) {
        double2 a = v[e];
        a.x *= s;
        a.y *= s;
        v[e] = a;
        q.x32[e] = make_float2(static_cast<float>(a.x), static_cast<float>(a.y));
    }
}

// The fp16 pair GEMM of the requests whose operator holds an fp16-term
// table, in three batched steps over the normalised V[j]: column maxima,
// the fp16 terms, and the GEMM (complex64 Q as the complex64 GEMM leaves it).
// Grids: per request, hs[] block ranges (find_owner), hidx[] its index.
__global__ __launch_bounds__(256) void serve_half_colmax_kernel(const BatchReq *__restrict__ reqs,
                                                                const ReqState *__restrict__ st,
                                                                const int *__restrict__ hs,
                                                                const int *__restrict__ hidx, int nh) {
    const int h = find_owner(hs, nh, blockIdx.x);
    const int r = uniform(hidx[h]);
    if (st[r].state != kArnoldi)
        return;
    const BatchReq &q = reqs[r];
    half_colmax_block(q.V + static_cast<long>(st[r].j) * q.n, q.N, q.nmax, q.colmax, blockIdx.x - hs[h]);
}

template <bool SPLIT>
__global__ __launch_bounds__(256) void serve_half_convert_kernel(const BatchReq *__restrict__ reqs,
                                                                 const ReqState *__restrict__ st,
                                                                 const int *__restrict__ hs,
                                                                 const int *__restrict__ hidx, int nh) {
    const int h = find_owner(hs, nh, blockIdx.x);
    const int r = uniform(hidx[h]);
    if (st[r].state != kArnoldi)
        return;
    const BatchReq &q = reqs[r];
    const long idx = (blockIdx.x - hs[h]) * 256L + threadIdx.x;
    if (idx >= static_cast<long>(q.nmax + 1) * q.N32 * 32)
        return;
    _Float16 *xlo = q.xh + 2 * half_x_plane(q.nmax, q.N32);
    half_convert_elem<SPLIT>(q.V + static_cast<long>(st[r].j) * q.n, q.N, q.N32, q.nmax, q.colmax, q.xexp, q.xh,
                             xlo, idx);
}

constexpr int kServeHalfWaves = 8;  // 2 orders per wave (the plan's fp16 jobs)

template <bool SPLIT>
__global__ __launch_bounds__(64 * kServeHalfWaves) void serve_half_gemm_kernel(const BatchReq *__restrict__ reqs,
                                                                               const ReqState *__restrict__ st,
                                                                               const int *__restrict__ hs,
                                                                               const int *__restrict__ hidx, int nh) {
    const int h = find_owner(hs, nh, blockIdx.x);
    const int r = uniform(hidx[h]);
    if (st[r].state != kArnoldi)
        return;
    const BatchReq &q = reqs[r];
    const int b = blockIdx.x - hs[h];
    const int bid = b / q.ntiles1;
    const int t = b - bid * q.ntiles1;
    const int P = 2 * q.lmax + 1;
    const size_t f0 = static_cast<size_t>(t) * P * P * q.N32 * 1024;
    const size_t term = static_cast<size_t>(q.ntiles1) * half_tile_halves(q.lmax, q.N32);
    const size_t xplane = half_x_plane(q.nmax, q.N32);
    float *Qt = q.Q + 2 * static_cast<size_t>(t) * q.nq * 256;
    half_gemm_block<kServeHalfWaves, kHalfJobTiles, SPLIT>(q.htab + f0, SPLIT ? q.htab + term + f0 : nullptr,
                                                           HalfOneSource{q.xh, q.xexp, Qt}, xplane, q.N32, q.nmax,
                                                           q.lmax, q.hbins + bid, q.colmap, q.dmoff, q.pexp);
}

// Requests r0 .. r0+n-1 of an operator group (ridx[]) as the sources of one
// fp16 GEMM block; qoff: the target tile's offset into each one's Q.
struct HalfGroupSource {
    static constexpr bool kMulti = true;
    const BatchReq *reqs;
    const ReqState *st;
    const int *ridx;
    int r0, n;
    size_t qoff;
    __device__ int req(int i) const { return uniform(ridx[r0 + i]); }
    __device__ int count() const { return n; }
    __device__ bool live(int i) const { return st[req(i)].state == kArnoldi; }
    __device__ const _Float16 *x(int i) const { return reqs[req(i)].xh; }
    __device__ const int *exps(int i) const { return reqs[req(i)].xexp; }
    __device__ float *out(int i) const { return reqs[req(i)].Q + qoff; }
};

// The same fp16 GEMM for the nr requests ridx[] of one operator over its
// fp16 table or a slab of it: target tiles t0 .. t0+ntl-1, tab at tile t0's
// hi term, the lo term term halves further.  A block computes one (bin,
// tile) of bins (the plan's bins of at most kHalfJobTiles / per column
// tiles) for up to per requests, each table fragment loaded once for all of
// them.  Logical block lb runs request group lb % ngrp of (bin, tile) lb /
// ngrp, so the groups' blocks of one table slice run together; the hardware
// hands block b to XCD b % 8, so block b takes lb = (b % 8) (gridDim / 8) +
// b / 8 and those blocks also share one XCD's L2 (the grid is padded to a
// multiple of 8; blocks past total return).
template <bool SPLIT>
__global__ __launch_bounds__(64 * kServeHalfWaves) void serve_half_shared_gemm_kernel(
    const BatchReq *__restrict__ reqs, const ReqState *__restrict__ st, const int *__restrict__ ridx, int nr,
    int per, const HalfBin *__restrict__ bins, const _Float16 *__restrict__ tab, size_t term, int t0, int ntl,
    int total) {
    const int lb = static_cast<int>(blockIdx.x % 8) * static_cast<int>(gridDim.x / 8) + static_cast<int>(blockIdx.x / 8);
    if (lb >= total)
        return;
    const int ngrp = (nr + per - 1) / per;
    const int grp = lb % ngrp;
    const int rest = lb / ngrp;
    const int tl = rest % ntl;
    const int bid = rest / ntl;
    const int r0 = grp * per, n = min(per, nr - r0);
    int live = 0;
    for (int i = 0; i < n; ++i)
        live |= st[uniform(ridx[r0 + i])].state == kArnoldi;
    if (!live)
        return;
    const BatchReq &q = reqs[uniform(ridx[r0])];  // operator data: the same in every request of the group
    const int P = 2 * q.lmax + 1;
    const size_t f0 = static_cast<size_t>(tl) * P * P * q.N32 * 1024;
    const size_t xplane = half_x_plane(q.nmax, q.N32);
    const HalfGroupSource src{reqs, st, ridx, r0, n, 2 * static_cast<size_t>(t0 + tl) * q.nq * 256};
    half_gemm_block<kServeHalfWaves, kHalfJobTiles, SPLIT>(tab + f0, SPLIT ? tab + term + f0 : nullptr, src, xplane,
                                                           q.N32, q.nmax, q.lmax, bins + bid, q.colmap, q.dmoff,
                                                           q.pexp);
}

// One (job, target tile t) block of a request's complex64 pair GEMM; tab
// points at target tile t of the request's pair table (or of a slab of it).
// Table GEMM shape of the server's FP64 products (residuals, C_sca): 8 waves
// computing one order at a time over jobs of up to 6 column tiles, the
// session default (each table slice streamed once).
constexpr int kServeShape64 = 2;
// Share of the arena an operator's resident inner table (complex64 or fp16
// terms) may take before the operator streams slabs instead: a
// 16,295-monomer aggregate's 172 GB table and its 16 requests (Q sized for
// ranges of FP64 products, fp64_range_tiles) fit together.
constexpr double kInnerShare = 0.875;
// Target tiles per FP64 range: a request's Q holds the complex64 products of
// all its tiles and the FP64 ones of this many, so its FP64 products (the
// residuals, the pair-sum C_sca) are formed and contracted range by range --
// half of the tiles where the FP64 products of all of them would take more
// than 256 MB, all of them below (one range: no extra launches)
inline int fp64_range_tiles(int ntiles1, int nq) {
    const size_t all = static_cast<size_t>(ntiles1) * nq * 256 * sizeof(double2);
    return all <= (size_t(256) << 20) ? ntiles1 : (ntiles1 + 1) / 2;
}
// completions whose C_sca quadratures share one launch, at most
constexpr int kMaxQuadBatch = 12;
// Table GEMM shape of a request's single complex64 product (the residual of
// x0 = b): 8 waves with jobs of up to 8 column tiles, the session default.
constexpr int kServeShape32 = 2;

// Orders per wave the serving GEMM instances hold: whole orders of blocks of
// 4 and 8 waves at lmax 8, and of 2-wave blocks up to lmax 4.
constexpr int kServeNip2 = 5;
constexpr int kServeNip4 = table_gemm_nipmax<4>(2 * kMaxLmax + 1);
constexpr int kServeNip8 = table_gemm_nipmax<8>(2 * kMaxLmax + 1);

template <int WAVES, int NIPMAX, int NTLMAX>
__device__ __forceinline__ void serve_gemm_block(const BatchReq &q, int jid, int t, const float2 *tabtile) {
    const int dm = uniform(q.jobs[jid].dm);
    const int tile0 = uniform(q.jobs[jid].tile0);
    const int ntile = uniform(q.jobs[jid].ntile);
    const int qbase = uniform(q.jobs[jid].qbase);
    const int tstride = uniform(q.jobs[jid].tstride);
    const int lmax = q.lmax;
    const int P = 2 * lmax + 1;
    const int R = P - (dm < 0 ? -dm : dm);
    const int d = dm + 2 * lmax;
    const int nord = uniform(q.jobs[jid].nord);  // orders computed (R is the table's order stride)
    const TableGemmOperands<float> op{tabtile + static_cast<size_t>(q.dmoff[d]) * q.N4 * 64,
                                      q.x32,
                                      q.N,
                                      q.N4,
                                      q.nmax,
                                      R,
                                      tile0,
                                      q.order_tiles + d * P + (dm < 0 ? -dm : dm),
                                      q.colmap + d * q.maxcols + tile0 * 16};
    float *Qt = q.Q + 2 * (static_cast<size_t>(t) * q.nq + qbase) * 256;
    table_gemm_block<float, WAVES, NIPMAX, NTLMAX>(op, nord, ntile, Qt, tstride);
}

// Complex64 pair GEMM over the resident tables of a group of requests that
// share one tile shape (WAVES waves per (job, target tile) block, jobs of at
// most NTLMAX column tiles).
template <int WAVES, int NIPMAX, int NTLMAX>
__global__ __launch_bounds__(64 * WAVES) void serve_gemm_kernel(const BatchReq *__restrict__ reqs,
                                                                const ReqState *__restrict__ st,
                                                                const int *__restrict__ gstart,
                                                                const int *__restrict__ gidx, int ngrp) {
    const int gi = find_owner(gstart, ngrp, blockIdx.x);
    const int r = uniform(gidx[gi]);
    if (st[r].state != kArnoldi)
        return;
    const BatchReq &q = reqs[r];
    const int b = blockIdx.x - gstart[gi];
    const int jid = b / q.ntiles1;
    const int t = b - jid * q.ntiles1;
    const int P = 2 * q.lmax + 1;
    serve_gemm_block<WAVES, NIPMAX, NTLMAX>(q, jid, t, q.table + static_cast<size_t>(t) * P * P * q.N4 * 64);
}

// The same GEMM for the nr requests ridx[] of one operator over a slab of its
// pair table (target tiles t0 .. t0+ntl-1, slab at tile t0).  Block order:
// request fastest, so the requests' blocks of one (job, tile) run together
// and share that slice of the slab in the caches.
template <int WAVES, int NIPMAX, int NTLMAX>
__global__ __launch_bounds__(64 * WAVES) void serve_slab_gemm_kernel(const BatchReq *__restrict__ reqs,
                                                                     const ReqState *__restrict__ st,
                                                                     const int *__restrict__ ridx, int nr,
                                                                     const float2 *__restrict__ slab, int t0,
                                                                     int ntl) {
    const int ri = blockIdx.x % nr;
    const int rest = blockIdx.x / nr;
    const int tl = rest % ntl;
    const int jid = rest / ntl;
    const int r = uniform(ridx[ri]);
    if (st[r].state != kArnoldi)
        return;
    const BatchReq &q = reqs[r];
    const int P = 2 * q.lmax + 1;
    serve_gemm_block<WAVES, NIPMAX, NTLMAX>(q, jid, t0 + tl, slab + static_cast<size_t>(tl) * P * P * q.N4 * 64);
}

// flag = 1 when any of the nr requests ridx[] is inside a cycle, else 0
// (gates the slab builds of an operator).
__global__ void serve_any_active_kernel(const ReqState *__restrict__ st, const int *__restrict__ ridx, int nr,
                                        int *__restrict__ flag) {
    int any = 0;
    for (int i = threadIdx.x; i < nr; i += blockDim.x)
        any |= st[ridx[i]].state == kArnoldi;
    any = __syncthreads_or(any);
    if (threadIdx.x == 0)
        *flag = any;
}

// V[j+1] = V[j] - t (W V[j]) from the complex64 Q, FP64 accumulation.
// Grid: (16-particle target tile, group of 16 target modes) of every request.
__global__ __launch_bounds__(256) void serve_contract_kernel(const BatchReq *__restrict__ reqs,
                                                             const ReqState *__restrict__ st,
                                                             const int *__restrict__ tstart, int nreq) {
    const int r = find_owner(tstart, nreq, blockIdx.x);
    if (st[r].state != kArnoldi)
        return;
    const BatchReq &q = reqs[r];
    const int j = st[r].j;
    const double2 *v = q.V + static_cast<long>(j) * q.n;
    double2 *w = q.V + static_cast<long>(j + 1) * q.n;
    const int groups = contract_groups(q.nmax);
    const int b = blockIdx.x - tstart[r];
    const int s1t = b / groups;
    contract_tile<float, double, double>(q.Q, q.nq, q.crow, q.cidx, q.cval, q.N, q.nmax,
                                         reinterpret_cast<const double *>(v), q.t, 1,
                                         reinterpret_cast<double *>(w), s1t, b - s1t * groups);
}

// Classical Gram-Schmidt passes over w = V[j+1], one block per chunk:
//   mode 0: partial[c][k] = <V_k, w>_c (k <= j), npartial[0][c] = ||w||_c^2
//   mode 1: w -= sum_k h1[k] V_k, then partial[c][k] = <V_k, w>_c
//   mode 2: w -= sum_k h2[k] V_k, then npartial[1][c] = ||w||_c^2
// (<a, b> = sum conj(a) b as np.vdot).
__global__ __launch_bounds__(kVecThreads) void serve_gs_kernel(const BatchReq *__restrict__ reqs,
                                                               const ReqState *__restrict__ st,
                                                               const int *__restrict__ cstart, int nreq,
                                                               int mode) {
    __shared__ double2 sh[kMaxRestart + 1][kVecThreads / 64];
    const int r = find_owner(cstart, nreq, blockIdx.x);
    if (st[r].state != kArnoldi)
        return;
    const BatchReq &q = reqs[r];
    const int c = blockIdx.x - cstart[r];
    const int j = st[r].j;
    const long n = q.n;
    const int R1 = q.restart + 1;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double2 *w = q.V + static_cast<long>(j + 1) * n;
    const long e0 = static_cast<long>(c) * kChunk;
    double2 wl[kPerThread];
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
        const long e = e0 + i * kVecThreads + threadIdx.x;
        wl[i] = e < n ? w[e] : make_double2(0.0, 0.0);
    }
    if (mode != 0) {
        const double2 *h = mode == 1 ? q.h1 : q.h2;
        for (int k = 0; k <= j; ++k) {
            const double2 hk = h[k];
            const double2 *vk = q.V + static_cast<long>(k) * n;
#pragma unroll
            for (int i = 0; i < kPerThread; ++i) {
                const long e = e0 + i * kVecThreads + threadIdx.x;
                if (e < n) {
                    const double2 a = vk[e];
                    wl[i].x -= hk.x * a.x - hk.y * a.y;
                    wl[i].y -= hk.x * a.y + hk.y * a.x;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kPerThread; ++i) {
            const long e = e0 + i * kVecThreads + threadIdx.x;
            if (e < n)
                w[e] = wl[i];
        }
    }
    if (mode != 2) {
        for (int k = 0; k <= j; ++k) {
            const double2 *vk = q.V + static_cast<long>(k) * n;
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int i = 0; i < kPerThread; ++i) {
                const long e = e0 + i * kVecThreads + threadIdx.x;
                if (e < n) {
                    const double2 a = vk[e];
                    acc.x += a.x * wl[i].x + a.y * wl[i].y;
                    acc.y += a.x * wl[i].y - a.y * wl[i].x;
                }
            }
            acc = wave_sum2(acc);
            if (lane == 0)
                sh[k][wv] = acc;
        }
    }
    if (mode != 1) {
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < kPerThread; ++i)
            acc.x += wl[i].x * wl[i].x + wl[i].y * wl[i].y;
        acc = wave_sum2(acc);
        if (lane == 0)
            sh[kMaxRestart][wv] = acc;
    }
    __syncthreads();
    if (mode != 2) {
        for (int k = threadIdx.x; k <= j; k += kVecThreads) {
            double2 s = make_double2(0.0, 0.0);
#pragma unroll
            for (int i = 0; i < kVecThreads / 64; ++i) {
                s.x += sh[k][i].x;
                s.y += sh[k][i].y;
            }
            q.partial[static_cast<long>(c) * R1 + k] = s;
        }
    }
    if (mode != 1 && threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < kVecThreads / 64; ++i)
            s += sh[kMaxRestart][i].x;
        q.npartial[(mode == 2 ? q.chunks : 0) + c] = s;
    }
}

// h[k] = sum_c partial[c][k] in chunk order (deterministic); one block per request.
__global__ __launch_bounds__(kMaxRestart + 1) void serve_reduce_kernel(const BatchReq *__restrict__ reqs,
                                                                      const ReqState *__restrict__ st,
                                                                      int nreq, int pass) {
    const int r = blockIdx.x;
    if (r >= nreq || st[r].state != kArnoldi)
        return;
    const BatchReq &q = reqs[r];
    const int j = st[r].j, R1 = q.restart + 1;
    for (int k = threadIdx.x; k <= j; k += blockDim.x) {
        double2 s = make_double2(0.0, 0.0);
        for (int c = 0; c < q.chunks; ++c) {
            const double2 p = q.partial[static_cast<long>(c) * R1 + k];
            s.x += p.x;
            s.y += p.y;
        }
        (pass == 0 ? q.h1 : q.h2)[k] = s;
    }
}

// Hessenberg column j, Givens rotation, residual estimate and the cycle's
// stopping decision (SciPy's inner loop, iterative.py:755-795).  One thread
// per request.
__global__ void serve_givens_kernel(const BatchReq *__restrict__ reqs, ReqState *__restrict__ st,
                                    int nreq) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nreq || st[r].state != kArnoldi)
        return;
    const BatchReq &q = reqs[r];
    ReqState s = st[r];
    const int j = s.j, R1 = q.restart + 1;
    double h0sq = 0.0, nsq = 0.0;
    for (int c = 0; c < q.chunks; ++c) {
        h0sq += q.npartial[c];
        nsq += q.npartial[q.chunks + c];
    }
    double2 *col = q.H + static_cast<long>(j) * R1;
    for (int k = 0; k <= j; ++k)
        col[k] = make_double2(q.h1[k].x + q.h2[k].x, q.h1[k].y + q.h2[k].y);
    const double hn = sqrt(nsq);
    bool breakdown = false;
    col[j + 1] = make_double2(hn, 0.0);
    if (hn <= 2.220446049250313e-16 * sqrt(h0sq)) {
        col[j + 1] = make_double2(0.0, 0.0);
        breakdown = true;
    }
    s.presid = givens_column(col, j, q.gc, q.gs, q.g);
    s.inv_norm = breakdown ? 1.0 : 1.0 / hn;
    s.breakdown = breakdown ? 1 : 0;
    s.j = j + 1;
    if (s.presid <= s.target || s.j == q.restart || breakdown)
        s.state = kCycleDone;
    st[r] = s;
}

// Cycle end: y = H^-1 g (upper triangular, m columns); one thread.
__global__ void serve_backsolve_kernel(const double2 *__restrict__ H, const double2 *__restrict__ g, int m,
                                       int R1, double2 *__restrict__ y) {
    backsolve_dev(H, g, m, R1, y);
}

// b = t a, x = b (x0 = b), at admission.
__global__ void serve_rhs_kernel(const double2 *__restrict__ t, const double2 *__restrict__ a,
                                 double2 *__restrict__ b, double2 *__restrict__ x, long n) {
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        const double2 tt = t[i], aa = a[i];
        const double2 v = make_double2(tt.x * aa.x - tt.y * aa.y, tt.x * aa.y + tt.y * aa.x);
        b[i] = v;
        x[i] = v;
    }
}

// g = [rnorm, 0, ...] at a cycle start.
__global__ void serve_cycle_init_kernel(double2 *__restrict__ g, int R1, double rnorm) {
    for (int k = threadIdx.x; k < R1; k += blockDim.x)
        g[k] = make_double2(k == 0 ? rnorm : 0.0, 0.0);
}

// ---------------------------------------------------------------- host side

// First-fit device arena over one allocation, offsets aligned to 256 bytes.
class Arena {
  public:
    void init(size_t bytes) {
        YH_CHECK(hipMalloc(&base_, bytes));
        cap_ = bytes;
        free_.clear();
        free_[0] = bytes;
    }
    ~Arena() {
        if (base_)
            (void)hipFree(base_);
    }
    void *alloc(size_t bytes) {
        bytes = (bytes + 255) / 256 * 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest)
                    free_[off + bytes] = rest;
                used_[off] = bytes;
                return static_cast<char *>(base_) + off;
            }
        }
        return nullptr;
    }
    void release(void *p) {
        const size_t off = static_cast<char *>(p) - static_cast<char *>(base_);
        auto u = used_.find(off);
        if (u == used_.end())
            return;
        size_t start = off, len = u->second;
        used_.erase(u);
        auto nx = free_.lower_bound(start);
        if (nx != free_.end() && nx->first == start + len) {
            len += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == start) {
                start = pv->first;
                len += pv->second;
                free_.erase(pv);
            }
        }
        free_[start] = len;
    }
    size_t capacity() const { return cap_; }
    size_t in_use() const {
        size_t s = 0;
        for (auto &u : used_)
            s += u.second;
        return s;
    }

  private:
    void *base_ = nullptr;
    size_t cap_ = 0;
    std::map<size_t, size_t> free_, used_;
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum Phase { kQueued, kRefresh, kInCycle, kFinished };

// Geometry, T-matrix and pair tables of one operator I - T W (one arena
// block), shared by the requests submitted with the same operator key, e.g.
// the incidence directions of one cluster at one wavelength.  The complex64
// table feeds the batched GEMM (without it the request runs the on-the-fly
// pair GEMM); the FP64 table, when present, the FP64 residuals.
struct SharedOp {
    long long key = 0;
    int N = 0, L = 0, ntiles1 = 0, refs = 0;
    long n = 0;
    double k = 0.0;
    void *block = nullptr;
    double4 *pos_d = nullptr;
    double2 *t_d = nullptr;
    double4 *sph_d = nullptr;  // (radius, relative index) per sphere, for the device Mie set-up
    bool device_setup = false;
    bool table32 = false, table64 = false;
    float2 *table = nullptr;
    double2 *tab64 = nullptr;
    // without tables: a slab of slab_tiles64 FP64 target tiles (2x as many
    // complex64 tiles), rebuilt range by range in every pair GEMM
    void *slab = nullptr;
    int slab_tiles64 = 0;
    double kR = 0.0;  // k times the largest distance of a sphere from the origin (far-field quadrature)
    // fp16 terms in place of the complex64 table or slab (inner precision
    // FP16X2 / FP16): the resident table, or slabs of hslab_tiles target
    // tiles in the slab buffer; hterm: halves from the hi to the lo term;
    // order exponents and scratch
    bool half = false, half_split = false;
    void *htab = nullptr;
    int hslab_tiles = 0;
    size_t hterm = 0;
    int *pexp = nullptr;
    unsigned long long *dmin = nullptr;
};

// The active requests of one operator (and tile shape) whose GEMMs run as one
// launch: a slab operator, or a table operator shared by several requests.
struct SlabGroup {
    SharedOp *op;
    int shape;         // tile shape of these requests (index into kShape)
    int first, count;  // d_index offset of the request indices, their number
    int flag;          // index of the operator's gate word in d_flags
};

struct Request {
    long long id = 0, key = 0;
    int N = 0, L = 0, nmax = 0, ntiles1 = 0, chunks = 0;
    long n = 0;
    double k = 0.0, rtol = 1e-6;
    int max_cycles = 100;
    bool observables = true;
    std::vector<double4> pos;
    std::vector<double2> t, a;
    // device set-up instead of host t and a: sph[s] = (radius, relative
    // index), the plane wave's angular coefficients and its geometry
    bool device_setup = false;
    std::vector<double4> sph;
    std::vector<double2> acoef;
    double polar = 0.0, azimuthal = 0.0, focal[3] = {0.0, 0.0, 0.0};
    SharedOp *op = nullptr;
    // device (one arena block per request besides the shared operator)
    void *block = nullptr;
    double2 *a_d = nullptr, *b_d = nullptr, *x_d = nullptr, *y_d = nullptr, *acoef_d = nullptr;
    float2 *x32 = nullptr;
    void *xh = nullptr;  // fp16 terms of the basis vector (operators with an fp16 table)
    unsigned *colmax = nullptr;
    int *xexp = nullptr;
    void *Q = nullptr;
    double2 *V = nullptr, *partial = nullptr, *H = nullptr, *h1 = nullptr, *h2 = nullptr, *g = nullptr,
            *gs = nullptr;
    double *npartial = nullptr, *gc = nullptr;
    // progress
    Phase phase = kQueued;
    bool cycle_ended = false, first_refresh = true;
    bool approx_residual = false;  // the last refresh's residual came from the complex64 operator
    int cycles = 0, iterations = 0, shape = 2, slot = -1;
    double bnorm = 0.0, rnorm = 0.0;
    double t_submit = 0.0, t_done = 0.0;
    // results
    std::vector<double2> x_out;
    int info = 0;
    double c_ext = 0.0, c_sca = 0.0;
};

struct PlanEntry {
    CouplingPlan plan;
    DevicePlan dp{};
};

}  // namespace

}  // namespace yasfhip

using namespace yasfhip;

struct yasfhip_server {
    hipStream_t stream = nullptr;
    Arena arena;
    int restart = 50;
    double inner_rtol = 1e-4;
    bool use_graph = true;
    long long next_id = 1;
    std::map<int, PlanEntry> plans;
    // index into kShape per lmax and request size class (fewer / at least
    // kLargeRequest particles: small GEMMs want more waves per block, large
    // ones fewer)
    int shape_for_lmax[kMaxLmax + 1][2];
    std::deque<Request *> queue;
    std::vector<Request *> active;
    std::map<long long, SharedOp *> ops;  // operators by key, while referenced
    struct KeyUse {
        int N, L;
        double k;
        int uses;
    };
    std::map<long long, KeyUse> key_uses;  // keys of queued or active requests
    std::vector<int> otf;                  // active requests without a complex64 table or slab
    std::vector<SlabGroup> slab_groups;    // requests sharing a slab operator, or a table one (>= 2)
    int *d_flags = nullptr;                // gate words of the slab groups
    int cap_flags = 0;
    // 1: resident tables where they fit, else slabs; 2: slabs only; 0: pair
    // factors on the fly only
    int table_policy = 1;
    int iters_per_step = 1;  // batched Arnoldi iterations between host syncs
    int slab_tiles_req = 0;  // FP64 target tiles per slab (0: an eighth of the arena)
    int csca_method = YASFHIP_CSCA_FAR_FIELD;  // how completions compute C_sca
    int inner_precision = YASFHIP_FP32;         // pair GEMM of the inner GMRES iterations
    std::vector<int> half_reqs;                 // active requests whose operator has fp16 terms
    int half_index = 0, half_colmax_blocks = 0, half_convert_blocks = 0, half_gemm_blocks = 0;
    int half_gemm_count = 0;  // of them, those with an unshared resident fp16 table
    std::deque<Request *> finished;
    long long completed = 0;
    // batched launch data
    BatchReq *d_reqs = nullptr;
    ReqState *d_state = nullptr, *h_state = nullptr;
    int *d_index = nullptr;  // [cstart | tstart | gstart x groups | gidx x groups]
    int cap_reqs = 0, cap_index = 0;
    std::vector<int> h_index;
    std::vector<BatchReq> h_reqs;  // host copies, kept alive across their async uploads
    int chunk_blocks = 0, tile_blocks = 0;
    int grp_blocks[kMaxGemmGroups] = {}, grp_count[kMaxGemmGroups] = {};
    bool set_changed = true;
    hipGraphExec_t graph = nullptr;
    // per-request scalar results of the refresh / completion syncs
    double *scal = nullptr, *scal_host = nullptr;
    int cap_scal = 0;
    double *partials = nullptr;
    unsigned *counter = nullptr;
    // FP64 slab shared by the residuals of operators that hold neither an
    // FP64 table nor a slab of their own (refreshes run one after another on
    // the stream, so one buffer serves them all); carved from large arenas
    void *res_pool = nullptr;
    size_t res_pool_bytes = 0;
    // FP64 residual pair GEMMs run so far: resident table, own slab, residual
    // pool, on the fly
    long long residual_forms[4] = {};
    hipEvent_t ev[2] = {nullptr, nullptr};

    ~yasfhip_server() {
        if (stream)
            (void)hipStreamSynchronize(stream);
        if (graph)
            (void)hipGraphExecDestroy(graph);
        std::vector<SharedOp *> live;  // operators still referenced (their memory is the arena's)
        for (auto *r : active)
            if (r->op && std::find(live.begin(), live.end(), r->op) == live.end())
                live.push_back(r->op);
        for (auto &kv : ops)
            if (std::find(live.begin(), live.end(), kv.second) == live.end())
                live.push_back(kv.second);
        for (auto *o : live)
            delete o;
        for (auto *r : queue)
            delete r;
        for (auto *r : active)
            delete r;
        for (auto *r : finished)
            delete r;
        for (auto &p : plans)
            free_device_plan(p.second.dp);
        void *bufs[] = {d_reqs, d_state, d_index, d_flags, scal, partials, counter};
        for (void *b : bufs)
            if (b)
                (void)hipFree(b);
        if (h_state)
            (void)hipHostFree(h_state);
        if (scal_host)
            (void)hipHostFree(scal_host);
        for (auto &e : ev)
            if (e)
                (void)hipEventDestroy(e);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    Reducer reducer() const { return Reducer{partials, counter}; }

    PlanEntry &plan_for(int L) {
        auto it = plans.find(L);
        if (it != plans.end())
            return it->second;
        PlanEntry &e = plans[L];
        e.plan = coupling_plan_cached(L);
        e.dp = upload_device_plan(e.plan, stream);
        return e;
    }

    static size_t aligned(size_t &off, size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    }

    // Bytes of an operator's arena block, with the carving of its buffers.
    size_t layout_op(SharedOp &o, bool carve) {
        const DevicePlan &dp = plan_for(o.L).dp;
        size_t off = 0;
        const size_t o_pos = aligned(off, sizeof(double4) * o.ntiles1 * 16);
        const size_t o_t = aligned(off, sizeof(double2) * o.n);
        const size_t o_sph = o.device_setup ? aligned(off, sizeof(double4) * o.N) : 0;
        const size_t tab_bytes = o.half ? half_table_term_bytes(dp, o.N) * (o.half_split ? 2 : 1)
                                        : pair_table_bytes(dp, o.N, true);
        const size_t o_tab = o.table32 ? aligned(off, tab_bytes) : 0;
        const size_t o_pexp = o.half ? aligned(off, sizeof(int) * (2 * o.L + 1)) : 0;
        const size_t o_dmin = o.half ? aligned(off, sizeof(unsigned long long)) : 0;
        const size_t o_tab64 = o.table64 ? aligned(off, pair_table_bytes(dp, o.N, false)) : 0;
        const size_t o_slab =
            o.slab_tiles64 ? aligned(off, pair_table_tile_bytes(dp, o.N, false) * o.slab_tiles64) : 0;
        if (carve) {
            char *b = static_cast<char *>(o.block);
            o.pos_d = reinterpret_cast<double4 *>(b + o_pos);
            o.t_d = reinterpret_cast<double2 *>(b + o_t);
            o.sph_d = o.device_setup ? reinterpret_cast<double4 *>(b + o_sph) : nullptr;
            o.table = o.table32 && !o.half ? reinterpret_cast<float2 *>(b + o_tab) : nullptr;
            o.htab = o.table32 && o.half ? b + o_tab : nullptr;
            o.pexp = o.half ? reinterpret_cast<int *>(b + o_pexp) : nullptr;
            o.dmin = o.half ? reinterpret_cast<unsigned long long *>(b + o_dmin) : nullptr;
            o.tab64 = o.table64 ? reinterpret_cast<double2 *>(b + o_tab64) : nullptr;
            o.slab = o.slab_tiles64 ? b + o_slab : nullptr;
        }
        return off;
    }

    // Bytes of one request's own arena block, with the carving of its buffers.
    size_t layout(Request &q, bool carve) {
        const DevicePlan &dp = plan_for(q.L).dp;
        const size_t R1 = restart + 1;
        size_t off = 0;
        auto take = [&](size_t bytes) { return aligned(off, bytes); };
        const size_t o_a = take(sizeof(double2) * q.n);
        const size_t o_b = take(sizeof(double2) * q.n), o_x = take(sizeof(double2) * q.n);
        const size_t o_y = take(sizeof(double2) * R1);
        const size_t o_ac = q.device_setup ? take(sizeof(double2) * q.nmax) : 0;
        const size_t o_x32 = take(sizeof(float2) * q.n);
        const bool half = q.op && q.op->half;
        const size_t o_xh = half ? take(half_x_bytes(dp, q.N)) : 0;
        const size_t o_cm = half ? take(sizeof(unsigned) * q.nmax) : 0;
        const size_t o_xe = half ? take(sizeof(int) * (q.nmax + 1)) : 0;
        // Q: the complex64 products of all target tiles, or the FP64 ones of
        // a range of fp64_range_tiles tiles (FP64 GEMMs run range by range)
        const size_t o_Q = take(sizeof(double2) * fp64_range_tiles(q.ntiles1, dp.nq) * dp.nq * 256);
        const size_t o_V = take(sizeof(double2) * R1 * q.n);
        const size_t o_part = take(sizeof(double2) * R1 * q.chunks);
        const size_t o_np = take(sizeof(double) * 2 * q.chunks);
        const size_t o_H = take(sizeof(double2) * R1 * restart);
        const size_t o_h1 = take(sizeof(double2) * R1), o_h2 = take(sizeof(double2) * R1);
        const size_t o_g = take(sizeof(double2) * R1);
        const size_t o_gc = take(sizeof(double) * R1), o_gs = take(sizeof(double2) * R1);
        if (carve) {
            char *b = static_cast<char *>(q.block);
            q.a_d = reinterpret_cast<double2 *>(b + o_a);
            q.b_d = reinterpret_cast<double2 *>(b + o_b);
            q.x_d = reinterpret_cast<double2 *>(b + o_x);
            q.y_d = reinterpret_cast<double2 *>(b + o_y);
            q.acoef_d = q.device_setup ? reinterpret_cast<double2 *>(b + o_ac) : nullptr;
            q.x32 = reinterpret_cast<float2 *>(b + o_x32);
            q.xh = half ? b + o_xh : nullptr;
            q.colmax = half ? reinterpret_cast<unsigned *>(b + o_cm) : nullptr;
            q.xexp = half ? reinterpret_cast<int *>(b + o_xe) : nullptr;
            q.Q = b + o_Q;
            q.V = reinterpret_cast<double2 *>(b + o_V);
            q.partial = reinterpret_cast<double2 *>(b + o_part);
            q.npartial = reinterpret_cast<double *>(b + o_np);
            q.H = reinterpret_cast<double2 *>(b + o_H);
            q.h1 = reinterpret_cast<double2 *>(b + o_h1);
            q.h2 = reinterpret_cast<double2 *>(b + o_h2);
            q.g = reinterpret_cast<double2 *>(b + o_g);
            q.gc = reinterpret_cast<double *>(b + o_gc);
            q.gs = reinterpret_cast<double2 *>(b + o_gs);
        }
        return off;
    }

    // A new operator for request q.  Policy 1 takes the first that fits the
    // arena's share: complex64 + FP64 tables within a quarter of the arena;
    // the inner table within kInnerShare of it, with an FP64 slab of up to a
    // sixteenth (else alone: FP64 residuals over the residual pool or on the
    // fly); else a slab of at most an eighth.  Null
    // when the chosen form cannot be allocated now: the request waits for
    // memory rather than running in a slower form -- except in an idle
    // arena, where smaller slabs and at last the tableless form (on-the-fly
    // pair GEMMs) are tried.
    SharedOp *create_op(const Request &q) {
        auto o = std::make_unique<SharedOp>();
        o->key = q.key;
        o->N = q.N;
        o->L = q.L;
        o->ntiles1 = q.ntiles1;
        o->n = q.n;
        o->k = q.k;
        o->device_setup = q.device_setup;
        o->half = inner_precision == YASFHIP_FP16X2 || inner_precision == YASFHIP_FP16;
        o->half_split = inner_precision == YASFHIP_FP16X2;
        double r2 = 0.0;
        for (const double4 &p : q.pos)
            r2 = std::max(r2, p.x * p.x + p.y * p.y + p.z * p.z);
        o->kR = q.k * std::sqrt(r2);
        const size_t cap = arena.capacity();
        const bool idle = active.empty();
        const DevicePlan &dpl = plan_for(q.L).dp;
        const int ntiles1 = (q.N + 15) / 16;
        const size_t tile64 = pair_table_tile_bytes(dpl, q.N, false);
        const int half = (ntiles1 + 1) / 2;  // FP64 tiles whose complex64 double covers all targets
        int slab64 = slab_tiles_req > 0
                         ? slab_tiles_req
                         : static_cast<int>(std::min<size_t>(cap / 8 / tile64, 1 << 20));
        slab64 = std::min(slab64, half);
        const int slab64s = slab_tiles_req > 0
                                ? slab64
                                : std::min(slab64, static_cast<int>(std::min<size_t>(cap / 16 / tile64, 1 << 20)));
        const int min_slab = idle ? 1 : std::min(16, slab64);  // smaller slabs starve the GEMM of blocks
        auto attempt = [&](bool t32, bool t64, int slab, size_t limit) {
            o->table32 = t32;
            o->table64 = t64;
            o->slab_tiles64 = slab;
            const size_t bytes = layout_op(*o, false);
            if (bytes <= limit)
                o->block = arena.alloc(bytes);
            return o->block != nullptr;
        };
        // the inner table with an FP64 slab for the residuals: the largest
        // slab that can be carved now, down to 16 tiles (smaller slabs
        // starve the FP64 GEMM of blocks; without a slab the residuals run
        // on the fly, 1.5x the slab form's time)
        auto table_with_slab = [&]() {
            for (int t = slab64s; t >= 1 && t >= std::min(16, slab64s); t /= 2)
                if (attempt(true, false, t, static_cast<size_t>(kInnerShare * cap) + tile64 * t))
                    return true;
            return false;
        };
        bool table_form = false;  // policy 1 and the inner table fits the arena's share
        if (table_policy == 1) {
            table_form = true;
            if (!attempt(true, true, 0, cap / 4) && !table_with_slab() &&
                !attempt(true, false, 0, static_cast<size_t>(kInnerShare * cap))) {
                o->table32 = true;
                o->table64 = false;
                o->slab_tiles64 = 0;
                table_form = layout_op(*o, false) <= static_cast<size_t>(kInnerShare * cap);
            }
        }
        if (!o->block && table_form && !idle)
            return nullptr;  // wait for memory: the table fits once others complete
        if (!o->block && table_policy != 0)
            for (int tiles = slab64; tiles >= min_slab && !o->block; tiles /= 2)
                attempt(false, false, tiles, cap);
        if (!o->block && (table_policy == 0 || idle))
            attempt(false, false, 0, cap);
        if (!o->block)
            return nullptr;
        if (o->half && !o->table32) {  // fp16 slabs in the slab buffer, as many tiles as fit
            const size_t tile = half_tile_halves(q.L, (q.N + 31) / 32) * sizeof(_Float16) * (o->half_split ? 2 : 1);
            const size_t bytes = o->slab_tiles64 ? tile64 * o->slab_tiles64 : 0;
            o->hslab_tiles = static_cast<int>(std::min<size_t>(bytes / tile, ntiles1));
            if (o->hslab_tiles < 1)
                o->half = false;  // no slab (pair factors on the fly): complex64
        }
        if (o->half)
            o->hterm = (o->table32 ? ntiles1 : o->hslab_tiles) * half_tile_halves(q.L, (q.N + 31) / 32);
        layout_op(*o, true);
        hipStream_t st = stream;
        YH_CHECK(hipMemcpyAsync(o->pos_d, q.pos.data(), sizeof(double4) * q.pos.size(), hipMemcpyHostToDevice,
                                st));
        const DevicePlan &dp = plan_for(q.L).dp;
        if (q.device_setup) {  // Mie T-matrix entries on the device
            YH_CHECK(hipMemcpyAsync(o->sph_d, q.sph.data(), sizeof(double4) * q.N, hipMemcpyHostToDevice, st));
            launch_mie(o->sph_d, q.N, q.L, q.k, o->t_d, st);
        } else {
            YH_CHECK(hipMemcpyAsync(o->t_d, q.t.data(), sizeof(double2) * q.n, hipMemcpyHostToDevice, st));
        }
        if (o->half)
            launch_half_exponents(dp, o->pos_d, o->N, o->k, o->dmin, o->pexp, st);
        if (o->table32 && o->half)
            launch_half_table_range(dp, o->pos_d, o->N, o->k, o->htab, o->hterm, o->half_split, 0, o->ntiles1,
                                    o->pexp, st, kAllPairs);
        else if (o->table32)
            launch_pair_table_build(dp, o->pos_d, o->N, o->k, true, o->table, st);
        if (o->table64)
            launch_pair_table_build(dp, o->pos_d, o->N, o->k, false, o->tab64, st);
        if (o->key != 0)
            ops[o->key] = o.get();
        return o.release();
    }

    // FP64 target tiles the residual pool holds for the operator, or 0 when
    // fewer than 16 fit (smaller slabs starve the GEMM of blocks)
    int pool_slab_tiles(const SharedOp &o) {
        if (!res_pool)
            return 0;
        const size_t tile64 = pair_table_tile_bytes(plan_for(o.L).dp, o.N, false);
        const int tiles = static_cast<int>(std::min<size_t>(res_pool_bytes / tile64, o.ntiles1));
        return tiles >= std::min(16, o.ntiles1) ? tiles : 0;
    }

    void release_op(SharedOp *o) {
        if (--o->refs > 0)
            return;
        if (o->key != 0)
            ops.erase(o->key);
        arena.release(o->block);
        delete o;
    }

    void ensure_scal(int count) {
        if (cap_scal >= count)
            return;
        if (scal)
            YH_CHECK(hipFree(scal));
        if (scal_host)
            YH_CHECK(hipHostFree(scal_host));
        cap_scal = std::max(count, 2 * cap_scal);
        YH_CHECK(hipMalloc(&scal, 8 * sizeof(double) * cap_scal));
        YH_CHECK(hipHostMalloc(reinterpret_cast<void **>(&scal_host), 8 * sizeof(double) * cap_scal,
                               hipHostMallocDefault));
    }

    // Whole orders per wave must not exceed the kernel's NIPMAX (blocks of
    // 4 and 8 waves cut the leftover orders along K, so any lmax fits).
    static bool shape_valid(int L, int shape) {
        if (shape < 0 || shape >= kShapes)
            return false;
        const int waves = kShape[shape].waves;
        return waves >= 4 || (2 * L + 1 + waves - 1) / waves <= kServeNip2;
    }

    static int size_class(int N) { return N >= kLargeRequest ? 1 : 0; }

    int shape_of(int L, int N) const {
        const int sh = shape_for_lmax[L][size_class(N)];
        YH_REQUIRE(shape_valid(L, sh), "invalid tile shape for this lmax");
        return sh;
    }

    // Admission: first-in first-out while the head request fits.
    void admit() {
        while (!queue.empty()) {
            Request &q = *queue.front();
            SharedOp *op = nullptr;
            bool created = false;
            if (q.key != 0) {
                auto it = ops.find(q.key);
                if (it != ops.end())
                    op = it->second;
            }
            if (!op) {
                op = create_op(q);
                created = op != nullptr;
            }
            if (op) {
                q.op = op;  // layout() reads the operator's form
                q.block = arena.alloc(layout(q, false));
                q.op = nullptr;
            }
            if (!op || !q.block) {
                if (created) {
                    ++op->refs;
                    release_op(op);
                }
                if (active.empty()) {  // can never fit: fail it
                    q.info = -1;
                    q.phase = kFinished;
                    q.t_done = now_s();
                    finished.push_back(&q);
                    queue.pop_front();
                    drop_key(q.key);
                    ++completed;
                    continue;
                }
                return;
            }
            ++op->refs;
            q.op = op;
            layout(q, true);
            queue.pop_front();
            hipStream_t st = stream;
            if (q.device_setup) {  // plane-wave coefficients on the device
                YH_CHECK(hipMemcpyAsync(q.acoef_d, q.acoef.data(), sizeof(double2) * q.nmax,
                                        hipMemcpyHostToDevice, st));
                launch_planewave(op->pos_d, q.N, q.L, q.k, q.polar, q.azimuthal, q.focal, q.acoef_d, q.a_d, st);
            } else {
                YH_CHECK(hipMemcpyAsync(q.a_d, q.a.data(), sizeof(double2) * q.n, hipMemcpyHostToDevice, st));
            }
            hipLaunchKernelGGL(serve_rhs_kernel, dim3(std::min<long>(1024, (q.n + 255) / 256)), dim3(256), 0,
                               st, op->t_d, q.a_d, q.b_d, q.x_d, q.n);
            YH_CHECK(hipGetLastError());
            q.shape = shape_of(q.L, q.N);
            q.phase = kRefresh;
            q.first_refresh = true;
            q.slot = static_cast<int>(active.size());
            active.push_back(&q);
            ensure_state(static_cast<int>(active.size()));
            h_state[q.slot] = ReqState{kIdle, 0, 0, 0, 1.0, 0.0, 0.0};
            set_changed = true;
        }
    }

    void drop_key(long long key) {
        if (key == 0)
            return;
        auto it = key_uses.find(key);
        if (it != key_uses.end() && --it->second.uses == 0)
            key_uses.erase(it);
    }

    // FP64 residual of every request waiting for one (after admission or a
    // finished cycle), convergence decisions and cycle starts.
    void refresh() {
        std::vector<Request *> todo;
        for (auto *q : active)
            if (q->phase == kRefresh)
                todo.push_back(q);
        if (todo.empty())
            return;
        ensure_scal(static_cast<int>(active.size()));
        hipStream_t st = stream;
        for (auto *q : todo) {
            if (q->cycle_ended) {
                const int m = h_state[q->slot].j;
                hipLaunchKernelGGL(serve_backsolve_kernel, dim3(1), dim3(1), 0, st, q->H, q->g, m, restart + 1,
                                   q->y_d);
                YH_CHECK(hipGetLastError());
                launch_update(q->x_d, q->V, q->y_d, m, q->n, q->n, st);
            }
        }
        // The residual of x0 = b only seeds the first cycle (convergence is
        // decided on FP64 residuals alone), so where the operator keeps a
        // complex64 table it is formed with the complex64 pair GEMM
        std::vector<char> gemm_done(todo.size(), 0);
        for (auto *q : todo) {
            q->approx_residual = q->first_refresh;
            if (q->approx_residual)
                launch_to_fp32(reinterpret_cast<float *>(q->x32), reinterpret_cast<const double *>(q->x_d), 2 * q->n,
                               st);
        }
        for (size_t i = 0; i < todo.size(); ++i) {
            Request *q = todo[i];
            if (!q->approx_residual || gemm_done[i])
                continue;
            const SharedOp &o = *q->op;
            const DevicePlan &dp = plan_for(q->L).dp;
            if (o.table32 && o.half) {
                launch_half_x(dp, q->N, q->x_d, q->xh, o.half_split, q->colmax, q->xexp, st);
                launch_half_gemm(dp, q->N, o.htab, q->xh, o.half_split, o.pexp, q->xexp, q->Q, st);
            } else if (o.slab && o.half) {  // fp16 slabs, one build for the operator's requests
                for (size_t j = i; j < todo.size(); ++j)
                    if (todo[j]->op == q->op && todo[j]->approx_residual)
                        launch_half_x(dp, q->N, todo[j]->x_d, todo[j]->xh, o.half_split, todo[j]->colmax,
                                      todo[j]->xexp, st);
                for (int t0 = 0; t0 < o.ntiles1; t0 += o.hslab_tiles) {
                    const int ntl = std::min(o.hslab_tiles, o.ntiles1 - t0);
                    launch_half_table_range(dp, o.pos_d, o.N, o.k, o.slab, o.hterm, o.half_split, t0, ntl, o.pexp,
                                            st, kAllPairs);
                    for (size_t j = i; j < todo.size(); ++j)
                        if (todo[j]->op == q->op && todo[j]->approx_residual)
                            launch_half_gemm(dp, q->N, o.slab, todo[j]->xh, o.half_split, o.pexp, todo[j]->xexp,
                                             todo[j]->Q, st, nullptr, t0, ntl, o.hterm);
                }
                for (size_t j = i; j < todo.size(); ++j)
                    if (todo[j]->op == q->op && todo[j]->approx_residual)
                        gemm_done[j] = 1;
            } else if (o.table32) {
                launch_pair_table_gemm(dp, q->N, o.table, q->x32, true, kServeShape32, q->Q, st);
            } else if (o.slab) {  // complex64 slabs (twice the FP64 tiles), one build for the operator's requests
                const int tiles = 2 * o.slab_tiles64;
                for (int t0 = 0; t0 < o.ntiles1; t0 += tiles) {
                    const int ntl = std::min(tiles, o.ntiles1 - t0);
                    launch_pair_table_build(dp, o.pos_d, o.N, o.k, true, o.slab, st, t0, ntl);
                    for (size_t j = i; j < todo.size(); ++j)
                        if (todo[j]->op == q->op && todo[j]->approx_residual)
                            launch_pair_table_gemm(dp, q->N, o.slab, todo[j]->x32, true, kServeShape32, todo[j]->Q,
                                                   st, t0, ntl);
                }
                for (size_t j = i; j < todo.size(); ++j)
                    if (todo[j]->op == q->op && todo[j]->approx_residual)
                        gemm_done[j] = 1;
            } else {
                launch_pair(dp, o.pos_d, q->N, q->k, q->x32, true, q->Q, st);
            }
            gemm_done[i] = 1;
        }
        // FP64 residuals r = b - A x into V[0], A x into V[1]: pair GEMMs over
        // the resident FP64 table, a slab (one build per slab for all of an
        // operator's requests here), the residual pool or on the fly, range
        // by range of target tiles (fp64_range_tiles: what Q holds in FP64),
        // each range contracted before the next overwrites Q
        auto contract64 = [&](Request *r, int t0, int ntl) {
            launch_contract(plan_for(r->L).dp, r->N, r->Q, r->x_d, r->op->t_d, 1, r->V + r->n, false, false, st,
                            t0, ntl);
        };
        for (size_t i = 0; i < todo.size(); ++i) {
            Request *q = todo[i];
            const SharedOp &o = *q->op;
            const DevicePlan &dp = plan_for(q->L).dp;
            if (gemm_done[i])
                continue;
            const int range = fp64_range_tiles(o.ntiles1, dp.nq);
            const int pool_tiles = o.table64 || o.slab ? 0 : std::min(pool_slab_tiles(o), range);
            if (o.table64) {
                const size_t tile_elems = pair_table_tile_bytes(dp, q->N, false) / sizeof(double2);
                for (int t0 = 0; t0 < o.ntiles1; t0 += range) {
                    const int ntl = std::min(range, o.ntiles1 - t0);
                    launch_pair_table_gemm(dp, q->N, static_cast<const double2 *>(o.tab64) + t0 * tile_elems,
                                           q->x_d, false, kServeShape64, q->Q, st, 0, ntl);
                    contract64(q, t0, ntl);
                }
                ++residual_forms[0];
            } else if (o.slab || pool_tiles) {  // one build per slab for all of the operator's requests
                void *slab = o.slab ? o.slab : res_pool;
                const int tiles = o.slab ? std::min(o.slab_tiles64, range) : pool_tiles;
                for (int t0 = 0; t0 < o.ntiles1; t0 += tiles) {
                    const int ntl = std::min(tiles, o.ntiles1 - t0);
                    launch_pair_table_build(dp, o.pos_d, o.N, o.k, false, slab, st, t0, ntl);
                    for (size_t j = i; j < todo.size(); ++j)
                        if (todo[j]->op == q->op && !todo[j]->approx_residual) {
                            launch_pair_table_gemm(dp, q->N, slab, todo[j]->x_d, false, kServeShape64, todo[j]->Q,
                                                   st, 0, ntl);
                            contract64(todo[j], t0, ntl);
                        }
                }
                for (size_t j = i; j < todo.size(); ++j)
                    if (todo[j]->op == q->op && !todo[j]->approx_residual) {
                        gemm_done[j] = 1;
                        ++residual_forms[o.slab ? 1 : 2];
                    }
            } else {
                for (int t0 = 0; t0 < o.ntiles1; t0 += range) {
                    const int ntl = std::min(range, o.ntiles1 - t0);
                    launch_pair(dp, o.pos_d, q->N, q->k, q->x_d, false, q->Q, st, nullptr, kAllPairs, t0, ntl);
                    contract64(q, t0, ntl);
                }
                ++residual_forms[3];
            }
            gemm_done[i] = 1;
        }
        for (auto *q : todo) {
            const DevicePlan &dp = plan_for(q->L).dp;
            double2 *ax = q->V + q->n;
            const SharedOp &o = *q->op;
            if (q->approx_residual)  // the inner operator's complex64 products of all tiles
                launch_contract(dp, q->N, q->Q, q->x_d, o.t_d, 1, ax, true, false, st);
            launch_residual(q->V, q->b_d, ax, q->n, reducer(), scal + 8 * q->slot, st);
            if (q->first_refresh)
                launch_norm2(q->b_d, q->n, reducer(), scal + 8 * q->slot + 1, st);
        }
        YH_CHECK(hipMemcpyAsync(scal_host, scal, 8 * sizeof(double) * active.size(), hipMemcpyDeviceToHost,
                                st));
        YH_CHECK(hipStreamSynchronize(st));
        std::vector<Request *> done;
        for (auto *q : todo) {
            q->rnorm = std::sqrt(scal_host[8 * q->slot]);
            if (q->first_refresh) {
                q->bnorm = std::sqrt(scal_host[8 * q->slot + 1]);
                q->first_refresh = false;
            }
            if (q->cycle_ended) {
                q->iterations += h_state[q->slot].j;
                ++q->cycles;
                q->cycle_ended = false;
            }
            const bool converged = !q->approx_residual && q->rnorm <= q->rtol * q->bnorm;
            if (converged || q->cycles >= q->max_cycles || q->bnorm == 0.0) {
                q->info = converged || q->bnorm == 0.0 ? 0 : 1;
                done.push_back(q);
                continue;
            }
            // next cycle: V[0] = r / ||r|| (applied by the prep kernel)
            hipLaunchKernelGGL(serve_cycle_init_kernel, dim3(1), dim3(128), 0, st, q->g, restart + 1, q->rnorm);
            YH_CHECK(hipGetLastError());
            ReqState &s = h_state[q->slot];
            s.state = kArnoldi;
            s.j = 0;
            s.breakdown = 0;
            s.inv_norm = 1.0 / q->rnorm;
            s.target = std::max(inner_rtol, 0.5 * q->rtol * q->bnorm / q->rnorm) * q->rnorm;
            s.presid = q->rnorm;
            q->phase = kInCycle;
        }
        if (!done.empty())
            complete(done);
    }

    // Observables, results to the host and release of the arena blocks.
    void complete(std::vector<Request *> &done) {
        hipStream_t st = stream;
        // C_sca by quadrature of the far field (the converged request's
        // Krylov basis is free and holds the per-block partial sums), or by
        // the regular (J) pair sum: pair GEMMs over slabs of the regular
        // table (built once per slab for an operator's completing requests),
        // else on the fly
        std::vector<char> regular_done(done.size(), 0);
        if (csca_method == YASFHIP_CSCA_FAR_FIELD) {
            // the completions of one operator share their quadrature launches
            // (phase factors computed once per batch)
            std::vector<char> batched(done.size(), 0);
            for (size_t i = 0; i < done.size(); ++i) {
                Request *q = done[i];
                if (!q->observables || batched[i])
                    continue;
                int nth = 0, nph = 0;
                csca_quadrature_grid(q->op->kR, q->L, &nth, &nph);
                const double2 *f[kMaxQuadBatch];
                double *part[kMaxQuadBatch], *out[kMaxQuadBatch];
                const int bmax = std::min(kMaxQuadBatch, csca_quadrature_batch_max(q->L));
                int nb = 0;
                for (size_t j = i; j < done.size() && nb < bmax; ++j) {
                    Request *r = done[j];
                    if (!r->observables || batched[j] || r->op != q->op)
                        continue;
                    YH_REQUIRE(csca_quadrature_partials(nth, nph) * sizeof(double) <=
                                   static_cast<size_t>(restart + 1) * r->n * sizeof(double2),
                               "far-field quadrature partials exceed the request's Krylov basis");
                    f[nb] = r->x_d;
                    part[nb] = reinterpret_cast<double *>(r->V);
                    out[nb] = scal + 8 * r->slot + 2;
                    ++nb;
                    batched[j] = 1;
                }
                launch_csca_quadrature_batch(q->op->pos_d, f, part, out, nb, q->N, q->L, q->k, nth, nph, st);
            }
            for (auto *q : done)
                if (q->observables)
                    launch_mgs_step(q->x_d, nullptr, nullptr, q->a_d, q->n, reducer(), scal + 8 * q->slot + 5, st);
        }
        // C_sca by the regular pair sum: W_J x into V[1] range by range of
        // target tiles (slabs, one build for all of an operator's requests,
        // or on the fly), each range contracted before the next overwrites Q
        auto contract_regular = [&](Request *r, int t0, int ntl) {
            launch_contract(plan_for(r->L).dp, r->N, r->Q, r->x_d, r->op->t_d, 0, r->V + r->n, false, false, st,
                            t0, ntl);
        };
        for (size_t i = 0; i < done.size() && csca_method == YASFHIP_CSCA_PAIR_SUM; ++i) {
            Request *q = done[i];
            if (!q->observables || regular_done[i])
                continue;
            const SharedOp &o = *q->op;
            const DevicePlan &dp = plan_for(q->L).dp;
            const int range = fp64_range_tiles(o.ntiles1, dp.nq);
            if (o.slab) {
                const int tiles = std::min(o.slab_tiles64, range);
                for (int t0 = 0; t0 < o.ntiles1; t0 += tiles) {
                    const int ntl = std::min(tiles, o.ntiles1 - t0);
                    launch_pair_table_build(dp, o.pos_d, o.N, o.k, false, o.slab, st, t0, ntl, nullptr, true);
                    for (size_t j = i; j < done.size(); ++j)
                        if (done[j]->op == q->op && done[j]->observables) {
                            launch_pair_table_gemm(dp, q->N, o.slab, done[j]->x_d, false, kServeShape64, done[j]->Q,
                                                   st, 0, ntl);
                            contract_regular(done[j], t0, ntl);
                        }
                }
                for (size_t j = i; j < done.size(); ++j)
                    if (done[j]->op == q->op)
                        regular_done[j] = 1;
            } else {
                for (int t0 = 0; t0 < o.ntiles1; t0 += range) {
                    const int ntl = std::min(range, o.ntiles1 - t0);
                    launch_pair_regular(dp, o.pos_d, q->N, q->k, q->x_d, q->Q, st, t0, ntl);
                    contract_regular(q, t0, ntl);
                }
            }
            regular_done[i] = 1;
        }
        for (auto *q : done) {
            if (!q->observables || csca_method != YASFHIP_CSCA_PAIR_SUM)
                continue;
            double2 *y = q->V + q->n;
            launch_mgs_step(y, nullptr, nullptr, q->x_d, q->n, reducer(), scal + 8 * q->slot + 2, st);
            launch_mgs_step(q->x_d, nullptr, nullptr, q->a_d, q->n, reducer(), scal + 8 * q->slot + 5, st);
        }
        for (auto *q : done) {
            q->x_out.resize(q->n);
            YH_CHECK(hipMemcpyAsync(q->x_out.data(), q->x_d, sizeof(double2) * q->n, hipMemcpyDeviceToHost, st));
        }
        YH_CHECK(hipMemcpyAsync(scal_host, scal, 8 * sizeof(double) * active.size(), hipMemcpyDeviceToHost,
                                st));
        YH_CHECK(hipStreamSynchronize(st));
        const double t = now_s();
        for (auto *q : done) {
            if (q->observables) {
                const double k2 = q->k * q->k;
                q->c_sca = std::hypot(scal_host[8 * q->slot + 2], scal_host[8 * q->slot + 3]) / k2 * M_PI;
                q->c_ext = -scal_host[8 * q->slot + 5] / k2 * M_PI;
            }
            q->phase = kFinished;
            q->t_done = t;
            arena.release(q->block);
            q->block = nullptr;
            release_op(q->op);
            q->op = nullptr;
            drop_key(q->key);
            finished.push_back(q);
            ++completed;
        }
        // compact the active list and its host states
        std::vector<Request *> keep;
        std::vector<ReqState> keep_state;
        for (auto *q : active)
            if (q->phase != kFinished) {
                keep_state.push_back(h_state[q->slot]);
                q->slot = static_cast<int>(keep.size());
                keep.push_back(q);
            }
        active.swap(keep);
        for (size_t i = 0; i < keep_state.size(); ++i)
            h_state[i] = keep_state[i];
        set_changed = true;
    }

    // Room for nreq descriptors and state words (device and pinned host).
    void ensure_state(int nreq) {
        if (nreq <= cap_reqs)
            return;
        YH_CHECK(hipStreamSynchronize(stream));
        if (graph) {
            YH_CHECK(hipGraphExecDestroy(graph));
            graph = nullptr;
        }
        if (d_reqs)
            YH_CHECK(hipFree(d_reqs));
        if (d_state)
            YH_CHECK(hipFree(d_state));
        ReqState *old = h_state;
        const int cap = std::max(nreq, 2 * cap_reqs);
        YH_CHECK(hipMalloc(&d_reqs, sizeof(BatchReq) * cap));
        YH_CHECK(hipMalloc(&d_state, sizeof(ReqState) * cap));
        YH_CHECK(hipHostMalloc(reinterpret_cast<void **>(&h_state), sizeof(ReqState) * cap,
                               hipHostMallocDefault));
        std::memset(h_state, 0, sizeof(ReqState) * cap);
        if (old) {
            std::memcpy(h_state, old, sizeof(ReqState) * cap_reqs);
            YH_CHECK(hipHostFree(old));
        }
        cap_reqs = cap;
        set_changed = true;
    }

    // Device descriptors and block partitions of the current active set.
    void rebuild_batch() {
        const int nreq = static_cast<int>(active.size());
        ensure_state(nreq);
        h_reqs.assign(nreq, BatchReq{});
        std::vector<BatchReq> &br = h_reqs;
        h_index.assign(2 * (nreq + 1) + kMaxGemmGroups * 2 * (nreq + 1), 0);
        int *cstart = h_index.data(), *tstart = cstart + nreq + 1;
        int cb = 0, tb = 0;
        for (int r = 0; r < nreq; ++r) {
            Request &q = *active[r];
            const DevicePlan &dp = plan_for(q.L).dp;
            BatchReq &b = br[r];
            b.pos = q.op->pos_d;
            b.table = q.op->table;
            b.t = q.op->t_d;
            b.V = q.V;
            b.x32 = q.x32;
            b.Q = static_cast<float *>(q.Q);
            b.partial = q.partial;
            b.npartial = q.npartial;
            b.H = q.H;
            b.h1 = q.h1;
            b.h2 = q.h2;
            b.g = q.g;
            b.gc = q.gc;
            b.gs = q.gs;
            b.jobs = dp.tjobs[kShape[q.shape].grouping];
            b.colmap = dp.colmap;
            b.crow = dp.crow;
            b.cidx = dp.cidx;
            b.dmoff = dp.dmoff;
            b.order_tiles = dp.order_tiles;
            b.cval = dp.cval;
            const bool half = q.op->half;
            b.htab = half ? static_cast<const _Float16 *>(q.op->htab) : nullptr;
            b.pexp = half ? q.op->pexp : nullptr;
            b.hbins = half ? dp.hbins[q.op->half_split ? 0 : 1][0] : nullptr;
            b.xh = half ? static_cast<_Float16 *>(q.xh) : nullptr;
            b.colmax = half ? q.colmax : nullptr;
            b.xexp = half ? q.xexp : nullptr;
            b.N32 = (q.N + 31) / 32;
            b.n = q.n;
            b.N = q.N;
            b.lmax = q.L;
            b.nmax = q.nmax;
            b.ntiles1 = q.ntiles1;
            b.N4 = (q.N + 3) / 4;
            b.nq = dp.nq;
            b.maxcols = dp.maxcols;
            b.chunks = q.chunks;
            b.restart = restart;
            cstart[r] = cb;
            tstart[r] = tb;
            cb += q.chunks;
            tb += q.ntiles1 * contract_groups(q.nmax);
        }
        cstart[nreq] = cb;
        tstart[nreq] = tb;
        chunk_blocks = cb;
        tile_blocks = tb;
        for (int gi = 0; gi < kMaxGemmGroups; ++gi) {
            int *gstart = h_index.data() + 2 * (nreq + 1) + gi * 2 * (nreq + 1);
            int *gidx = gstart + nreq + 1;
            int cnt = 0, blocks = 0;
            for (int r = 0; r < nreq; ++r) {
                Request &q = *active[r];
                if (q.shape != gi || !q.op->table32 || q.op->half || op_group_size(r) > 1)
                    continue;
                gstart[cnt] = blocks;
                gidx[cnt] = r;
                blocks += plan_for(q.L).dp.ntjobs[kShape[gi].grouping] * q.ntiles1;
                ++cnt;
            }
            gstart[cnt] = blocks;
            grp_count[gi] = cnt;
            grp_blocks[gi] = blocks;
        }
        // requests over fp16 terms: block ranges of the batched column maxima
        // and fp16 terms (all of them) and of the per-request GEMM (those
        // with an unshared resident table; the others run in operator
        // groups below), with their indices
        half_reqs.clear();
        half_gemm_count = 0;
        for (int r = 0; r < nreq; ++r)
            if (active[r]->op->half) {
                half_reqs.push_back(r);
                half_gemm_count += active[r]->op->table32 && op_group_size(r) < 2;
            }
        {
            const int nh = static_cast<int>(half_reqs.size()), ng = half_gemm_count;
            half_index = static_cast<int>(h_index.size());
            h_index.resize(h_index.size() + 3 * (nh + 1) + 2 * (ng + 1), 0);
            int *hc = h_index.data() + half_index, *hv = hc + nh + 1, *hi = hv + nh + 1, *hg = hi + nh + 1,
                *hgi = hg + ng + 1;
            int bc = 0, bv = 0, bg = 0, g = 0;
            for (int h = 0; h < nh; ++h) {
                const Request &q = *active[half_reqs[h]];
                const int N32 = (q.N + 31) / 32;
                hc[h] = bc;
                hv[h] = bv;
                hi[h] = half_reqs[h];
                bc += (q.N + 255) / 256;
                bv += static_cast<int>((static_cast<long>(q.nmax + 1) * N32 * 32 + 255) / 256);
                if (q.op->table32 && op_group_size(half_reqs[h]) < 2) {
                    hg[g] = bg;
                    hgi[g++] = half_reqs[h];
                    bg += plan_for(q.L).dp.nhbins[q.op->half_split ? 0 : 1][0] * q.ntiles1;
                }
            }
            hc[nh] = bc;
            hv[nh] = bv;
            hg[ng] = bg;
            half_colmax_blocks = bc;
            half_convert_blocks = bv;
            half_gemm_blocks = bg;
        }
        otf.clear();
        slab_groups.clear();
        for (int r = 0; r < nreq; ++r) {
            SharedOp *o = active[r]->op;
            if (o->table32 && op_group_size(r) < 2)
                continue;  // in the batched GEMMs above
            if (!o->table32 && !o->slab) {
                otf.push_back(r);
                continue;
            }
            const int shape = active[r]->shape;
            bool seen = false;
            for (const auto &g : slab_groups)
                seen |= g.op == o && g.shape == shape;
            if (seen)
                continue;
            SlabGroup g{o, shape, static_cast<int>(h_index.size()), 0, static_cast<int>(slab_groups.size())};
            for (int r2 = r; r2 < nreq; ++r2)
                if (active[r2]->op == o && active[r2]->shape == shape) {
                    h_index.push_back(r2);
                    ++g.count;
                }
            slab_groups.push_back(g);
        }
        if (static_cast<int>(slab_groups.size()) > cap_flags) {
            if (d_flags)
                YH_CHECK(hipFree(d_flags));
            cap_flags = 2 * static_cast<int>(slab_groups.size());
            YH_CHECK(hipMalloc(&d_flags, sizeof(int) * cap_flags));
        }
        if (static_cast<int>(h_index.size()) > cap_index) {
            if (d_index)
                YH_CHECK(hipFree(d_index));
            cap_index = static_cast<int>(2 * h_index.size());
            YH_CHECK(hipMalloc(&d_index, sizeof(int) * cap_index));
        }
        YH_CHECK(hipMemcpyAsync(d_reqs, br.data(), sizeof(BatchReq) * nreq, hipMemcpyHostToDevice, stream));
        YH_CHECK(hipMemcpyAsync(d_index, h_index.data(), sizeof(int) * h_index.size(), hipMemcpyHostToDevice,
                                stream));
        if (graph) {
            YH_CHECK(hipGraphExecDestroy(graph));
            graph = nullptr;
        }
        set_changed = false;
    }

    // The batched complex64 pair GEMM of one tile-shape group.
    void launch_gemm(int shape, const int *gstart, const int *gidx, int count, int blocks) {
        hipStream_t st = stream;
        static_assert(kShapes == 6, "one instance per entry of kShape");
        switch (shape) {
        case 0:
            hipLaunchKernelGGL((serve_gemm_kernel<2, kServeNip2, 2>), dim3(blocks), dim3(128), 0, st, d_reqs, d_state,
                               gstart, gidx, count);
            break;
        case 1:
            hipLaunchKernelGGL((serve_gemm_kernel<4, kServeNip4, 2>), dim3(blocks), dim3(256), 0, st, d_reqs, d_state,
                               gstart, gidx, count);
            break;
        case 2:
            hipLaunchKernelGGL((serve_gemm_kernel<8, kServeNip8, 2>), dim3(blocks), dim3(512), 0, st, d_reqs, d_state,
                               gstart, gidx, count);
            break;
        case 3:
            hipLaunchKernelGGL((serve_gemm_kernel<4, kServeNip4, 4>), dim3(blocks), dim3(256), 0, st, d_reqs, d_state,
                               gstart, gidx, count);
            break;
        case 4:
            hipLaunchKernelGGL((serve_gemm_kernel<8, kServeNip8, 4>), dim3(blocks), dim3(512), 0, st, d_reqs, d_state,
                               gstart, gidx, count);
            break;
        default:
            hipLaunchKernelGGL((serve_gemm_kernel<8, kServeNip8, 8>), dim3(blocks), dim3(512), 0, st, d_reqs, d_state,
                               gstart, gidx, count);
            break;
        }
        YH_CHECK(hipGetLastError());
    }

    // Active requests that share active request r's operator and tile shape.
    int op_group_size(int r) const {
        int c = 0;
        for (const Request *q : active)
            c += q->op == active[r]->op && q->shape == active[r]->shape;
        return c;
    }

    // The fp16 GEMM of an operator group's requests over target tiles [t0,
    // t0 + ntl) of its resident fp16 table (tab, all tiles) or of a slab.
    void launch_half_shared_gemm(const SlabGroup &g, const void *tab, int t0, int ntl) {
        const SharedOp &o = *g.op;
        const PlanEntry &pe = plan_for(o.L);
        const DevicePlan &dp = pe.dp;
        // bin capacity and requests per block: cap column tiles per request,
        // kHalfJobTiles / cap requests per block.  A block costs a fixed part
        // (about 4 column tiles of one wave's MFMAs) plus its busiest wave's
        // tiles for each of its requests; the cheapest capacity wins.
        const int f = o.half_split ? 0 : 1;
        int cbest = 0, per = 1;
        double best = 0.0;
        for (int c = 0; c < kHalfBinCaps; ++c) {
            const int pc = std::max(1, std::min(g.count, kHalfJobTiles / kHalfBinCap[c]));
            const double cost = (4.0 * dp.nhbins[f][c] + static_cast<double>(pc) * dp.hbin_tiles[f][c]) *
                                ((g.count + pc - 1) / pc);
            if (c == 0 || cost < best) {
                best = cost;
                cbest = c;
                per = pc;
            }
        }
        YH_REQUIRE(per * kHalfBinCap[cbest] <= kHalfJobTiles, "fp16 group blocks exceed the column tiles");
        const long total = static_cast<long>(dp.nhbins[f][cbest]) * ntl * ((g.count + per - 1) / per);
        YH_REQUIRE(total < (1L << 31) - 8, "fp16 operator group grid too large");
        const int blocks = static_cast<int>((total + 7) / 8 * 8);
        const int *ridx = d_index + g.first;
        const _Float16 *h = static_cast<const _Float16 *>(tab);
        const HalfBin *bins = dp.hbins[f][cbest];
        if (o.half_split)
            hipLaunchKernelGGL(serve_half_shared_gemm_kernel<true>, dim3(blocks), dim3(64 * kServeHalfWaves), 0, stream,
                               d_reqs, d_state, ridx, g.count, per, bins, h, o.hterm, t0, ntl,
                               static_cast<int>(total));
        else
            hipLaunchKernelGGL(serve_half_shared_gemm_kernel<false>, dim3(blocks), dim3(64 * kServeHalfWaves), 0,
                               stream, d_reqs, d_state, ridx, g.count, per, bins, h, o.hterm, t0, ntl,
                               static_cast<int>(total));
        YH_CHECK(hipGetLastError());
    }

    // The complex64 GEMM of an operator group's requests over target tiles
    // [t0, t0 + ntl) of its resident table (all tiles) or of its slab.
    void launch_slab_gemm(const SlabGroup &g, int t0, int ntl) {
        hipStream_t st = stream;
        const SharedOp &o = *g.op;
        const int shape = g.shape;
        const int blocks = plan_for(o.L).dp.ntjobs[kShape[shape].grouping] * ntl * g.count;
        const int *ridx = d_index + g.first;
        const float2 *slab = o.table32 ? o.table : static_cast<const float2 *>(o.slab);
        switch (shape) {
        case 0:
            hipLaunchKernelGGL((serve_slab_gemm_kernel<2, kServeNip2, 2>), dim3(blocks), dim3(128), 0, st, d_reqs, d_state,
                               ridx, g.count, slab, t0, ntl);
            break;
        case 1:
            hipLaunchKernelGGL((serve_slab_gemm_kernel<4, kServeNip4, 2>), dim3(blocks), dim3(256), 0, st, d_reqs, d_state,
                               ridx, g.count, slab, t0, ntl);
            break;
        case 2:
            hipLaunchKernelGGL((serve_slab_gemm_kernel<8, kServeNip8, 2>), dim3(blocks), dim3(512), 0, st, d_reqs, d_state,
                               ridx, g.count, slab, t0, ntl);
            break;
        case 3:
            hipLaunchKernelGGL((serve_slab_gemm_kernel<4, kServeNip4, 4>), dim3(blocks), dim3(256), 0, st, d_reqs, d_state,
                               ridx, g.count, slab, t0, ntl);
            break;
        case 4:
            hipLaunchKernelGGL((serve_slab_gemm_kernel<8, kServeNip8, 4>), dim3(blocks), dim3(512), 0, st, d_reqs, d_state,
                               ridx, g.count, slab, t0, ntl);
            break;
        default:
            hipLaunchKernelGGL((serve_slab_gemm_kernel<8, kServeNip8, 8>), dim3(blocks), dim3(512), 0, st, d_reqs, d_state,
                               ridx, g.count, slab, t0, ntl);
            break;
        }
        YH_CHECK(hipGetLastError());
    }

    // The batched Arnoldi iteration (see the file comment).
    void launch_iteration() {
        const int nreq = static_cast<int>(active.size());
        const int *cstart = d_index, *tstart = d_index + nreq + 1;
        hipStream_t st = stream;
        hipLaunchKernelGGL(serve_prep_kernel, dim3(chunk_blocks), dim3(kVecThreads), 0, st, d_reqs, d_state,
                           cstart, nreq);
        for (int gi = 0; gi < kMaxGemmGroups; ++gi) {
            if (!grp_count[gi])
                continue;
            const int *gstart = d_index + 2 * (nreq + 1) + gi * 2 * (nreq + 1);
            const int *gidx = gstart + nreq + 1;
            launch_gemm(gi, gstart, gidx, grp_count[gi], grp_blocks[gi]);
        }
        if (!half_reqs.empty()) {  // fp16 terms of the requests over fp16 operators, and the GEMMs of unshared tables
            const int nh = static_cast<int>(half_reqs.size()), ng = half_gemm_count;
            const int *hc = d_index + half_index, *hv = hc + nh + 1, *hi = hv + nh + 1, *hg = hi + nh + 1,
                      *hgi = hg + ng + 1;
            const bool split = inner_precision == YASFHIP_FP16X2;
            hipLaunchKernelGGL(serve_half_colmax_kernel, dim3(half_colmax_blocks), dim3(256), 0, st, d_reqs, d_state,
                               hc, hi, nh);
            if (split)
                hipLaunchKernelGGL(serve_half_convert_kernel<true>, dim3(half_convert_blocks), dim3(256), 0, st,
                                   d_reqs, d_state, hv, hi, nh);
            else
                hipLaunchKernelGGL(serve_half_convert_kernel<false>, dim3(half_convert_blocks), dim3(256), 0, st,
                                   d_reqs, d_state, hv, hi, nh);
            if (ng > 0 && split)
                hipLaunchKernelGGL(serve_half_gemm_kernel<true>, dim3(half_gemm_blocks), dim3(64 * kServeHalfWaves),
                                   0, st, d_reqs, d_state, hg, hgi, ng);
            else if (ng > 0)
                hipLaunchKernelGGL(serve_half_gemm_kernel<false>, dim3(half_gemm_blocks), dim3(64 * kServeHalfWaves),
                                   0, st, d_reqs, d_state, hg, hgi, ng);
            YH_CHECK(hipGetLastError());
        }
        // operator groups: the requests' GEMMs block-interleaved over the
        // shared table, or over each slab, built once for all of them
        for (const auto &g : slab_groups) {
            const SharedOp &o = *g.op;
            const DevicePlan &dp = plan_for(o.L).dp;
            if (o.table32) {
                if (o.half)
                    launch_half_shared_gemm(g, o.htab, 0, o.ntiles1);
                else
                    launch_slab_gemm(g, 0, o.ntiles1);
                continue;
            }
            int *flag = d_flags + g.flag;
            hipLaunchKernelGGL(serve_any_active_kernel, dim3(1), dim3(64), 0, st, d_state, d_index + g.first,
                               g.count, flag);
            YH_CHECK(hipGetLastError());
            if (o.half) {
                for (int t0 = 0; t0 < o.ntiles1; t0 += o.hslab_tiles) {
                    const int ntl = std::min(o.hslab_tiles, o.ntiles1 - t0);
                    launch_half_table_range(dp, o.pos_d, o.N, o.k, o.slab, o.hterm, o.half_split, t0, ntl, o.pexp,
                                            st, kAllPairs, flag);
                    launch_half_shared_gemm(g, o.slab, t0, ntl);
                }
                continue;
            }
            const int tiles = 2 * o.slab_tiles64;
            for (int t0 = 0; t0 < o.ntiles1; t0 += tiles) {
                const int ntl = std::min(tiles, o.ntiles1 - t0);
                launch_pair_table_build(dp, o.pos_d, o.N, o.k, true, o.slab, st, t0, ntl, flag);
                launch_slab_gemm(g, t0, ntl);
            }
        }
        for (int r : otf) {  // requests without a complex64 table: on-the-fly pair GEMM, gated
            const Request &q = *active[r];
            launch_pair(plan_for(q.L).dp, q.op->pos_d, q.N, q.k, q.x32, true, q.Q, st,
                        reinterpret_cast<const int *>(d_state + r));
        }
        hipLaunchKernelGGL(serve_contract_kernel, dim3(tile_blocks), dim3(256), 0, st, d_reqs, d_state, tstart,
                           nreq);
        hipLaunchKernelGGL(serve_gs_kernel, dim3(chunk_blocks), dim3(kVecThreads), 0, st, d_reqs, d_state,
                           cstart, nreq, 0);
        hipLaunchKernelGGL(serve_reduce_kernel, dim3(nreq), dim3(kMaxRestart + 1), 0, st, d_reqs, d_state, nreq,
                           0);
        hipLaunchKernelGGL(serve_gs_kernel, dim3(chunk_blocks), dim3(kVecThreads), 0, st, d_reqs, d_state,
                           cstart, nreq, 1);
        hipLaunchKernelGGL(serve_reduce_kernel, dim3(nreq), dim3(kMaxRestart + 1), 0, st, d_reqs, d_state, nreq,
                           1);
        hipLaunchKernelGGL(serve_gs_kernel, dim3(chunk_blocks), dim3(kVecThreads), 0, st, d_reqs, d_state,
                           cstart, nreq, 2);
        hipLaunchKernelGGL(serve_givens_kernel, dim3((nreq + 63) / 64), dim3(64), 0, st, d_reqs, d_state, nreq);
        YH_CHECK(hipGetLastError());
    }

    // Times the batched complex64 pair GEMM of one synthetic request (N spheres
    // on a jittered cubic lattice, lmax L) for every valid tile shape and
    // keeps the fastest for L.  Needs an idle server.
    int autotune(int L, int N, int reps, double *ms) {
        YH_REQUIRE(active.empty() && queue.empty(), "autotune needs an idle server");
        Request q;
        q.N = N;
        q.L = L;
        q.nmax = 2 * L * (L + 2);
        q.n = static_cast<long>(N) * q.nmax;
        q.ntiles1 = (N + 15) / 16;
        q.chunks = static_cast<int>((q.n + kChunk - 1) / kChunk);
        q.k = 2.0 * M_PI / 500.0;
        const int side = static_cast<int>(std::ceil(std::cbrt(static_cast<double>(N))));
        q.pos.assign(q.ntiles1 * 16, make_double4(0, 0, 0, 0));
        unsigned rng = 12345u;
        auto jitter = [&] {
            rng = rng * 1664525u + 1013904223u;
            return 20.0 * ((rng >> 8) / 16777216.0 - 0.5);
        };
        for (int i = 0; i < N; ++i)
            q.pos[i] = make_double4(250.0 * (i % side) + jitter(), 250.0 * ((i / side) % side) + jitter(),
                                    250.0 * (i / (side * side)) + jitter(), 0.0);
        q.t.assign(q.n, make_double2(0.0, 0.0));
        const int policy = table_policy;
        table_policy = 1;  // the tuner times the resident-table GEMM
        SharedOp *op = nullptr;
        try {
            op = create_op(q);
        } catch (...) {
            table_policy = policy;
            throw;
        }
        table_policy = policy;
        if (op && !op->table32) {
            ++op->refs;
            release_op(op);
            op = nullptr;
        }
        YH_REQUIRE(op, "autotune request does not fit in the arena");
        ++op->refs;
        q.op = op;
        q.block = arena.alloc(layout(q, false));
        if (!q.block) {
            release_op(op);
            YH_REQUIRE(false, "autotune request does not fit in the arena");
        }
        layout(q, true);
        hipStream_t st = stream;
        std::vector<float2> xh(q.n);
        for (long i = 0; i < q.n; ++i)
            xh[i] = make_float2(static_cast<float>(std::sin(0.37 * i)), static_cast<float>(std::cos(0.21 * i)));
        YH_CHECK(hipMemcpyAsync(q.x32, xh.data(), sizeof(float2) * q.n, hipMemcpyHostToDevice, st));
        int best = -1;
        double best_ms = 1e300;
        q.slot = 0;
        active.push_back(&q);
        try {
            for (int gi = 0; gi < kShapes; ++gi) {
                ms[gi] = -1.0;
                if (!shape_valid(L, gi))
                    continue;
                q.shape = gi;
                set_changed = true;
                rebuild_batch();
                h_state[0] = ReqState{kArnoldi, 0, 0, 0, 1.0, 0.0, 0.0};
                YH_CHECK(hipMemcpyAsync(d_state, h_state, sizeof(ReqState), hipMemcpyHostToDevice, st));
                const int *gstart = d_index + 2 * 2 + gi * 2 * 2;
                launch_gemm(gi, gstart, gstart + 2, 1, grp_blocks[gi]);  // warm-up
                YH_CHECK(hipEventRecord(ev[0], st));
                for (int i = 0; i < reps; ++i)
                    launch_gemm(gi, gstart, gstart + 2, 1, grp_blocks[gi]);
                YH_CHECK(hipEventRecord(ev[1], st));
                YH_CHECK(hipEventSynchronize(ev[1]));
                float t = 0.f;
                YH_CHECK(hipEventElapsedTime(&t, ev[0], ev[1]));
                ms[gi] = t / reps;
                if (ms[gi] < best_ms) {
                    best_ms = ms[gi];
                    best = gi;
                }
            }
        } catch (...) {
            active.clear();
            arena.release(q.block);
            release_op(op);
            set_changed = true;
            throw;
        }
        YH_CHECK(hipStreamSynchronize(st));
        active.clear();
        arena.release(q.block);
        release_op(op);
        set_changed = true;
        shape_for_lmax[L][size_class(N)] = best;
        return best;
    }

    // One engine step; returns the number of requests inside a cycle.
    int step() {
        admit();
        refresh();
        admit();  // completions may have freed memory
        refresh();
        if (set_changed)
            rebuild_batch();
        int in_cycle = 0;
        for (auto *q : active)
            if (q->phase == kInCycle)
                ++in_cycle;
        if (!in_cycle)
            return 0;
        const int nreq = static_cast<int>(active.size());
        YH_CHECK(hipMemcpyAsync(d_state, h_state, sizeof(ReqState) * nreq, hipMemcpyHostToDevice, stream));
        // iters_per_step Arnoldi iterations per step: every kernel of an
        // iteration skips the requests whose device state word left
        // kArnoldi, so a cycle that ends inside the step idles its request
        // until the host refreshes it
        if (use_graph) {
            if (!graph) {
                hipGraph_t g = nullptr;
                YH_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < iters_per_step; ++i)
                    launch_iteration();
                YH_CHECK(hipStreamEndCapture(stream, &g));
                YH_CHECK(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0));
                YH_CHECK(hipGraphDestroy(g));
            }
            YH_CHECK(hipGraphLaunch(graph, stream));
        } else {
            for (int i = 0; i < iters_per_step; ++i)
                launch_iteration();
        }
        YH_CHECK(hipMemcpyAsync(h_state, d_state, sizeof(ReqState) * nreq, hipMemcpyDeviceToHost, stream));
        YH_CHECK(hipStreamSynchronize(stream));
        for (auto *q : active)
            if (q->phase == kInCycle && h_state[q->slot].state == kCycleDone) {
                q->phase = kRefresh;
                q->cycle_ended = true;
                h_state[q->slot].state = kIdle;
            }
        return in_cycle;
    }
};

namespace {

template <typename F>
int sguard(F &&f) {
    try {
        f();
        set_last_error("", kOk);
        return kOk;
    } catch (const yasfhip::Error &e) {
        return set_last_error(e.what(), e.code);
    } catch (const std::exception &e) {
        return set_last_error(e.what(), kInternal);
    }
}

}  // namespace

namespace {

// Validation, key bookkeeping and queueing shared by the submit entry points;
// fill() sets the request's operator and incidence data.
template <typename F>
void submit_request(yasfhip_server *s, const double *positions, int particle_number, int lmax, double k_medium,
                    double rtol, int max_cycles, int observables, long long operator_key, long long *request_id,
                    F &&fill) {
    YH_REQUIRE(s && positions && request_id, "NULL argument");
    YH_REQUIRE(particle_number >= 1, "particle_number must be >= 1");
    YH_REQUIRE(lmax >= 1 && lmax <= kMaxLmax, "lmax must be in [1, 8]");
    YH_REQUIRE(std::isfinite(k_medium) && k_medium > 0.0, "k_medium must be real, finite and positive");
    YH_REQUIRE(rtol > 0.0, "rtol must be > 0");
    YH_REQUIRE(max_cycles >= 1, "max_cycles must be >= 1");
    YH_REQUIRE(table_gemm_extent_ok(particle_number, 2 * lmax * (lmax + 2), lmax, sizeof(double2)),
               "particle_number too large for the pair GEMM's 32-bit offsets");
    if (operator_key != 0) {
        auto it = s->key_uses.find(operator_key);
        if (it != s->key_uses.end())
            YH_REQUIRE(it->second.N == particle_number && it->second.L == lmax && it->second.k == k_medium,
                       "operator_key already names an operator with a different size, lmax or wave number");
    }
    auto q = std::make_unique<Request>();
    q->key = operator_key;
    q->N = particle_number;
    q->L = lmax;
    q->nmax = 2 * lmax * (lmax + 2);
    q->n = static_cast<long>(q->N) * q->nmax;
    q->ntiles1 = (q->N + 15) / 16;
    q->chunks = static_cast<int>((q->n + kChunk - 1) / kChunk);
    q->k = k_medium;
    q->rtol = rtol;
    q->max_cycles = max_cycles;
    q->observables = observables != 0;
    q->pos.assign(q->ntiles1 * 16, make_double4(0, 0, 0, 0));
    for (int i = 0; i < q->N; ++i)
        q->pos[i] = make_double4(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], 0.0);
    fill(*q);
    q->id = s->next_id++;
    q->t_submit = now_s();
    *request_id = q->id;
    if (operator_key != 0) {
        auto &u = s->key_uses[operator_key];
        if (u.uses == 0)
            u = yasfhip_server::KeyUse{particle_number, lmax, k_medium, 0};
        ++u.uses;
    }
    s->queue.push_back(q.release());
}

}  // namespace

extern "C" {

int yasfhip_server_create(size_t memory_bytes, int restart, double inner_rtol, int use_graph,
                          yasfhip_server **out) {
    return sguard([&] {
        YH_REQUIRE(out, "out is NULL");
        YH_REQUIRE(restart >= 1 && restart <= kMaxRestart, "restart must be in [1, 128]");
        YH_REQUIRE(inner_rtol > 0.0 && inner_rtol < 1.0, "inner_rtol must be in (0, 1)");
        auto s = std::make_unique<yasfhip_server>();
        s->restart = restart;
        s->inner_rtol = inner_rtol;
        s->use_graph = use_graph != 0;
        for (int L = 0; L <= kMaxLmax; ++L)
            s->shape_for_lmax[L][0] = s->shape_for_lmax[L][1] = 2;  // 8 waves, jobs of up to 2 column tiles
        YH_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        for (auto &e : s->ev)
            YH_CHECK(hipEventCreate(&e));
        if (memory_bytes == 0) {
            size_t free_b = 0, total_b = 0;
            YH_CHECK(hipMemGetInfo(&free_b, &total_b));
            memory_bytes = free_b / 10 * 7;
        }
        s->arena.init(memory_bytes);
        if (memory_bytes >= (size_t(32) << 30)) {
            s->res_pool_bytes = memory_bytes / 32 / 256 * 256;
            s->res_pool = s->arena.alloc(s->res_pool_bytes);
        }
        YH_CHECK(hipMalloc(&s->partials, kRedPartials * sizeof(double)));
        YH_CHECK(hipMalloc(&s->counter, sizeof(unsigned)));
        YH_CHECK(hipMemsetAsync(s->counter, 0, sizeof(unsigned), s->stream));
        s->ensure_scal(64);
        YH_CHECK(hipStreamSynchronize(s->stream));
        *out = s.release();
    });
}

int yasfhip_server_destroy(yasfhip_server *s) {
    return sguard([&] { delete s; });
}

int yasfhip_server_submit(yasfhip_server *s, const double *positions, int particle_number, int lmax,
                          double k_medium, const double *t, const double *a_inc, double rtol, int max_cycles,
                          int observables, long long operator_key, long long *request_id) {
    return sguard([&] {
        YH_REQUIRE(t && a_inc, "NULL argument");
        submit_request(s, positions, particle_number, lmax, k_medium, rtol, max_cycles, observables, operator_key,
                       request_id, [&](Request &q) {
                           q.t.resize(q.n);
                           q.a.resize(q.n);
                           std::memcpy(q.t.data(), t, sizeof(double2) * q.n);
                           std::memcpy(q.a.data(), a_inc, sizeof(double2) * q.n);
                       });
    });
}

int yasfhip_server_submit_spheres(yasfhip_server *s, const double *positions, const double *radius,
                                  const double *relative_index, int particle_number, int lmax, double k_medium,
                                  double polar, double azimuthal, int polarization, double amplitude,
                                  const double *focal_point, double rtol, int max_cycles, int observables,
                                  long long operator_key, long long *request_id) {
    return sguard([&] {
        YH_REQUIRE(radius && relative_index, "NULL argument");
        YH_REQUIRE(std::isfinite(polar) && std::isfinite(azimuthal) && std::isfinite(amplitude),
                   "incidence angles and amplitude must be finite");
        submit_request(s, positions, particle_number, lmax, k_medium, rtol, max_cycles, observables, operator_key,
                       request_id, [&](Request &q) {
                           q.device_setup = true;
                           q.sph.resize(q.N);
                           for (int i = 0; i < q.N; ++i) {
                               YH_REQUIRE(radius[i] > 0.0 && std::isfinite(radius[i]),
                                          "radii must be positive and finite");
                               q.sph[i] = make_double4(radius[i], relative_index[2 * i],
                                                       relative_index[2 * i + 1], 0.0);
                           }
                           q.acoef.resize(q.nmax);
                           planewave_angular(lmax, polar, azimuthal, polarization, amplitude, q.acoef.data());
                           q.polar = polar;
                           q.azimuthal = azimuthal;
                           if (focal_point)
                               for (int c = 0; c < 3; ++c)
                                   q.focal[c] = focal_point[c];
                       });
    });
}

int yasfhip_server_step(yasfhip_server *s, int *active, int *queued, long long *completed) {
    return sguard([&] {
        YH_REQUIRE(s, "NULL argument");
        s->step();
        if (active)
            *active = static_cast<int>(s->active.size());
        if (queued)
            *queued = static_cast<int>(s->queue.size());
        if (completed)
            *completed = s->completed;
    });
}

int yasfhip_server_autotune(yasfhip_server *s, int lmax, int particle_number, int reps, int *waves,
                            int *column_tiles, double *ms) {
    return sguard([&] {
        YH_REQUIRE(s && waves && column_tiles && ms, "NULL argument");
        YH_REQUIRE(lmax >= 1 && lmax <= kMaxLmax, "lmax must be in [1, 8]");
        YH_REQUIRE(particle_number >= 16, "particle_number must be >= 16");
        YH_REQUIRE(reps >= 1, "reps must be >= 1");
        const int sh = s->autotune(lmax, particle_number, reps, ms);
        *waves = kShape[sh].waves;
        *column_tiles = kTableGrouping[kShape[sh].grouping];
    });
}

int yasfhip_server_set_tile_shape(yasfhip_server *s, int lmax, int waves, int column_tiles) {
    return sguard([&] {
        YH_REQUIRE(s, "NULL argument");
        YH_REQUIRE(lmax >= 1 && lmax <= kMaxLmax, "lmax must be in [1, 8]");
        int found = -1;
        for (int i = 0; i < kShapes; ++i)
            if (kShape[i].waves == waves && kTableGrouping[kShape[i].grouping] == column_tiles)
                found = i;
        YH_REQUIRE(found >= 0 && yasfhip_server::shape_valid(lmax, found),
                   "unsupported tile shape for this lmax (waves 2/4/8 x column tiles 2/4/8 as listed in "
                   "yasfhip.h, with ceil((2 lmax + 1) / waves) orders per wave within range)");
        s->shape_for_lmax[lmax][0] = s->shape_for_lmax[lmax][1] = found;
    });
}

int yasfhip_server_set_inner_precision(yasfhip_server *s, int precision) {
    return sguard([&] {
        YH_REQUIRE(s, "server is NULL");
        YH_REQUIRE(precision == YASFHIP_FP32 || precision == YASFHIP_FP16X2 || precision == YASFHIP_FP16,
                   "inner precision must be YASFHIP_FP32, YASFHIP_FP16X2 or YASFHIP_FP16");
        YH_REQUIRE(s->queue.empty() && s->active.empty(), "the inner precision changes on an idle server only");
        s->inner_precision = precision;
    });
}

int yasfhip_server_set_inner_rtol(yasfhip_server *s, double inner_rtol) {
    return sguard([&] {
        YH_REQUIRE(s, "server is NULL");
        YH_REQUIRE(inner_rtol > 0.0 && inner_rtol < 1.0, "inner_rtol must be in (0, 1)");
        s->inner_rtol = inner_rtol;
    });
}

int yasfhip_server_set_scattering_method(yasfhip_server *s, int method) {
    return sguard([&] {
        YH_REQUIRE(s, "server is NULL");
        YH_REQUIRE(method == YASFHIP_CSCA_FAR_FIELD || method == YASFHIP_CSCA_PAIR_SUM,
                   "method must be YASFHIP_CSCA_FAR_FIELD or YASFHIP_CSCA_PAIR_SUM");
        s->csca_method = method;
    });
}

int yasfhip_server_set_table_policy(yasfhip_server *s, int policy, int slab_tiles) {
    return sguard([&] {
        YH_REQUIRE(s, "NULL argument");
        YH_REQUIRE(policy >= 0 && policy <= 2, "unknown table policy");
        YH_REQUIRE(slab_tiles >= 0, "slab_tiles must be >= 0");
        s->table_policy = policy;
        s->slab_tiles_req = slab_tiles;
    });
}

int yasfhip_server_set_iterations_per_step(yasfhip_server *s, int iterations) {
    return sguard([&] {
        YH_REQUIRE(s, "NULL argument");
        YH_REQUIRE(iterations >= 1 && iterations <= 16, "iterations per step must be in [1, 16]");
        if (iterations != s->iters_per_step) {
            s->iters_per_step = iterations;
            s->set_changed = true;  // the graph is captured again
        }
    });
}

int yasfhip_server_set_residual_pool(yasfhip_server *s, size_t bytes) {
    return sguard([&] {
        YH_REQUIRE(s, "NULL argument");
        YH_REQUIRE(s->queue.empty() && s->active.empty(), "the residual pool needs an idle server");
        YH_CHECK(hipStreamSynchronize(s->stream));
        if (s->res_pool)
            s->arena.release(s->res_pool);
        s->res_pool = nullptr;
        s->res_pool_bytes = 0;
        if (bytes) {
            s->res_pool = s->arena.alloc(bytes);
            YH_REQUIRE(s->res_pool, "the residual pool does not fit the arena's free memory");
            s->res_pool_bytes = bytes;
        }
    });
}

int yasfhip_server_residual_forms(yasfhip_server *s, long long *forms) {
    return sguard([&] {
        YH_REQUIRE(s && forms, "NULL argument");
        std::copy(s->residual_forms, s->residual_forms + 4, forms);
    });
}

int yasfhip_server_poll(yasfhip_server *s, long long *request_id, double *x_out, long long x_capacity,
                        int *info, int *cycles, int *iterations, double *residual, double *c_ext,
                        double *c_sca, double *latency_s) {
    return sguard([&] {
        YH_REQUIRE(s && request_id, "NULL argument");
        *request_id = -1;
        if (s->finished.empty())
            return;
        std::unique_ptr<Request> q(s->finished.front());
        s->finished.pop_front();
        *request_id = q->id;
        if (x_out && x_capacity >= q->n && !q->x_out.empty())
            std::memcpy(x_out, q->x_out.data(), sizeof(double2) * q->n);
        if (info)
            *info = q->info;
        if (cycles)
            *cycles = q->cycles;
        if (iterations)
            *iterations = q->iterations;
        if (residual)
            *residual = q->bnorm > 0 ? q->rnorm / q->bnorm : 0.0;
        if (c_ext)
            *c_ext = q->c_ext;
        if (c_sca)
            *c_sca = q->c_sca;
        if (latency_s)
            *latency_s = q->t_done - q->t_submit;
    });
}

}  // extern "C"
