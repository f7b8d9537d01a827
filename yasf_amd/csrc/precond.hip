// Block-diagonal preconditioner on the device.
//
// Intended semantics of the reference's BlockDiagonalPreconditioner
// (yasfpy/preconditioner.py:1-14 and 92-219 docstrings): particles are grouped
// into spatial boxes; for each box the block of M = I - T W restricted to
// intra-box pairs is LU-factored with partial pivoting and applied as M_box^-1
// on the box's slice of the vector.  Unlike the reference's code this uses
// particle-major indices, the receiver's T, r_s1 - r_s2 geometry, FP64 blocks
// and P^T on the right-hand side.
//
// Storage: boxes are packed back to back, each an mp x mp row-major complex
// block (m = particles_in_box * nmax rounded up to mp, a multiple of 16, with
// the identity in the padding).
#include <algorithm>
#include <numeric>
#include <type_traits>

#include "device_common.hpp"

namespace yasfhip {

struct BlockPrecond {
    int nbox = 0, nmax = 0, lmax = 0;
    long n = 0;
    int *box_first = nullptr;    // [nbox+1] offsets into the particle order
    int *particles = nullptr;    // particle ids ordered by box
    long long *blk_off = nullptr;  // [nbox] element offset of each block
    int *piv_off = nullptr;      // [nbox] offset of each block's pivots
    int *dinv_off = nullptr;     // [nbox] first 16-row diagonal block of each box
    double2 *blocks = nullptr;   // FP64 blocks and their factors (freed once blocks32 holds the factors)
    float2 *blocks32 = nullptr;  // the factors rounded to complex64, column-major per box, when chosen
    int *pivots = nullptr;       // after the build: the row permutation of each box (see below)
    double2 *dinv = nullptr;     // per 16-row diagonal block: L11^-1 then U11^-1 (2 x 16 x 16)
    long long total_elems = 0;
    int max_m = 0;  // largest padded block size
};

// Largest padded block (rows); the apply keeps one block's right-hand side in LDS.
constexpr int kMaxBlock = 3840;
// The apply's LDS: the box's right-hand side (up to kMaxBlock complex128 rows,
// dynamic) plus yb and Dsh; above 64 KiB per workgroup it needs the
// dynamic-LDS opt-in, which the build sets.
static_assert(kMaxBlock * 16 + (16 + 256) * 16 <= 160 * 1024,
              "lu_apply_kernel exceeds the 160 KiB of LDS per workgroup");

namespace {

__device__ __forceinline__ int padded16(int m) { return (m + 15) & ~15; }

// One block per (box, target particle a1): for each source particle a2 of the
// box the pair's factors G[p, dm] (all (2L+1)^2 of them) are generated once
// into LDS, then the nmax x nmax sub-block
//   M[(a1,n1),(a2,n2)] = -t[s1,n1] * sum_p ab5[n2,n1,p] G[p, m2-m1]
// is written through a CSR over (n1, n2) of the nonzero ab5 entries
// (the reference's coupling formula, cpu_numba.py:74-91).  Diagonal
// sub-blocks are the identity.
__global__ __launch_bounds__(256) void build_blocks_kernel(
    const double4 *__restrict__ pos, const int *__restrict__ box_first, const int *__restrict__ particles,
    const long long *__restrict__ blk_off, const int2 *__restrict__ rowblocks, int nmax, int lmax,
    double k, const double2 *__restrict__ t, const int *__restrict__ erow, const int *__restrict__ eg,
    const double2 *__restrict__ ev, const int *__restrict__ dmoff, const double *__restrict__ la,
    const double *__restrict__ lb, const double *__restrict__ ld, double2 *__restrict__ blocks) {
    __shared__ double2 G[(2 * kMaxLmax + 1) * (2 * kMaxLmax + 1)];
    const int b = rowblocks[blockIdx.x].x, a1 = rowblocks[blockIdx.x].y;
    const int first = box_first[b];
    const int np_box = box_first[b + 1] - first;
    const int mp = padded16(np_box * nmax);
    double2 *A = blocks + blk_off[b] + static_cast<long>(a1) * nmax * mp;
    const int P = 2 * lmax + 1, ndm = 4 * lmax + 1;
    const int s1 = particles[first + a1];
    const double4 p1 = pos[s1];
    const int nn = nmax * nmax;
    for (int a2 = 0; a2 < np_box; ++a2) {
        if (a2 == a1) {
            for (int e = threadIdx.x; e < nn; e += blockDim.x) {
                const int n1 = e / nmax, n2 = e - n1 * nmax;
                A[static_cast<long>(n1) * mp + a2 * nmax + n2] = make_double2(n1 == n2 ? 1.0 : 0.0, 0.0);
            }
            continue;
        }
        const double4 p2 = pos[particles[first + a2]];
        if (threadIdx.x < ndm) {
            const int dm = static_cast<int>(threadIdx.x) - 2 * lmax;
            pair_series<double2>(p1.x - p2.x, p1.y - p2.y, p1.z - p2.z, k, dm, P, true, la, lb, ld,
                                 G + dmoff[threadIdx.x], 1);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            const int n1 = e / nmax, n2 = e - n1 * nmax;
            double sr = 0.0, si = 0.0;
            for (int q = erow[e]; q < erow[e + 1]; ++q) {
                const double2 a = ev[q], g = G[eg[q]];
                sr += a.x * g.x - a.y * g.y;
                si += a.x * g.y + a.y * g.x;
            }
            const double2 tt = t[static_cast<long>(s1) * nmax + n1];
            A[static_cast<long>(n1) * mp + a2 * nmax + n2] =
                make_double2(-(tt.x * sr - tt.y * si), -(tt.x * si + tt.y * sr));
        }
        __syncthreads();
    }
}

// Identity in the padding of each box block (rows/columns m..mp-1).
__global__ __launch_bounds__(256) void pad_blocks_kernel(const int *__restrict__ box_first, int nmax,
                                                         const long long *__restrict__ blk_off,
                                                         double2 *__restrict__ blocks) {
    const int b = blockIdx.x;
    const int m = (box_first[b + 1] - box_first[b]) * nmax;
    const int mp = padded16(m);
    if (mp == m)
        return;
    double2 *A = blocks + blk_off[b];
    for (long e = threadIdx.x; e < static_cast<long>(mp) * mp; e += blockDim.x) {
        const int i = static_cast<int>(e / mp), j = static_cast<int>(e % mp);
        if (i >= m || j >= m)
            A[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
}

__device__ __forceinline__ double cabs1(double2 z) { return fabs(z.x) + fabs(z.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cinv(double2 d) {
    const double dd = d.x * d.x + d.y * d.y;
    return dd > 0.0 ? make_double2(d.x / dd, -d.y / dd) : make_double2(0.0, 0.0);
}

typedef double v4d __attribute__((ext_vector_type(4)));

// Blocked right-looking LU with partial pivoting (pivot = largest |re|+|im|
// as LAPACK's izamax), one block of kLuWaves waves per box, in place on the
// padded mp x mp block, with two levels of blocking:
// * outer panels of kLuOuter columns; inside one, 16-column blocks: unblocked
//   factorisation of the 16-column panel (pivot search, rank-1 updates within
//   the 16 columns; in registers, one row per thread, with up to
//   kLuExtraRows more rows in LDS, else in global memory), the row swaps
//   outside the panel in one deferred pass, the 16x16 diagonal inverses,
//   then the 16-row U = L^-1 A and the rank-16 update restricted to the
//   outer panel's columns;
// * then, for the columns right of the outer panel, U12 = L11^-1 A12 by block
//   forward substitution (16-row blocks, stored inverses) and one rank-64
//   update A22 -= L21 U12, with U12 staged in LDS kLuChunk columns at a time.
// The trailing matrix is read and written once per outer panel instead of
// once per 16 columns.  Complex products are four real FP64 MFMAs
// (16x16x4); the C tile layout of the MFMA (rows 4 i + (lane >> 4), column
// lane & 15) is also its B operand layout, so solved tiles feed the next
// products from registers.  The phases run as separate kernels (below).
constexpr int kLuWaves = 8;     // waves per box
// Rows of a 16-column panel the register form holds: one per thread.
constexpr int kLuPanelRows = 64 * kLuWaves;

// Argmax of |A(i, c)|_1 over the block, ties to the lower row: every thread
// returns the same row.  (bv, bi): the thread's candidate.
__device__ __forceinline__ int block_argmax(double bv, int bi, double *red_v, int *red_i) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_down(bv, off, 64);
        const int oi = __shfl_down(bi, off, 64);
        if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
        }
    }
    if (lane == 0) {
        red_v[wave] = bv;
        red_i[wave] = bi;
    }
    __syncthreads();
    double v = red_v[0];
    int ii = red_i[0];
#pragma unroll
    for (int q = 1; q < kLuWaves; ++q)
        if (red_v[q] > v || (red_v[q] == v && red_i[q] < ii)) {
            v = red_v[q];
            ii = red_i[q];
        }
    return ii;
}

// Partial-pivoting LU of the 16-column panel [kb, kb + 16) of an mp x mp
// row-major block, rows kb .. mp - 1 (at most kLuPanelRows of them) held in
// registers: thread t owns row kb + t.  Per column one barrier publishes the
// wave maxima (every thread then picks the same pivot row p) and one
// publishes rows p and c, which swap; the elimination runs on the registers.
// The row swaps outside the panel follow afterwards in one pass (each
// thread applies the 16 swaps in order to its own columns), and piv[c]
// records them (row c swapped with piv[c]).  (Two rows per thread in
// registers spilled to scratch and ran slower than the global form.)
// EXTRA: up to kLuExtraRows more rows, kb + NT + t for thread t, live in LDS
// (xrow, rows kLuExtraStride double2 apart so that a wave's rows fall on
// distinct banks) and are eliminated there four columns at a time.
constexpr int kLuExtraRows = 256;
constexpr int kLuExtraStride = 17;
template <bool EXTRA>
__device__ void panel_registers(double2 *__restrict__ A, int mp, int kb, int *__restrict__ piv, double *red_v,
                                int *red_i, double2 *prow_sh, double2 *crow_sh, double2 *xrow) {
    const int tid = threadIdx.x;
    constexpr int NT = 64 * kLuWaves;
    auto at = [&](int i, int j) -> double2 & { return A[static_cast<long>(i) * mp + j]; };
    const int r = kb + tid;
    const int r2 = kb + NT + tid;  // this thread's LDS row (EXTRA)
    const bool has2 = EXTRA && tid < kLuExtraRows && r2 < mp;
    double2 *xr = EXTRA ? xrow + tid * kLuExtraStride : nullptr;
    if (has2)
#pragma unroll
        for (int j = 0; j < 16; ++j)
            xr[j] = at(r2, kb + j);
    double2 pr[16];
#pragma unroll
    for (int j = 0; j < 16; ++j)
        pr[j] = r < mp ? at(r, kb + j) : make_double2(0.0, 0.0);
    int pv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int c = kb + j;
        double bv = -1.0;
        int bi = c;
        if (r >= c && r < mp) {
            bv = cabs1(pr[j]);
            bi = r;
        }
        if (has2) {  // r2 > r: it wins only on a strictly larger value
            const double v2 = cabs1(xr[j]);
            if (v2 > bv) {
                bv = v2;
                bi = r2;
            }
        }
        const int p = __builtin_amdgcn_readfirstlane(block_argmax(bv, bi, red_v, red_i));  // same in all threads
        pv[j] = p;
        // publish rows p (the pivot row) and c; p >= c, both in the panel
        if (r == p) {
#pragma unroll
            for (int jj = 0; jj < 16; ++jj)
                prow_sh[jj] = pr[jj];
        }
        if (has2 && r2 == p) {
#pragma unroll
            for (int jj = 0; jj < 16; ++jj)
                prow_sh[jj] = xr[jj];
        }
        if (r == c) {
#pragma unroll
            for (int jj = 0; jj < 16; ++jj)
                crow_sh[jj] = pr[jj];
        }
        __syncthreads();
        const double2 inv = cinv(prow_sh[j]);
        if (r == c) {
#pragma unroll
            for (int jj = 0; jj < 16; ++jj)
                pr[jj] = prow_sh[jj];
        } else if (r == p) {
#pragma unroll
            for (int jj = 0; jj < 16; ++jj)
                pr[jj] = crow_sh[jj];
        }
        if (has2) {  // the LDS row (below every register row): swap, then eliminate in LDS
            if (r2 == p) {
#pragma unroll
                for (int jj = 0; jj < 16; ++jj)
                    xr[jj] = crow_sh[jj];
            }
            const double2 l2 = cmul(xr[j], inv);
            xr[j] = l2;
#pragma unroll
            for (int j4 = (j + 1) & ~3; j4 < 16; j4 += 4) {
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int jj = j4; jj < j4 + 4; ++jj)
                    if (jj > j)
                        xr[jj] = csub(xr[jj], cmul(l2, prow_sh[jj]));
            }
        }
        const bool below = r > c && r < mp;
        const double2 l = below ? cmul(pr[j], inv) : make_double2(0.0, 0.0);
        if (below)
            pr[j] = l;
        // pr[jj] -= l prow[jj] (l = 0 for rows at or above c), the pivot row
        // read four columns at a time: the scheduling barriers keep the
        // compiler from holding all of it in registers beside the panel
#pragma unroll
        for (int j4 = (j + 1) & ~3; j4 < 16; j4 += 4) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int jj = j4; jj < j4 + 4; ++jj)
                if (jj > j)
                    pr[jj] = csub(pr[jj], cmul(l, prow_sh[jj]));
        }
        // the next column's maxima go to red_v / red_i, whose reads all ended
        // before the barrier above; prow_sh / crow_sh are rewritten only after
        // the next column's first barrier
    }
    if (r < mp)
#pragma unroll
        for (int j = 0; j < 16; ++j)
            at(r, kb + j) = pr[j];
    if (has2)
#pragma unroll
        for (int j = 0; j < 16; ++j)
            at(r2, kb + j) = xr[j];
    if (tid < 16)
        piv[kb + tid] = pv[tid];
    // deferred row swaps: columns outside the panel, each thread its own
    for (int jj = tid; jj < mp; jj += NT) {
        if (jj >= kb && jj < kb + 16)
            continue;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int c = kb + j, p = pv[j];
            if (p != c) {
                const double2 t0 = at(c, jj);
                at(c, jj) = at(p, jj);
                at(p, jj) = t0;
            }
        }
    }
    __syncthreads();
}

// The same panel for blocks of more rows, in global memory: per column a
// block-wide pivot search, the row swap inside the panel and the elimination
// within the panel, each followed by a barrier; the swaps outside the panel
// are deferred as above.
__device__ void panel_global(double2 *__restrict__ A, int mp, int kb, int *__restrict__ piv, double *red_v,
                             int *red_i) {
    const int tid = threadIdx.x;
    auto at = [&](int i, int j) -> double2 & { return A[static_cast<long>(i) * mp + j]; };
    int pv[16];
    for (int j = 0; j < 16; ++j) {
        const int c = kb + j;
        double bv = -1.0;
        int bi = c;
        for (int i = c + tid; i < mp; i += blockDim.x) {
            const double v = cabs1(at(i, c));
            if (v > bv) {
                bv = v;
                bi = i;
            }
        }
        const int p = __builtin_amdgcn_readfirstlane(block_argmax(bv, bi, red_v, red_i));
        pv[j] = p;
        if (p != c && tid < 16) {  // the swap inside the panel; the rest is deferred
            const double2 t0 = at(c, kb + tid);
            at(c, kb + tid) = at(p, kb + tid);
            at(p, kb + tid) = t0;
        }
        __syncthreads();
        const double2 inv = cinv(at(c, c));
        const int bend = kb + 16;
        for (int i = c + 1 + tid; i < mp; i += blockDim.x) {
            const double2 l = cmul(at(i, c), inv);
            at(i, c) = l;
            for (int jj = c + 1; jj < bend; ++jj)
                at(i, jj) = csub(at(i, jj), cmul(l, at(c, jj)));
        }
        __syncthreads();
    }
    if (tid < 16)
        piv[kb + tid] = pv[tid];
    for (int jj = tid; jj < mp; jj += blockDim.x) {
        if (jj >= kb && jj < kb + 16)
            continue;
        for (int j = 0; j < 16; ++j) {
            const int c = kb + j, p = pv[j];
            if (p != c) {
                const double2 t0 = at(c, jj);
                at(c, jj) = at(p, jj);
                at(p, jj) = t0;
            }
        }
    }
    __syncthreads();
}

constexpr int kLuOuter = 64;    // columns per outer panel
constexpr int kLuChunk = 128;   // trailing columns per LDS stage of U12
constexpr size_t kLuLds = static_cast<size_t>(kLuOuter) * kLuChunk * sizeof(double2);  // 128 KB
constexpr size_t kLuPanelLds = static_cast<size_t>(kLuExtraRows) * kLuExtraStride * sizeof(double2);

// The LU runs as three kernels per step (one block per box each), so that
// each phase gets its own register allocation: in one kernel the phases'
// live ranges met and spilled (196 VGPRs, and 1,165 with the panel's LDS
// rows), while each phase alone fits.  Per 16-column block kb of an outer
// panel K0: lu_panel_kernel (pivoting and elimination of the 16 columns),
// then lu_block_kernel (diagonal inverses, U rows inside the outer panel,
// rank-16 update); per outer panel: lu_trailing_kernel (U12 and the rank-64
// update of the trailing matrix).  Boxes smaller than the step return.

// The 16-column panel at kb: in registers (and LDS rows) up to
// kLuPanelRows + kLuExtraRows rows, else in global memory.
__global__ __launch_bounds__(64 * kLuWaves) void lu_panel_kernel(const int *__restrict__ box_first, int nmax,
                                                       const long long *__restrict__ blk_off,
                                                       const int *__restrict__ piv_off,
                                                       double2 *__restrict__ blocks, int *__restrict__ pivots,
                                                       int kb) {
    __shared__ double red_v[kLuWaves];
    __shared__ int red_i[kLuWaves];
    __shared__ double2 prow_sh[16], crow_sh[16];
    extern __shared__ double2 xrow[];  // [kLuExtraRows][kLuExtraStride]
    const int b = blockIdx.x;
    const int mp = padded16((box_first[b + 1] - box_first[b]) * nmax);
    if (kb >= mp)
        return;
    double2 *A = blocks + blk_off[b];
    int *piv = pivots + piv_off[b];
    if (mp - kb <= kLuPanelRows + kLuExtraRows)
        panel_registers<true>(A, mp, kb, piv, red_v, red_i, prow_sh, crow_sh, xrow);
    else
        panel_global(A, mp, kb, piv, red_v, red_i);
}

// After the panel at kb (in the outer panel from K0): the inverses of its
// diagonal 16x16 factors, its U rows for the columns inside the outer panel
// and the rank-16 update of those columns below it.
__global__ __launch_bounds__(64 * kLuWaves) void lu_block_kernel(const int *__restrict__ box_first, int nmax,
                                                       const long long *__restrict__ blk_off,
                                                       double2 *__restrict__ blocks,
                                                       const int *__restrict__ dinv_off,
                                                       double2 *__restrict__ dinv, int K0, int kb) {
    __shared__ double2 Lsh[16][16];
    const int b = blockIdx.x;
    const int m = (box_first[b + 1] - box_first[b]) * nmax;
    const int mp = padded16(m);
    if (kb >= mp)
        return;
    double2 *A = blocks + blk_off[b];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lo = lane & 15, hi = lane >> 4;
    auto at = [&](int i, int j) -> double2 & { return A[static_cast<long>(i) * mp + j]; };
    double2 *Dbox = dinv + static_cast<long>(dinv_off[b]) * 512;
    const int pend = min(K0 + kLuOuter, mp);  // outer panel columns [K0, pend)

    // Inverses of the diagonal 16x16 factors, kept for the
    // substitutions: thread c < 16 solves L11 y = e_c (unit lower),
    // thread 16 + c solves U11 z = e_c (upper).
    double2 *Dinv = Dbox + static_cast<long>(kb / 16) * 512;
    // (compile-time indices throughout, so yv / zv stay in registers;
    // the terms outside a column's range are multiplied by zeros)
    if (tid < 16) {
        const int c = tid;
        double2 yv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            double2 v = make_double2(r == c ? 1.0 : 0.0, 0.0);
#pragma unroll
            for (int q = 0; q < r; ++q)
                v = csub(v, cmul(at(kb + r, kb + q), yv[q]));
            yv[r] = r < c ? make_double2(0.0, 0.0) : v;
            Lsh[r][c] = yv[r];
            Dinv[r * 16 + c] = yv[r];
        }
    } else if (tid < 32) {
        const int c = tid - 16;
        double2 zv[16];
#pragma unroll
        for (int r = 15; r >= 0; --r) {
            double2 v = make_double2(r == c ? 1.0 : 0.0, 0.0);
#pragma unroll
            for (int q = r + 1; q < 16; ++q)
                v = csub(v, cmul(at(kb + r, kb + q), zv[q]));
            zv[r] = r > c ? make_double2(0.0, 0.0) : cmul(v, cinv(at(kb + r, kb + r)));
            Dinv[256 + r * 16 + c] = zv[r];
        }
    }
    __syncthreads();
    const int ntp = (pend - kb - 16) >> 4;  // column tiles right of the block, inside the panel
    if (ntp == 0)
        return;
    // U = Linv A for the block's 16 rows (each wave owns whole column
    // tiles, so reading and overwriting a tile needs no barrier).
    for (int cb = wave; cb < ntp; cb += kLuWaves) {
        const int col0 = kb + 16 + cb * 16;
        double2 bv[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
            bv[kk] = at(kb + 4 * kk + hi, col0 + lo);
        v4d cre = {0.0, 0.0, 0.0, 0.0}, cim = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const double2 a = Lsh[lo][4 * kk + hi];
            cre = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bv[kk].x, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, bv[kk].y, cre, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bv[kk].y, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, bv[kk].x, cim, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
            at(kb + 4 * i + hi, col0 + lo) = make_double2(cre[i], cim[i]);
    }
    __syncthreads();
    // rank-16 update of the panel's remaining columns, all rows below
    const int nrb = (mp - kb - 16) >> 4;
    for (int rb = wave; rb < nrb; rb += kLuWaves) {
        const int row0 = kb + 16 + rb * 16;
        double ar[4], ai[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const double2 v = at(row0 + lo, kb + 4 * kk + hi);
            ar[kk] = v.x;
            ai[kk] = v.y;
        }
        for (int cb = 0; cb < ntp; ++cb) {
            const int col0 = kb + 16 + cb * 16;
            v4d cre, cim;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double2 v = at(row0 + 4 * i + hi, col0 + lo);
                cre[i] = v.x;
                cim[i] = v.y;
            }
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const double2 bv2 = at(kb + 4 * kk + hi, col0 + lo);
                cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar[kk], bv2.x, cre, 0, 0, 0);
                cre = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[kk], bv2.y, cre, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar[kk], bv2.y, cim, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[kk], bv2.x, cim, 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
                at(row0 + 4 * i + hi, col0 + lo) = make_double2(cre[i], cim[i]);
        }
    }

}

// After the outer panel [K0, K0 + kLuOuter): U12 = L11^-1 A12 and
// A22 -= L21 U12 for the trailing matrix.
__global__ __launch_bounds__(64 * kLuWaves) void lu_trailing_kernel(const int *__restrict__ box_first, int nmax,
                                                          const long long *__restrict__ blk_off,
                                                          double2 *__restrict__ blocks,
                                                          const int *__restrict__ dinv_off,
                                                          const double2 *__restrict__ dinv, int K0) {
    extern __shared__ double2 Ush[];  // [kLuOuter][kLuChunk]
    const int b = blockIdx.x;
    const int m = (box_first[b + 1] - box_first[b]) * nmax;
    const int mp = padded16(m);
    const int pend = min(K0 + kLuOuter, mp);
    if (pend >= mp)
        return;  // no trailing matrix (also K0 >= mp)
    double2 *A = blocks + blk_off[b];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lo = lane & 15, hi = lane >> 4;
    auto at = [&](int i, int j) -> double2 & { return A[static_cast<long>(i) * mp + j]; };
    const double2 *Dbox = dinv + static_cast<long>(dinv_off[b]) * 512;

    const int nbi = (pend - K0) >> 4;  // 16-row blocks of the outer panel (<= 4)
    // U12 = L11^-1 A12 for the columns right of the panel: per column
    // tile, block forward substitution t_i = Linv_i (A12_i - sum_{j<i}
    // L_ij t_j), the tiles held in registers (element kk of a tile is
    // row 4 kk + hi: also the B operand of K-substep kk).
    const int ntc = (mp - pend) >> 4;
    for (int cb = wave; cb < ntc; cb += kLuWaves) {
        const int col0 = pend + cb * 16;
        double2 tv[kLuOuter / 16][4];
#pragma unroll
        for (int i = 0; i < kLuOuter / 16; ++i)
            if (i < nbi)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
                    tv[i][kk] = at(K0 + 16 * i + 4 * kk + hi, col0 + lo);
#pragma unroll
        for (int i = 0; i < kLuOuter / 16; ++i) {
            if (i >= nbi)
                break;
            v4d cre, cim;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                cre[kk] = tv[i][kk].x;
                cim[kk] = tv[i][kk].y;
            }
#pragma unroll
            for (int j = 0; j < i; ++j)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const double2 a = at(K0 + 16 * i + lo, K0 + 16 * j + 4 * kk + hi);
                    const double2 bt = tv[j][kk];
                    cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.x, bt.x, cre, 0, 0, 0);
                    cre = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, bt.y, cre, 0, 0, 0);
                    cim = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.x, bt.y, cim, 0, 0, 0);
                    cim = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, bt.x, cim, 0, 0, 0);
                }
            const double2 *Li = Dbox + static_cast<long>(K0 / 16 + i) * 512;  // unit-lower inverse
            v4d ure = {0.0, 0.0, 0.0, 0.0}, uim = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const double2 a = Li[lo * 16 + 4 * kk + hi];
                ure = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, cre[kk], ure, 0, 0, 0);
                ure = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, cim[kk], ure, 0, 0, 0);
                uim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, cim[kk], uim, 0, 0, 0);
                uim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, cre[kk], uim, 0, 0, 0);
            }
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                tv[i][kk] = make_double2(ure[kk], uim[kk]);
        }
#pragma unroll
        for (int i = 0; i < kLuOuter / 16; ++i)
            if (i < nbi)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
                    at(K0 + 16 * i + 4 * kk + hi, col0 + lo) = tv[i][kk];
    }
    __syncthreads();
    // A22 -= L21 U12 (K = pend - K0), U12 staged in LDS kLuChunk columns at
    // a time; each wave holds the L21 fragments of one row block.
    const int nks = (pend - K0) >> 2;  // K-substeps of 4
    const int nrb = (mp - pend) >> 4;
    for (int c0 = pend; c0 < mp; c0 += kLuChunk) {
        const int cw = min(kLuChunk, mp - c0);
        for (int e = tid; e < (pend - K0) * cw; e += blockDim.x) {
            const int r = e / cw, c = e - r * cw;
            Ush[r * kLuChunk + c] = at(K0 + r, c0 + c);
        }
        __syncthreads();
        const int nct = cw >> 4;
        for (int rb = wave; rb < nrb; rb += kLuWaves) {
            const int row0 = pend + rb * 16;
            // L21 fragments (ar, ai) and their sums for the three-MFMA
            // complex product (Gauss): P1 = Ar Br, P2 = Ai Bi, P3 =
            // (Ar + Ai)(Br + Bi) from zero, so the MFMAs need not wait for
            // the C tile, which is loaded first and combined at the end:
            // C -= (P1 - P2) + i (P3 - P1 - P2)
            double ar[kLuOuter / 4], ai[kLuOuter / 4];
#pragma unroll
            for (int kk = 0; kk < kLuOuter / 4; ++kk)
                if (kk < nks) {
                    const double2 v = at(row0 + lo, K0 + 4 * kk + hi);
                    ar[kk] = v.x;
                    ai[kk] = v.y;
                }
            for (int ct = 0; ct < nct; ++ct) {
                const int col0 = c0 + ct * 16;
                double2 cv[4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    cv[i] = at(row0 + 4 * i + hi, col0 + lo);
                v4d p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
                // (the Gauss sums Ar + Ai are formed per MFMA: holding them
                // for all K-substeps pushed this loop past the registers)
#pragma unroll 4
                for (int kk = 0; kk < kLuOuter / 4; ++kk)
                    if (kk < nks) {
                        const double2 bv2 = Ush[(4 * kk + hi) * kLuChunk + ct * 16 + lo];
                        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[kk], bv2.x, p1, 0, 0, 0);
                        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[kk], bv2.y, p2, 0, 0, 0);
                        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[kk] + ai[kk], bv2.x + bv2.y, p3, 0, 0, 0);
                    }
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    at(row0 + 4 * i + hi, col0 + lo) =
                        make_double2(cv[i].x - (p1[i] - p2[i]), cv[i].y - (p3[i] - p1[i] - p2[i]));
            }
        }
        __syncthreads();
    }
}

// Replaces each box's pivot sequence (row c swapped with piv[c], c = 0, 1,
// ...) by the permutation it amounts to: after the kernel, piv[i] is the row
// of the right-hand side that lands in row i, so the apply gathers in
// parallel instead of swapping in sequence.  One block per box.
__global__ __launch_bounds__(256) void pivots_to_permutation_kernel(const int *__restrict__ box_first, int nmax,
                                                                    const int *__restrict__ piv_off,
                                                                    int *__restrict__ pivots) {
    extern __shared__ int idx[];
    const int b = blockIdx.x;
    const int mp = padded16((box_first[b + 1] - box_first[b]) * nmax);
    int *piv = pivots + piv_off[b];
    for (int i = threadIdx.x; i < mp; i += blockDim.x)
        idx[i] = i;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < mp; ++c) {
            const int p = piv[c];
            if (p != c) {
                const int t0 = idx[c];
                idx[c] = idx[p];
                idx[p] = t0;
            }
        }
    __syncthreads();
    for (int i = threadIdx.x; i < mp; i += blockDim.x)
        piv[i] = idx[i];
}

// Barrier for LDS traffic only: waits for this wave's LDS operations, not
// for its global loads, so loads issued before it stay in flight across it
// (__syncthreads also waits for every outstanding global load).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// z[box] = U^-1 L^-1 P r[box] with 16-row blocked substitutions; the
// right-hand side lives in LDS (mp complex).  Each 16-row diagonal step is a
// 16x16 product with the stored L11^-1 / U11^-1, the off-diagonal updates use
// all threads (one row per thread at a time, 16 contiguous entries).  A
// thread's first kApplyRows rows of the next step are loaded as soon as the
// current update is done, so the loads run under the two barriers and the
// next diagonal product (rows past kApplyRows x blockDim.x of a step are
// loaded when used).
constexpr int kApplyThreads = 768;  // one block per box (at least 320: the diagonal staging)
constexpr int kApplyRows = 1;       // rows per thread loaded a step ahead
__device__ __forceinline__ double2 to_c128(double2 v) { return v; }
__device__ __forceinline__ double2 to_c128(float2 v) { return make_double2(v.x, v.y); }
// TF: the type the off-diagonal factors are stored in: complex128 row-major
// (the LU's own storage), or complex64 column-major for half the bytes and
// coalesced reads (a wave's 64 rows of one column are 512 contiguous bytes,
// where row-major storage has every lane on its own cache line).  The
// substitution arithmetic is FP64 either way, with the FP64 diagonal inverses.
template <typename TF>
__global__ __launch_bounds__(kApplyThreads) void lu_apply_kernel(const int *__restrict__ box_first,
                                                       const int *__restrict__ particles, int nmax,
                                                       const long long *__restrict__ blk_off,
                                                       const int *__restrict__ piv_off,
                                                       const TF *__restrict__ blocks,
                                                       const int *__restrict__ pivots,
                                                       const int *__restrict__ dinv_off,
                                                       const double2 *__restrict__ dinv,
                                                       const double2 *__restrict__ r,
                                                       double2 *__restrict__ z,
                                                       const int *__restrict__ gate) {
    if (gate && *gate != 1)
        return;
    extern __shared__ double2 y[];
    __shared__ double2 yb[16];
    __shared__ double2 Dsh[256];
    const int b = blockIdx.x;
    const int first = box_first[b];
    const int m = (box_first[b + 1] - first) * nmax;
    const int mp = padded16(m);
    const TF *A = blocks + blk_off[b];
    const int *piv = pivots + piv_off[b];
    const double2 *D = dinv + static_cast<long>(dinv_off[b]) * 512;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    for (int i = tid; i < mp; i += nthr) {  // y = P r: row i of y is row piv[i] of r
        const int src = piv[i];
        y[i] = src < m ? r[static_cast<long>(particles[first + src / nmax]) * nmax + src % nmax]
                       : make_double2(0.0, 0.0);
    }
    const int nb = mp >> 4;
    // entry (i, base + q) of the factors: stride ld between the 16 columns
    constexpr bool kColMajor = std::is_same<TF, float2>::value;
    const long ld = kColMajor ? mp : 1;
    auto entries = [&](int i, int base) {
        return kColMajor ? A + static_cast<long>(base) * mp + i : A + static_cast<long>(i) * mp + base;
    };
    // pan[k]: the 16 entries of row lo + tid + k nthr in the step's column
    // block, loaded one step ahead
    TF pan[kApplyRows][16];  // converted at use, so the loads stay in flight
    auto load_rows = [&](int lo, int hi, int base) {
#pragma unroll
        for (int k = 0; k < kApplyRows; ++k) {
            const int i = lo + tid + k * nthr;
            if (i < hi) {
                const TF *row = entries(i, base);
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    pan[k][q] = row[q * ld];
            }
        }
    };
    // one 16-row step: diagonal product into y[base..base+16), then the rows
    // [lo, hi) less the 16 new values times their columns of A
    // The step's 16 x 16 diagonal inverse goes through LDS (Dsh): threads
    // 64..319 load it a step ahead, before their row loads, so the wait for it
    // does not wait for those.
    double2 dv = make_double2(0.0, 0.0);
    auto stage_load = [&](const double2 *Dblk) {
        if (tid >= 64 && tid < 64 + 256)
            dv = Dblk[tid - 64];
    };
    auto stage_store = [&]() {
        if (tid >= 64 && tid < 64 + 256)
            Dsh[tid - 64] = dv;
    };
    auto step = [&](int base, int lo, int hi) {
        if (tid < 64) {  // wave 0: row tid / 4 of the 16 x 16 product, a quarter per lane
            const int rr = tid >> 2, q0 = (tid & 3) * 4;
            const double2 *Di = Dsh + rr * 16 + q0;
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double2 v = cmul(Di[q], y[base + q0 + q]);
                acc.x += v.x;
                acc.y += v.y;
            }
#pragma unroll
            for (int off = 1; off < 4; off <<= 1) {
                acc.x += __shfl_xor(acc.x, off, 64);
                acc.y += __shfl_xor(acc.y, off, 64);
            }
            if ((tid & 3) == 0)
                yb[rr] = acc;
        }
        lds_barrier();
        if (tid < 16)
            y[base + tid] = yb[tid];
        double2 yv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q)
            yv[q] = yb[q];
#pragma unroll
        for (int k = 0; k < kApplyRows; ++k) {
            const int i = lo + tid + k * nthr;
            if (i < hi) {
                double2 acc = y[i];
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    acc = csub(acc, cmul(to_c128(pan[k][q]), yv[q]));
                y[i] = acc;
            }
        }
        for (int i = lo + tid + kApplyRows * nthr; i < hi; i += nthr) {
            double2 acc = y[i];
            const TF *row = entries(i, base);
#pragma unroll
            for (int q = 0; q < 16; ++q)
                acc = csub(acc, cmul(to_c128(row[q * ld]), yv[q]));
            y[i] = acc;
        }
        asm volatile("" ::: "memory");  // the next step's row loads stay below this step's rows
    };
    stage_load(D);
    load_rows(16, mp, 0);  // unit lower: step bb updates rows [base + 16, mp)
    stage_store();
    __syncthreads();
    for (int bb = 0; bb < nb; ++bb) {
        const int base = bb * 16;
        step(base, base + 16, mp);
        if (bb + 1 < nb) {
            stage_load(D + static_cast<long>(bb + 1) * 512);
            load_rows(base + 32, mp, base + 16);
        } else {  // the upper pass starts at the last block
            stage_load(D + static_cast<long>(bb) * 512 + 256);
            load_rows(0, base, base);
        }
        stage_store();
        lds_barrier();
    }
    for (int bb = nb - 1; bb >= 0; --bb) {  // upper with diagonal: step bb updates rows [0, base)
        const int base = bb * 16;
        step(base, 0, base);
        if (bb > 0) {
            stage_load(D + static_cast<long>(bb - 1) * 512 + 256);
            load_rows(0, base - 16, base - 16);
            stage_store();
        }
        lds_barrier();
    }
    for (int i = tid; i < m; i += nthr)
        z[static_cast<long>(particles[first + i / nmax]) * nmax + i % nmax] = y[i];
}

// blocks32 = blocks rounded to complex64 and transposed per box (the
// column-major factors the apply reads).  Block (box, y) moves the 16 x 16
// tiles y, y + gridDim.y, ... of its box through LDS, reading rows and
// writing columns contiguously; mp is a multiple of 16.
__global__ __launch_bounds__(256) void round_blocks_kernel(const int *__restrict__ box_first, int nmax,
                                                           const long long *__restrict__ blk_off,
                                                           const double2 *__restrict__ src,
                                                           float2 *__restrict__ dst) {
    __shared__ float2 tile[16][17];
    const int b = blockIdx.x;
    const int mp = padded16((box_first[b + 1] - box_first[b]) * nmax);
    const int nt = mp >> 4;
    const double2 *S = src + blk_off[b];
    float2 *D = dst + blk_off[b];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    for (int t = blockIdx.y; t < nt * nt; t += gridDim.y) {
        const int r0 = (t / nt) * 16, c0 = (t % nt) * 16;
        const double2 v = S[static_cast<long>(r0 + ty) * mp + c0 + tx];
        tile[ty][tx] = make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
        __syncthreads();
        D[static_cast<long>(c0 + ty) * mp + r0 + tx] = tile[tx][ty];
        __syncthreads();
    }
}

template <typename T>
T *dalloc(size_t count) {
    void *p = nullptr;
    YH_CHECK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    return static_cast<T *>(p);
}

template <typename T>
T *dupload(const std::vector<T> &v, hipStream_t s) {
    T *d = dalloc<T>(v.size());
    if (!v.empty())
        YH_CHECK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return d;
}

}  // namespace

void precond_destroy(BlockPrecond *p);

BlockPrecond *precond_create(const CouplingPlan &plan, const std::vector<cplx> &ab5,
                             const double *pos_host, int N, const int *box_of_particle, int num_boxes,
                             double k, const double2 *t_dev, bool fp32_factors, hipStream_t s) {
    auto *bp = new BlockPrecond();
    try {
        bp->nmax = plan.nmax;
        bp->lmax = plan.lmax;
        bp->n = static_cast<long>(N) * plan.nmax;
        std::vector<std::vector<int>> members(num_boxes);
        for (int i = 0; i < N; ++i)
            members[box_of_particle[i]].push_back(i);
        std::vector<int> first{0}, order;
        std::vector<long long> off;
        std::vector<int> poff;
        long long elems = 0;
        int pv = 0;
        for (auto &mb : members) {
            if (mb.empty())
                continue;
            const long long m = static_cast<long long>(mb.size()) * plan.nmax;
            const long long mp = (m + 15) / 16 * 16;
            if (mp > kMaxBlock)
                throw Error(kInvalidArgument, "a preconditioner box holds " + std::to_string(mb.size()) +
                                                  " particles; its block exceeds " +
                                                  std::to_string(kMaxBlock) + " rows (use smaller boxes)");
            off.push_back(elems);
            poff.push_back(pv);
            elems += mp * mp;
            pv += static_cast<int>(mp);
            order.insert(order.end(), mb.begin(), mb.end());
            first.push_back(static_cast<int>(order.size()));
            bp->max_m = std::max<int>(bp->max_m, static_cast<int>(mp));
        }
        bp->nbox = static_cast<int>(off.size());
        bp->total_elems = elems;
        bp->box_first = dupload(first, s);
        bp->particles = dupload(order, s);
        bp->blk_off = dupload(off, s);
        bp->piv_off = dupload(poff, s);
        bp->blocks = dalloc<double2>(static_cast<size_t>(elems));
        bp->pivots = dalloc<int>(static_cast<size_t>(pv));

        std::vector<int> doff;
        std::vector<int2> rowblocks;
        int nbtot = 0;
        for (int bx = 0; bx < bp->nbox; ++bx) {
            doff.push_back(nbtot);
            const int cnt = first[bx + 1] - first[bx];
            nbtot += (cnt * plan.nmax + 15) / 16;
            for (int a1 = 0; a1 < cnt; ++a1)
                rowblocks.push_back(make_int2(bx, a1));
        }
        bp->dinv_off = dupload(doff, s);
        bp->dinv = dalloc<double2>(static_cast<size_t>(nbtot) * 512);

        // CSR over (n1, n2) of the nonzero ab5[n2][n1][p]; the G index of an
        // entry is dmoff[m2 - m1 + 2L] + p - |m2 - m1|.
        const int L = plan.lmax, nmax = plan.nmax, P = 2 * L + 1;
        const auto modes = modes_of(L);
        std::vector<int> dmoff(4 * L + 1);
        for (int d = 0, acc = 0; d < 4 * L + 1; ++d) {
            dmoff[d] = acc;
            acc += P - std::abs(d - 2 * L);
        }
        std::vector<int> erow(1, 0), eg;
        std::vector<double2> ev;
        for (int n1 = 0; n1 < nmax; ++n1)
            for (int n2 = 0; n2 < nmax; ++n2) {
                const int dm = modes[n2].m - modes[n1].m;
                for (int p = 0; p < P; ++p) {
                    const cplx a = ab5[(static_cast<size_t>(n2) * nmax + n1) * P + p];
                    if (a == cplx(0.0, 0.0))
                        continue;
                    eg.push_back(dmoff[dm + 2 * L] + p - std::abs(dm));
                    ev.push_back(make_double2(a.real(), a.imag()));
                }
                erow.push_back(static_cast<int>(eg.size()));
            }
        std::vector<double4> pos(N);
        for (int i = 0; i < N; ++i)
            pos[i] = make_double4(pos_host[3 * i], pos_host[3 * i + 1], pos_host[3 * i + 2], 0.0);
        double4 *dpos = dupload(pos, s);
        int2 *drb = dupload(rowblocks, s);
        int *derow = dupload(erow, s), *deg = dupload(eg, s), *ddmoff = dupload(dmoff, s);
        double2 *dev = dupload(ev, s);
        double *la = dupload(plan.leg_alpha, s), *lb = dupload(plan.leg_beta, s),
               *ld = dupload(plan.leg_diag, s);
        hipLaunchKernelGGL(build_blocks_kernel, dim3(static_cast<unsigned>(rowblocks.size())), dim3(256), 0,
                           s, dpos, bp->box_first, bp->particles, bp->blk_off, drb, nmax, L, k, t_dev,
                           derow, deg, dev, ddmoff, la, lb, ld, bp->blocks);
        YH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(pad_blocks_kernel, dim3(bp->nbox), dim3(256), 0, s, bp->box_first, nmax,
                           bp->blk_off, bp->blocks);
        YH_CHECK(hipGetLastError());
        static bool lu_lds_opt_in = false;  // gfx950: up to 160 KB of LDS per workgroup
        if (!lu_lds_opt_in) {  // set here, at the build, so that later applies can be graph-captured
            YH_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&lu_trailing_kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kLuLds)));
            YH_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&lu_panel_kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         static_cast<int>(kLuPanelLds)));
            const int dyn = static_cast<int>(kMaxBlock * sizeof(double2));
            YH_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&lu_apply_kernel<float2>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
            YH_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&lu_apply_kernel<double2>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
            lu_lds_opt_in = true;
        }
        const int mpmax = bp->max_m;  // padded to 16 rows
        for (int K0 = 0; K0 < mpmax; K0 += kLuOuter) {
            for (int kb = K0; kb < std::min(K0 + kLuOuter, mpmax); kb += 16) {
                hipLaunchKernelGGL(lu_panel_kernel, dim3(bp->nbox), dim3(64 * kLuWaves), kLuPanelLds, s,
                                   bp->box_first, nmax, bp->blk_off, bp->piv_off, bp->blocks, bp->pivots, kb);
                hipLaunchKernelGGL(lu_block_kernel, dim3(bp->nbox), dim3(64 * kLuWaves), 0, s, bp->box_first, nmax,
                                   bp->blk_off, bp->blocks, bp->dinv_off, bp->dinv, K0, kb);
            }
            hipLaunchKernelGGL(lu_trailing_kernel, dim3(bp->nbox), dim3(64 * kLuWaves), kLuLds, s, bp->box_first,
                               nmax, bp->blk_off, bp->blocks, bp->dinv_off, bp->dinv, K0);
        }
        YH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(pivots_to_permutation_kernel, dim3(bp->nbox), dim3(256), bp->max_m * sizeof(int), s,
                           bp->box_first, nmax, bp->piv_off, bp->pivots);
        YH_CHECK(hipGetLastError());
        if (fp32_factors) {  // the factors rounded to complex64 (column-major); the FP64 copy is not needed after this
            bp->blocks32 = dalloc<float2>(static_cast<size_t>(elems));
            hipLaunchKernelGGL(round_blocks_kernel, dim3(bp->nbox, 16), dim3(256), 0, s, bp->box_first, nmax,
                               bp->blk_off, bp->blocks, bp->blocks32);
            YH_CHECK(hipGetLastError());
        }
        YH_CHECK(hipStreamSynchronize(s));
        if (fp32_factors) {
            YH_CHECK(hipFree(bp->blocks));
            bp->blocks = nullptr;
        }
        for (void *p : {static_cast<void *>(dpos), static_cast<void *>(drb), static_cast<void *>(derow),
                        static_cast<void *>(deg), static_cast<void *>(ddmoff), static_cast<void *>(dev),
                        static_cast<void *>(la), static_cast<void *>(lb), static_cast<void *>(ld)})
            YH_CHECK(hipFree(p));
    } catch (...) {
        precond_destroy(bp);
        throw;
    }
    return bp;
}

void precond_destroy(BlockPrecond *p) {
    if (!p)
        return;
    for (void *b : {static_cast<void *>(p->box_first), static_cast<void *>(p->particles),
                    static_cast<void *>(p->blk_off), static_cast<void *>(p->piv_off),
                    static_cast<void *>(p->blocks), static_cast<void *>(p->blocks32),
                    static_cast<void *>(p->pivots), static_cast<void *>(p->dinv_off), static_cast<void *>(p->dinv)})
        if (b)
            (void)hipFree(b);
    delete p;
}

void precond_apply(BlockPrecond *p, const double2 *r, double2 *z, hipStream_t s, const int *gate) {
    if (p->blocks32)
        hipLaunchKernelGGL(lu_apply_kernel<float2>, dim3(p->nbox), dim3(kApplyThreads), p->max_m * sizeof(double2),
                           s, p->box_first, p->particles, p->nmax, p->blk_off, p->piv_off, p->blocks32, p->pivots,
                           p->dinv_off, p->dinv, r, z, gate);
    else
        hipLaunchKernelGGL(lu_apply_kernel<double2>, dim3(p->nbox), dim3(kApplyThreads),
                           p->max_m * sizeof(double2), s, p->box_first, p->particles, p->nmax, p->blk_off,
                           p->piv_off, p->blocks, p->pivots, p->dinv_off, p->dinv, r, z, gate);
    YH_CHECK(hipGetLastError());
}

}  // namespace yasfhip
