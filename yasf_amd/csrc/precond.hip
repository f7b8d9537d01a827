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

#include "internal.hpp"

namespace yasfhip {

struct BlockPrecond {
    int nbox = 0, nmax = 0, lmax = 0;
    long n = 0;
    int *box_first = nullptr;    // [nbox+1] offsets into the particle order
    int *particles = nullptr;    // particle ids ordered by box
    long long *blk_off = nullptr;  // [nbox] element offset of each block
    int *piv_off = nullptr;      // [nbox] offset of each block's pivots
    double2 *blocks = nullptr;
    int *pivots = nullptr;
    long long total_elems = 0;
    int max_m = 0;  // largest padded block size
};

// Largest padded block (rows); the apply keeps one block's right-hand side in LDS.
constexpr int kMaxBlock = 3840;

namespace {

// Coupling entry W[(s1,n1),(s2,n2)] (reference cpu_numba.py:74-91 with the
// tables of misc.py:149-229 evaluated for this one pair).
__device__ double2 coupling_entry(double dx, double dy, double dz, double k, int l1, int m1, int tau1,
                                  int n1, int l2, int m2, int tau2, int n2, int nmax, int P,
                                  const double2 *__restrict__ ab5, const double *__restrict__ la,
                                  const double *__restrict__ lb, const double *__restrict__ ld) {
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    double ct = d > 0.0 ? dz / d : 0.0;
    ct = fmin(1.0, fmax(-1.0, ct));
    const double st = sqrt(1.0 - ct * ct);
    const double rho2 = dx * dx + dy * dy;
    double c1 = 1.0, s1 = 0.0;
    if (rho2 > 0.0) {
        const double ir = 1.0 / sqrt(rho2);
        c1 = dx * ir;
        s1 = dy * ir;
    }
    const int dm = m2 - m1;
    const int adm = dm < 0 ? -dm : dm;
    double er = 1.0, ei = 0.0, br = c1, bi = s1;
    for (int q = adm; q; q >>= 1) {
        if (q & 1) {
            const double t = er * br - ei * bi;
            ei = er * bi + ei * br;
            er = t;
        }
        const double t = br * br - bi * bi;
        bi = 2.0 * br * bi;
        br = t;
    }
    if (dm < 0)
        ei = -ei;
    const double x = k * d;
    double sx, cx;
    sincos(x, &sx, &cx);
    const double inv = 1.0 / x;
    double ar = sx * inv, ai = -cx * inv;          // h_p
    double hr = (ar - cx) * inv, hi = (ai - sx) * inv;  // h_{p+1}
    double cur = ld[adm], prv = 0.0;                // P_p^adm from p = adm
    for (int i = 0; i < adm; ++i)
        cur *= st;
    const int dl = l1 > l2 ? l1 - l2 : l2 - l1;
    const int dt = tau1 > tau2 ? tau1 - tau2 : tau2 - tau1;
    const int p_start = max(adm, dl + dt);
    double sr = 0.0, si = 0.0;
    for (int p = 0; p <= l1 + l2; ++p) {
        if (p >= p_start) {
            const double2 a = ab5[(static_cast<long>(n2) * nmax + n1) * P + p];
            const double phr = cur * ar, phi = cur * ai;
            sr += a.x * phr - a.y * phi;
            si += a.x * phi + a.y * phr;
        }
        const double f = (2 * p + 3) * inv;
        const double nr = f * hr - ar, ni = f * hi - ai;
        ar = hr;
        ai = hi;
        hr = nr;
        hi = ni;
        if (p >= adm && p + 1 < P) {
            const double nx = la[p * P + adm] * ct * cur - lb[p * P + adm] * prv;
            prv = cur;
            cur = nx;
        }
    }
    return make_double2(sr * er - si * ei, sr * ei + si * er);
}

__device__ __forceinline__ int padded16(int m) { return (m + 15) & ~15; }

// One block per box; threads stride over the box's mp x mp elements
// (mp = m rounded up to 16; the padding is the identity).
__global__ __launch_bounds__(256) void build_blocks_kernel(
    const double4 *__restrict__ pos, const int *__restrict__ box_first, const int *__restrict__ particles,
    const long long *__restrict__ blk_off, const int4 *__restrict__ modes, int nmax, int lmax, double k,
    const double2 *__restrict__ t, const double2 *__restrict__ ab5, const double *__restrict__ la,
    const double *__restrict__ lb, const double *__restrict__ ld, double2 *__restrict__ blocks) {
    const int b = blockIdx.x;
    const int first = box_first[b];
    const int np_box = box_first[b + 1] - first;
    const long m = static_cast<long>(np_box) * nmax;
    const long mp = padded16(static_cast<int>(m));
    double2 *A = blocks + blk_off[b];
    const int P = 2 * lmax + 1;
    for (long e = threadIdx.x; e < mp * mp; e += blockDim.x) {
        const long i = e / mp, j = e % mp;
        double2 v;
        if (i >= m || j >= m) {
            v = make_double2(i == j ? 1.0 : 0.0, 0.0);
        } else {
            const int a1 = static_cast<int>(i / nmax), n1 = static_cast<int>(i % nmax);
            const int a2 = static_cast<int>(j / nmax), n2 = static_cast<int>(j % nmax);
            if (a1 == a2) {
                v = make_double2(i == j ? 1.0 : 0.0, 0.0);
            } else {
                const int s1 = particles[first + a1], s2 = particles[first + a2];
                const double4 p1 = pos[s1], p2 = pos[s2];
                const int4 md1 = modes[n1], md2 = modes[n2];
                const double2 w = coupling_entry(p1.x - p2.x, p1.y - p2.y, p1.z - p2.z, k, md1.y, md1.z,
                                                 md1.x, n1, md2.y, md2.z, md2.x, n2, nmax, P, ab5, la,
                                                 lb, ld);
                const double2 tt = t[static_cast<long>(s1) * nmax + n1];
                v = make_double2(-(tt.x * w.x - tt.y * w.y), -(tt.x * w.y + tt.y * w.x));
            }
        }
        A[e] = v;
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
// as LAPACK's izamax), one 256-thread block per box, in place on the padded
// mp x mp block.  Per 16-column panel: unblocked panel factorisation, the
// 16-row U12 = L11^-1 A12 solve, then A22 -= L21 U12 as 16x16 complex tiles on
// the FP64 MFMA (16x16x4, four real products per complex product).
__global__ __launch_bounds__(256) void lu_blocked_kernel(const int *__restrict__ box_first, int nmax,
                                                         const long long *__restrict__ blk_off,
                                                         const int *__restrict__ piv_off,
                                                         double2 *__restrict__ blocks,
                                                         int *__restrict__ pivots) {
    __shared__ double red_v[4];
    __shared__ int red_i[4];
    __shared__ int piv_sh;
    __shared__ double2 Lsh[16][16];
    const int b = blockIdx.x;
    const int m = (box_first[b + 1] - box_first[b]) * nmax;
    const int mp = padded16(m);
    double2 *A = blocks + blk_off[b];
    int *piv = pivots + piv_off[b];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lo = lane & 15, hi = lane >> 4;
    auto at = [&](int i, int j) -> double2 & { return A[static_cast<long>(i) * mp + j]; };

    for (int kb = 0; kb < mp; kb += 16) {
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
            if (tid == 0) {
                double v = red_v[0];
                int ii = red_i[0];
                for (int q = 1; q < 4; ++q)
                    if (red_v[q] > v || (red_v[q] == v && red_i[q] < ii)) {
                        v = red_v[q];
                        ii = red_i[q];
                    }
                piv_sh = ii;
                piv[c] = ii;
            }
            __syncthreads();
            const int p = piv_sh;
            if (p != c)
                for (int jj = tid; jj < mp; jj += blockDim.x) {
                    const double2 t0 = at(c, jj);
                    at(c, jj) = at(p, jj);
                    at(p, jj) = t0;
                }
            __syncthreads();
            const double2 inv = cinv(at(c, c));
            const int pend = kb + 16;
            for (int i = c + 1 + tid; i < mp; i += blockDim.x) {
                const double2 l = cmul(at(i, c), inv);
                at(i, c) = l;
                for (int jj = c + 1; jj < pend; ++jj)
                    at(i, jj) = csub(at(i, jj), cmul(l, at(c, jj)));
            }
            __syncthreads();
        }
        if (kb + 16 >= mp)
            break;
        // Linv = L11^-1 (unit lower); thread c solves L11 y = e_c.
        if (tid < 16) {
            const int c = tid;
            double2 yv[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                double2 v = make_double2(r == c ? 1.0 : 0.0, 0.0);
                for (int q = c; q < r; ++q)
                    v = csub(v, cmul(at(kb + r, kb + q), yv[q]));
                yv[r] = r < c ? make_double2(0.0, 0.0) : v;
                Lsh[r][c] = yv[r];
            }
        }
        __syncthreads();
        const int nt = (mp - kb - 16) >> 4;
        // U12 = Linv A12, 16x16 complex tiles on the MFMA (each wave owns whole
        // column tiles, so reading and overwriting a tile needs no barrier).
        for (int cb = wave; cb < nt; cb += 4) {
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
        for (int rb = wave; rb < nt; rb += 4) {
            const int row0 = kb + 16 + rb * 16;
            double ar[4], ai[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const double2 v = at(row0 + lo, kb + 4 * kk + hi);
                ar[kk] = v.x;
                ai[kk] = v.y;
            }
            for (int cb = 0; cb < nt; ++cb) {
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
        __syncthreads();
    }
}

// z[box] = U^-1 L^-1 P r[box] with 16-row blocked substitutions; the
// right-hand side lives in LDS (mp complex), diagonal 16x16 blocks are staged
// in LDS and solved by one thread, off-diagonal updates use all threads.
__global__ __launch_bounds__(256) void lu_apply_kernel(const int *__restrict__ box_first,
                                                       const int *__restrict__ particles, int nmax,
                                                       const long long *__restrict__ blk_off,
                                                       const int *__restrict__ piv_off,
                                                       const double2 *__restrict__ blocks,
                                                       const int *__restrict__ pivots,
                                                       const double2 *__restrict__ r,
                                                       double2 *__restrict__ z) {
    extern __shared__ double2 y[];
    __shared__ double2 D[16][16];
    const int b = blockIdx.x;
    const int first = box_first[b];
    const int m = (box_first[b + 1] - first) * nmax;
    const int mp = padded16(m);
    const double2 *A = blocks + blk_off[b];
    const int *piv = pivots + piv_off[b];
    const int tid = threadIdx.x;
    for (int i = tid; i < mp; i += blockDim.x)
        y[i] = i < m ? r[static_cast<long>(particles[first + i / nmax]) * nmax + i % nmax]
                     : make_double2(0.0, 0.0);
    __syncthreads();
    if (tid == 0)
        for (int c = 0; c < mp; ++c) {
            const int p = piv[c];
            if (p != c) {
                const double2 t0 = y[c];
                y[c] = y[p];
                y[p] = t0;
            }
        }
    __syncthreads();
    const int nb = mp >> 4;
    for (int bb = 0; bb < nb; ++bb) {  // unit lower
        const int base = bb * 16;
        D[tid >> 4][tid & 15] = A[static_cast<long>(base + (tid >> 4)) * mp + base + (tid & 15)];
        __syncthreads();
        if (tid == 0)
            for (int rr = 1; rr < 16; ++rr) {
                double2 v = y[base + rr];
                for (int q = 0; q < rr; ++q)
                    v = csub(v, cmul(D[rr][q], y[base + q]));
                y[base + rr] = v;
            }
        __syncthreads();
        double2 yb[16];
#pragma unroll
        for (int q = 0; q < 16; ++q)
            yb[q] = y[base + q];
        for (int i = base + 16 + tid; i < mp; i += blockDim.x) {
            double2 acc = y[i];
            const double2 *row = A + static_cast<long>(i) * mp + base;
#pragma unroll
            for (int q = 0; q < 16; ++q)
                acc = csub(acc, cmul(row[q], yb[q]));
            y[i] = acc;
        }
        __syncthreads();
    }
    for (int bb = nb - 1; bb >= 0; --bb) {  // upper with diagonal
        const int base = bb * 16;
        D[tid >> 4][tid & 15] = A[static_cast<long>(base + (tid >> 4)) * mp + base + (tid & 15)];
        __syncthreads();
        if (tid == 0)
            for (int rr = 15; rr >= 0; --rr) {
                double2 v = y[base + rr];
                for (int q = rr + 1; q < 16; ++q)
                    v = csub(v, cmul(D[rr][q], y[base + q]));
                y[base + rr] = cmul(v, cinv(D[rr][rr]));
            }
        __syncthreads();
        double2 yb[16];
#pragma unroll
        for (int q = 0; q < 16; ++q)
            yb[q] = y[base + q];
        for (int i = tid; i < base; i += blockDim.x) {
            double2 acc = y[i];
            const double2 *row = A + static_cast<long>(i) * mp + base;
#pragma unroll
            for (int q = 0; q < 16; ++q)
                acc = csub(acc, cmul(row[q], yb[q]));
            y[i] = acc;
        }
        __syncthreads();
    }
    for (int i = tid; i < m; i += blockDim.x)
        z[static_cast<long>(particles[first + i / nmax]) * nmax + i % nmax] = y[i];
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
                             double k, const double2 *t_dev, hipStream_t s) {
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

        std::vector<double4> pos(N);
        for (int i = 0; i < N; ++i)
            pos[i] = make_double4(pos_host[3 * i], pos_host[3 * i + 1], pos_host[3 * i + 2], 0.0);
        const auto modes = modes_of(plan.lmax);
        std::vector<int4> md(modes.size());
        for (size_t i = 0; i < modes.size(); ++i)
            md[i] = make_int4(modes[i].tau, modes[i].l, modes[i].m, 0);
        std::vector<double2> ab(ab5.size());
        for (size_t i = 0; i < ab5.size(); ++i)
            ab[i] = make_double2(ab5[i].real(), ab5[i].imag());
        double4 *dpos = dupload(pos, s);
        int4 *dmd = dupload(md, s);
        double2 *dab = dupload(ab, s);
        double *la = dupload(plan.leg_alpha, s), *lb = dupload(plan.leg_beta, s),
               *ld = dupload(plan.leg_diag, s);
        hipLaunchKernelGGL(build_blocks_kernel, dim3(bp->nbox), dim3(256), 0, s, dpos, bp->box_first,
                           bp->particles, bp->blk_off, dmd, plan.nmax, plan.lmax, k, t_dev, dab, la, lb,
                           ld, bp->blocks);
        YH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(lu_blocked_kernel, dim3(bp->nbox), dim3(256), 0, s, bp->box_first, plan.nmax,
                           bp->blk_off, bp->piv_off, bp->blocks, bp->pivots);
        YH_CHECK(hipGetLastError());
        YH_CHECK(hipStreamSynchronize(s));
        for (void *p : {static_cast<void *>(dpos), static_cast<void *>(dmd), static_cast<void *>(dab),
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
                    static_cast<void *>(p->blocks), static_cast<void *>(p->pivots)})
        if (b)
            (void)hipFree(b);
    delete p;
}

void precond_apply(BlockPrecond *p, const double2 *r, double2 *z, hipStream_t s) {
    hipLaunchKernelGGL(lu_apply_kernel, dim3(p->nbox), dim3(256), p->max_m * sizeof(double2), s,
                       p->box_first, p->particles, p->nmax, p->blk_off, p->piv_off, p->blocks, p->pivots,
                       r, z);
    YH_CHECK(hipGetLastError());
}

}  // namespace yasfhip
