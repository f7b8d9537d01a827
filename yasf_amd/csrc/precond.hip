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
// Storage: boxes are packed back to back, each an m x m row-major complex
// block (m = particles_in_box * nmax).
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
    double2 *work = nullptr;     // [n] permuted right-hand sides
    long long total_elems = 0;
    int max_m = 0;
};

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

// One block per box; threads stride over the box's m x m elements.
__global__ __launch_bounds__(256) void build_blocks_kernel(
    const double4 *__restrict__ pos, const int *__restrict__ box_first, const int *__restrict__ particles,
    const long long *__restrict__ blk_off, const int4 *__restrict__ modes, int nmax, int lmax, double k,
    const double2 *__restrict__ t, const double2 *__restrict__ ab5, const double *__restrict__ la,
    const double *__restrict__ lb, const double *__restrict__ ld, double2 *__restrict__ blocks) {
    const int b = blockIdx.x;
    const int first = box_first[b];
    const int np_box = box_first[b + 1] - first;
    const long m = static_cast<long>(np_box) * nmax;
    double2 *A = blocks + blk_off[b];
    const int P = 2 * lmax + 1;
    for (long e = threadIdx.x; e < m * m; e += blockDim.x) {
        const long i = e / m, j = e % m;
        const int a1 = static_cast<int>(i / nmax), n1 = static_cast<int>(i % nmax);
        const int a2 = static_cast<int>(j / nmax), n2 = static_cast<int>(j % nmax);
        double2 v;
        if (a1 == a2) {
            v = make_double2(i == j ? 1.0 : 0.0, 0.0);
        } else {
            const int s1 = particles[first + a1], s2 = particles[first + a2];
            const double4 p1 = pos[s1], p2 = pos[s2];
            const int4 md1 = modes[n1], md2 = modes[n2];
            const double2 w = coupling_entry(p1.x - p2.x, p1.y - p2.y, p1.z - p2.z, k, md1.y, md1.z,
                                             md1.x, n1, md2.y, md2.z, md2.x, n2, nmax, P, ab5, la, lb, ld);
            const double2 tt = t[static_cast<long>(s1) * nmax + n1];
            v = make_double2(-(tt.x * w.x - tt.y * w.y), -(tt.x * w.y + tt.y * w.x));
        }
        A[e] = v;
    }
}

__device__ __forceinline__ double cabs1(double2 z) { return fabs(z.x) + fabs(z.y); }

// Right-looking LU with partial pivoting (pivot = largest |re|+|im| as in
// LAPACK's izamax), one block per box, in place.
__global__ __launch_bounds__(1024) void lu_kernel(const int *__restrict__ box_first, int nmax,
                                                  const long long *__restrict__ blk_off,
                                                  const int *__restrict__ piv_off,
                                                  double2 *__restrict__ blocks, int *__restrict__ pivots) {
    __shared__ double best_v[1024 / 64];
    __shared__ int best_i[1024 / 64];
    __shared__ int piv_sh;
    const int b = blockIdx.x;
    const int m = (box_first[b + 1] - box_first[b]) * nmax;
    double2 *A = blocks + blk_off[b];
    int *piv = pivots + piv_off[b];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int kc = 0; kc < m; ++kc) {
        double bv = -1.0;
        int bi = kc;
        for (int i = kc + threadIdx.x; i < m; i += blockDim.x) {
            const double v = cabs1(A[static_cast<long>(i) * m + kc]);
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
            best_v[w] = bv;
            best_i[w] = bi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double v = best_v[0];
            int ii = best_i[0];
            for (int q = 1; q < nw; ++q)
                if (best_v[q] > v || (best_v[q] == v && best_i[q] < ii)) {
                    v = best_v[q];
                    ii = best_i[q];
                }
            piv_sh = ii;
            piv[kc] = ii;
        }
        __syncthreads();
        const int p = piv_sh;
        if (p != kc)
            for (int j = threadIdx.x; j < m; j += blockDim.x) {
                const double2 t0 = A[static_cast<long>(kc) * m + j];
                A[static_cast<long>(kc) * m + j] = A[static_cast<long>(p) * m + j];
                A[static_cast<long>(p) * m + j] = t0;
            }
        __syncthreads();
        const double2 d = A[static_cast<long>(kc) * m + kc];
        const double dd = d.x * d.x + d.y * d.y;
        const double2 dinv = dd > 0.0 ? make_double2(d.x / dd, -d.y / dd) : make_double2(0.0, 0.0);
        for (int i = kc + 1 + threadIdx.x; i < m; i += blockDim.x) {
            const double2 a = A[static_cast<long>(i) * m + kc];
            A[static_cast<long>(i) * m + kc] =
                make_double2(a.x * dinv.x - a.y * dinv.y, a.x * dinv.y + a.y * dinv.x);
        }
        __syncthreads();
        const int rem = m - kc - 1;
        for (long e = threadIdx.x; e < static_cast<long>(rem) * rem; e += blockDim.x) {
            const int i = kc + 1 + static_cast<int>(e / rem), j = kc + 1 + static_cast<int>(e % rem);
            const double2 l = A[static_cast<long>(i) * m + kc];
            const double2 u = A[static_cast<long>(kc) * m + j];
            double2 a = A[static_cast<long>(i) * m + j];
            a.x -= l.x * u.x - l.y * u.y;
            a.y -= l.x * u.y + l.y * u.x;
            A[static_cast<long>(i) * m + j] = a;
        }
        __syncthreads();
    }
}

// z[box] = U^-1 L^-1 P r[box], one block per box.
__global__ __launch_bounds__(256) void lu_apply_kernel(const int *__restrict__ box_first,
                                                       const int *__restrict__ particles, int nmax,
                                                       const long long *__restrict__ blk_off,
                                                       const int *__restrict__ piv_off,
                                                       const double2 *__restrict__ blocks,
                                                       const int *__restrict__ pivots,
                                                       const double2 *__restrict__ r,
                                                       double2 *__restrict__ z, double2 *__restrict__ work) {
    const int b = blockIdx.x;
    const int first = box_first[b];
    const int m = (box_first[b + 1] - first) * nmax;
    const double2 *A = blocks + blk_off[b];
    const int *piv = pivots + piv_off[b];
    double2 *y = work + static_cast<long>(first) * nmax;
    for (int i = threadIdx.x; i < m; i += blockDim.x)
        y[i] = r[static_cast<long>(particles[first + i / nmax]) * nmax + i % nmax];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int kc = 0; kc < m; ++kc) {
            const int p = piv[kc];
            if (p != kc) {
                const double2 t0 = y[kc];
                y[kc] = y[p];
                y[p] = t0;
            }
        }
    __syncthreads();
    for (int kc = 0; kc < m; ++kc) {  // unit lower
        const double2 yk = y[kc];
        for (int i = kc + 1 + threadIdx.x; i < m; i += blockDim.x) {
            const double2 l = A[static_cast<long>(i) * m + kc];
            double2 v = y[i];
            v.x -= l.x * yk.x - l.y * yk.y;
            v.y -= l.x * yk.y + l.y * yk.x;
            y[i] = v;
        }
        __syncthreads();
    }
    for (int kc = m - 1; kc >= 0; --kc) {  // upper
        if (threadIdx.x == 0) {
            const double2 d = A[static_cast<long>(kc) * m + kc];
            const double dd = d.x * d.x + d.y * d.y;
            const double2 v = y[kc];
            y[kc] = make_double2((v.x * d.x + v.y * d.y) / dd, (v.y * d.x - v.x * d.y) / dd);
        }
        __syncthreads();
        const double2 yk = y[kc];
        for (int i = threadIdx.x; i < kc; i += blockDim.x) {
            const double2 u = A[static_cast<long>(i) * m + kc];
            double2 v = y[i];
            v.x -= u.x * yk.x - u.y * yk.y;
            v.y -= u.x * yk.y + u.y * yk.x;
            y[i] = v;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < m; i += blockDim.x)
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
            off.push_back(elems);
            poff.push_back(pv);
            elems += m * m;
            pv += static_cast<int>(m);
            order.insert(order.end(), mb.begin(), mb.end());
            first.push_back(static_cast<int>(order.size()));
            bp->max_m = std::max<int>(bp->max_m, static_cast<int>(m));
        }
        bp->nbox = static_cast<int>(off.size());
        bp->total_elems = elems;
        bp->box_first = dupload(first, s);
        bp->particles = dupload(order, s);
        bp->blk_off = dupload(off, s);
        bp->piv_off = dupload(poff, s);
        bp->blocks = dalloc<double2>(static_cast<size_t>(elems));
        bp->pivots = dalloc<int>(static_cast<size_t>(pv));
        bp->work = dalloc<double2>(static_cast<size_t>(bp->n));

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
        hipLaunchKernelGGL(lu_kernel, dim3(bp->nbox), dim3(1024), 0, s, bp->box_first, plan.nmax,
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
                    static_cast<void *>(p->blocks), static_cast<void *>(p->pivots),
                    static_cast<void *>(p->work)})
        if (b)
            (void)hipFree(b);
    delete p;
}

void precond_apply(BlockPrecond *p, const double2 *r, double2 *z, hipStream_t s) {
    hipLaunchKernelGGL(lu_apply_kernel, dim3(p->nbox), dim3(256), 0, s, p->box_first, p->particles,
                       p->nmax, p->blk_off, p->piv_off, p->blocks, p->pivots, r, z, p->work);
    YH_CHECK(hipGetLastError());
}

}  // namespace yasfhip
