// Coupling operator W x on CDNA4 (gfx950), matrix-free.
//
// The reference evaluates (cpu_numba.py:21-95, cuda_numba.py:278-468)
//   (W x)[s1,n1] = sum_{s2!=s1} sum_{n2} e^{i(m2-m1)phi} x[s2,n2]
//                   sum_p ab5[n2,n1,p] Ptilde_p^{|m1-m2|}(cos th) h_p(k d)
// from O(N^2) stored tables.  Here the sum is regrouped as
//   Q[s1,p,dm,n2] = sum_{s2} G_{s1 s2}[p,dm] x[s2,n2],
//   G_{s1 s2}[p,dm] = Ptilde_p^{|dm|}(cos th) h_p(k d) e^{i dm phi}     (pair_kernel)
//   (W x)[s1,n1] = sum_{p,n2} ab5[n2,n1,p] Q[s1,p,m2-m1,n2]              (contract_kernel)
// For fixed (dm, p) the first line is a complex GEMM over the source
// particles s2 whose A operand is generated on the fly.  pair_kernel runs it
// on the 16x16x4 MFMA: lane (r, c) = (lane&15, lane>>4) of a wave owns exactly
// one (target s1 = tile*16 + r, source s2 = k0 + c) pair, evaluates that pair's
// geometry, Hankel and Legendre recurrences in FP64 registers and feeds them
// straight in as the A operand -- no LDS, no barriers.  The B operand is the
// source coefficients x[s2, n2] of the wave's 16 columns.  Each wave owns one
// PairJob (dm, <=4 orders p, <=2 column tiles) and sweeps all sources.
#include "internal.hpp"

namespace yasfhip {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// A operand: lane holds A[lane & 15][lane >> 4]; B: B[lane >> 4][lane & 15].
// D: accumulator element i of a lane is column lane & 15 and row
// row(i, lane >> 4), which differs between the f64 and f32 instructions.
template <>
struct Mfma<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int i, int hi) { return 4 * i + hi; }
};
template <>
struct Mfma<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int i, int hi) { return 4 * hi + i; }
};

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// G_{s1 s2}[p, dm] for p in [p_lo, p_lo + NP) (entries past np are unused).
// Geometry follows the reference's mutual_lookup (misc.py:149-229):
// d = r_s1 - r_s2, cos th = d_z/|d| clipped, sin th = sqrt(1 - cos^2),
// phi = atan2(d_y, d_x), h_p = j_p + i y_p; Legendre per
// legendre_normalized_trigon.py:59-91.  h_p comes from the upward recurrence
// h_{p+1} = (2p+1)/x h_p - h_{p-1} seeded with h_0, h_1 (stable for the y_p
// part that dominates |h_p|).
template <int NP>
__device__ __forceinline__ void pair_series(double dx, double dy, double dz, double k, int dm, int p_lo,
                                            int P, const double *__restrict__ la,
                                            const double *__restrict__ lb,
                                            const double *__restrict__ ld, double (&gre)[NP],
                                            double (&gim)[NP]) {
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const bool zero_dist = !(d > 0.0);
    double ct = zero_dist ? 0.0 : dz / d;
    ct = fmin(1.0, fmax(-1.0, ct));
    const double st = sqrt(1.0 - ct * ct);
    const double rho2 = dx * dx + dy * dy;
    double c1 = 1.0, s1 = 0.0;
    if (rho2 > 0.0) {
        const double ir = 1.0 / sqrt(rho2);
        c1 = dx * ir;
        s1 = dy * ir;
    }
    const int adm = dm < 0 ? -dm : dm;

    // e^{i dm phi} by binary powering of e^{i phi}.
    double er = 1.0, ei = 0.0, br = c1, bi = s1;
    for (int n = adm; n; n >>= 1) {
        if (n & 1) {
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

    // Hankel h_p(x): (ar, ai) = h_p, (br2, bi2) = h_{p+1}.
    const double x = zero_dist ? 1.0 : k * d;
    double sx, cx;
    sincos(x, &sx, &cx);
    const double inv = 1.0 / x;
    double ar = sx * inv, ai = -cx * inv;
    double hr = (ar - cx) * inv, hi = (ai - sx) * inv;
    int p = 0;
    for (; p < p_lo; ++p) {
        const double f = (2 * p + 3) * inv;
        const double nr = f * hr - ar, ni = f * hi - ai;
        ar = hr;
        ai = hi;
        hr = nr;
        hi = ni;
    }
    // Legendre P_l^adm: cur = P_l, prv = P_{l-1}, starting at l = adm.
    double cur = ld[adm], prv = 0.0;
    for (int i = 0; i < adm; ++i)
        cur *= st;
    for (int l = adm; l < p_lo; ++l) {
        const double nx = la[l * P + adm] * ct * cur - lb[l * P + adm] * prv;
        prv = cur;
        cur = nx;
    }
#pragma unroll
    for (int ip = 0; ip < NP; ++ip) {
        const int pp = p_lo + ip;
        // G = P h e
        const double phr = cur * ar, phi = cur * ai;
        double gr = phr * er - phi * ei;
        double gi = phr * ei + phi * er;
        if (zero_dist) {  // h_p(0) = j_p(0) = delta_p0, phi = 0, cos th = 0
            gr = (pp == 0) ? cur : 0.0;
            gi = 0.0;
        }
        gre[ip] = gr;
        gim[ip] = gi;
        if (pp + 1 < P) {
            const double f = (2 * pp + 3) * inv;
            const double nr = f * hr - ar, ni = f * hi - ai;
            ar = hr;
            ai = hi;
            hr = nr;
            hi = ni;
            const double nx = la[pp * P + adm] * ct * cur - lb[pp * P + adm] * prv;
            prv = cur;
            cur = nx;
        }
    }
}

// One wave per (job, target tile).  Block = 4 independent waves.
// Q layout: [target tile][q][col 16][row 16] complex T (interleaved).
template <typename T>
__global__ __launch_bounds__(256) void pair_kernel(const double4 *__restrict__ pos, const T *__restrict__ x,
                                                   int N, int nmax, double k,
                                                   const PairJob *__restrict__ jobs, int njobs,
                                                   int ntiles1, const int *__restrict__ colmap,
                                                   int maxcols, int lmax, const double *__restrict__ la,
                                                   const double *__restrict__ lb,
                                                   const double *__restrict__ ld, T *__restrict__ Q,
                                                   int nq) {
    typedef typename Mfma<T>::acc_t acc_t;
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= njobs * ntiles1)
        return;
    const int jid = uniform(wave / ntiles1);
    const int s1t = uniform(wave % ntiles1);
    const int dm = uniform(jobs[jid].dm);
    const int p_lo = uniform(jobs[jid].p_lo);
    const int np = uniform(jobs[jid].np);
    const int tile0 = uniform(jobs[jid].tile0);
    const int ntile = uniform(jobs[jid].ntile);
    const int qbase = uniform(jobs[jid].qbase);
    const int P = 2 * lmax + 1;

    const int lo = lane & 15;  // A row (target in tile) / B,D column
    const int hi = lane >> 4;  // A,B k index (source in K-step)
    const int s1 = s1t * 16 + lo;
    const double4 p1 = pos[s1];  // pos is padded to a multiple of 16

    int n2c[kJobMaxTiles];
#pragma unroll
    for (int it = 0; it < kJobMaxTiles; ++it) {
        const int col = (tile0 + it) * 16 + lo;
        n2c[it] = (it < ntile) ? colmap[(dm + 2 * lmax) * maxcols + col] : -1;
    }

    acc_t acc[kJobMaxP][kJobMaxTiles][2];
#pragma unroll
    for (int ip = 0; ip < kJobMaxP; ++ip)
#pragma unroll
        for (int it = 0; it < kJobMaxTiles; ++it) {
            acc[ip][it][0] = acc_t{0, 0, 0, 0};
            acc[ip][it][1] = acc_t{0, 0, 0, 0};
        }

    for (int k0 = 0; k0 < N; k0 += 4) {
        const int s2 = k0 + hi;
        const bool valid = (s2 < N) && (s1 < N) && (s2 != s1);
        const double4 p2 = pos[s2 < N ? s2 : 0];
        double gre[kJobMaxP], gim[kJobMaxP];
        pair_series<kJobMaxP>(p1.x - p2.x, p1.y - p2.y, p1.z - p2.z, k, dm, p_lo, P, la, lb, ld, gre,
                              gim);
        T bre[kJobMaxTiles], bim[kJobMaxTiles];
#pragma unroll
        for (int it = 0; it < kJobMaxTiles; ++it) {
            const bool ok = (n2c[it] >= 0) && (s2 < N);
            const size_t o = 2 * (static_cast<size_t>(ok ? s2 : 0) * nmax + (ok ? n2c[it] : 0));
            bre[it] = ok ? x[o] : T(0);
            bim[it] = ok ? x[o + 1] : T(0);
        }
#pragma unroll
        for (int ip = 0; ip < kJobMaxP; ++ip) {
            if (ip < np) {
                const T are = valid ? static_cast<T>(gre[ip]) : T(0);
                const T aim = valid ? static_cast<T>(gim[ip]) : T(0);
#pragma unroll
                for (int it = 0; it < kJobMaxTiles; ++it) {
                    if (it < ntile) {
                        acc[ip][it][0] = Mfma<T>::mac(are, bre[it], acc[ip][it][0]);
                        acc[ip][it][0] = Mfma<T>::mac(-aim, bim[it], acc[ip][it][0]);
                        acc[ip][it][1] = Mfma<T>::mac(are, bim[it], acc[ip][it][1]);
                        acc[ip][it][1] = Mfma<T>::mac(aim, bre[it], acc[ip][it][1]);
                    }
                }
            }
        }
    }

    // Accumulator element i: target row Mfma<T>::row(i, hi), column lo.
#pragma unroll
    for (int ip = 0; ip < kJobMaxP; ++ip)
#pragma unroll
        for (int it = 0; it < kJobMaxTiles; ++it)
            if (ip < np && it < ntile) {
                const int q = qbase + ip * ntile + it;
                T *dst = Q + 2 * (((static_cast<size_t>(s1t) * nq + q) * 16 + lo) * 16);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = Mfma<T>::row(i, hi);
                    dst[2 * r] = acc[ip][it][0][i];
                    dst[2 * r + 1] = acc[ip][it][1][i];
                }
            }
}

// (W x)[s1, n1] = sum_e ab5_e Q[e], then either y = W x (mode 0) or the
// master operator y = x - t (W x) (mode 1, matrix_apply.py:104-110).
// One block per 16-particle target tile; lane&15 = target in tile.
template <typename TQ, typename TX, typename TY>
__global__ __launch_bounds__(256) void contract_kernel(const TQ *__restrict__ Q, int nq,
                                                       const int *__restrict__ crow,
                                                       const int *__restrict__ cidx,
                                                       const double2 *__restrict__ cval, int N, int nmax,
                                                       const TX *__restrict__ x,
                                                       const double2 *__restrict__ t, int mode,
                                                       TY *__restrict__ y) {
    const int s1t = blockIdx.x;
    const int r = threadIdx.x & 15;
    const int s1 = s1t * 16 + r;
    const size_t qrow = static_cast<size_t>(s1t) * nq * 16;
    for (int n1 = threadIdx.x >> 4; n1 < nmax; n1 += blockDim.x >> 4) {
        double sr = 0.0, si = 0.0;
        const int e1 = crow[n1 + 1];
        for (int e = crow[n1]; e < e1; ++e) {
            const double2 a = cval[e];
            const size_t qi = 2 * ((qrow + cidx[e]) * 16 + r);
            const double qr = static_cast<double>(Q[qi]);
            const double qim = static_cast<double>(Q[qi + 1]);
            sr += a.x * qr - a.y * qim;
            si += a.x * qim + a.y * qr;
        }
        if (s1 < N) {
            const size_t j = static_cast<size_t>(s1) * nmax + n1;
            if (mode == 1) {
                const double2 tj = t[j];
                const double xr = static_cast<double>(x[2 * j]);
                const double xi = static_cast<double>(x[2 * j + 1]);
                y[2 * j] = static_cast<TY>(xr - (tj.x * sr - tj.y * si));
                y[2 * j + 1] = static_cast<TY>(xi - (tj.x * si + tj.y * sr));
            } else {
                y[2 * j] = static_cast<TY>(sr);
                y[2 * j + 1] = static_cast<TY>(si);
            }
        }
    }
}

// Device-side views of a CouplingPlan.
struct DevicePlan {
    int lmax, nmax, maxcols, nq, njobs;
    PairJob *jobs;
    int *colmap, *crow, *cidx;
    double2 *cval;
    double *la, *lb, *ld;
};

void launch_pair(const DevicePlan &dp, const double4 *pos, int N, double k, const void *x, bool fp32,
                 void *Q, hipStream_t s) {
    const int ntiles1 = (N + 15) / 16;
    const long waves = static_cast<long>(dp.njobs) * ntiles1;
    const int blocks = static_cast<int>((waves + 3) / 4);
    if (fp32)
        hipLaunchKernelGGL(pair_kernel<float>, dim3(blocks), dim3(256), 0, s, pos,
                           static_cast<const float *>(x), N, dp.nmax, k, dp.jobs, dp.njobs, ntiles1,
                           dp.colmap, dp.maxcols, dp.lmax, dp.la, dp.lb, dp.ld, static_cast<float *>(Q),
                           dp.nq);
    else
        hipLaunchKernelGGL(pair_kernel<double>, dim3(blocks), dim3(256), 0, s, pos,
                           static_cast<const double *>(x), N, dp.nmax, k, dp.jobs, dp.njobs, ntiles1,
                           dp.colmap, dp.maxcols, dp.lmax, dp.la, dp.lb, dp.ld,
                           static_cast<double *>(Q), dp.nq);
    YH_CHECK(hipGetLastError());
}

// fp32: Q, x and y are complex64; otherwise complex128.
void launch_contract(const DevicePlan &dp, int N, const void *Q, const void *x, const double2 *t,
                     int mode, void *y, bool fp32, hipStream_t s) {
    const int ntiles1 = (N + 15) / 16;
    if (fp32)
        hipLaunchKernelGGL((contract_kernel<float, float, float>), dim3(ntiles1), dim3(256), 0, s,
                           static_cast<const float *>(Q), dp.nq, dp.crow, dp.cidx, dp.cval, N, dp.nmax,
                           static_cast<const float *>(x), t, mode, static_cast<float *>(y));
    else
        hipLaunchKernelGGL((contract_kernel<double, double, double>), dim3(ntiles1), dim3(256), 0, s,
                           static_cast<const double *>(Q), dp.nq, dp.crow, dp.cidx, dp.cval, N,
                           dp.nmax, static_cast<const double *>(x), t, mode, static_cast<double *>(y));
    YH_CHECK(hipGetLastError());
}

}  // namespace yasfhip
