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

template <typename T>
struct Cplx;
template <>
struct Cplx<double> {
    typedef double2 type;
};
template <>
struct Cplx<float> {
    typedef float2 type;
};

// G_{s1 s2}[p, dm] for every p in [|dm|, 2L], stored to dst[(p - |dm|) * stride].
// Geometry follows the reference's mutual_lookup (misc.py:149-229):
// d = r_s1 - r_s2, cos th = d_z/|d| clipped, sin th = sqrt(1 - cos^2),
// phi = atan2(d_y, d_x), h_p = j_p + i y_p; Legendre per
// legendre_normalized_trigon.py:59-91.  h_p comes from the upward recurrence
// h_{p+1} = (2p+1)/x h_p - h_{p-1} seeded with h_0, h_1 (stable for the y_p
// part that dominates |h_p|).  All arithmetic is FP64; invalid pairs store 0.
template <typename T2>
__device__ __forceinline__ void pair_series(double dx, double dy, double dz, double k, int dm, int P,
                                            bool valid, const double *__restrict__ la,
                                            const double *__restrict__ lb,
                                            const double *__restrict__ ld, T2 *dst, int stride) {
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
    // Hankel: (ar, ai) = h_p, (hr, hi) = h_{p+1}.
    const double x = zero_dist ? 1.0 : k * d;
    double sx, cx;
    sincos(x, &sx, &cx);
    const double inv = 1.0 / x;
    double ar = sx * inv, ai = -cx * inv;
    double hr = (ar - cx) * inv, hi = (ai - sx) * inv;
    // Legendre P_l^adm: cur = P_l, prv = P_{l-1}, starting at l = adm.
    double cur = ld[adm], prv = 0.0;
    for (int i = 0; i < adm; ++i)
        cur *= st;
    for (int p = 0; p < P; ++p) {
        if (p >= adm) {
            const double phr = cur * ar, phi = cur * ai;
            double gr = phr * er - phi * ei;
            double gi = phr * ei + phi * er;
            if (zero_dist) {  // h_p(0) = j_p(0) = delta_p0, phi = 0, cos th = 0
                gr = (p == 0) ? cur : 0.0;
                gi = 0.0;
            }
            if (!valid) {
                gr = 0.0;
                gi = 0.0;
            }
            T2 g;
            g.x = gr;
            g.y = gi;
            dst[(p - adm) * stride] = g;
            const double nx = la[p * P + adm] * ct * cur - lb[p * P + adm] * prv;
            prv = cur;
            cur = nx;
        }
        const double f = (2 * p + 3) * inv;
        const double nr = f * hr - ar, ni = f * hi - ai;
        ar = hr;
        ai = hi;
        hr = nr;
        hi = ni;
    }
}

// One block (4 waves) per (job, 16-particle target tile); jobs are sorted by
// cost and the block index is job-major so the heaviest jobs start first.
// Per chunk of 4*KS sources: (1) every pair's G[p] for all of the job's
// orders is generated once into LDS in the MFMA A-operand layout and the
// job's source columns of x are staged as the B operand; (2) each wave runs
// the complex 16x16x4 MFMAs of its C tiles (c = wave, wave+4, ...).
// Q layout: [target tile][q][col 16][row 16] complex T (interleaved).
template <typename T, int KS>
__global__ __launch_bounds__(256) void pair_kernel(const double4 *__restrict__ pos,
                                                   const typename Cplx<T>::type *__restrict__ x, int N,
                                                   int nmax, double k, const PairJob *__restrict__ jobs,
                                                   int ntiles1, const int *__restrict__ colmap,
                                                   int maxcols, int lmax, const double *__restrict__ la,
                                                   const double *__restrict__ lb,
                                                   const double *__restrict__ ld, T *__restrict__ Q,
                                                   int nq) {
    typedef typename Mfma<T>::acc_t acc_t;
    typedef typename Cplx<T>::type T2;
    constexpr int SRC = 4 * KS;  // sources per chunk
    extern __shared__ unsigned char lds_raw[];
    T2 *Ash = reinterpret_cast<T2 *>(lds_raw);  // [R][64*KS]

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int jid = blockIdx.x / ntiles1;
    const int s1t = blockIdx.x - jid * ntiles1;
    const int dm = uniform(jobs[jid].dm);
    const int tile0 = uniform(jobs[jid].tile0);
    const int ntile = uniform(jobs[jid].ntile);
    const int qbase = uniform(jobs[jid].qbase);
    const int P = 2 * lmax + 1;
    const int R = P - (dm < 0 ? -dm : dm);
    const int ncols = ntile * 16;
    const int total = R * ntile;
    T2 *Bsh = Ash + R * 64 * KS;  // [SRC][ncols]
    const int *cmap = colmap + (dm + 2 * lmax) * maxcols + tile0 * 16;

    const int lo = lane & 15, hi = lane >> 4;
    const int gr = tid & 15, gk = tid >> 4;  // generation: target row, source in chunk
    const int s1 = s1t * 16 + gr;
    const double4 p1 = pos[s1];  // pos is padded to a multiple of 16

    acc_t acc[kWaveMaxTiles][2];
#pragma unroll
    for (int j = 0; j < kWaveMaxTiles; ++j) {
        acc[j][0] = acc_t{0, 0, 0, 0};
        acc[j][1] = acc_t{0, 0, 0, 0};
    }

    for (int c0 = 0; c0 < N; c0 += SRC) {
        if (gk < SRC) {
            const int s2 = c0 + gk;
            const bool valid = (s2 < N) && (s1 < N) && (s2 != s1);
            const double4 p2 = pos[s2 < N ? s2 : 0];
            pair_series<T2>(p1.x - p2.x, p1.y - p2.y, p1.z - p2.z, k, dm, P, valid, la, lb, ld,
                            Ash + tid, 64 * KS);
        }
        for (int e = tid; e < SRC * ncols; e += blockDim.x) {
            const int kk = e / ncols, c = e - kk * ncols;
            const int n2 = cmap[c];
            const int s2 = c0 + kk;
            T2 v;
            v.x = T(0);
            v.y = T(0);
            if (n2 >= 0 && s2 < N)
                v = x[static_cast<size_t>(s2) * nmax + n2];
            Bsh[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kWaveMaxTiles; ++j) {
            const int c = wave + kWavesPerJob * j;
            if (c < total) {
                const int ip = c / ntile, it = c - ip * ntile;
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    const T2 a = Ash[ip * 64 * KS + ks * 64 + lane];
                    const T2 bv = Bsh[(ks * 4 + hi) * ncols + it * 16 + lo];
                    acc[j][0] = Mfma<T>::mac(a.x, bv.x, acc[j][0]);
                    acc[j][0] = Mfma<T>::mac(-a.y, bv.y, acc[j][0]);
                    acc[j][1] = Mfma<T>::mac(a.x, bv.y, acc[j][1]);
                    acc[j][1] = Mfma<T>::mac(a.y, bv.x, acc[j][1]);
                }
            }
        }
        __syncthreads();
    }

    // Accumulator element i: target row Mfma<T>::row(i, hi), column lo.
#pragma unroll
    for (int j = 0; j < kWaveMaxTiles; ++j) {
        const int c = wave + kWavesPerJob * j;
        if (c < total) {
            T *dst = Q + 2 * (((static_cast<size_t>(s1t) * nq + qbase + c) * 16 + lo) * 16);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = Mfma<T>::row(i, hi);
                dst[2 * r] = acc[j][0][i];
                dst[2 * r + 1] = acc[j][1][i];
            }
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
    int lmax, nmax, maxcols, nq, njobs, max_job_cols;
    int ks;  // source K-steps per LDS chunk: 2, 4, or 0 = largest that fits
    PairJob *jobs;
    int *colmap, *crow, *cidx;
    double2 *cval;
    double *la, *lb, *ld;
};

template <typename T, int KS>
static size_t pair_lds_bytes(const DevicePlan &dp) {
    return (static_cast<size_t>(2 * dp.lmax + 1) * 64 * KS + 4 * KS * dp.max_job_cols) *
           sizeof(typename Cplx<T>::type);
}

template <typename T, int KS>
static void launch_pair_t(const DevicePlan &dp, const double4 *pos, int N, double k, const void *x,
                          void *Q, hipStream_t s) {
    const int ntiles1 = (N + 15) / 16;
    const size_t lds = pair_lds_bytes<T, KS>(dp);
    hipLaunchKernelGGL((pair_kernel<T, KS>), dim3(dp.njobs * ntiles1), dim3(256), lds, s, pos, static_cast<const typename Cplx<T>::type *>(x), N, dp.nmax, k, dp.jobs,
                       ntiles1, dp.colmap, dp.maxcols, dp.lmax, dp.la, dp.lb, dp.ld, static_cast<T *>(Q),
                       dp.nq);
    YH_CHECK(hipGetLastError());
}

constexpr size_t kPairLdsLimit = 64 * 1024;

void launch_pair(const DevicePlan &dp, const double4 *pos, int N, double k, const void *x, bool fp32,
                 void *Q, hipStream_t s) {
    const bool ks4_fits = fp32 ? pair_lds_bytes<float, 4>(dp) <= kPairLdsLimit
                               : pair_lds_bytes<double, 4>(dp) <= kPairLdsLimit;
    const bool use4 = ks4_fits && dp.ks != 2;
    if (fp32) {
        if (use4)
            launch_pair_t<float, 4>(dp, pos, N, k, x, Q, s);
        else
            launch_pair_t<float, 2>(dp, pos, N, k, x, Q, s);
    } else {
        if (use4)
            launch_pair_t<double, 4>(dp, pos, N, k, x, Q, s);
        else
            launch_pair_t<double, 2>(dp, pos, N, k, x, Q, s);
    }
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
