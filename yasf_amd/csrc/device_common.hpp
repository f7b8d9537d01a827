// Device helpers shared by the pair kernels: MFMA wrappers with their
// operand/result layouts, complex element types and the per-pair series of
// Legendre x Hankel x azimuthal factors.
#pragma once

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

// Raw buffer resource over [base, base + bytes) (bytes < 2^32): a load whose
// per-lane offset lies at or past `bytes` returns zero instead of touching
// memory.  Word 3 = 0x00020000: 32-bit data format, the value the CDNA
// raw-buffer descriptors use.  base and bytes must be the same in every lane:
// they are read from the first active lane, so the descriptor lives in
// scalar registers (a descriptor the compiler cannot prove uniform turns every
// load into a loop over the distinct values).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void *base, size_t bytes) {
    const unsigned long long p = reinterpret_cast<unsigned long long>(base);
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(p));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(p >> 32));
    void *u = reinterpret_cast<void *>((static_cast<unsigned long long>(hi) << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(u, 0, __builtin_amdgcn_readfirstlane(static_cast<int>(bytes)),
                                             0x00020000);
}
// Offset that is out of range for every resource made by buffer_rsrc.
constexpr unsigned kBufferNull = 0x80000000u;

// One complex element per lane from byte offset voff (per lane, range-checked)
// plus soff (wave-uniform, not range-checked).
__device__ __forceinline__ double2 buffer_load_cplx(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff,
                                                    double2 *) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
__device__ __forceinline__ float2 buffer_load_cplx(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff,
                                                   float2 *) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

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

// Pair geometry of the reference's mutual_lookup (misc.py:149-229):
constexpr int kMaxOrders = 2 * kMaxLmax + 1;  // p = 0 .. 2 lmax

// d = r_s1 - r_s2, cos th = d_z/|d| clipped (0 at d = 0), sin th =
// sqrt(1 - cos^2), phi = atan2(d_y, d_x); (er, ei) = e^{i dm phi} by binary
// powering of e^{i phi}; x = k |d| (1 at d = 0, unused there).
struct PairGeometry {
    double ct, st, er, ei, x;
    bool zero_dist;
};

__device__ __forceinline__ PairGeometry pair_geometry(double dx, double dy, double dz, double k, int dm) {
    PairGeometry g;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    g.zero_dist = !(d > 0.0);
    double ct = g.zero_dist ? 0.0 : dz / d;
    g.ct = fmin(1.0, fmax(-1.0, ct));
    g.st = sqrt(1.0 - g.ct * g.ct);
    const double rho2 = dx * dx + dy * dy;
    double c1 = 1.0, s1 = 0.0;
    if (rho2 > 0.0) {
        const double ir = 1.0 / sqrt(rho2);
        c1 = dx * ir;
        s1 = dy * ir;
    }
    const int adm = dm < 0 ? -dm : dm;
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
    g.er = er;
    g.ei = dm < 0 ? -ei : ei;
    g.x = g.zero_dist ? 1.0 : k * d;
    return g;
}

// G_{s1 s2}[p, dm] for every p in [|dm|, 2L], handed to sink(p - |dm|, re, im):
// Ptilde_p^{|dm|}(cos th) h_p(k d) e^{i dm phi} with h_p = j_p + i y_p and the
// Legendre functions of legendre_normalized_trigon.py:59-91.  h_p comes from
// the upward recurrence h_{p+1} = (2p+1)/x h_p - h_{p-1} seeded with h_0, h_1
// (stable for the y_p part that dominates |h_p|).  All arithmetic is FP64;
// invalid pairs give 0.
template <typename Sink>
__device__ __forceinline__ void pair_series_to(double dx, double dy, double dz, double k, int dm, int P,
                                               bool valid, const double *__restrict__ la,
                                               const double *__restrict__ lb,
                                               const double *__restrict__ ld, Sink &&sink) {
    const PairGeometry g = pair_geometry(dx, dy, dz, k, dm);
    const int adm = dm < 0 ? -dm : dm;
    // Hankel: (ar, ai) = h_p, (hr, hi) = h_{p+1}.
    double sx, cx;
    sincos(g.x, &sx, &cx);
    const double inv = 1.0 / g.x;
    double ar = sx * inv, ai = -cx * inv;
    double hr = (ar - cx) * inv, hi = (ai - sx) * inv;
    // Legendre P_l^adm: cur = P_l, prv = P_{l-1}, starting at l = adm.
    double cur = ld[adm], prv = 0.0;
    for (int i = 0; i < adm; ++i)
        cur *= g.st;
    for (int p = 0; p < P; ++p) {
        if (p >= adm) {
            const double phr = cur * ar, phi = cur * ai;
            double gr = phr * g.er - phi * g.ei;
            double gi = phr * g.ei + phi * g.er;
            if (g.zero_dist) {  // h_p(0) = j_p(0) = delta_p0, phi = 0, cos th = 0
                gr = (p == 0) ? cur : 0.0;
                gi = 0.0;
            }
            if (!valid) {
                gr = 0.0;
                gi = 0.0;
            }
            sink(p - adm, gr, gi);
            const double nx = la[p * P + adm] * g.ct * cur - lb[p * P + adm] * prv;
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

// G_{s1 s2}[p, dm] for every dm in [-2L, 2L] and p in [|dm|, 2L], handed to
// sink(dm + 2L, p - |dm|, re, im): pair_series_to's values for each dm (the
// same operations; the compiler may fuse a multiply-add differently, so the
// last bit can differ), with the distance, angles and Hankel functions
// formed once per pair instead of once per dm.
template <typename Sink>
__device__ __forceinline__ void pair_series_all(double dx, double dy, double dz, double k, int lmax, bool valid,
                                                const double *__restrict__ la, const double *__restrict__ lb,
                                                const double *__restrict__ ld, Sink &&sink) {
    const int P = 2 * lmax + 1;
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
    const double x = zero_dist ? 1.0 : k * d;
    // h_p, p < P, by pair_series_to's recurrence
    double hre[kMaxOrders], him[kMaxOrders];
    {
        double sx, cx;
        sincos(x, &sx, &cx);
        const double inv = 1.0 / x;
        double ar = sx * inv, ai = -cx * inv;
        double hr = (ar - cx) * inv, hi = (ai - sx) * inv;
#pragma unroll
        for (int p = 0; p < kMaxOrders; ++p) {
            hre[p] = ar;
            him[p] = ai;
            const double f = (2 * p + 3) * inv;
            const double nr = f * hr - ar, ni = f * hi - ai;
            ar = hr;
            ai = hi;
            hr = nr;
            hi = ni;
        }
    }
    for (int dmi = 0; dmi < 4 * lmax + 1; ++dmi) {
        const int dm = dmi - 2 * lmax;
        const int adm = dm < 0 ? -dm : dm;
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
        double cur = ld[adm], prv = 0.0;
        for (int i = 0; i < adm; ++i)
            cur *= st;
#pragma unroll
        for (int p = 0; p < kMaxOrders; ++p) {
            if (p < P && p >= adm) {
                const double phr = cur * hre[p], phi = cur * him[p];
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
                sink(dmi, p - adm, gr, gi);
                const double nx = la[p * P + adm] * ct * cur - lb[p * P + adm] * prv;
                prv = cur;
                cur = nx;
            }
        }
    }
}

// As above, stored to dst[(p - |dm|) * stride] (float2 or double2).
template <typename T2>
__device__ __forceinline__ void pair_series(double dx, double dy, double dz, double k, int dm, int P,
                                            bool valid, const double *__restrict__ la,
                                            const double *__restrict__ lb,
                                            const double *__restrict__ ld, T2 *dst, int stride) {
    pair_series_to(dx, dy, dz, k, dm, P, valid, la, lb, ld, [&](int ip, double gr, double gi) {
        T2 v;
        v.x = gr;
        v.y = gi;
        dst[ip * stride] = v;
    });
}

// Spherical Bessel j_p(x), p < kMaxOrders, x > 0.  The upward recurrence is
// stable while p < x, so it serves x >= kMaxOrders; below that Miller's
// downward recurrence from order kMaxOrders + 24 (where j_p/y_p is below
// 1e-18 of its value at p <= 2 lmax for x < kMaxOrders) is normalised to the
// closed form of j_0 or j_1, whichever is larger; x < 1e-3 uses the series
// j_p = x^p/(2p+1)!! (1 - x^2/(2(2p+3))) (relative error below 1e-13).
__device__ __forceinline__ void sph_bessel_j_all(double x, double (&jv)[kMaxOrders]) {
    double sx, cx;
    sincos(x, &sx, &cx);
    const double inv = 1.0 / x;
    const double j0 = sx * inv, j1 = (sx * inv - cx) * inv;
    if (x >= kMaxOrders) {
        jv[0] = j0;
        jv[1] = j1;
#pragma unroll
        for (int p = 1; p + 1 < kMaxOrders; ++p)
            jv[p + 1] = (2 * p + 1) * inv * jv[p] - jv[p - 1];
        return;
    }
    if (x < 1e-3) {
        double term = 1.0;  // x^p / (2p+1)!!
#pragma unroll
        for (int p = 0; p < kMaxOrders; ++p) {
            jv[p] = term * (1.0 - x * x / (2.0 * (2 * p + 3)));
            term *= x / (2 * p + 3);
        }
        return;
    }
    constexpr int kStart = kMaxOrders + 24;
    double jp1 = 0.0, jc = 1e-280;  // j_{p+1}, j_p up to a common factor, p = kStart
#pragma unroll
    for (int p = kStart; p >= 1; --p) {
        const double jm = (2 * p + 1) * inv * jc - jp1;
        jp1 = jc;
        jc = jm;
        if (p - 1 < kMaxOrders)
            jv[p - 1] = jm;
    }
    const double scale = fabs(j0) >= fabs(j1) ? j0 / jv[0] : j1 / jv[1];
#pragma unroll
    for (int p = 0; p < kMaxOrders; ++p)
        jv[p] *= scale;
}

// As pair_series with the regular radial function j_p(k d) in place of h_p:
// the J-translation factors of the scattering cross section
// (cpu_numba.py:410-457 with the sph_j lookup).  Coincident positions give
// j_p(0) = delta_p0.
template <typename T2>
__device__ __forceinline__ void pair_series_regular(double dx, double dy, double dz, double k, int dm,
                                                    int P, bool valid, const double *__restrict__ la,
                                                    const double *__restrict__ lb,
                                                    const double *__restrict__ ld, T2 *dst, int stride) {
    const PairGeometry g = pair_geometry(dx, dy, dz, k, dm);
    const int adm = dm < 0 ? -dm : dm;
    double jv[kMaxOrders];
    if (g.zero_dist) {
#pragma unroll
        for (int p = 0; p < kMaxOrders; ++p)
            jv[p] = p == 0 ? 1.0 : 0.0;
    } else {
        sph_bessel_j_all(g.x, jv);
    }
    double cur = ld[adm], prv = 0.0;
    for (int i = 0; i < adm; ++i)
        cur *= g.st;
#pragma unroll
    for (int p = 0; p < kMaxOrders; ++p) {
        if (p < P && p >= adm) {
            const double ph = valid ? cur * jv[p] : 0.0;
            T2 v;
            v.x = ph * g.er;
            v.y = ph * g.ei;
            dst[(p - adm) * stride] = v;
            const double nx = la[p * P + adm] * g.ct * cur - lb[p * P + adm] * prv;
            prv = cur;
            cur = nx;
        }
    }
}

// (W x)[s1, n1] = sum_e ab5_e Q[e] for the 16 targets of tile s1t and the 16
// target modes of contraction rows 16 g .. 16 g + 15 (mode group g; the
// plan's crow gives each row's mode, rows grouped by m1), then either y = W x
// (mode 0) or the master operator y = x - t (W x) (mode 1,
// matrix_apply.py:104-110); FP64 accumulation.  256 threads: lane&15 =
// target in tile, so the 16 targets of one Q slot are read as one coalesced
// row.  A launch covers (target tile, mode group) pairs: contract_groups(nmax)
// blocks per tile.

template <typename TQ, typename TX, typename TY>
__device__ __forceinline__ void contract_tile(const TQ *__restrict__ Q, int nq, const int *__restrict__ crow,
                                              const int *__restrict__ cidx, const double2 *__restrict__ cval,
                                              int N, int nmax, const TX *__restrict__ x,
                                              const double2 *__restrict__ t, int mode, TY *__restrict__ y,
                                              int s1t, int group, int qt0 = 0) {
    // Q holds the target tiles from qt0 on (qt0 = 0: all of them)
    const int r = threadIdx.x & 15;
    const int s1 = s1t * 16 + r;
    const int row = group * kContractModes + (threadIdx.x >> 4);
    const int n1 = crow[contract_groups(nmax) * kContractModes + 1 + row];
    if (n1 < 0)
        return;
    const size_t qrow = static_cast<size_t>(s1t - qt0) * nq * 16;
    double sr = 0.0, si = 0.0;
    const int e1 = crow[row + 1];
    // batches of kBatch entries: all their index, coefficient and Q loads are
    // in flight together (one entry at a time waits two memory round trips
    // per entry); entries past the row's end repeat its first with a zero
    // coefficient, and the sum runs in entry order as before
    constexpr int kBatch = 8;
    for (int e = crow[row]; e < e1; e += kBatch) {
        const int cnt = min(kBatch, e1 - e);
        int ci[kBatch];
        double2 a[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int ee = u < cnt ? e + u : e;
            ci[u] = cidx[ee];
            a[u] = cval[ee];
        }
        double qr[kBatch], qim[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const size_t qi = 2 * ((qrow + ci[u]) * 16 + r);
            qr[u] = static_cast<double>(Q[qi]);
            qim[u] = static_cast<double>(Q[qi + 1]);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            if (u < cnt) {
                sr += a[u].x * qr[u] - a[u].y * qim[u];
                si += a[u].x * qim[u] + a[u].y * qr[u];
            }
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

}  // namespace yasfhip
