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

}  // namespace yasfhip
