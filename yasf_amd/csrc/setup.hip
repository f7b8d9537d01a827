// Per-request set-up of the linear system on the device: the Mie T-matrix
// diagonal t and the plane-wave incident coefficients a_inc, so a served
// request needs only its geometry, refractive indices and incidence.
//
// * Mie (yasfpy/mie_coefficients.py:88-161): per sphere, with x = k r and
//   z = rel x (rel = sphere index / medium index),
//     b_l = -(j_l(z) [x j_l(x)]' - j_l(x) [z j_l(z)]') / (j_l(z) [x h_l(x)]' - h_l(x) [z j_l(z)]')
//     a_l = the same with rel^2 on the j_l(z) [.]' terms,
//   t = b_l on tau = 1 and a_l on tau = 2, for every m of the order l.
//   [w f_l(w)]' = w f_{l-1}(w) - l f_l(w).  j_l of the complex argument by
//   its power series below |z| = 1, else Miller's downward recurrence
//   normalised to the closed form of j_0 or j_1; y_l(x) by the upward
//   recurrence.
// * Plane wave (yasfpy/initial_field_coefficients.py:99-142):
//     a[s, n] = c[n] e^{i k (r_s - r_f) . d},  c[n] = 4 A e^{-i m alpha} B^dagger_{tau l m}(beta)
//   with the angular part c (O(nmax), host) and the per-sphere phase (device).
#include <cmath>

#include "device_common.hpp"

namespace yasfhip {

namespace {

constexpr int kMieOrders = kMaxLmax + 2;  // l = 0 .. lmax + 1

struct C {
    double re, im;
};
__device__ __forceinline__ C csub(C a, C b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ C cmul(C a, C b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ C cscale(C a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ C cdiv(C a, C b) {
    // Smith's algorithm
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__device__ __forceinline__ double cabs2(C a) { return a.re * a.re + a.im * a.im; }

// j_l(z), l = 0 .. nl-1 (nl <= kMieOrders), complex z != 0.  |z| < 1: the
// power series z^l/(2l+1)!! sum_k (-z^2/2)^k / (k! (2l+3)(2l+5)...(2l+2k+1))
// to full precision (the Mie quotients cancel to O(x^2) at small x, so j_l
// must be exact there); else Miller's downward recurrence.
__device__ void sph_bessel_j_complex(C z, int nl, C (&jv)[kMieOrders]) {
    const double az = sqrt(cabs2(z));
    if (az < 1.0) {
        C lead = {1.0, 0.0};  // z^l / (2l+1)!!
        const C mz2h = cscale(cmul(z, z), -0.5);
        for (int l = 0; l < nl; ++l) {
            C sum = {1.0, 0.0}, term = {1.0, 0.0};
            for (int q = 1; q <= 30; ++q) {
                term = cscale(cmul(term, mz2h), 1.0 / (q * (2.0 * l + 2 * q + 1)));
                sum = {sum.re + term.re, sum.im + term.im};
                if (cabs2(term) < 1e-36 * cabs2(sum))
                    break;
            }
            jv[l] = cmul(lead, sum);
            lead = cscale(cmul(lead, z), 1.0 / (2 * l + 3));
        }
        return;
    }
    double sr, cr;
    sincos(z.re, &sr, &cr);
    const double ch = cosh(z.im), sh = sinh(z.im);
    const C sz = {sr * ch, cr * sh}, cz = {cr * ch, -sr * sh};
    const C inv = cdiv(C{1.0, 0.0}, z);
    const C j0 = cmul(sz, inv);
    const C j1 = cmul(csub(cmul(sz, inv), cz), inv);
    // downward from an order well above |z| and lmax + 1
    const int start = kMieOrders + 30 + static_cast<int>(az);
    C jp1 = {0.0, 0.0}, jc = {1e-250, 0.0};
    for (int l = start; l >= 1; --l) {
        const C jm = csub(cmul(cscale(inv, 2 * l + 1), jc), jp1);
        jp1 = jc;
        jc = jm;
        if (l - 1 < nl)
            jv[l - 1] = jm;
        // keep the unnormalised sequence inside the exponent range
        const double m2 = cabs2(jc);
        if (m2 > 1e200) {
            jc = cscale(jc, 1e-200);
            jp1 = cscale(jp1, 1e-200);
            for (int q = l - 1; q < nl; ++q)
                jv[q] = cscale(jv[q], 1e-200);
        }
    }
    const C scale = cabs2(j0) >= cabs2(j1) ? cdiv(j0, jv[0]) : cdiv(j1, jv[1]);
    for (int l = 0; l < nl; ++l)
        jv[l] = cmul(jv[l], scale);
}

// One thread per sphere: t[s * nmax + n] for all modes.
// sph[s] = (radius, rel.re, rel.im, unused).
__global__ __launch_bounds__(128) void mie_kernel(const double4 *__restrict__ sph, int N, int lmax,
                                                  double k, double2 *__restrict__ t) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N)
        return;
    const double4 p = sph[s];
    const double x = k * p.x;
    const C rel = {p.y, p.z};
    const C z = cscale(rel, x);
    const int nl = lmax + 1;
    C jx[kMieOrders], jz[kMieOrders];
    sph_bessel_j_complex(C{x, 0.0}, nl, jx);
    sph_bessel_j_complex(z, nl, jz);
    double yv[kMieOrders];
    {
        double sx, cx;
        sincos(x, &sx, &cx);
        yv[0] = -cx / x;
        yv[1] = -cx / (x * x) - sx / x;
        for (int l = 1; l + 1 < nl; ++l)
            yv[l + 1] = (2 * l + 1) / x * yv[l] - yv[l - 1];
    }
    const C rel2 = cmul(rel, rel);
    const int nmax = 2 * lmax * (lmax + 2);
    const int half = lmax * (lmax + 2);
    for (int l = 1; l <= lmax; ++l) {
        const C jxl = jx[l], jxm = jx[l - 1];
        const C hxl = {jxl.re, yv[l]}, hxm = {jxm.re, yv[l - 1]};
        const C djx = csub(cscale(jxm, x), cscale(jxl, l));
        const C dhx = csub(cscale(hxm, x), cscale(hxl, l));
        const C djz = csub(cmul(z, jz[l - 1]), cscale(jz[l], l));
        const C jzl = jz[l];
        const C b = cdiv(csub(cmul(jxl, djz), cmul(jzl, djx)), csub(cmul(jzl, dhx), cmul(hxl, djz)));
        const C r2jz = cmul(rel2, jzl);
        const C a = cdiv(csub(cmul(jxl, djz), cmul(r2jz, djx)), csub(cmul(r2jz, dhx), cmul(hxl, djz)));
        double2 *row = t + static_cast<size_t>(s) * nmax;
        for (int m = -l; m <= l; ++m) {
            const int n1 = l * l - 1 + l + m;  // (l-1)(l+1) + l + m
            row[n1] = make_double2(b.re, b.im);
            row[half + n1] = make_double2(a.re, a.im);
        }
    }
}

// a[s * nmax + n] = c[n] e^{i k (r_s - r_f) . d}; one thread per entry.
__global__ __launch_bounds__(256) void planewave_kernel(const double4 *__restrict__ pos, long count, int nmax,
                                                        double k, double dx, double dy, double dz, double fx,
                                                        double fy, double fz, const double2 *__restrict__ c,
                                                        double2 *__restrict__ a) {
    for (long e = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; e < count;
         e += static_cast<long>(gridDim.x) * blockDim.x) {
        const long s = e / nmax;
        const int n = static_cast<int>(e - s * nmax);
        const double4 p = pos[s];
        const double ph = k * ((p.x - fx) * dx + (p.y - fy) * dy + (p.z - fz) * dz);
        double sp, cp;
        sincos(ph, &sp, &cp);
        const double2 cn = c[n];
        a[e] = make_double2(cn.x * cp - cn.y * sp, cn.x * sp + cn.y * cp);
    }
}

}  // namespace

void launch_mie(const double4 *sph, int N, int lmax, double k, double2 *t, hipStream_t s) {
    hipLaunchKernelGGL(mie_kernel, dim3((N + 127) / 128), dim3(128), 0, s, sph, N, lmax, k, t);
    YH_CHECK(hipGetLastError());
}

void launch_planewave(const double4 *pos, int N, int lmax, double k, double polar, double azimuthal,
                      const double *focal, const double2 *c, double2 *a, hipStream_t s) {
    const int nmax = 2 * lmax * (lmax + 2);
    const long count = static_cast<long>(N) * nmax;
    const double sb = std::sin(polar), cb = std::cos(polar);
    const int blocks = static_cast<int>(std::min<long>(2048, (count + 255) / 256));
    hipLaunchKernelGGL(planewave_kernel, dim3(blocks), dim3(256), 0, s, pos, count, nmax, k,
                       sb * std::cos(azimuthal), sb * std::sin(azimuthal), cb, focal[0], focal[1], focal[2], c,
                       a);
    YH_CHECK(hipGetLastError());
}

// c[n] = 4 A e^{-i m alpha} B^dagger_{tau l m}(beta) (initial_field_coefficients.py:120-142,
// transformation_coefficients misc.py:70-114 with dagger, angular functions
// spherical_functions_trigon.py:39-93), normalised pi_lm and tau_lm.
void planewave_angular(int lmax, double polar, double azimuthal, int pol, double amplitude, double2 *c) {
    YH_REQUIRE(pol == 1 || pol == 2, "polarization must be 1 (TE) or 2 (TM)");
    const int L1 = lmax + 1;
    const double ct = std::cos(polar), st = std::sin(polar);
    std::vector<double> pilm(L1 * L1, 0.0), taulm(L1 * L1, 0.0), plm(L1 * L1, 0.0), dp0(L1, 0.0);
    auto at = [L1](std::vector<double> &v, int l, int m) -> double & { return v[l * L1 + m]; };
    at(plm, 0, 0) = std::sqrt(0.5);
    if (lmax >= 1) {
        at(plm, 1, 0) = std::sqrt(1.5) * ct;
        dp0[1] = std::sqrt(3.0) * at(plm, 0, 0);
        at(taulm, 1, 0) = -st * dp0[1];
    }
    for (int l = 1; l < lmax; ++l) {
        at(plm, l + 1, 0) = std::sqrt((2.0 * l + 1) * (2.0 * l + 3)) / (l + 1) * ct * at(plm, l, 0) -
                            static_cast<double>(l) / (l + 1) * std::sqrt((2.0 * l + 3) / (2.0 * l - 1)) *
                                at(plm, l - 1, 0);
        dp0[l + 1] = std::sqrt((2.0 * l + 3) / (2.0 * l + 1)) * ((l + 1) * at(plm, l, 0) + ct * dp0[l]);
        at(taulm, l + 1, 0) = -st * dp0[l + 1];
    }
    for (int m = 1; m <= lmax; ++m) {
        // sqrt((2m+1)/2/(2m)!) (2m-1)!!
        double c0 = std::sqrt((2.0 * m + 1) / 2.0);
        for (int i = 1; i <= 2 * m; ++i)
            c0 /= std::sqrt(static_cast<double>(i));
        for (int i = 1; i < 2 * m; i += 2)
            c0 *= i;
        at(plm, m, m) = c0 * std::pow(st, m);
        at(pilm, m, m) = c0 * std::pow(st, m - 1);
        at(taulm, m, m) = m * ct * at(pilm, m, m);
        for (int l = m; l < lmax; ++l) {
            const double a = std::sqrt((2.0 * l + 1) * (2.0 * l + 3) / ((l + 1.0 - m) * (l + 1.0 + m))) * ct;
            const double b = std::sqrt((2.0 * l + 3) * (l - m) * (l + m) /
                                       ((2.0 * l - 1) * (l + 1.0 - m) * (l + 1.0 + m)));
            at(plm, l + 1, m) = a * at(plm, l, m) - b * at(plm, l - 1, m);
            at(pilm, l + 1, m) = a * at(pilm, l, m) - b * at(pilm, l - 1, m);
            at(taulm, l + 1, m) = (l + 1) * ct * at(pilm, l + 1, m) -
                                  (l + 1 + m) * std::sqrt((2.0 * l + 3) * (l + 1 - m) / ((2.0 * l + 1) * (l + 1 + m))) *
                                      at(pilm, l, m);
        }
    }
    const std::complex<double> ifac(0.0, -1.0);
    const int half = lmax * (lmax + 2);
    for (int m = -lmax; m <= lmax; ++m) {
        const std::complex<double> ph = 4.0 * amplitude * std::exp(std::complex<double>(0.0, -m * azimuthal));
        const int am = m < 0 ? -m : m;
        for (int tau = 1; tau <= 2; ++tau)
            for (int l = std::max(1, am); l <= lmax; ++l) {
                const double f = tau == pol ? at(taulm, l, am) : m * at(pilm, l, am);
                const std::complex<double> pre = -1.0 / std::pow(ifac, l + 1) / std::sqrt(2.0 * l * (l + 1));
                const std::complex<double> v = pre * (pol == 1 ? ifac : std::complex<double>(1.0, 0.0)) * f * ph;
                const int n = (tau - 1) * half + (l - 1) * (l + 1) + l + m;
                c[n] = make_double2(v.real(), v.imag());
            }
    }
}

}  // namespace yasfhip
