// Far field of the scattered-field coefficients on the device, for the
// phase function (yasfpy/optics.py:177-266, 269-551 with
// compute_electric_field_angle_components, cpu_numba.py:535-606):
//   F[a, n]   = sum_s e^{-i k r_s . e_a} f[s, n]                  (far_field_gemm_kernel)
//   t         = i^{tau-l-2} F[a, n] / sqrt(2 l (l+1)) e^{i m phi_a}
//   tau = 1:  E_theta += t pi_l^|m| i m,  E_phi -= t tau_l^|m|
//   tau = 2:  E_theta += t tau_l^|m|,     E_phi += t pi_l^|m| i m  (far_field_assemble_kernel)
// with e_a = (sin th cos ph, sin th sin ph, cos th) (numerics.py:314-330) and
// the angular functions of spherical_functions_trigon.py:39-93.  The first
// line is a complex GEMM (directions x sources) x (sources x modes) whose A
// operand is generated in registers; it runs on the FP64 16x16x4 MFMA.
#include "device_common.hpp"

namespace yasfhip {

namespace {

constexpr int kFarColTiles = 4;  // 16-mode column tiles per wave

// One wave per (16-direction tile, group of kFarColTiles mode tiles).
// dir: unit vectors padded to a multiple of 16 directions; F: [A][nmax].
__global__ __launch_bounds__(256) void far_field_gemm_kernel(const double4 *__restrict__ pos,
                                                             const double2 *__restrict__ f, int N,
                                                             int nmax, double k,
                                                             const double4 *__restrict__ dir, int A,
                                                             int ncol_groups, double2 *__restrict__ F) {
    typedef Mfma<double>::acc_t acc_t;
    const int lane = threadIdx.x & 63;
    const int gw = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int atile = gw / ncol_groups;
    const int cg = gw - atile * ncol_groups;
    if (atile * 16 >= A)
        return;
    const int lo = lane & 15, hi = lane >> 4;
    const double4 e = dir[atile * 16 + lo];
    const int ntiles = (nmax + 15) / 16;
    const int ct0 = cg * kFarColTiles;
    const int nct = min(kFarColTiles, ntiles - ct0);
    acc_t acc[kFarColTiles][2];
#pragma unroll
    for (int t = 0; t < kFarColTiles; ++t) {
        acc[t][0] = acc_t{0, 0, 0, 0};
        acc[t][1] = acc_t{0, 0, 0, 0};
    }
    for (int k0 = 0; k0 < N; k0 += 4) {
        const int s = k0 + hi;
        double ar = 0.0, ai = 0.0;
        if (s < N) {
            const double4 r = pos[s];
            sincos(-k * (r.x * e.x + r.y * e.y + r.z * e.z), &ai, &ar);
        }
#pragma unroll
        for (int t = 0; t < kFarColTiles; ++t) {
            if (t < nct) {
                const int n = (ct0 + t) * 16 + lo;
                double2 b = make_double2(0.0, 0.0);
                if (s < N && n < nmax)
                    b = f[static_cast<size_t>(s) * nmax + n];
                acc[t][0] = Mfma<double>::mac(ar, b.x, acc[t][0]);
                acc[t][0] = Mfma<double>::mac(-ai, b.y, acc[t][0]);
                acc[t][1] = Mfma<double>::mac(ar, b.y, acc[t][1]);
                acc[t][1] = Mfma<double>::mac(ai, b.x, acc[t][1]);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < kFarColTiles; ++t) {
        if (t < nct) {
            const int n = (ct0 + t) * 16 + lo;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int a = atile * 16 + Mfma<double>::row(i, hi);
                if (a < A && n < nmax)
                    F[static_cast<size_t>(a) * nmax + n] = make_double2(acc[t][0][i], acc[t][1][i]);
            }
        }
    }
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// i^e for any integer e
__device__ __forceinline__ double2 ipow(int e) {
    switch (((e % 4) + 4) % 4) {
    case 0:
        return make_double2(1.0, 0.0);
    case 1:
        return make_double2(0.0, 1.0);
    case 2:
        return make_double2(-1.0, 0.0);
    default:
        return make_double2(0.0, -1.0);
    }
}

// Adds the tau = 1, 2 terms of modes (l, mm) with angular values pi, tau.
__device__ __forceinline__ void far_terms(const double2 *__restrict__ Fa, int lmax, int l, int mm,
                                          double pil, double taul, double phi, double2 &eth,
                                          double2 &eph) {
    const int LL = lmax * (lmax + 2);
    double sp, cp;
    sincos(mm * phi, &sp, &cp);
    const double norm = 1.0 / sqrt(2.0 * l * (l + 1));
    const double2 em = make_double2(cp * norm, sp * norm);
    const int n1 = l * (l + 1) - 1 + mm;
    // tau = 1
    {
        const double2 t = cmul(cmul(ipow(-1 - l), Fa[n1]), em);
        // t * pi * i mm
        eth.x += -t.y * pil * mm;
        eth.y += t.x * pil * mm;
        eph.x -= t.x * taul;
        eph.y -= t.y * taul;
    }
    // tau = 2
    {
        const double2 t = cmul(cmul(ipow(-l), Fa[LL + n1]), em);
        eth.x += t.x * taul;
        eth.y += t.y * taul;
        eph.x += -t.y * pil * mm;
        eph.y += t.x * pil * mm;
    }
}

// One thread per direction; pi_l^m, tau_l^m are generated per m by the
// recurrences of spherical_functions_trigon.py:39-93 while they are consumed.
__global__ __launch_bounds__(256) void far_field_assemble_kernel(const double2 *__restrict__ F, int A,
                                                                 int lmax, int nmax,
                                                                 const double *__restrict__ polar,
                                                                 const double *__restrict__ azimuthal,
                                                                 double2 *__restrict__ e_theta,
                                                                 double2 *__restrict__ e_phi) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A)
        return;
    const double th = polar[a], ph = azimuthal[a];
    const double ct = cos(th), st = sin(th);
    const double2 *Fa = F + static_cast<size_t>(a) * nmax;
    double2 eth = make_double2(0.0, 0.0), eph = make_double2(0.0, 0.0);
    // m = 0: pi = 0, tau_l^0 = -sin th P'_l with P' by its own recurrence.
    {
        double p_prev = sqrt(0.5), p_cur = sqrt(1.5) * ct;  // P_0^0, P_1^0
        double d_cur = sqrt(3.0) * sqrt(0.5);               // P'_1
        for (int l = 1; l <= lmax; ++l) {
            far_terms(Fa, lmax, l, 0, 0.0, -st * d_cur, ph, eth, eph);
            // advance to l + 1
            const double p_next = sqrt((2.0 * l + 1) * (2.0 * l + 3)) / (l + 1) * ct * p_cur -
                                  static_cast<double>(l) / (l + 1) * sqrt((2.0 * l + 3) / (2.0 * l - 1)) * p_prev;
            d_cur = sqrt((2.0 * l + 3) / (2.0 * l + 1)) * ((l + 1) * p_cur + ct * d_cur);
            p_prev = p_cur;
            p_cur = p_next;
        }
    }
    double cm = 1.0;  // (2m-1)!! / sqrt((2m)!)
    for (int m = 1; m <= lmax; ++m) {
        cm *= sqrt((2.0 * m - 1) / (2.0 * m));
        const double coeff = sqrt((2.0 * m + 1) / 2.0) * cm;
        double pi_prev = 0.0, pi_cur = coeff;  // pi_{m-1}^m, pi_m^m (times st^(m-1))
        for (int i = 1; i < m; ++i)
            pi_cur *= st;
        double tau_cur = m * ct * pi_cur;
        for (int l = m; l <= lmax; ++l) {
            far_terms(Fa, lmax, l, m, pi_cur, tau_cur, ph, eth, eph);
            far_terms(Fa, lmax, l, -m, pi_cur, tau_cur, ph, eth, eph);
            const double c1 = sqrt((2.0 * l + 1) * (2.0 * l + 3) / ((l + 1.0 - m) * (l + 1.0 + m))) * ct;
            const double c2 = sqrt((2.0 * l + 3) * (l - m) * (l + m) / ((2.0 * l - 1) * (l + 1.0 - m) * (l + 1.0 + m)));
            const double pi_next = c1 * pi_cur - c2 * pi_prev;
            tau_cur = (l + 1) * ct * pi_next -
                      (l + 1.0 + m) * sqrt((2.0 * l + 3) * (l + 1.0 - m) / ((2.0 * l + 1) * (l + 1.0 + m))) * pi_cur;
            pi_prev = pi_cur;
            pi_cur = pi_next;
        }
    }
    e_theta[a] = eth;
    e_phi[a] = eph;
}

}  // namespace

void launch_far_field(const double4 *pos, const double2 *f, int N, int lmax, double k,
                      const double4 *dir, const double *polar, const double *azimuthal, int A,
                      double2 *F, double2 *e_theta, double2 *e_phi, hipStream_t s) {
    const int nmax = 2 * lmax * (lmax + 2);
    const int ncol_groups = ((nmax + 15) / 16 + kFarColTiles - 1) / kFarColTiles;
    const int waves = (A + 15) / 16 * ncol_groups;
    hipLaunchKernelGGL(far_field_gemm_kernel, dim3((waves + 3) / 4), dim3(256), 0, s, pos, f, N, nmax, k,
                       dir, A, ncol_groups, F);
    YH_CHECK(hipGetLastError());
    hipLaunchKernelGGL(far_field_assemble_kernel, dim3((A + 255) / 256), dim3(256), 0, s, F, A, lmax, nmax,
                       polar, azimuthal, e_theta, e_phi);
    YH_CHECK(hipGetLastError());
}

}  // namespace yasfhip
