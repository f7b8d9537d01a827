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

// E_theta, E_phi of one direction (cos th, sin th, ph) from that direction's
// row Fa of F; pi_l^m, tau_l^m are generated per m by the recurrences of
// spherical_functions_trigon.py:39-93 while they are consumed.
__device__ __forceinline__ void assemble_direction(const double2 *Fa, int lmax, double ct, double st, double ph,
                                                   double2 &eth, double2 &eph) {
    eth = make_double2(0.0, 0.0);
    eph = make_double2(0.0, 0.0);
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
}

// One thread per direction.
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
    double2 eth, eph;
    assemble_direction(F + static_cast<size_t>(a) * nmax, lmax, cos(th), sin(th), ph, eth, eph);
    e_theta[a] = eth;
    e_phi[a] = eph;
}

// Node x_i and weight w_i of the n-point Gauss-Legendre rule on [-1, 1]
// (Newton's iteration on P_n from x = cos(pi (i + 3/4) / (n + 1/2))).
__device__ inline void gauss_legendre_node(int i, int n, double &x, double &w) {
    x = cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
        double p0 = 1.0, p1 = x;
        for (int j = 2; j <= n; ++j) {
            const double p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / j;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (x * p1 - p0) / (x * x - 1.0);
        const double dx = p1 / dp;
        x -= dx;
        if (fabs(dx) <= 1e-16)
            break;
    }
    // P_n'(x) at the converged node
    double p0 = 1.0, p1 = x;
    for (int j = 2; j <= n; ++j) {
        const double p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
    }
    dp = n * (x * p1 - p0) / (x * x - 1.0);
    w = 2.0 / ((1.0 - x * x) * dp * dp);
}

// Scattering cross section by quadrature of the far field:
//   C_sca k^2 = int |E|^2 dOmega = sum_{i,j} w_i (2 pi / n_phi) |E(x_i, phi_j)|^2,
// Gauss-Legendre in cos th (n_theta nodes) and the trapezoid rule in phi
// (n_phi points), exact for the band-limited far field of the cluster.  One
// block (4 waves) per 16 directions and up to kQuadMaxReq solutions f of one
// geometry (their mode tiles, at most kQuadMaxTiles, dealt to the waves in
// turn): per chunk of 16 sources the waves compute the 16 x 16 phase factors
// once into LDS (each wave 4 sources), then every wave runs the chunk's four
// K-steps of the far-field GEMM (far_field_gemm_kernel's) for its tiles.
// The F rows of the block's directions go to LDS, 16 threads per solution
// assemble E, and the block leaves the weighted sum of its 16 directions (in
// direction order) in partials[i][block] of solution i.
constexpr int kQuadMaxTiles = 12;  // mode tiles of a block: 3 per wave
constexpr int kQuadMaxReq = 12;
struct QuadBatch {
    const double2 *f[kQuadMaxReq];
    double *partials[kQuadMaxReq];
};
__global__ __launch_bounds__(256) void csca_quadrature_kernel(const double4 *__restrict__ pos, QuadBatch batch,
                                                              int nreq, int N, int lmax, int nmax, double k,
                                                              int n_theta, int n_phi) {
    typedef Mfma<double>::acc_t acc_t;
    constexpr int kPerWave = kQuadMaxTiles / 4;
    __shared__ double2 Fs[16][kQuadMaxTiles * 16];
    __shared__ double2 Ph[16][16];  // phase factor of direction r and source k0 + c
    __shared__ double Is[kQuadMaxReq][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lo = lane & 15, hi = lane >> 4;
    const int A = n_theta * n_phi;
    const int a = blockIdx.x * 16 + lo;
    const int ith = (a < A ? a : 0) / n_phi, jph = (a < A ? a : 0) - ith * n_phi;
    double x, w;
    gauss_legendre_node(ith, n_theta, x, w);
    const double st = sqrt(fmax(0.0, 1.0 - x * x));
    const double ph = 2.0 * M_PI * jph / n_phi;
    double sph, cph;
    sincos(ph, &sph, &cph);
    const double ex = st * cph, ey = st * sph, ez = x;
    const int ntiles = (nmax + 15) / 16;
    const int ntt = nreq * ntiles;  // the block's mode tiles: solution ct / ntiles, tile ct % ntiles
    acc_t acc[kPerWave][2];
#pragma unroll
    for (int t = 0; t < kPerWave; ++t) {
        acc[t][0] = acc_t{0, 0, 0, 0};
        acc[t][1] = acc_t{0, 0, 0, 0};
    }
    for (int k0 = 0; k0 < N; k0 += 16) {
        {
            const int s = k0 + 4 * wave + hi;
            double ar = 0.0, ai = 0.0;
            if (s < N) {
                const double4 r = pos[s];
                sincos(-k * (r.x * ex + r.y * ey + r.z * ez), &ai, &ar);
            }
            Ph[lo][4 * wave + hi] = make_double2(ar, ai);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int s = k0 + 4 * kk + hi;
            const double2 e = Ph[lo][4 * kk + hi];
            const double ar = e.x, ai = e.y;
#pragma unroll
            for (int t = 0; t < kPerWave; ++t) {
                const int ct = wave + 4 * t;
                if (ct < ntt) {
                    const int rq = ct / ntiles;
                    const int n = (ct - rq * ntiles) * 16 + lo;
                    double2 b = make_double2(0.0, 0.0);
                    if (s < N && n < nmax)
                        b = batch.f[rq][static_cast<size_t>(s) * nmax + n];
                    acc[t][0] = Mfma<double>::mac(ar, b.x, acc[t][0]);
                    acc[t][0] = Mfma<double>::mac(-ai, b.y, acc[t][0]);
                    acc[t][1] = Mfma<double>::mac(ar, b.y, acc[t][1]);
                    acc[t][1] = Mfma<double>::mac(ai, b.x, acc[t][1]);
                }
            }
        }
        __syncthreads();  // the phases of this chunk are read
    }
#pragma unroll
    for (int t = 0; t < kPerWave; ++t) {
        const int ct = wave + 4 * t;
        if (ct < ntt) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                Fs[Mfma<double>::row(i, hi)][ct * 16 + lo] = make_double2(acc[t][0][i], acc[t][1][i]);
        }
    }
    __syncthreads();
    if (threadIdx.x < 16 * nreq) {
        const int r = threadIdx.x & 15, rq = threadIdx.x >> 4;  // direction r of the tile, solution rq
        double v = 0.0;
        if (a < A) {  // lanes 0..15 and their copies in each group of 16 hold direction r's node
            double2 eth, eph;
            assemble_direction(Fs[r] + rq * ntiles * 16, lmax, x, st, ph, eth, eph);
            v = w * (2.0 * M_PI / n_phi) * (eth.x * eth.x + eth.y * eth.y + eph.x * eph.x + eph.y * eph.y);
        }
        Is[rq][r] = v;
    }
    __syncthreads();
    if (threadIdx.x < nreq) {
        double sum = 0.0;
        for (int r = 0; r < 16; ++r)
            sum += Is[threadIdx.x][r];
        batch.partials[threadIdx.x][blockIdx.x] = sum;
    }
}

// out[0] = scale * sum of count partials, summed in a fixed order (strided
// per thread, then a fixed tree), out[1] = 0.
__global__ __launch_bounds__(256) void sum_partials_kernel(const double *__restrict__ partials, int count,
                                                           double scale, double *__restrict__ out) {
    __shared__ double red[256];
    double v = 0.0;
    for (int i = threadIdx.x; i < count; i += 256)
        v += partials[i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = scale * red[0];
        out[1] = 0.0;
    }
}

}  // namespace

// Quadrature order M for the far field of spheres within radius R of the
// origin (k R) at expansion order lmax: n_theta = M + 1 Gauss-Legendre nodes
// and n_phi = 2 M + 2 azimuths integrate |E|^2 (degree <= 2 M in the
// direction) exactly; M = k R + lmax + 15^(2/3) (k R + 1)^(1/3) leaves the
// plane-wave expansion's truncation error below 1e-15 relative.
void csca_quadrature_grid(double kR, int lmax, int *n_theta, int *n_phi) {
    const int M = static_cast<int>(std::ceil(kR + lmax + std::pow(15.0, 2.0 / 3.0) * std::cbrt(kR + 1.0)));
    *n_theta = M + 1;
    *n_phi = 2 * M + 2;
}

size_t csca_quadrature_partials(int n_theta, int n_phi) {
    return (static_cast<size_t>(n_theta) * n_phi + 15) / 16;
}

// out[0] = C_sca k^2 / pi (so that the pair-sum's |out[0] + i out[1]| pi / k^2
// convention gives C_sca), out[1] = 0.  partials: csca_quadrature_partials
// doubles of device scratch.
void launch_csca_quadrature(const double4 *pos, const double2 *f, int N, int lmax, double k, int n_theta,
                            int n_phi, double *partials, double *out, hipStream_t s) {
    launch_csca_quadrature_batch(pos, &f, &partials, &out, 1, N, lmax, k, n_theta, n_phi, s);
}

int csca_quadrature_batch_max(int lmax) {
    const int ntiles = (2 * lmax * (lmax + 2) + 15) / 16;
    return std::max(1, std::min(kQuadMaxReq, kQuadMaxTiles / ntiles));
}

void launch_csca_quadrature_batch(const double4 *pos, const double2 *const *f, double *const *partials,
                                  double *const *out, int nreq, int N, int lmax, double k, int n_theta, int n_phi,
                                  hipStream_t s) {
    const int nmax = 2 * lmax * (lmax + 2);
    YH_REQUIRE(nreq >= 1 && nreq <= csca_quadrature_batch_max(lmax), "C_sca quadrature batch too large");
    const int blocks = static_cast<int>(csca_quadrature_partials(n_theta, n_phi));
    QuadBatch batch{};
    for (int i = 0; i < nreq; ++i) {
        batch.f[i] = f[i];
        batch.partials[i] = partials[i];
    }
    hipLaunchKernelGGL(csca_quadrature_kernel, dim3(blocks), dim3(256), 0, s, pos, batch, nreq, N, lmax, nmax, k,
                       n_theta, n_phi);
    YH_CHECK(hipGetLastError());
    for (int i = 0; i < nreq; ++i)
        hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, partials[i], blocks, 1.0 / M_PI, out[i]);
    YH_CHECK(hipGetLastError());
}

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
