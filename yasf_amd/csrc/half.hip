// Resident pair-factor table in fp16 terms and the pair GEMM on the fp16
// matrix cores (see half_gemm.hpp for the number format and layouts).
//
// The table holds the same factors G_{s1 s2}[p, dm] as the complex64 table
// (generated in FP64 by pair_series_to), scaled per order and split into a
// hi and a lo fp16 term: the two terms take the bytes of one complex64 table
// and the GEMM's products run at the fp16 MFMA rate; the hi term alone is
// half of that.
#include "half_gemm.hpp"

namespace yasfhip {

namespace {

__global__ void init_min_kernel(unsigned long long *v) { *v = 0x7ff0000000000000ull; }  // +inf

// Smallest squared distance between two spheres (as double bits, atomic min;
// non-negative doubles order like their bit patterns).  One thread per s1.
__global__ __launch_bounds__(256) void min_distance_kernel(const double4 *__restrict__ pos, int N,
                                                           unsigned long long *__restrict__ out) {
    const int s1 = blockIdx.x * blockDim.x + threadIdx.x;
    double best = HUGE_VAL;
    if (s1 < N) {
        const double4 p1 = pos[s1];
        for (int s2 = s1 + 1; s2 < N; ++s2) {
            const double4 p2 = pos[s2];
            const double dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
            best = fmin(best, dx * dx + dy * dy + dz * dz);
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        best = fmin(best, __shfl_down(best, off, 64));
    if ((threadIdx.x & 63) == 0 && best < HUGE_VAL)
        atomicMin(out, static_cast<unsigned long long>(__double_as_longlong(best)));
}

// The order exponents pexp[p] (p < P) from the smallest squared pair
// distance (double bits in dmin2).  One block.
__global__ void half_exponents_kernel(const unsigned long long *__restrict__ dmin2, double k, int P,
                                      int *__restrict__ pexp) {
    const double d2 = __longlong_as_double(static_cast<long long>(*dmin2));
    for (int p = threadIdx.x; p < P; p += blockDim.x)
        pexp[p] = half_order_exponent(p, k, d2);
}

// One block per (target tile, 16-source chunk); thread = one pair, all of
// its dm (the pair's geometry and Hankel functions formed once), with the 8
// sources of a lane's A fragment on consecutive threads so the fp16 stores of
// a wave are contiguous.  Builds target tiles t0.. (table points at
// tile t0); sources up to N32 * 32 (zero past N); pexp: the order exponents.
// gate: as the GEMM's (returns unless *gate == 1).
template <bool SPLIT>
__global__ __launch_bounds__(256) void half_table_build_kernel(
    const double4 *__restrict__ pos, int N, int N32, double k, int lmax, int ndm, const int *__restrict__ dmoff,
    const double *__restrict__ la, const double *__restrict__ lb, const double *__restrict__ ld,
    _Float16 *__restrict__ hi_tab, _Float16 *__restrict__ lo_tab, int t0, const int *__restrict__ pexp,
    double near2, const int *__restrict__ gate) {
    if (gate && *gate != 1)
        return;
    __shared__ int es[kMaxOrders];
    __shared__ int dms[4 * kMaxLmax + 1];
    const int P = 2 * lmax + 1;
    if (threadIdx.x < P)
        es[threadIdx.x] = pexp[threadIdx.x];
    if (threadIdx.x < ndm)
        dms[threadIdx.x] = dmoff[threadIdx.x];
    __syncthreads();
    const int nchunk = 2 * N32;
    const int c = blockIdx.x % nchunk;
    const int t = blockIdx.x / nchunk;
    const int SR = P * P;
    const int i = threadIdx.x & 7, r = (threadIdx.x >> 3) & 15, g = threadIdx.x >> 7;
    const int s1 = (t0 + t) * 16 + r, s2 = c * 16 + g * 8 + i;
    const int ks = s2 >> 5, lane = ((s2 & 31) >> 3) * 16 + r;
    const double4 p1 = pos[s1];  // pos is padded to a multiple of 16
    const double4 p2 = pos[s2 < N ? s2 : 0];
    const double dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
    const bool valid = (s1 < N) && (s2 < N) && s1 != s2 && dx * dx + dy * dy + dz * dz <= near2;
    const size_t tbase = static_cast<size_t>(t) * SR * N32 * 1024 + lane * 8 + i;
    // fp16 terms through FP32: two conversion instructions instead of the
    // correctly rounded FP64 -> fp16 integer sequence (the empty asm keeps
    // the compiler from folding the two roundings back into that one).  hi
    // may differ from the nearest fp16 in a rare double-rounding tie; lo =
    // v - hi, formed in FP64, takes up the difference.
    auto to_half = [](double v) {
        float f = static_cast<float>(v);
        asm volatile("" : "+v"(f));
        return static_cast<_Float16>(f);
    };
    pair_series_all(dx, dy, dz, k, lmax, valid, la, lb, ld, [&](int d, int ip, double gr, double gi) {
        const int adm = d < 2 * lmax ? 2 * lmax - d : d - 2 * lmax;
        const int R = P - adm;
        const int e = es[adm + ip];
        const double vr = ldexp(gr, e), vi = ldexp(gi, e);
        const _Float16 hr = to_half(vr), hm = to_half(vi);
        const size_t o = tbase + (static_cast<size_t>(dms[d]) * N32 + static_cast<size_t>(ks) * R + ip) * 1024;
        hi_tab[o] = hr;
        hi_tab[o + 512] = hm;
        if constexpr (SPLIT) {
            lo_tab[o] = to_half(vr - static_cast<double>(hr));
            lo_tab[o + 512] = to_half(vi - static_cast<double>(hm));
        }
    });
}

__global__ void zero_u32_kernel(unsigned *v, int n, const int *__restrict__ gate) {
    if (gate && *gate != 1)
        return;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        v[i] = 0u;
}

__global__ __launch_bounds__(256) void half_colmax_kernel(const double2 *__restrict__ x, int N, int nmax,
                                                          unsigned *__restrict__ colmax,
                                                          const int *__restrict__ gate) {
    if (gate && *gate != 1)
        return;
    half_colmax_block(x, N, nmax, colmax, blockIdx.x);
}

template <bool SPLIT>
__global__ __launch_bounds__(256) void half_convert_kernel(const double2 *__restrict__ x, int N, int N32, int nmax,
                                                           const unsigned *__restrict__ colmax,
                                                           int *__restrict__ xexp, _Float16 *__restrict__ xhi,
                                                           _Float16 *__restrict__ xlo,
                                                           const int *__restrict__ gate) {
    if (gate && *gate != 1)
        return;
    const long idx = blockIdx.x * 256L + threadIdx.x;
    if (idx < static_cast<long>(nmax + 1) * N32 * 32)
        half_convert_elem<SPLIT>(x, N, N32, nmax, colmax, xexp, xhi, xlo, idx);
}

// One block per (bin, target tile) of tiles [t0, t0 + ntl); hi / lo point at
// the table's tile t0.
template <int WAVES, int NTL, bool SPLIT>
__global__ __launch_bounds__(64 * WAVES) void half_gemm_kernel(
    const _Float16 *__restrict__ hi_tab, const _Float16 *__restrict__ lo_tab, const _Float16 *__restrict__ xhi,
    const int *__restrict__ pexp, const int *__restrict__ xexp, int N32, int nmax, int lmax,
    const HalfBin *__restrict__ bins, int ntl, int t0, const int *__restrict__ colmap,
    const int *__restrict__ dmoff, float *__restrict__ Q, int nq, const int *__restrict__ gate) {
    if (gate && *gate != 1)
        return;
    const int bid = blockIdx.x / ntl;
    const int t = blockIdx.x - bid * ntl;
    const int P = 2 * lmax + 1;
    const size_t f0 = static_cast<size_t>(t) * P * P * N32 * 1024;
    const size_t xplane = half_x_plane(nmax, N32);
    float *Qt = Q + 2 * static_cast<size_t>(t0 + t) * nq * 256;
    half_gemm_block<WAVES, NTL, SPLIT>(hi_tab + f0, SPLIT ? lo_tab + f0 : nullptr, HalfOneSource{xhi, xexp, Qt},
                                       xplane, N32, nmax, lmax, bins + bid, colmap, dmoff, pexp);
}

constexpr int kHalfWaves = kHalfBinWaves;

}  // namespace

size_t half_table_term_bytes(const DevicePlan &dp, int N) {
    const size_t ntiles1 = (N + 15) / 16;
    return ntiles1 * half_tile_halves(dp.lmax, (N + 31) / 32) * sizeof(_Float16);
}

size_t half_x_bytes(const DevicePlan &dp, int N) {
    return 4 * half_x_plane(dp.nmax, (N + 31) / 32) * sizeof(_Float16);  // 2 terms x 2 planes
}

// The order exponents pexp[2L+1] of a cluster: its smallest pair distance
// (into the scratch word dmin), then the bound of each order.
void launch_half_exponents(const DevicePlan &dp, const double4 *pos, int N, double k, unsigned long long *dmin,
                           int *pexp, hipStream_t s) {
    hipLaunchKernelGGL(init_min_kernel, dim3(1), dim3(1), 0, s, dmin);
    hipLaunchKernelGGL(min_distance_kernel, dim3((N + 255) / 256), dim3(256), 0, s, pos, N, dmin);
    hipLaunchKernelGGL(half_exponents_kernel, dim3(1), dim3(64), 0, s, dmin, k, 2 * dp.lmax + 1, pexp);
    YH_CHECK(hipGetLastError());
}

// Target tiles [t0, t0 + ntl) of the fp16 table into tab (hi term of tile t0
// at tab, lo term term_halves further when split), scaled by the order
// exponents pexp (launch_half_exponents).
void launch_half_table_range(const DevicePlan &dp, const double4 *pos, int N, double k, void *tab,
                             size_t term_halves, bool split, int t0, int ntl, const int *pexp, hipStream_t s,
                             double near2, const int *gate) {
    const int ntiles1 = (N + 15) / 16, N32 = (N + 31) / 32;
    YH_REQUIRE(t0 >= 0 && ntl >= 1 && t0 + ntl <= ntiles1, "fp16 table range outside the target tiles");
    const int ndm = 4 * dp.lmax + 1;
    _Float16 *hi = static_cast<_Float16 *>(tab);
    _Float16 *lo = split ? hi + term_halves : nullptr;
    const dim3 grid(static_cast<unsigned>(ntl) * 2 * N32), block(256);
    if (split)
        hipLaunchKernelGGL(half_table_build_kernel<true>, grid, block, 0, s, pos, N, N32, k, dp.lmax, ndm, dp.dmoff,
                           dp.la, dp.lb, dp.ld, hi, lo, t0, pexp, near2, gate);
    else
        hipLaunchKernelGGL(half_table_build_kernel<false>, grid, block, 0, s, pos, N, N32, k, dp.lmax, ndm, dp.dmoff,
                           dp.la, dp.lb, dp.ld, hi, lo, t0, pexp, near2, gate);
    YH_CHECK(hipGetLastError());
}

// Builds the fp16 table (hi term at table, lo term at table + term bytes when
// split) and the order exponents pexp[2L+1]; dmin is device scratch.
void launch_half_table_build(const DevicePlan &dp, const double4 *pos, int N, double k, void *table, bool split,
                             int *pexp, unsigned long long *dmin, hipStream_t s, double near2) {
    launch_half_exponents(dp, pos, N, k, dmin, pexp, s);
    launch_half_table_range(dp, pos, N, k, table, half_table_term_bytes(dp, N) / sizeof(_Float16), split, 0,
                            (N + 15) / 16, pexp, s, near2, nullptr);
}

// x (N x nmax complex128) to its fp16 terms in xh (half_x_bytes) with the
// column exponents xexp[nmax + 1]; colmax: nmax unsigned of device scratch.
void launch_half_x(const DevicePlan &dp, int N, const double2 *x, void *xh, bool split, unsigned *colmax, int *xexp,
                   hipStream_t s, const int *gate) {
    const int N32 = (N + 31) / 32;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(256), 0, s, colmax, dp.nmax, gate);
    hipLaunchKernelGGL(half_colmax_kernel, dim3((N + 255) / 256), dim3(256), 0, s, x, N, dp.nmax, colmax, gate);
    _Float16 *xhi = static_cast<_Float16 *>(xh);
    _Float16 *xlo = xhi + 2 * half_x_plane(dp.nmax, N32);
    const long count = static_cast<long>(dp.nmax + 1) * N32 * 32;
    const dim3 grid(static_cast<unsigned>((count + 255) / 256)), block(256);
    if (split)
        hipLaunchKernelGGL(half_convert_kernel<true>, grid, block, 0, s, x, N, N32, dp.nmax, colmax, xexp, xhi, xlo, gate);
    else
        hipLaunchKernelGGL(half_convert_kernel<false>, grid, block, 0, s, x, N, N32, dp.nmax, colmax, xexp, xhi, xlo, gate);
    YH_CHECK(hipGetLastError());
}

// Q (complex64, the table GEMM's layout) = G x over the fp16 table and x
// terms, for target tiles [t0, t0 + ntl) (ntl < 0: to the last); table holds
// tile t0's hi term first and the lo terms term_halves further (0: the
// whole table's term).
void launch_half_gemm(const DevicePlan &dp, int N, const void *table, const void *xh, bool split, const int *pexp,
                      const int *xexp, void *Q, hipStream_t s, const int *gate, int t0, int ntl,
                      size_t term_halves) {
    const int ntiles1 = (N + 15) / 16, N32 = (N + 31) / 32;
    if (ntl < 0)
        ntl = ntiles1 - t0;
    YH_REQUIRE(t0 >= 0 && ntl >= 1 && t0 + ntl <= ntiles1, "fp16 GEMM range outside the target tiles");
    const _Float16 *hi = static_cast<const _Float16 *>(table);
    const _Float16 *lo = hi + (term_halves ? term_halves : half_table_term_bytes(dp, N) / sizeof(_Float16));
    const _Float16 *xhi = static_cast<const _Float16 *>(xh);
    const int f = split ? 0 : 1;
    const dim3 grid(static_cast<unsigned>(dp.nhbins[f][0]) * ntl), block(64 * kHalfWaves);
    if (split)
        hipLaunchKernelGGL((half_gemm_kernel<kHalfWaves, kHalfJobTiles, true>), grid, block, 0, s, hi, lo, xhi,
                           pexp, xexp, N32, dp.nmax, dp.lmax, dp.hbins[f][0], ntl, t0, dp.colmap, dp.dmoff,
                           static_cast<float *>(Q), dp.nq, gate);
    else
        hipLaunchKernelGGL((half_gemm_kernel<kHalfWaves, kHalfJobTiles, false>), grid, block, 0, s, hi, lo,
                           xhi, pexp, xexp, N32, dp.nmax, dp.lmax, dp.hbins[f][0], ntl, t0, dp.colmap, dp.dmoff,
                           static_cast<float *>(Q), dp.nq, gate);
    YH_CHECK(hipGetLastError());
}

}  // namespace yasfhip
