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
#include "device_common.hpp"

namespace yasfhip {

// One block of 8 waves per (job, 16-particle target tile); jobs are sorted by
// cost and the block index is job-major so the heaviest jobs start first.
// Waves 4..7 produce, waves 0..3 consume, over chunks of 16 sources:
//   producers: chunk i+1 -> LDS buffer (i+1)&1: every pair's G[p] for all of
//              the job's orders (FP64 series, stored in the MFMA A layout) and
//              the job's source columns of x (the B operand);
//   consumers: complex 16x16x4 MFMAs of their C tiles (c = wave, wave+4, ...)
//              on chunk i from buffer i&1.
// One barrier per chunk hands the buffers over, so generation runs on the
// VALU while the matrix cores work.
// Q layout: [target tile][slot][col 16][row 16] complex T (interleaved); C tile
// (ip, column tile it) of a dm goes to slot qbase + ip * tstride + it.
// Regular: the J-translation variant (j_p in place of h_p, coincident pairs
// included) used by the scattering cross section.
template <typename T, bool Regular>
__global__ __launch_bounds__(512) void pair_kernel(const double4 *__restrict__ pos,
                                                   const typename Cplx<T>::type *__restrict__ x, int N,
                                                   int nmax, double k, const PairJob *__restrict__ jobs,
                                                   int ntiles1, const int *__restrict__ colmap,
                                                   int maxcols, int lmax, const double *__restrict__ la,
                                                   const double *__restrict__ lb,
                                                   const double *__restrict__ ld, T *__restrict__ Q, int nq,
                                                   int stage_elems, const int *__restrict__ gate,
                                                   double near2, int t0) {
    typedef typename Mfma<T>::acc_t acc_t;
    typedef typename Cplx<T>::type T2;
    constexpr int SRC = 16;  // sources per chunk (4 MFMA K-steps)
    extern __shared__ unsigned char lds_raw[];
    T2 *stage = reinterpret_cast<T2 *>(lds_raw);  // [2][stage_elems]: A [R][256] then B [16][ncols]

    if (gate && *gate != 1)  // batched callers: run only while the request's state word is 1
        return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool producer = wave >= kWavesPerJob;
    // ntiles1 target tiles from t0 on; Q holds them from its start
    const int jid = blockIdx.x / ntiles1;
    const int s1t = t0 + static_cast<int>(blockIdx.x) - jid * ntiles1;
    const int dm = uniform(jobs[jid].dm);
    const int tile0 = uniform(jobs[jid].tile0);
    const int ntile = uniform(jobs[jid].ntile);
    const int qbase = uniform(jobs[jid].qbase);
    const int tstride = uniform(jobs[jid].tstride);
    const int P = 2 * lmax + 1;
    const int R = P - (dm < 0 ? -dm : dm);
    const int ncols = ntile * 16;
    const int nord = uniform(jobs[jid].nord);
    const int total = nord * ntile;  // C tiles: the first nord orders x the job's column tiles
    const int boff = R * 256;
    const int *cmap = colmap + (dm + 2 * lmax) * maxcols + tile0 * 16;
    const int nchunks = (N + SRC - 1) / SRC;

    // Producers (waves 4..7): one (target gr, source gk) pair per thread; they
    // generate G for chunk ch+1 while the consumers multiply chunk ch, and
    // prefetch the next source position one chunk ahead.
    // Consumers (waves 0..3): besides their MFMAs they stage the B operand:
    // thread (source bk, column bc + 16 j) keeps chunk ch+1's x values in
    // registers, loaded before the MFMAs of chunk ch and stored after them.
    const int ptid = tid - kWavesPerJob * 64;
    const int gr = ptid & 15, gk = ptid >> 4;
    const int s1 = s1t * 16 + gr;
    const int bc = tid & 15, bk = tid >> 4;
    double4 p1 = make_double4(0, 0, 0, 0), p2next = make_double4(0, 0, 0, 0);
    int n2reg[kJobMaxColTiles];
    T2 breg[kJobMaxColTiles];
    if (producer) {
        p1 = pos[s1];  // pos is padded to a multiple of 16
        p2next = pos[gk < N ? gk : 0];
    } else {
#pragma unroll
        for (int j = 0; j < kJobMaxColTiles; ++j)
            n2reg[j] = (j < ntile) ? cmap[j * 16 + bc] : -1;
    }
    auto load_b = [&](int chunk) {
        const int s2 = chunk * SRC + bk;
#pragma unroll
        for (int j = 0; j < kJobMaxColTiles; ++j) {
            breg[j].x = T(0);
            breg[j].y = T(0);
            if (n2reg[j] >= 0 && s2 < N)
                breg[j] = x[static_cast<size_t>(s2) * nmax + n2reg[j]];
        }
    };
    auto store_b = [&](T2 *buf) {
#pragma unroll
        for (int j = 0; j < kJobMaxColTiles; ++j)
            if (j < ntile)
                buf[boff + bk * ncols + j * 16 + bc] = breg[j];
    };
    auto produce = [&](int chunk, T2 *buf) {
        const int s2 = chunk * SRC + gk;
        const double4 p2 = p2next;
        const int s2n = s2 + SRC;
        p2next = pos[s2n < N ? s2n : 0];
        const double dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
        // near2: pairs farther apart than the near-field radius do not couple
        const bool valid = (s2 < N) && (s1 < N) && (Regular || (s2 != s1 && dx * dx + dy * dy + dz * dz <= near2));
        if constexpr (Regular)
            pair_series_regular<T2>(dx, dy, dz, k, dm, P, valid, la, lb, ld, buf + ptid, 256);
        else
            pair_series<T2>(dx, dy, dz, k, dm, P, valid, la, lb, ld, buf + ptid, 256);
    };

    acc_t acc[kWaveMaxTiles][2];
#pragma unroll
    for (int j = 0; j < kWaveMaxTiles; ++j) {
        acc[j][0] = acc_t{0, 0, 0, 0};
        acc[j][1] = acc_t{0, 0, 0, 0};
    }
    const int lo = lane & 15, hi = lane >> 4;

    if (producer) {
        produce(0, stage);
    } else {
        load_b(0);
        store_b(stage);
        if (nchunks > 1)
            load_b(1);
    }
    __syncthreads();
    for (int ch = 0; ch < nchunks; ++ch) {
        T2 *cur = stage + (ch & 1) * stage_elems;
        T2 *nxt = stage + ((ch + 1) & 1) * stage_elems;
        if (producer) {
            if (ch + 1 < nchunks)
                produce(ch + 1, nxt);
        } else {
#pragma unroll
            for (int j = 0; j < kWaveMaxTiles; ++j) {
                const int c = wave + kWavesPerJob * j;
                if (c < total) {
                    const int ip = c / ntile, it = c - ip * ntile;
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) {
                        const T2 a = cur[ip * 256 + ks * 64 + lane];
                        const T2 bv = cur[boff + (ks * 4 + hi) * ncols + it * 16 + lo];
                        acc[j][0] = Mfma<T>::mac(a.x, bv.x, acc[j][0]);
                        acc[j][0] = Mfma<T>::mac(-a.y, bv.y, acc[j][0]);
                        acc[j][1] = Mfma<T>::mac(a.x, bv.y, acc[j][1]);
                        acc[j][1] = Mfma<T>::mac(a.y, bv.x, acc[j][1]);
                    }
                }
            }
            if (ch + 1 < nchunks) {
                store_b(nxt);
                if (ch + 2 < nchunks)
                    load_b(ch + 2);
            }
        }
        __syncthreads();
    }
    if (producer)
        return;

    // Accumulator element i: target row Mfma<T>::row(i, hi), column lo.
#pragma unroll
    for (int j = 0; j < kWaveMaxTiles; ++j) {
        const int c = wave + kWavesPerJob * j;
        if (c < total) {
            const int ip = c / ntile, it = c - ip * ntile;
            const size_t slot = static_cast<size_t>(s1t - t0) * nq + qbase + ip * tstride + tile0 + it;
            T *dst = Q + 2 * ((slot * 16 + lo) * 16);
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
// One block per (16-particle target tile, group of 16 target modes).  The
// mode groups of a tile read the same Q rows, so they share an XCD (and its
// L2): blocks are handed to the 8 XCDs in turn, and block 8 q + x runs tile
// 8 (q / groups) + x, group q % groups (the grid is padded to a multiple of
// 8 tiles; blocks past the last tile return).
template <typename TQ, typename TX, typename TY>
__global__ __launch_bounds__(256) void contract_kernel(const TQ *__restrict__ Q, int nq,
                                                       const int *__restrict__ crow,
                                                       const int *__restrict__ cidx,
                                                       const double2 *__restrict__ cval, int N, int nmax,
                                                       const TX *__restrict__ x,
                                                       const double2 *__restrict__ t, int mode,
                                                       TY *__restrict__ y, int t0, int t1,
                                                       const int *__restrict__ gate) {
    // target tiles [t0, t1); Q holds them from its start
    if (gate && *gate != 1)
        return;
    const int groups = contract_groups(nmax);
    const int q = blockIdx.x >> 3, xcd = blockIdx.x & 7;
    const int s1t = t0 + (q / groups) * 8 + xcd;
    if (s1t >= t1 || s1t * 16 >= N)
        return;
    contract_tile<TQ, TX, TY>(Q, nq, crow, cidx, cval, N, nmax, x, t, mode, y, s1t, q % groups, t0);
}

template <typename T, bool Regular>
static void launch_pair_t(const DevicePlan &dp, const double4 *pos, int N, double k, const void *x,
                          void *Q, hipStream_t s, const int *gate = nullptr, double near2 = kAllPairs,
                          int t0 = 0, int ntl = -1) {
    const int ntiles1 = ntl < 0 ? (N + 15) / 16 - t0 : ntl;
    const size_t lds = 2 * static_cast<size_t>(dp.stage_elems) * sizeof(typename Cplx<T>::type);
    static bool lds_opt_in = false;  // gfx950: up to 160 KB of LDS per workgroup
    if (!lds_opt_in) {
        YH_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&pair_kernel<T, Regular>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        lds_opt_in = true;
    }
    hipLaunchKernelGGL((pair_kernel<T, Regular>), dim3(dp.njobs * ntiles1), dim3(2 * kWavesPerJob * 64), lds, s,
                       pos, static_cast<const typename Cplx<T>::type *>(x), N, dp.nmax, k, dp.jobs,
                       ntiles1, dp.colmap, dp.maxcols, dp.lmax, dp.la, dp.lb, dp.ld, static_cast<T *>(Q),
                       dp.nq, dp.stage_elems, gate, near2, t0);
    YH_CHECK(hipGetLastError());
}

void launch_pair(const DevicePlan &dp, const double4 *pos, int N, double k, const void *x, bool fp32,
                 void *Q, hipStream_t s, const int *gate, double near2, int t0, int ntl) {
    if (fp32)
        launch_pair_t<float, false>(dp, pos, N, k, x, Q, s, gate, near2, t0, ntl);
    else
        launch_pair_t<double, false>(dp, pos, N, k, x, Q, s, gate, near2, t0, ntl);
}

void launch_pair_regular(const DevicePlan &dp, const double4 *pos, int N, double k, const void *x,
                         void *Q, hipStream_t s, int t0, int ntl) {
    launch_pair_t<double, true>(dp, pos, N, k, x, Q, s, nullptr, kAllPairs, t0, ntl);
}

// Largest LDS the pair kernel asks for, bytes (2 stages of complex128).
size_t pair_lds_bytes(const DevicePlan &dp) { return 2 * static_cast<size_t>(dp.stage_elems) * 16; }

// q_fp32: Q is complex64 (the FP32 pair GEMM), otherwise complex128; the
// contraction always accumulates in FP64.  io_fp32: x and y are complex64,
// otherwise complex128 (the mixed-precision operator reads FP32 Q with FP64
// x and y, so x - t W x is formed in FP64).
void launch_contract(const DevicePlan &dp, int N, const void *Q, const void *x, const double2 *t,
                     int mode, void *y, bool q_fp32, bool io_fp32, hipStream_t s, int t0, int ntl,
                     const int *gate) {
    const int ntiles = ntl < 0 ? (N + 15) / 16 - t0 : ntl;
    const int t1 = t0 + ntiles;
    const int nblocks = (ntiles + 7) / 8 * 8 * contract_groups(dp.nmax);
    YH_REQUIRE(q_fp32 || !io_fp32, "complex64 vectors need the complex64 pair GEMM");
    if (io_fp32)
        hipLaunchKernelGGL((contract_kernel<float, float, float>), dim3(nblocks), dim3(256), 0, s,
                           static_cast<const float *>(Q), dp.nq, dp.crow, dp.cidx, dp.cval, N, dp.nmax,
                           static_cast<const float *>(x), t, mode, static_cast<float *>(y), t0, t1, gate);
    else if (q_fp32)
        hipLaunchKernelGGL((contract_kernel<float, double, double>), dim3(nblocks), dim3(256), 0, s,
                           static_cast<const float *>(Q), dp.nq, dp.crow, dp.cidx, dp.cval, N, dp.nmax,
                           static_cast<const double *>(x), t, mode, static_cast<double *>(y), t0, t1, gate);
    else
        hipLaunchKernelGGL((contract_kernel<double, double, double>), dim3(nblocks), dim3(256), 0, s,
                           static_cast<const double *>(Q), dp.nq, dp.crow, dp.cidx, dp.cval, N,
                           dp.nmax, static_cast<const double *>(x), t, mode, static_cast<double *>(y), t0, t1, gate);
    YH_CHECK(hipGetLastError());
}

}  // namespace yasfhip
