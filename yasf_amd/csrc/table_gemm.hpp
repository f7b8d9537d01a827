// The streaming complex GEMM of one wave over a resident pair-factor table
// (layout table[target tile][dm][K-step][order][lane], MFMA A operand), shared
// by the single-session kernel (pairtable.hip) and the batched serving kernel
// (server.hip).
#pragma once

#include "device_common.hpp"

namespace yasfhip {

// One wave's GEMM with NIP orders (ip = wave + WAVES j) and NTL column tiles,
// all compile-time, so the MFMAs of a K-step form one straight-line block the
// scheduler can interleave.  tab points at the wave's target tile and dm
// (N4 K-steps x R orders of 64-lane fragments); n2c[it] is the source mode of
// the lane's column in column tile it (-1: padding).  otile[ip]: column tiles
// (from tile 0 of the dm) that order ip feeds -- a prefix, since the dm's
// columns are sorted by l2 -- so the C tiles (ip, it) past it are skipped
// (left zero).  C tile (ip, tile0 + it) goes to slot ip * tstride + tile0 + it
// past Qt.  Byte extents: the table slice and x below 2^31 (buffer offsets),
// checked by the launchers with table_gemm_extent_ok.
template <typename T, int NIP, int NTL, int WAVES, int DEPTH>
__device__ __forceinline__ void table_gemm_wave(const typename Cplx<T>::type *__restrict__ tab,
                                                const typename Cplx<T>::type *__restrict__ x, int N,
                                                int N4, int nmax, int R, int wave,
                                                const int (&n2c)[kTableMaxColTiles], T *__restrict__ Qt,
                                                int tstride, int tile0, const int *__restrict__ otile) {
    typedef typename Mfma<T>::acc_t acc_t;
    typedef typename Cplx<T>::type T2;
    const int lane = threadIdx.x & 63, lo = lane & 15, hi = lane >> 4;
    int tn[NIP];  // active column tiles of each of the wave's orders (uniform)
#pragma unroll
    for (int j = 0; j < NIP; ++j) {
        const int t = uniform(otile[wave + WAVES * j]) - tile0;
        tn[j] = t < 0 ? 0 : (t > NTL ? NTL : t);
    }
    // Complex128: the complex product with three real MFMAs (Gauss): P1 =
    // Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); Re = P1 - P2, Im = P3 - P1 -
    // P2 at the end -- a quarter fewer FP64 MFMAs (measured: 7.04 -> 6.07 ms
    // at 2,000 spheres, lmax 6).  Complex64 keeps four (the third accumulator
    // costs it more occupancy than the saved MFMAs gain).
    constexpr bool kGaussComplex = sizeof(T) == sizeof(double);
    // Complex64 skips the C tiles past each order's column prefix (the
    // padding of the job's order range: -12% to -19% time at lmax 4..8);
    // in complex128 the branches cost what the skipped MFMAs save.
    constexpr bool kTrimTiles = sizeof(T) == sizeof(float);
    constexpr int NACC = kGaussComplex ? 3 : 2;
    acc_t acc[NIP][NTL][NACC];
#pragma unroll
    for (int j = 0; j < NIP; ++j)
#pragma unroll
        for (int it = 0; it < NTL; ++it)
#pragma unroll
            for (int c = 0; c < NACC; ++c)
                acc[j][it][c] = acc_t{0, 0, 0, 0};
    // DEPTH (3) operand buffers in rotation: the loads of K-step ks+2 are
    // issued before the MFMAs of step ks and only the oldest batch is waited
    // for (measured: two buffers are up to 7% slower, four no faster).
    // Loads are unconditional, so the wait counts are static.  Both operands
    // come through buffer loads: A at a wave-uniform offset (scalar) plus the
    // lane's, B at the lane's offset into x, where lanes past the last source
    // or in padding columns fall outside x and read zero -- no 64-bit
    // address registers per load, no masking at use.
    T2 aA[NIP], bA[NTL], aB[NIP], bB[NTL];
    const int wv = uniform(wave);
    const __amdgpu_buffer_rsrc_t ra = buffer_rsrc(tab, static_cast<size_t>(R) * N4 * 64 * sizeof(T2));
    const __amdgpu_buffer_rsrc_t rx = buffer_rsrc(x, static_cast<size_t>(N) * nmax * sizeof(T2));
    const unsigned va = lane * sizeof(T2);
    unsigned vb[NTL];
#pragma unroll
    for (int it = 0; it < NTL; ++it)
        vb[it] = n2c[it] >= 0 ? static_cast<unsigned>((hi * nmax + n2c[it]) * sizeof(T2)) : kBufferNull;
    auto load = [&](int ks, T2 (&an)[NIP], T2 (&bn)[NTL]) {
        const int kc = ks < N4 ? ks : N4 - 1;
#pragma unroll
        for (int j = 0; j < NIP; ++j)
            an[j] = buffer_load_cplx(ra, va, static_cast<unsigned>((kc * R + wv + WAVES * j) * 64 * sizeof(T2)),
                                     static_cast<T2 *>(nullptr));
        const unsigned kb = static_cast<unsigned>(kc * 4 * nmax * sizeof(T2));
#pragma unroll
        for (int it = 0; it < NTL; ++it)
            bn[it] = buffer_load_cplx(rx, vb[it] + kb, 0, static_cast<T2 *>(nullptr));
    };
    auto mma = [&](int ks, const T2 (&a)[NIP], const T2 (&b)[NTL]) {
        // one accumulator after another: consecutive MFMAs never update the
        // same accumulator
        if constexpr (kGaussComplex) {
            T as[NIP], bs[NTL];
#pragma unroll
            for (int j = 0; j < NIP; ++j)
                as[j] = a[j].x + a[j].y;
#pragma unroll
            for (int it = 0; it < NTL; ++it)
                bs[it] = b[it].x + b[it].y;
#pragma unroll
            for (int j = 0; j < NIP; ++j)
#pragma unroll
                for (int it = 0; it < NTL; ++it)
                    acc[j][it][0] = Mfma<T>::mac(a[j].x, b[it].x, acc[j][it][0]);
#pragma unroll
            for (int j = 0; j < NIP; ++j)
#pragma unroll
                for (int it = 0; it < NTL; ++it)
                    acc[j][it][1] = Mfma<T>::mac(a[j].y, b[it].y, acc[j][it][1]);
#pragma unroll
            for (int j = 0; j < NIP; ++j)
#pragma unroll
                for (int it = 0; it < NTL; ++it)
                    acc[j][it][2] = Mfma<T>::mac(as[j], bs[it], acc[j][it][2]);
        } else {
#pragma unroll
            for (int j = 0; j < NIP; ++j)
#pragma unroll
                for (int it = 0; it < NTL; ++it)
                    if (!kTrimTiles || it < tn[j]) {
                        acc[j][it][0] = Mfma<T>::mac(a[j].x, b[it].x, acc[j][it][0]);
                        acc[j][it][1] = Mfma<T>::mac(a[j].x, b[it].y, acc[j][it][1]);
                    }
#pragma unroll
            for (int j = 0; j < NIP; ++j)
#pragma unroll
                for (int it = 0; it < NTL; ++it)
                    if (!kTrimTiles || it < tn[j]) {
                        acc[j][it][0] = Mfma<T>::mac(-a[j].y, b[it].y, acc[j][it][0]);
                        acc[j][it][1] = Mfma<T>::mac(a[j].y, b[it].x, acc[j][it][1]);
                    }
        }
    };
    static_assert(DEPTH == 3, "a three-deep operand ring");
    T2 aC[NIP], bC[NTL];
    {
        load(0, aA, bA);
        load(1, aB, bB);
        int ks = 0;
        for (; ks + 2 < N4; ks += 3) {
            load(ks + 2, aC, bC);
            mma(ks, aA, bA);
            load(ks + 3, aA, bA);
            mma(ks + 1, aB, bB);
            load(ks + 4, aB, bB);
            mma(ks + 2, aC, bC);
        }
        if (ks < N4)
            mma(ks, aA, bA);
        if (ks + 1 < N4)
            mma(ks + 1, aB, bB);
    }
#pragma unroll
    for (int j = 0; j < NIP; ++j)
#pragma unroll
        for (int it = 0; it < NTL; ++it) {
            const size_t slot = static_cast<size_t>(wave + WAVES * j) * tstride + tile0 + it;
            T *dst = Qt + 2 * ((slot * 16 + lo) * 16);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = Mfma<T>::row(i, hi);
                if constexpr (kGaussComplex) {
                    const T p1 = acc[j][it][0][i], p2 = acc[j][it][1][i];
                    dst[2 * r] = p1 - p2;
                    dst[2 * r + 1] = acc[j][it][2][i] - p1 - p2;
                } else {
                    dst[2 * r] = acc[j][it][0][i];
                    dst[2 * r + 1] = acc[j][it][1][i];
                }
            }
        }
}

// Host check of the 32-bit buffer offsets table_gemm_wave uses: the
// (target tile, dm) slice of the table and x must stay below 2^31 bytes.
inline bool table_gemm_extent_ok(int N, int nmax, int lmax, size_t elem_bytes) {
    const size_t N4 = (static_cast<size_t>(N) + 3) / 4;
    const size_t slice = static_cast<size_t>(2 * lmax + 1) * N4 * 64 * elem_bytes;
    const size_t xb = static_cast<size_t>(N) * nmax * elem_bytes;
    return slice < (size_t{1} << 31) && xb < (size_t{1} << 31);
}

// Runs the compile-time instance for a wave's (nip, ntile), nip <= NIPMAX and
// ntile <= NTLMAX (uniform values: scalar branches).
template <typename T, int NIP, int NTL, int NTLMAX, int WAVES, int DEPTH>
__device__ __forceinline__ void table_gemm_ntl(const typename Cplx<T>::type *__restrict__ tab,
                                               const typename Cplx<T>::type *__restrict__ x, int N, int N4,
                                               int nmax, int R, int wave, int ntile,
                                               const int (&n2c)[kTableMaxColTiles], T *__restrict__ Qt,
                                               int tstride, int tile0, const int *__restrict__ otile) {
    if constexpr (NTL <= NTLMAX) {
        if (ntile == NTL) {
            table_gemm_wave<T, NIP, NTL, WAVES, DEPTH>(tab, x, N, N4, nmax, R, wave, n2c, Qt, tstride, tile0,
                                                       otile);
            return;
        }
        table_gemm_ntl<T, NIP, NTL + 1, NTLMAX, WAVES, DEPTH>(tab, x, N, N4, nmax, R, wave, ntile, n2c, Qt,
                                                              tstride, tile0, otile);
    }
}

template <typename T, int NIP, int NIPMAX, int NTLMAX, int WAVES, int DEPTH>
__device__ __forceinline__ void table_gemm_dispatch(const typename Cplx<T>::type *__restrict__ tab,
                                                    const typename Cplx<T>::type *__restrict__ x, int N,
                                                    int N4, int nmax, int R, int wave, int nip, int ntile,
                                                    const int (&n2c)[kTableMaxColTiles], T *__restrict__ Qt,
                                                    int tstride, int tile0, const int *__restrict__ otile) {
    if constexpr (NIP <= NIPMAX) {
        if (nip == NIP) {
            table_gemm_ntl<T, NIP, 1, NTLMAX, WAVES, DEPTH>(tab, x, N, N4, nmax, R, wave, ntile, n2c, Qt,
                                                            tstride, tile0, otile);
            return;
        }
        table_gemm_dispatch<T, NIP + 1, NIPMAX, NTLMAX, WAVES, DEPTH>(tab, x, N, N4, nmax, R, wave, nip, ntile,
                                                                      n2c, Qt, tstride, tile0, otile);
    }
}

}  // namespace yasfhip
