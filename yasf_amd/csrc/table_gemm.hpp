// The streaming complex GEMM of one wave over a resident pair-factor table
// (layout table[target tile][dm][K-step][order][lane], MFMA A operand), shared
// by the single-session kernel (pairtable.hip) and the batched serving kernel
// (server.hip).
#pragma once

#include "device_common.hpp"

namespace yasfhip {

// Operands of a block's GEMM: tab points at the target tile and dm (N4
// K-steps x R orders of 64-lane fragments); x is the source vector (N x nmax);
// otile[p]: column tiles (from tile 0 of the dm) that order p feeds -- a
// prefix, since the dm's columns are sorted by l2 -- so the C tiles past it
// are skipped (left zero); the job's column tiles start at tile0 and
// cols[16 it + c] is the source mode of column c of its tile it (-1:
// padding).  Byte extents: the table slice and x below 2^31 (buffer
// offsets), checked by the launchers with table_gemm_extent_ok.
template <typename T>
struct TableGemmOperands {
    const typename Cplx<T>::type *tab, *x;
    int N, N4, nmax, R, tile0;
    const int *otile, *cols;
};

// One wave's GEMM over K-steps [k0, k1) for NIP orders (p = ip0 + ipstep j)
// and NTL column tiles, all compile-time, so the MFMAs of a K-step form one
// straight-line block the scheduler can interleave.  The result goes out
// through emit(j, it, i, re, im): element i of the lane's accumulator
// (row Mfma<T>::row(i, lane >> 4), column lane & 15) of C tile (j, it).
template <typename T, int NIP, int NTL, int DEPTH, bool GAUSS, typename Emit>
__device__ __forceinline__ void table_gemm_wave(const TableGemmOperands<T> &op, int ip0, int ipstep, int k0,
                                                int k1, Emit &emit) {
    typedef typename Mfma<T>::acc_t acc_t;
    typedef typename Cplx<T>::type T2;
    const int lane = threadIdx.x & 63, hi = lane >> 4;
    const int R = op.R, nmax = op.nmax;
    int tn[NIP];  // active column tiles of each of the wave's orders (uniform)
#pragma unroll
    for (int j = 0; j < NIP; ++j) {
        const int t = uniform(op.otile[ip0 + ipstep * j]) - op.tile0;
        tn[j] = t < 0 ? 0 : (t > NTL ? NTL : t);
    }
    // GAUSS: the complex product with three real MFMAs (Gauss): P1 = Ar Br,
    // P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); Re = P1 - P2, Im = P3 - P1 - P2 at
    // the end -- a quarter fewer MFMAs for a third accumulator.  Complex128
    // always (measured: 7.04 -> 6.07 ms at 2,000 spheres, lmax 6); complex64
    // only in the widest jobs (8 column tiles: -4% at lmax 6, -6% at lmax 8),
    // where the MFMAs saved outweigh the occupancy the accumulator costs.
    constexpr bool kGaussComplex = GAUSS;
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
    const __amdgpu_buffer_rsrc_t ra = buffer_rsrc(op.tab, static_cast<size_t>(R) * op.N4 * 64 * sizeof(T2));
    const __amdgpu_buffer_rsrc_t rx = buffer_rsrc(op.x, static_cast<size_t>(op.N) * nmax * sizeof(T2));
    const unsigned va = lane * sizeof(T2);
    unsigned vb[NTL];
#pragma unroll
    for (int it = 0; it < NTL; ++it) {
        const int n2 = op.cols[it * 16 + (lane & 15)];
        vb[it] = n2 >= 0 ? static_cast<unsigned>((hi * nmax + n2) * sizeof(T2)) : kBufferNull;
    }
    auto load = [&](int ks, T2 (&an)[NIP], T2 (&bn)[NTL]) {
        const int kc = ks < k1 ? ks : k1 - 1;
#pragma unroll
        for (int j = 0; j < NIP; ++j)
            an[j] = buffer_load_cplx(ra, va, static_cast<unsigned>((kc * R + ip0 + ipstep * j) * 64 * sizeof(T2)),
                                     static_cast<T2 *>(nullptr));
        const unsigned kb = static_cast<unsigned>(kc * 4 * nmax * sizeof(T2));
#pragma unroll
        for (int it = 0; it < NTL; ++it)
            bn[it] = buffer_load_cplx(rx, vb[it] + kb, 0, static_cast<T2 *>(nullptr));
    };
    auto mma = [&](const T2 (&a)[NIP], const T2 (&b)[NTL]) {
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
                    if (!kTrimTiles || it < tn[j])
                        acc[j][it][0] = Mfma<T>::mac(a[j].x, b[it].x, acc[j][it][0]);
#pragma unroll
            for (int j = 0; j < NIP; ++j)
#pragma unroll
                for (int it = 0; it < NTL; ++it)
                    if (!kTrimTiles || it < tn[j])
                        acc[j][it][1] = Mfma<T>::mac(a[j].y, b[it].y, acc[j][it][1]);
#pragma unroll
            for (int j = 0; j < NIP; ++j)
#pragma unroll
                for (int it = 0; it < NTL; ++it)
                    if (!kTrimTiles || it < tn[j])
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
    static_assert(DEPTH == 2 || DEPTH == 3, "a two- or three-deep operand ring");
    T2 aC[NIP], bC[NTL];
    if (DEPTH == 2 && k0 < k1) {
        load(k0, aA, bA);
        int ks = k0;
        for (; ks + 1 < k1; ks += 2) {
            load(ks + 1, aB, bB);
            mma(aA, bA);
            load(ks + 2, aA, bA);
            mma(aB, bB);
        }
        if (ks < k1)
            mma(aA, bA);
    } else if (k0 < k1) {
        load(k0, aA, bA);
        load(k0 + 1, aB, bB);
        int ks = k0;
        for (; ks + 2 < k1; ks += 3) {
            load(ks + 2, aC, bC);
            mma(aA, bA);
            load(ks + 3, aA, bA);
            mma(aB, bB);
            load(ks + 4, aB, bB);
            mma(aC, bC);
        }
        if (ks < k1)
            mma(aA, bA);
        if (ks + 1 < k1)
            mma(aB, bB);
    }
#pragma unroll
    for (int j = 0; j < NIP; ++j)
#pragma unroll
        for (int it = 0; it < NTL; ++it)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if constexpr (kGaussComplex) {
                    const T p1 = acc[j][it][0][i], p2 = acc[j][it][1][i];
                    emit(j, it, i, p1 - p2, acc[j][it][2][i] - p1 - p2);
                } else {
                    emit(j, it, i, acc[j][it][0][i], acc[j][it][1][i]);
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
template <typename T, int NIP, int NTL, int NTLMAX, typename Emit>
__device__ __forceinline__ void table_gemm_ntl(const TableGemmOperands<T> &op, int ip0, int ipstep, int k0,
                                               int k1, int ntile, Emit &emit) {
    if constexpr (NTL <= NTLMAX) {
        if (ntile == NTL) {
            // three operand stages unless the accumulators and operands leave
            // no room (a complex128 wave with 2 x 4 C tiles holds 192
            // accumulator VGPRs; 5 or 6 tiles of one order, 144 and 24 of x)
            constexpr int kDepth = (sizeof(T) == sizeof(double) && (NIP * NTL >= 8 || NTL >= 5)) ? 2 : 3;
            constexpr bool kGauss = sizeof(T) == sizeof(double) || NTLMAX >= 8;
            table_gemm_wave<T, NIP, NTL, kDepth, kGauss>(op, ip0, ipstep, k0, k1, emit);
            return;
        }
        table_gemm_ntl<T, NIP, NTL + 1, NTLMAX>(op, ip0, ipstep, k0, k1, ntile, emit);
    }
}

template <typename T, int NIP, int NIPMAX, int NTLMAX, typename Emit>
__device__ __forceinline__ void table_gemm_dispatch(const TableGemmOperands<T> &op, int ip0, int ipstep, int nip,
                                                    int k0, int k1, int ntile, Emit &emit) {
    if constexpr (NIP <= NIPMAX) {
        if (nip == NIP) {
            table_gemm_ntl<T, NIP, 1, NTLMAX>(op, ip0, ipstep, k0, k1, ntile, emit);
            return;
        }
        table_gemm_dispatch<T, NIP + 1, NIPMAX, NTLMAX>(op, ip0, ipstep, nip, k0, k1, ntile, emit);
    }
}

// Whole orders per wave of table_gemm_block at most (orders < 2 lmax + 2).
template <int WAVES>
constexpr int table_gemm_nipmax(int P) {
    return WAVES >= 4 ? ((P / 4) + WAVES / 4 - 1) / (WAVES / 4) : (P + WAVES - 1) / WAVES;
}

// One (job, target tile) block of WAVES waves: the job's nord orders times
// its ntile column tiles, C tile (p, it) to Q slot p * tstride + tile0 + it
// past Qt.  Waves w and w + 4 share SIMD w % 4, and a block holds its CU
// until its slowest SIMD is done, so the orders are dealt out for equal work
// per SIMD: with nord = 4a + b, orders p < 4a whole to wave p % WAVES (a per
// SIMD); the b left over are cut along K -- on a line of b orders measured in
// quarters, SIMD s takes quarters [s b, s b + b) -- at most two pieces per
// SIMD, run after the whole orders.  The pieces' partial C tiles are summed
// in LDS in SIMD (= K) order, one SIMD per round, so the result does not
// depend on timing.  (Order-granular dealing leaves the SIMDs 80% busy at
// lmax 4, 86% at lmax 6.)  Blocks of fewer than 4 waves deal whole orders.
// SEQ: a wave computes its whole orders one after another (one order's
// accumulators at a time; the x fragments are read again per order, the
// table fragments still once) instead of all at once, each over only the
// column tiles that order feeds (otile): C tiles past that prefix hold no
// term the contraction reads, so they are neither computed nor written.
template <typename T, int WAVES, int NIPMAX, int NTLMAX, bool SEQ = false>
__device__ __forceinline__ void table_gemm_block(const TableGemmOperands<T> &op, int nord, int ntile,
                                                 T *__restrict__ Qt, int tstride) {
    typedef typename Cplx<T>::type T2;
    const int lane = threadIdx.x & 63, lo = lane & 15, hi = lane >> 4;
    const int wave = uniform(threadIdx.x >> 6);
    auto to_q = [&](int ip0, int ipstep) {
        return [=](int j, int it, int i, T re, T im) {
            const size_t slot = static_cast<size_t>(ip0 + ipstep * j) * tstride + op.tile0 + it;
            T *dst = Qt + 2 * ((slot * 16 + lo) * 16 + Mfma<T>::row(i, hi));
            dst[0] = re;
            dst[1] = im;
        };
    };
    if constexpr (WAVES < 4) {
        const int nip = nord > wave ? (nord - wave + WAVES - 1) / WAVES : 0;
        auto emit = to_q(wave, WAVES);
        if (nip > 0)
            table_gemm_dispatch<T, 1, NIPMAX, NTLMAX>(op, wave, WAVES, nip, 0, op.N4, ntile, emit);
    } else {
        const int a = nord / 4, b = nord - 4 * a;
        const int nfull = 4 * a > wave ? (4 * a - wave + WAVES - 1) / WAVES : 0;
        // column tiles order ip feeds within the job
        auto order_tiles_of = [&](int ip) {
            const int t = uniform(op.otile[ip]) - op.tile0;
            return t < 0 ? 0 : (t > ntile ? ntile : t);
        };
        if constexpr (SEQ) {
            for (int j = 0; j < nfull; ++j) {
                const int ip = wave + WAVES * j;
                const int tn = order_tiles_of(ip);
                auto emit = to_q(ip, WAVES);
                if (tn > 0)
                    table_gemm_dispatch<T, 1, 1, NTLMAX>(op, ip, WAVES, 1, 0, op.N4, tn, emit);
            }
        } else {
            auto emit = to_q(wave, WAVES);
            if (nfull > 0)
                table_gemm_dispatch<T, 1, NIPMAX, NTLMAX>(op, wave, WAVES, nfull, 0, op.N4, ntile, emit);
        }
        if (b == 0)
            return;  // block-uniform
        // This SIMD's quarters [q0, q1) of the leftover line: piece 0 in
        // order i0 = q0 / 4, piece 1 (if the range crosses into the next
        // order) at the start of order i0 + 1; piece c covers quarters
        // [u0, u1) of its order's K range.  8 waves: piece 0 to wave simd + 4
        // (the one with fewer whole orders), piece 1 to wave simd; 4 waves:
        // both to the SIMD's one wave.
        const int simd = wave & 3;
        const int q0 = simd * b, q1 = q0 + b, i0 = q0 / 4;
        const bool two = q1 > 4 * i0 + 4;
        auto piece_order = [&](int pc) { return i0 + pc; };
        auto piece_u0 = [&](int pc) { return pc == 0 ? q0 - 4 * i0 : 0; };
        auto piece_u1 = [&](int pc) { return pc == 0 ? min(q1, 4 * i0 + 4) - 4 * i0 : q1 - 4 * (i0 + 1); };
        constexpr int kMine = WAVES == 8 ? 1 : 2;  // pieces per wave at most
        const int pfirst = WAVES == 8 ? (wave >= 4 ? 0 : 1) : 0;
        const int npc = two ? 2 : 1;
        T hold[kMine][NTLMAX][4][2];
#pragma unroll
        for (int c = 0; c < kMine; ++c) {
            const int pc = pfirst + c;
#pragma unroll
            for (int it = 0; it < NTLMAX; ++it)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    hold[c][it][i][0] = hold[c][it][i][1] = T(0);
            if (pc < npc) {
                const int k0 = piece_u0(pc) * op.N4 / 4, k1 = piece_u1(pc) * op.N4 / 4;
                auto emit = [&](int j, int it, int i, T re, T im) {
                    hold[c][it][i][0] = re;
                    hold[c][it][i][1] = im;
                };
                const int ip = 4 * a + piece_order(pc);
                const int tn = SEQ ? order_tiles_of(ip) : ntile;  // tiles past tn stay zero in hold
                if (tn > 0)
                    table_gemm_dispatch<T, 1, 1, NTLMAX>(op, ip, 0, 1, k0, k1, tn, emit);
            }
        }
        __shared__ T2 part[3][NTLMAX][256];
        for (int r = 0; r < 4; ++r) {
            if (simd == r) {
#pragma unroll
                for (int c = 0; c < kMine; ++c) {
                    const int pc = pfirst + c;
                    if (pc < npc) {
                        const bool first = piece_u0(pc) == 0;
                        T2(*dst)[256] = part[piece_order(pc)];
#pragma unroll
                        for (int it = 0; it < NTLMAX; ++it)
                            if (it < ntile)
#pragma unroll
                                for (int i = 0; i < 4; ++i) {
                                    T2 &d = dst[it][lo * 16 + Mfma<T>::row(i, hi)];
                                    d.x = first ? hold[c][it][i][0] : d.x + hold[c][it][i][0];
                                    d.y = first ? hold[c][it][i][1] : d.y + hold[c][it][i][1];
                                }
                    }
                }
            }
            __syncthreads();
        }
        for (int u = wave; u < b * ntile; u += WAVES) {
            const int i = u / ntile, it = u - i * ntile;
            const size_t slot = static_cast<size_t>(4 * a + i) * tstride + op.tile0 + it;
            T2 *q = reinterpret_cast<T2 *>(Qt) + slot * 256;
            for (int e = lane; e < 256; e += 64)
                q[e] = part[i][it][e];
        }
    }
}

}  // namespace yasfhip
