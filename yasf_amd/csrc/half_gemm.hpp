// The pair GEMM on the fp16 matrix cores (v_mfma_f32_16x16x32_f16, FP32
// accumulation), shared by the session kernels (half.hip) and the batched
// server (server.hip).
//
// Each operand value v is carried as v ~ (hi + lo) 2^-e with hi = fp16(v 2^e)
// and lo = fp16(v 2^e - hi): a 22-bit significand, so the product of two such
// sums (hi hi + hi lo + lo hi; lo lo is below the accumulator's rounding)
// keeps the accuracy of the complex64 GEMM while the table stream and the
// MFMA rate are those of fp16.  The one-term form (hi only) is offered too
// (relative error ~1e-4).  Scales are powers of two, so scaling is exact:
//   A (pair factors): per order p, 2^e_p with e_p from a bound of
//     |P_p^m| |h_p(k d)| <= sqrt((2p+1)/2) |h_p(k d_min)| (|h_p| falls with
//     its argument; the normalised Legendre functions obey sum_m |Y_p^m|^2 =
//     (2p+1)/4pi);
//   B (source coefficients): per mode column n2, 2^f from the column's
//     largest |Re|, |Im|;
// so scaled magnitudes stay below 2^kHalfTop and fp16 keeps them in its
// normal range over 2^-14..2^16.  The GEMM's C tile is scaled back by
// 2^-(e_p + f_n2) in FP32 and stored as complex64, like the complex64 GEMM's,
// so the ab5 contraction that follows is unchanged.
//
// Table layout (fp16, two arrays for the hi and the lo terms, lo optional):
//   frag(t, d, ks, ip) = (t * (2L+1)^2 + dmoff[d]) * N32 + ks * R + ip
//   half index         = (frag * 2 + plane) * 512 + lane * 8 + i
// plane 0 = Re, 1 = Im; ks = K-step of 32 sources; lane = (source % 32 / 8) *
// 16 + target % 16 and i = source % 8: the A-operand layout of
// v_mfma_f32_16x16x32_f16 (lane l holds A[l % 16][8 (l / 16) + i]).
// Source vectors: xh[term][plane][n2 (0..nmax, row nmax zero)][N32 * 32]
// halves, so a lane's 8 sources of one column are 16 contiguous bytes.
#pragma once

#include "device_common.hpp"

namespace yasfhip {

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));

constexpr int kHalfTop = 14;  // scaled magnitudes stay below 2^kHalfTop

// Column tiles of an fp16 GEMM block's B stage (the widest bins).
constexpr int kHalfJobTiles = kHalfBinTiles;

// Exponent e with bound * 2^e < 2^kHalfTop (0 for a zero or non-finite bound).
__host__ __device__ inline int half_exponent(double bound) {
    if (!(bound > 0.0) || !(bound < HUGE_VAL))
        return 0;
    int e = 0;
    frexp(bound, &e);  // bound < 2^e
    return kHalfTop - e;
}

// Halves per (term, plane) of a source-vector copy: rows 0..nmax, N32 * 32 sources.
__host__ __device__ inline size_t half_x_plane(int nmax, int N32) {
    return static_cast<size_t>(nmax + 1) * N32 * 32;
}

// Halves of one target tile of the table, one term.
__host__ __device__ inline size_t half_tile_halves(int lmax, int N32) {
    const size_t P = 2 * lmax + 1;
    return P * P * N32 * 2 * 512;
}

__device__ __forceinline__ f4_t mfma_h16(h8_t a, h8_t b, f4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

typedef unsigned u4_t __attribute__((ext_vector_type(4)));

// -a by flipping the sign bits (four 32-bit XORs)
__device__ __forceinline__ h8_t neg_h8(h8_t a) {
    return __builtin_bit_cast(h8_t, __builtin_bit_cast(u4_t, a) ^ u4_t{0x80008000u, 0x80008000u, 0x80008000u,
                                                                        0x80008000u});
}

// The source copies and outputs a block works on: one request's, or those
// of several requests over one operator (Src::count() of them, each with the
// bin's ntile column tiles).  x(i): request i's copy (its hi term's plane 0,
// the lo term 2 xplane halves further); exps(i): its column exponents;
// out(i): its Q at the target tile (complex64 slots); live(i): false for a
// request whose copy is not current (its tiles read zeros and are not
// stored).  Every i is wave-uniform where these are called.
struct HalfOneSource {
    static constexpr bool kMulti = false;  // one copy: one buffer descriptor for all B pieces
    const _Float16 *xh;
    const int *xexp;
    float *Qt;
    __device__ int count() const { return 1; }
    __device__ bool live(int) const { return true; }
    __device__ const _Float16 *x(int) const { return xh; }
    __device__ const int *exps(int) const { return xexp; }
    __device__ float *out(int) const { return Qt; }
};

// One (bin, target tile) block of WAVES = kHalfBinWaves waves.  hi / lo
// point at the target tile of the two terms' tables (lo unused unless
// SPLIT), laid out as above; src: the source copies (plane 1 = Im follows
// plane 0 at xplane halves), whose column tiles follow one another (request
// i's bin tile k is block tile i ntile + k; src.count() ntile <= NTL); bin:
// the block's jobs (see HalfBin); colmap, dmoff: the plan's; pexp: the order
// exponents.  Wave w computes its orders wip over its bin tiles [wtb, wte)
// for every request of src; C tile (order, k) of request i goes to Q slot
// wq + k past out(i) (complex64).  Each table fragment is loaded once and
// used for every request of src.
//
// Pipelining: the table stream (the operand that comes from HBM) and the B
// fragments of a K-step (the column tiles of the bin's requests, shared by
// all waves, loaded by the whole block and stored to one of two LDS buffers
// a K-step ahead) each have three register sets and two K-steps to arrive.
// (Table fragments three K-steps ahead with the x pieces one step in flight
// were 5-10% slower.)  Operands come through raw buffer loads: a wave's
// order slot, when it has none, and the B pieces of absent column tiles read
// out of range, which returns zero without touching memory, so every load is
// unconditional and the compiler's waits stay static.  One barrier per
// K-step hands the LDS buffers over.
template <int WAVES, int NTL, bool SPLIT, class Src>
__device__ __forceinline__ void half_gemm_block(const _Float16 *__restrict__ hi_tab, const _Float16 *__restrict__ lo_tab,
                                                const Src &src, size_t xplane, int N32, int nmax, int lmax,
                                                const HalfBin *__restrict__ bin, const int *__restrict__ colmap,
                                                const int *__restrict__ dmoff, const int *__restrict__ pexp) {
    static_assert(WAVES == kHalfBinWaves && NTL == kHalfBinTiles, "fp16 GEMM blocks are bins");
    constexpr int NIP = SPLIT ? 1 : 2;  // orders per wave
    constexpr int NPL = SPLIT ? 4 : 2;  // planes: hi Re, hi Im, lo Re, lo Im
    constexpr int THREADS = WAVES * 64;
    constexpr int CHUNKS = NTL * NPL * 64;  // 16-byte B pieces of one K-step
    constexpr int PER = (CHUNKS + THREADS - 1) / THREADS;
    __shared__ h8_t bsh[2][NTL][NPL][64];
    const int tid = threadIdx.x;
    const int lane = tid & 63, lo = lane & 15, hi = lane >> 4;
    const int wave = uniform(tid >> 6);
    const size_t S32 = static_cast<size_t>(N32) * 32;
    const int nreq = src.count();
    const int ntile = uniform(bin->ntile);
    const int ntt = nreq * ntile;  // column tiles of the block
    // this wave's dm, orders and tiles: bit it of mine is set when block
    // tile it (request it / ntile, bin tile it % ntile) is the wave's
    const int d = uniform(bin->wd[wave]);
    const int adm = d < 0 ? 0 : abs(d - 2 * lmax);
    const int R = 2 * lmax + 1 - adm;
    const int tb = uniform(bin->wtb[wave]), te = uniform(bin->wte[wave]);
    int wip[NIP];
#pragma unroll
    for (int j = 0; j < NIP; ++j)
        wip[j] = d < 0 ? -1 : uniform(bin->wip[wave][j]);
    unsigned mine = 0;
    if (d >= 0)
        for (int it = 0; it < ntt; ++it) {
            const int k = it % ntile;
            if (k >= tb && k < te)
                mine |= 1u << it;
        }
    mine = uniform(mine);
    // B pieces of this thread: the copy they come from (a wave's 64 pieces
    // of one u share a column tile, hence a request), byte offset (K-step 0)
    // into that copy (4 planes) and LDS slot
    constexpr int NRX = Src::kMulti ? PER : 1;
    __amdgpu_buffer_rsrc_t rx[NRX];
    unsigned boff[PER];
    int bslot[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int q = tid + u * THREADS;
        const int itw = (wave * 64 + u * THREADS) / (NPL * 64);  // = q / (NPL * 64), wave-uniform
        const int rq = min(itw / ntile, nreq - 1), k = itw - rq * ntile;
        if (Src::kMulti || u == 0)
            rx[Src::kMulti ? u : 0] = buffer_rsrc(src.x(rq), 4 * xplane * sizeof(_Float16));
        const int pl = (q / 64) % NPL, ln = q % 64;
        const int tc = (q < CHUNKS && itw < ntt && src.live(rq)) ? uniform(bin->tcol[k]) : -1;
        const int n2 = tc >= 0 ? colmap[tc + (ln & 15)] : -1;
        const size_t plane = (pl < 2 ? 0 : 2 * xplane) + (pl & 1) * xplane;
        boff[u] = n2 < 0 ? kBufferNull
                         : static_cast<unsigned>((plane + static_cast<size_t>(n2) * S32 + (ln >> 4) * 8) *
                                                 sizeof(_Float16));
        bslot[u] = q < CHUNKS && itw < ntt ? q : -1;  // LDS tiles past the block's are never read
    }
    auto load_b = [&](int ks, h8_t (&breg)[PER]) {
        const unsigned so = static_cast<unsigned>(ks) * 32 * sizeof(_Float16);
#pragma unroll
        for (int u = 0; u < PER; ++u)
            breg[u] = __builtin_bit_cast(
                h8_t, __builtin_amdgcn_raw_buffer_load_b128(rx[Src::kMulti ? u : 0], boff[u], so, 0));
    };
    auto store_b = [&](int buf, const h8_t (&breg)[PER]) {
#pragma unroll
        for (int u = 0; u < PER; ++u)
            if (bslot[u] >= 0)
                (&bsh[buf][0][0][0])[bslot[u]] = breg[u];
    };
    // the wave's dm slice of each term (an idle wave: an empty range)
    const size_t f0 = d < 0 ? 0 : static_cast<size_t>(dmoff[d]) * N32 * 1024;
    const size_t slice = d < 0 ? 0 : static_cast<size_t>(N32) * R * 1024 * sizeof(_Float16);
    const __amdgpu_buffer_rsrc_t rh = buffer_rsrc(hi_tab + f0, slice);
    const __amdgpu_buffer_rsrc_t rl = buffer_rsrc((SPLIT ? lo_tab : hi_tab) + f0, slice);
    const unsigned kstride = static_cast<unsigned>(R) * 1024 * sizeof(_Float16);  // bytes per K-step
    auto kc = [&](int k) { return k < N32 ? k : N32 - 1; };
    unsigned aoff[NIP];
#pragma unroll
    for (int j = 0; j < NIP; ++j)
        aoff[j] = wip[j] >= 0 ? static_cast<unsigned>((wip[j] * 1024 + lane * 8) * sizeof(_Float16)) : kBufferNull;
    auto load_a = [&](int ks, h8_t (&a)[NIP][NPL]) {
        const unsigned so = static_cast<unsigned>(ks) * kstride;
#pragma unroll
        for (int j = 0; j < NIP; ++j) {
            a[j][0] = __builtin_bit_cast(h8_t, __builtin_amdgcn_raw_buffer_load_b128(rh, aoff[j], so, 0));
            a[j][1] = __builtin_bit_cast(h8_t, __builtin_amdgcn_raw_buffer_load_b128(rh, aoff[j] + 1024, so, 0));
            if constexpr (SPLIT) {
                a[j][2] = __builtin_bit_cast(h8_t, __builtin_amdgcn_raw_buffer_load_b128(rl, aoff[j], so, 0));
                a[j][3] = __builtin_bit_cast(h8_t, __builtin_amdgcn_raw_buffer_load_b128(rl, aoff[j] + 1024, so, 0));
            }
        }
    };
    f4_t acc[NIP][NTL][2];
#pragma unroll
    for (int j = 0; j < NIP; ++j)
#pragma unroll
        for (int it = 0; it < NTL; ++it)
            acc[j][it][0] = acc[j][it][1] = f4_t{0.f, 0.f, 0.f, 0.f};
    // MFMAs of one K-step from the A fragments a and buffer buf of the B
    // fragments (Im planes negated per column tile, for the real parts)
    auto step = [&](const h8_t (&a)[NIP][NPL], int buf) {
#pragma unroll
        for (int it = 0; it < NTL; ++it) {
            if (!((mine >> it) & 1u))
                continue;
            h8_t b[NPL], nb[NPL / 2];
#pragma unroll
            for (int pl = 0; pl < NPL; ++pl)
                b[pl] = bsh[buf][it][pl][lane];
            nb[0] = neg_h8(b[1]);
            if constexpr (SPLIT)
                nb[1] = neg_h8(b[3]);
#pragma unroll
            for (int j = 0; j < NIP; ++j) {
                if (wip[j] < 0)
                    continue;
                f4_t cr = acc[j][it][0], ci = acc[j][it][1];
                cr = mfma_h16(a[j][0], b[0], cr);
                ci = mfma_h16(a[j][0], b[1], ci);
                cr = mfma_h16(a[j][1], nb[0], cr);
                ci = mfma_h16(a[j][1], b[0], ci);
                if constexpr (SPLIT) {
                    cr = mfma_h16(a[j][0], b[2], cr);
                    ci = mfma_h16(a[j][0], b[3], ci);
                    cr = mfma_h16(a[j][1], nb[1], cr);
                    ci = mfma_h16(a[j][1], b[2], ci);
                    cr = mfma_h16(a[j][2], b[0], cr);
                    ci = mfma_h16(a[j][2], b[1], ci);
                    cr = mfma_h16(a[j][3], nb[0], cr);
                    ci = mfma_h16(a[j][3], b[0], ci);
                }
                acc[j][it][0] = cr;
                acc[j][it][1] = ci;
            }
        }
    };
    // K-step ks computes with A set ks % 3 and LDS buffer ks % 2; it loads
    // the table fragments of ks + 2 (set (ks + 2) % 3) before its MFMAs, then
    // stores the x pieces of ks + 1 (B set (ks + 1) % 3) to LDS and reloads
    // that set with ks + 4.  The vector-memory counter completes in order, so
    // waiting for A of ks forces the x pieces of ks + 1 (issued before it):
    // both have two K-steps to arrive.  K-steps past the last reload the last
    // one (the loads stay unconditional).
    h8_t a0[NIP][NPL], a1[NIP][NPL], a2[NIP][NPL];
    h8_t b0[PER], b1[PER], b2[PER];
    load_b(0, b0);
    store_b(0, b0);
    load_b(kc(1), b1);
    load_b(kc(2), b2);
    load_a(0, a0);
    load_a(kc(1), a1);
    load_b(kc(3), b0);
    __syncthreads();
    auto kstep = [&](int ks, const h8_t (&acur)[NIP][NPL], h8_t (&anext)[NIP][NPL], h8_t (&bnext)[PER]) {
        load_a(kc(ks + 2), anext);
        step(acur, ks & 1);
        store_b((ks + 1) & 1, bnext);
        load_b(kc(ks + 4), bnext);
        __syncthreads();
    };
    for (int ks = 0; ks < N32; ks += 3) {
        kstep(ks, a0, a2, b1);
        if (ks + 1 >= N32)
            break;
        kstep(ks + 1, a1, a0, b2);
        if (ks + 2 >= N32)
            break;
        kstep(ks + 2, a2, a1, b0);
    }
#pragma unroll
    for (int j = 0; j < NIP; ++j) {
        if (wip[j] < 0)
            continue;
        const int wq = uniform(bin->wq[wave][j]);
        const int eo = pexp[adm + wip[j]];
#pragma unroll
        for (int it = 0; it < NTL; ++it)
            if ((mine >> it) & 1u) {
                const int rq = it / ntile, k = it - rq * ntile;
                if (!src.live(rq))
                    continue;
                const int n2 = colmap[uniform(bin->tcol[k]) + lo];
                const int e = eo + src.exps(rq)[n2 < 0 ? nmax : n2];
                const size_t qs = static_cast<size_t>(wq + k);
                float *dst = src.out(rq) + 2 * ((qs * 16 + lo) * 16 + 4 * hi);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    dst[2 * i] = ldexpf(acc[j][it][0][i], -e);
                    dst[2 * i + 1] = ldexpf(acc[j][it][1][i], -e);
                }
            }
    }
}

// Source-vector copy for the fp16 GEMM, step 1: per column n2 the largest
// |Re| or |Im| over the N sources (as float bits, atomic max into colmax,
// which the caller zeroes).  One block per 256 rows.
__device__ __forceinline__ void half_colmax_block(const double2 *__restrict__ x, int N, int nmax,
                                                  unsigned *__restrict__ colmax, int block) {
    __shared__ unsigned sm[2 * kMaxLmax * (kMaxLmax + 2)];
    for (int c = threadIdx.x; c < nmax; c += blockDim.x)
        sm[c] = 0u;
    __syncthreads();
    const long e0 = static_cast<long>(block) * 256 * nmax;
    const long e1 = min(static_cast<long>(N) * nmax, e0 + 256L * nmax);
    for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const double2 v = x[e];
        const float m = static_cast<float>(fmax(fabs(v.x), fabs(v.y)));
        atomicMax(&sm[e % nmax], __float_as_uint(m));
    }
    __syncthreads();
    for (int c = threadIdx.x; c < nmax; c += blockDim.x)
        if (sm[c])
            atomicMax(&colmax[c], sm[c]);
}

// Step 2: xh terms of x 2^f per column (f = half_exponent of the column's
// max), rows n2 = 0..nmax (row nmax and sources past N zero); xexp[n2] = f.
// One thread per (row, source), sources fastest.
template <bool SPLIT>
__device__ __forceinline__ void half_convert_elem(const double2 *__restrict__ x, int N, int N32, int nmax,
                                                  const unsigned *__restrict__ colmax, int *__restrict__ xexp,
                                                  _Float16 *__restrict__ xhi, _Float16 *__restrict__ xlo,
                                                  long idx) {
    const long S32 = static_cast<long>(N32) * 32;
    const int n2 = static_cast<int>(idx / S32);
    const long s = idx - n2 * S32;
    if (n2 > nmax)
        return;
    const int f = n2 < nmax ? half_exponent(__uint_as_float(colmax[n2])) : 0;
    if (s == 0)
        xexp[n2] = f;
    double re = 0.0, im = 0.0;
    if (n2 < nmax && s < N) {
        const double2 v = x[s * nmax + n2];
        re = ldexp(v.x, f);
        im = ldexp(v.y, f);
    }
    const size_t plane = half_x_plane(nmax, N32);
    const size_t o = static_cast<size_t>(n2) * S32 + s;
    const _Float16 hr = static_cast<_Float16>(re), hm = static_cast<_Float16>(im);
    xhi[o] = hr;
    xhi[plane + o] = hm;
    if constexpr (SPLIT) {
        xlo[o] = static_cast<_Float16>(re - static_cast<double>(hr));
        xlo[plane + o] = static_cast<_Float16>(im - static_cast<double>(hm));
    }
}

// Order exponents e_p (p < P) of a cluster whose closest pair is
// sqrt(dmin2) apart: |G_p| <= sqrt((2p+1)/2) |h_p(k d_min)|.
__device__ __forceinline__ int half_order_exponent(int p, double k, double dmin2) {
    const double x = k * sqrt(dmin2);
    if (!(x > 0.0))
        return 0;
    double sx, cx;
    sincos(x, &sx, &cx);
    const double inv = 1.0 / x;
    double ar = sx * inv, ai = -cx * inv;                      // h_0
    double hr = (ar - cx) * inv, hi = (ai - sx) * inv;         // h_1
    for (int q = 0; q < p; ++q) {
        const double f = (2 * q + 3) * inv;
        const double nr = f * hr - ar, ni = f * hi - ai;
        ar = hr;
        ai = hi;
        hr = nr;
        hi = ni;
    }
    return half_exponent(sqrt((2.0 * p + 1.0) / 2.0) * sqrt(ar * ar + ai * ai));
}

}  // namespace yasfhip
