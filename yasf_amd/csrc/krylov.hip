// Vector kernels of the device GMRES: the Arnoldi step kernels (below),
// dot products, norms, scaling, the solution update and the residual.  Reductions are
// deterministic: a fixed grid writes one partial per block and the last block
// to finish sums the partials in block order.
#include "givens.hpp"
#include "internal.hpp"

namespace yasfhip {

constexpr int kRedThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_down(v, off, 64);
    return v;
}

// Block-reduce up to 3 values, publish the partial and let the last block
// combine all partials into out[0..nv).
template <int NV>
__device__ void finish_reduction(double (&v)[NV], Reducer red, double *out) {
    __shared__ double sh[kRedThreads / 64][NV];
    __shared__ bool last;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double s = wave_sum(v[i]);
        if (lane == 0)
            sh[w][i] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int j = 0; j < kRedThreads / 64; ++j)
                s += sh[j][i];
            red.partials[blockIdx.x * 4 + i] = s;
        }
        __threadfence();
        const unsigned done = atomicAdd(red.counter, 1u);
        last = (done == gridDim.x - 1);
    }
    __syncthreads();
    if (!last)
        return;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i)
        acc[i] = 0.0;
    for (int b = threadIdx.x; b < static_cast<int>(gridDim.x); b += blockDim.x)
#pragma unroll
        for (int i = 0; i < NV; ++i)
            acc[i] += __hip_atomic_load(&red.partials[b * 4 + i], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double s = wave_sum(acc[i]);
        if (lane == 0)
            sh[w][i] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int j = 0; j < kRedThreads / 64; ++j)
                s += sh[j][i];
            out[i] = s;
        }
        *red.counter = 0u;
    }
}

// w -= h_prev * v_prev (when v_prev), then out = [Re, Im] <v_next, w>
// (np.vdot convention: conj(v) . w) and, when want_norm, out[2] = ||w||^2.
__global__ __launch_bounds__(kRedThreads) void mgs_step_kernel(double2 *__restrict__ w,
                                                               const double2 *__restrict__ v_prev,
                                                               const double2 *__restrict__ h_prev,
                                                               const double2 *__restrict__ v_next,
                                                               long n, Reducer red, double *out) {
    double acc[3] = {0.0, 0.0, 0.0};
    const double2 h = v_prev ? *h_prev : make_double2(0.0, 0.0);
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        double2 wi = w[i];
        if (v_prev) {
            const double2 vp = v_prev[i];
            wi.x -= h.x * vp.x - h.y * vp.y;
            wi.y -= h.x * vp.y + h.y * vp.x;
            w[i] = wi;
        }
        if (v_next) {
            const double2 vn = v_next[i];
            acc[0] += vn.x * wi.x + vn.y * wi.y;
            acc[1] += vn.x * wi.y - vn.y * wi.x;
        }
        acc[2] += wi.x * wi.x + wi.y * wi.y;
    }
    finish_reduction<3>(acc, red, out);
}

// y = a * x (real scale), in place allowed.
__global__ void scale_kernel(double2 *__restrict__ y, const double2 *__restrict__ x, double a, long n) {
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        const double2 v = x[i];
        y[i] = make_double2(a * v.x, a * v.y);
    }
}

// x += d
__global__ void add_kernel(double2 *__restrict__ x, const double2 *__restrict__ d, long n) {
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        const double2 a = x[i], b = d[i];
        x[i] = make_double2(a.x + b.x, a.y + b.y);
    }
}

// x += sum_k y_k V[k]  (k < m)
__global__ void update_kernel(double2 *__restrict__ x, const double2 *__restrict__ V,
                              const double2 *__restrict__ y, int m, long n, long ld) {
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < m; ++k) {
            const double2 yk = y[k];
            const double2 v = V[k * ld + i];
            sr += yk.x * v.x - yk.y * v.y;
            si += yk.x * v.y + yk.y * v.x;
        }
        double2 xi = x[i];
        x[i] = make_double2(xi.x + sr, xi.y + si);
    }
}

// r = b - ax; out[0] = ||r||^2.
__global__ __launch_bounds__(kRedThreads) void residual_kernel(double2 *__restrict__ r,
                                                               const double2 *__restrict__ b,
                                                               const double2 *__restrict__ ax, long n,
                                                               Reducer red, double *out) {
    double acc[1] = {0.0};
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        const double2 bi = b[i], ai = ax[i];
        const double2 ri = make_double2(bi.x - ai.x, bi.y - ai.y);
        r[i] = ri;
        acc[0] += ri.x * ri.x + ri.y * ri.y;
    }
    finish_reduction<1>(acc, red, out);
}

// ||x||^2 into out[0]
__global__ __launch_bounds__(kRedThreads) void norm2_kernel(const double2 *__restrict__ x, long n,
                                                            Reducer red, double *out) {
    double acc[1] = {0.0};
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        const double2 v = x[i];
        acc[0] += v.x * v.x + v.y * v.y;
    }
    finish_reduction<1>(acc, red, out);
}

static int grid_for(long n) {
    const long b = (n + kRedThreads - 1) / kRedThreads;
    return static_cast<int>(b < kRedBlocks ? (b > 0 ? b : 1) : kRedBlocks);
}

void launch_mgs_step(double2 *w, const double2 *v_prev, const double2 *h_prev, const double2 *v_next,
                     long n, Reducer red, double *out, hipStream_t s) {
    hipLaunchKernelGGL(mgs_step_kernel, dim3(grid_for(n)), dim3(kRedThreads), 0, s, w, v_prev, h_prev,
                       v_next, n, red, out);
    YH_CHECK(hipGetLastError());
}

void launch_scale(double2 *y, const double2 *x, double a, long n, hipStream_t s) {
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(kRedThreads), 0, s, y, x, a, n);
    YH_CHECK(hipGetLastError());
}

void launch_update(double2 *x, const double2 *V, const double2 *y, int m, long n, long ld,
                   hipStream_t s) {
    hipLaunchKernelGGL(update_kernel, dim3(grid_for(n)), dim3(kRedThreads), 0, s, x, V, y, m, n, ld);
    YH_CHECK(hipGetLastError());
}

void launch_add(double2 *x, const double2 *d, long n, hipStream_t s) {
    hipLaunchKernelGGL(add_kernel, dim3(grid_for(n)), dim3(kRedThreads), 0, s, x, d, n);
    YH_CHECK(hipGetLastError());
}

void launch_residual(double2 *r, const double2 *b, const double2 *ax, long n, Reducer red, double *out,
                     hipStream_t s) {
    hipLaunchKernelGGL(residual_kernel, dim3(grid_for(n)), dim3(kRedThreads), 0, s, r, b, ax, n, red,
                       out);
    YH_CHECK(hipGetLastError());
}

void launch_norm2(const double2 *x, long n, Reducer red, double *out, hipStream_t s) {
    hipLaunchKernelGGL(norm2_kernel, dim3(grid_for(n)), dim3(kRedThreads), 0, s, x, n, red, out);
    YH_CHECK(hipGetLastError());
}

__global__ void to_fp32_kernel(float *__restrict__ dst, const double *__restrict__ src, long count,
                               const int *__restrict__ gate) {
    if (gate && *gate != 1)
        return;
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        dst[i] = static_cast<float>(src[i]);
}

__global__ void to_fp64_kernel(double *__restrict__ dst, const float *__restrict__ src, long count) {
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        dst[i] = static_cast<double>(src[i]);
}

void launch_to_fp32(float *dst, const double *src, long count, hipStream_t s, const int *gate) {
    hipLaunchKernelGGL(to_fp32_kernel, dim3(grid_for(count)), dim3(kRedThreads), 0, s, dst, src, count, gate);
    YH_CHECK(hipGetLastError());
}

// dst[i] = a deterministic pseudo-random value in [-1, 1) (a hash of i):
// timing inputs, since the GEMMs' speed depends on their data (all-zero
// operands draw less power and run faster than real ones).
__global__ void fill_test_kernel(double *__restrict__ dst, long count) {
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        unsigned long long h = static_cast<unsigned long long>(i) * 0x9E3779B97F4A7C15ull;
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
        dst[i] = static_cast<double>(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
}

void launch_fill_test(double *dst, long count, hipStream_t s) {
    hipLaunchKernelGGL(fill_test_kernel, dim3(grid_for(count)), dim3(kRedThreads), 0, s, dst, count);
    YH_CHECK(hipGetLastError());
}

void launch_to_fp64(double *dst, const float *src, long count, hipStream_t s) {
    hipLaunchKernelGGL(to_fp64_kernel, dim3(grid_for(count)), dim3(kRedThreads), 0, s, dst, src, count);
    YH_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ device Arnoldi
//
// Step j of a GMRES cycle of the single-solve driver (w = V[j+1] holds the
// operator applied to V[j]) with two grid-wide reductions and no host round
// trip:
//   arnoldi_dots_kernel     per-block partials of h1_k = <V_k, w> (k <= j)
//                           and of ||w||^2
//   arnoldi_reduce_kernel   one block per coefficient sums the partials in
//                           block order (deterministic)
//   arnoldi_project_kernel  w -= sum_k h1_k V_k, then partials of
//                           h2_k = <V_k, w> and of ||w||^2 again
//   arnoldi_reduce_kernel   the same for h2 and ||w||^2
//   arnoldi_givens_kernel   one block: the Hessenberg column h1 + h2 with
//                           h[j+1] = sqrt(||w||^2 - sum_k |h2_k|^2), the
//                           breakdown test, the Givens update, the residual
//                           estimate and the stop decision
//   arnoldi_finish_kernel   V[j+1] = (w - sum_k h2_k V_k) / h[j+1]
// That is classical Gram-Schmidt with one reorthogonalisation (CGS2) whose
// final norm comes from the Pythagorean identity ||w - V h2||^2 = ||w||^2 -
// ||h2||^2 for orthonormal V, the low-synchronisation form of Swirydowicz,
// Langou, Ananthan, Yang and Thomas (Numer. Linear Algebra Appl. 28, e2343,
// 2021).  <a, b> = sum conj(a) b as np.vdot.  Every kernel returns at once
// unless the cycle's state word is 1, so the host can queue steps ahead of
// the device's stop decision.

namespace {

constexpr int kArnoldiThreads = 256;
constexpr int kArnoldiPerThread = 2;
constexpr int kArnoldiChunk = kArnoldiThreads * kArnoldiPerThread;  // vector elements per block
constexpr int kArnoldiWaves = kArnoldiThreads / 64;
// The project kernel holds 5 (restart + 2) complex128 values in LDS; 800
// keeps that within 64 KiB.
constexpr int kArnoldiMaxRestart = 800;

__device__ __forceinline__ double2 wave_sum2(double2 v) { return make_double2(wave_sum(v.x), wave_sum(v.y)); }

__device__ __forceinline__ bool cycle_active(const ArnoldiState *st) {
    return __builtin_amdgcn_readfirstlane(st->active) == 1;
}

__device__ __forceinline__ void load_chunk(const double2 *__restrict__ w, long e0, long n,
                                           double2 (&wl)[kArnoldiPerThread]) {
#pragma unroll
    for (int i = 0; i < kArnoldiPerThread; ++i) {
        const long e = e0 + i * kArnoldiThreads + threadIdx.x;
        wl[i] = e < n ? w[e] : make_double2(0.0, 0.0);
    }
}

__device__ __forceinline__ void store_chunk(double2 *__restrict__ w, long e0, long n,
                                            const double2 (&wl)[kArnoldiPerThread]) {
#pragma unroll
    for (int i = 0; i < kArnoldiPerThread; ++i) {
        const long e = e0 + i * kArnoldiThreads + threadIdx.x;
        if (e < n)
            w[e] = wl[i];
    }
}

// wl -= sum_k h[k] V_k over this block's chunk (k <= j).
__device__ __forceinline__ void chunk_subtract(const ArnoldiBuffers &a, int j, const double2 *h, long e0,
                                               double2 (&wl)[kArnoldiPerThread]) {
#pragma unroll 4
    for (int k = 0; k <= j; ++k) {
        const double2 hk = h[k];
        const double2 *vk = a.V + static_cast<long>(k) * a.n;
#pragma unroll
        for (int i = 0; i < kArnoldiPerThread; ++i) {
            const long e = e0 + i * kArnoldiThreads + threadIdx.x;
            if (e < a.n) {
                const double2 v = vk[e];
                wl[i].x -= hk.x * v.x - hk.y * v.y;
                wl[i].y -= hk.x * v.y + hk.y * v.x;
            }
        }
    }
}

// conj(V_k) . wl over this thread's elements.
__device__ __forceinline__ double2 thread_dot(const ArnoldiBuffers &a, int k, const double2 (&wl)[kArnoldiPerThread],
                                              long e0) {
    const double2 *vk = a.V + static_cast<long>(k) * a.n;
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < kArnoldiPerThread; ++i) {
        const long e = e0 + i * kArnoldiThreads + threadIdx.x;
        if (e < a.n) {
            const double2 v = vk[e];
            acc.x += v.x * wl[i].x + v.y * wl[i].y;
            acc.y += v.x * wl[i].y - v.y * wl[i].x;
        }
    }
    return acc;
}

// This block's partials of <V_k, wl> (k <= j) to part[k][blk] and of
// ||wl||^2 to part[restart + 1][blk].x; sh: 4 (restart + 2) complex128 of LDS.
__device__ void chunk_dots(const ArnoldiBuffers &a, int j, const double2 (&wl)[kArnoldiPerThread], long e0,
                           double2 *sh, double2 *__restrict__ part) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int R2 = a.restart + 2;
    int k = 0;
    for (; k + 3 <= j; k += 4) {  // four coefficients at a time: their loads in flight together
        double2 acc[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            acc[u] = thread_dot(a, k + u, wl, e0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc[u] = wave_sum2(acc[u]);
            if (lane == 0)
                sh[(k + u) * kArnoldiWaves + wv] = acc[u];
        }
    }
    for (; k <= j; ++k) {
        const double2 acc = wave_sum2(thread_dot(a, k, wl, e0));
        if (lane == 0)
            sh[k * kArnoldiWaves + wv] = acc;
    }
    double nrm = 0.0;
#pragma unroll
    for (int i = 0; i < kArnoldiPerThread; ++i)
        nrm += wl[i].x * wl[i].x + wl[i].y * wl[i].y;
    nrm = wave_sum(nrm);
    if (lane == 0)
        sh[(R2 - 1) * kArnoldiWaves + wv] = make_double2(nrm, 0.0);
    __syncthreads();
    for (int slot = threadIdx.x; slot <= j + 1; slot += kArnoldiThreads) {
        const int kk = slot <= j ? slot : R2 - 1;
        double2 v = make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < kArnoldiWaves; ++i) {
            v.x += sh[kk * kArnoldiWaves + i].x;
            v.y += sh[kk * kArnoldiWaves + i].y;
        }
        part[static_cast<long>(kk) * a.nblk + blockIdx.x] = v;
    }
}

__global__ __launch_bounds__(kArnoldiThreads) void arnoldi_start_kernel(ArnoldiBuffers a,
                                                                       const double *__restrict__ nrm2,
                                                                       double target) {
    const double nr = sqrt(*nrm2), inv = 1.0 / nr;
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < a.n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        const double2 v = a.V[i];
        a.V[i] = make_double2(inv * v.x, inv * v.y);
    }
    if (blockIdx.x == 0) {
        for (int k = threadIdx.x; k <= a.restart; k += blockDim.x)
            a.g[k] = make_double2(k == 0 ? nr : 0.0, 0.0);
        if (threadIdx.x == 0) {
            ArnoldiState s{};
            s.active = 1;
            s.target = target;
            s.inv_norm = 1.0;
            *a.st = s;
        }
    }
}

__global__ __launch_bounds__(kArnoldiThreads) void arnoldi_dots_kernel(ArnoldiBuffers a, int j) {
    if (!cycle_active(a.st))
        return;
    extern __shared__ double2 lds[];
    const long e0 = static_cast<long>(blockIdx.x) * kArnoldiChunk;
    double2 wl[kArnoldiPerThread];
    load_chunk(a.V + static_cast<long>(j + 1) * a.n, e0, a.n, wl);
    chunk_dots(a, j, wl, e0, lds, a.partA);
}

// Block s sums slot s of part over the nblk blocks in a fixed order: slots
// 0..j to h[s], slot j + 1 (the norm row) to *nrm.
__global__ __launch_bounds__(kArnoldiThreads) void arnoldi_reduce_kernel(ArnoldiBuffers a, int j,
                                                                        const double2 *__restrict__ part,
                                                                        double2 *__restrict__ h,
                                                                        double *__restrict__ nrm) {
    if (!cycle_active(a.st))
        return;
    __shared__ double2 sh[kArnoldiWaves];
    const int slot = blockIdx.x;
    const int kk = slot <= j ? slot : a.restart + 1;
    const double2 *row = part + static_cast<long>(kk) * a.nblk;
    double2 v = make_double2(0.0, 0.0);
    for (int b = threadIdx.x; b < a.nblk; b += kArnoldiThreads) {
        const double2 p = row[b];
        v.x += p.x;
        v.y += p.y;
    }
    v = wave_sum2(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        sh[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 t = sh[0];
#pragma unroll
        for (int i = 1; i < kArnoldiWaves; ++i) {
            t.x += sh[i].x;
            t.y += sh[i].y;
        }
        if (slot <= j)
            h[slot] = t;
        else
            *nrm = t.x;
    }
}

__global__ __launch_bounds__(kArnoldiThreads) void arnoldi_project_kernel(ArnoldiBuffers a, int j) {
    if (!cycle_active(a.st))
        return;
    extern __shared__ double2 lds[];
    const int R2 = a.restart + 2;
    double2 *h1 = lds + R2 * kArnoldiWaves;  // after chunk_dots' scratch
    for (int k = threadIdx.x; k <= j; k += kArnoldiThreads)
        h1[k] = a.h1[k];
    __syncthreads();
    const long e0 = static_cast<long>(blockIdx.x) * kArnoldiChunk;
    double2 *w = a.V + static_cast<long>(j + 1) * a.n;
    double2 wl[kArnoldiPerThread];
    load_chunk(w, e0, a.n, wl);
    chunk_subtract(a, j, h1, e0, wl);
    store_chunk(w, e0, a.n, wl);
    chunk_dots(a, j, wl, e0, lds, a.partB);
}

// One block: the column, rotations and right-hand side staged in LDS for the
// serial Givens recurrence of thread 0.
__global__ __launch_bounds__(kArnoldiThreads) void arnoldi_givens_kernel(ArnoldiBuffers a, int j) {
    if (!cycle_active(a.st))
        return;
    extern __shared__ double2 lds[];
    const int R1 = a.restart + 1;
    double2 *col = lds;            // [j + 2]
    double2 *gs = col + R1 + 1;    // [j + 1]
    double2 *g = gs + R1;          // [j + 2]
    double *gc = reinterpret_cast<double *>(g + R1 + 1);  // [j + 1]
    __shared__ double ssq_w[kArnoldiWaves];
    double ssq = 0.0;
    for (int k = threadIdx.x; k <= j; k += kArnoldiThreads) {
        const double2 h1 = a.h1[k], h2 = a.h2[k];
        col[k] = make_double2(h1.x + h2.x, h1.y + h2.y);
        ssq += h2.x * h2.x + h2.y * h2.y;
        if (k < j) {
            gs[k] = a.gs[k];
            gc[k] = a.gc[k];
        }
    }
    for (int k = threadIdx.x; k <= j + 1; k += kArnoldiThreads)
        g[k] = a.g[k];
    ssq = wave_sum(ssq);
    if ((threadIdx.x & 63) == 0)
        ssq_w[threadIdx.x >> 6] = ssq;
    __syncthreads();
    if (threadIdx.x == 0) {
        ArnoldiState s = *a.st;
        double tot = 0.0;
#pragma unroll
        for (int i = 0; i < kArnoldiWaves; ++i)
            tot += ssq_w[i];
        const double hn2 = s.wsq - tot;
        const double hn = hn2 > 0.0 ? sqrt(hn2) : 0.0;
        // SciPy: breakdown when h[j+1, j] <= eps ||w|| (w before orthogonalisation)
        const bool breakdown = hn <= 2.220446049250313e-16 * sqrt(s.h0sq);
        col[j + 1] = make_double2(breakdown ? 0.0 : hn, 0.0);
        s.presid = givens_column(col, j, gc, gs, g);
        a.gc[j] = gc[j];
        a.gs[j] = gs[j];
        a.g[j] = g[j];
        a.g[j + 1] = g[j + 1];
        a.hist[j] = s.presid;
        s.inv_norm = breakdown ? 1.0 : 1.0 / hn;
        s.breakdown = breakdown ? 1 : 0;
        s.j = j + 1;
        if (s.presid <= s.target || s.j == a.restart || breakdown)
            s.active = 0;
        *a.st = s;
    }
    __syncthreads();
    double2 *H = a.H + static_cast<long>(j) * R1;
    for (int k = threadIdx.x; k <= j + 1; k += kArnoldiThreads)
        H[k] = col[k];
}

__global__ __launch_bounds__(kArnoldiThreads) void arnoldi_finish_kernel(ArnoldiBuffers a, int j) {
    if (!cycle_active(a.st))
        return;
    extern __shared__ double2 lds[];
    for (int k = threadIdx.x; k <= j; k += kArnoldiThreads)
        lds[k] = a.h2[k];
    __syncthreads();
    const double inv = a.st->inv_norm;
    const long e0 = static_cast<long>(blockIdx.x) * kArnoldiChunk;
    double2 *w = a.V + static_cast<long>(j + 1) * a.n;
    double2 wl[kArnoldiPerThread];
    load_chunk(w, e0, a.n, wl);
    chunk_subtract(a, j, lds, e0, wl);
#pragma unroll
    for (int i = 0; i < kArnoldiPerThread; ++i)
        wl[i] = make_double2(inv * wl[i].x, inv * wl[i].y);
    store_chunk(w, e0, a.n, wl);
}

// y = H^-1 g over the cycle's steps; one thread.
__global__ void arnoldi_backsolve_kernel(ArnoldiBuffers a) { backsolve_dev(a.H, a.g, a.st->j, a.restart + 1, a.y); }

// x += sum_{k < m} y_k V_k with m the cycle's steps.
__global__ void arnoldi_update_kernel(ArnoldiBuffers a, double2 *__restrict__ x) {
    const int m = __builtin_amdgcn_readfirstlane(a.st->j);
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < a.n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < m; ++k) {
            const double2 yk = a.y[k];
            const double2 v = a.V[static_cast<long>(k) * a.n + i];
            sr += yk.x * v.x - yk.y * v.y;
            si += yk.x * v.y + yk.y * v.x;
        }
        const double2 xi = x[i];
        x[i] = make_double2(xi.x + sr, xi.y + si);
    }
}

}  // namespace

int arnoldi_blocks(long n) { return static_cast<int>((n + kArnoldiChunk - 1) / kArnoldiChunk); }
int arnoldi_max_restart() { return kArnoldiMaxRestart; }

void launch_arnoldi_start(const ArnoldiBuffers &a, const double *nrm2, double target, hipStream_t s) {
    hipLaunchKernelGGL(arnoldi_start_kernel, dim3(a.nblk), dim3(kArnoldiThreads), 0, s, a, nrm2, target);
    YH_CHECK(hipGetLastError());
}

void launch_arnoldi_step(const ArnoldiBuffers &a, int j, hipStream_t s) {
    YH_REQUIRE(a.restart >= 1 && a.restart <= kArnoldiMaxRestart && j >= 0 && j < a.restart,
               "Arnoldi step outside the cycle");
    YH_REQUIRE(a.nblk == arnoldi_blocks(a.n), "Arnoldi partials sized for another vector length");
    const size_t R2 = a.restart + 2;
    const size_t dots_lds = R2 * kArnoldiWaves * sizeof(double2);
    hipLaunchKernelGGL(arnoldi_dots_kernel, dim3(a.nblk), dim3(kArnoldiThreads), dots_lds, s, a, j);
    hipLaunchKernelGGL(arnoldi_reduce_kernel, dim3(j + 2), dim3(kArnoldiThreads), 0, s, a, j, a.partA, a.h1,
                       &a.st->h0sq);
    hipLaunchKernelGGL(arnoldi_project_kernel, dim3(a.nblk), dim3(kArnoldiThreads), dots_lds + R2 * sizeof(double2),
                       s, a, j);
    hipLaunchKernelGGL(arnoldi_reduce_kernel, dim3(j + 2), dim3(kArnoldiThreads), 0, s, a, j, a.partB, a.h2,
                       &a.st->wsq);
    // col, gs, g: 3 R2 complex128, gc: R2 doubles
    hipLaunchKernelGGL(arnoldi_givens_kernel, dim3(1), dim3(kArnoldiThreads),
                       3 * R2 * sizeof(double2) + R2 * sizeof(double), s, a, j);
    hipLaunchKernelGGL(arnoldi_finish_kernel, dim3(a.nblk), dim3(kArnoldiThreads), (j + 1) * sizeof(double2), s, a,
                       j);
    YH_CHECK(hipGetLastError());
}

void launch_arnoldi_end(const ArnoldiBuffers &a, double2 *x, hipStream_t s) {
    hipLaunchKernelGGL(arnoldi_backsolve_kernel, dim3(1), dim3(1), 0, s, a);
    hipLaunchKernelGGL(arnoldi_update_kernel, dim3(grid_for(a.n)), dim3(kRedThreads), 0, s, a, x);
    YH_CHECK(hipGetLastError());
}

}  // namespace yasfhip
