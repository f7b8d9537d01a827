// Vector kernels of the device GMRES: fused modified-Gram-Schmidt steps,
// norms, scaling, the solution update and the residual.  Reductions are
// deterministic: a fixed grid writes one partial per block and the last block
// to finish sums the partials in block order.
#include "internal.hpp"

namespace yasfhip {

constexpr int kRedThreads = 256;

struct Reducer {
    double *partials;   // [kRedBlocks][4]
    unsigned *counter;  // zero between launches
};

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

// Sum of the per-block partials of one launch (nblk rows of 4 doubles, the
// first 3 used) in a fixed order, left in every thread.
__device__ __forceinline__ void sum_partials3(const double *__restrict__ part, int nblk, double (&tot)[3]) {
    __shared__ double sh[kRedThreads / 64][3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double a[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            a[i] += part[b * 4 + i];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double v = wave_sum(a[i]);
        if (lane == 0)
            sh[w][i] = v;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double v = 0.0;
        for (int j = 0; j < kRedThreads / 64; ++j)
            v += sh[j][i];
        tot[i] = v;
    }
    __syncthreads();
}

// One modified-Gram-Schmidt step without a grid-wide reduction inside the
// launch: the previous step's partials (part_prev, nprev blocks) are summed
// by every block at the start, giving h = <v_prev, w> of that step (block 0
// stores it with that step's ||w||^2 to h_out), then w -= h v_prev, and this
// step's partials of <v_next, w> and ||w||^2 go to part_out (one row per
// block).  A chain of these needs no fence or atomic: each launch boundary
// orders the partials (a device-scope release per block would write back
// the L2 of every XCD).
__global__ __launch_bounds__(kRedThreads) void mgs_chain_kernel(double2 *__restrict__ w,
                                                                const double2 *__restrict__ v_prev,
                                                                const double *__restrict__ part_prev, int nprev,
                                                                const double2 *__restrict__ v_next, long n,
                                                                double *__restrict__ part_out,
                                                                double *__restrict__ h_out) {
    double2 h = make_double2(0.0, 0.0);
    if (v_prev) {
        double tot[3];
        sum_partials3(part_prev, nprev, tot);
        h = make_double2(tot[0], tot[1]);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            h_out[0] = tot[0];
            h_out[1] = tot[1];
            h_out[2] = tot[2];
        }
    }
    double acc[3] = {0.0, 0.0, 0.0};
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
    __shared__ double sh[kRedThreads / 64][3];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double v = wave_sum(acc[i]);
        if (lane == 0)
            sh[wv][i] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double v = 0.0;
        for (int j = 0; j < kRedThreads / 64; ++j)
            v += sh[j][threadIdx.x];
        part_out[blockIdx.x * 4 + threadIdx.x] = v;
    }
}

// out[0..3) = the summed partials of a launch (one block).
__global__ __launch_bounds__(kRedThreads) void partials_sum_kernel(const double *__restrict__ part, int nblk,
                                                                   double *__restrict__ out) {
    double tot[3];
    sum_partials3(part, nblk, tot);
    if (threadIdx.x < 3)
        out[threadIdx.x] = tot[threadIdx.x];
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

// Partials buffers of launch_mgs_chain hold kRedBlocks rows of 4 doubles.
void launch_mgs_chain(double2 *w, const double2 *v_prev, const double *part_prev, const double2 *v_next,
                      long n, double *part_out, double *h_out, hipStream_t s) {
    const int g = grid_for(n);
    hipLaunchKernelGGL(mgs_chain_kernel, dim3(g), dim3(kRedThreads), 0, s, w, v_prev, part_prev, g, v_next, n,
                       part_out, h_out);
    YH_CHECK(hipGetLastError());
}

void launch_partials_sum(const double *part, long n, double *out, hipStream_t s) {
    hipLaunchKernelGGL(partials_sum_kernel, dim3(1), dim3(kRedThreads), 0, s, part, grid_for(n), out);
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

__global__ void to_fp32_kernel(float *__restrict__ dst, const double *__restrict__ src, long count) {
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        dst[i] = static_cast<float>(src[i]);
}

__global__ void to_fp64_kernel(double *__restrict__ dst, const float *__restrict__ src, long count) {
    for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        dst[i] = static_cast<double>(src[i]);
}

void launch_to_fp32(float *dst, const double *src, long count, hipStream_t s) {
    hipLaunchKernelGGL(to_fp32_kernel, dim3(grid_for(count)), dim3(kRedThreads), 0, s, dst, src, count);
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

}  // namespace yasfhip
