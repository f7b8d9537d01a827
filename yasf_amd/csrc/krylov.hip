// Vector kernels of the device GMRES: fused modified-Gram-Schmidt steps,
// norms, scaling, the solution update and the residual.  Reductions are
// deterministic: a fixed grid writes one partial per block and the last block
// to finish sums the partials in block order.
#include "internal.hpp"

namespace yasfhip {

constexpr int kRedBlocks = 512;
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
