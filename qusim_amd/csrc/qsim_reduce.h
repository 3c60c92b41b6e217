// qsim_reduce.h -- grid-stride maps and compensated (double-double)
// reductions shared by the Device ops and the solver jobs.  A reduction's
// partial sums live in the per-device scratch block (no allocation per
// call); the result is copied to the host, which synchronises the stream.
#pragma once

#include <cstring>

#include "qsim_common.h"

namespace qsim {
namespace red {

constexpr int kBlock = 256;

template <typename F>
__global__ void __launch_bounds__(kBlock) k_map(size_t n, F f)
{
    size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) f(i);
}

inline size_t grid_for(size_t n)
{
    size_t blocks = (n + kBlock - 1) / kBlock;
    size_t maxb = (size_t)device_cu_count() * 16;
    return blocks < maxb ? (blocks ? blocks : 1) : maxb;
}

template <typename F>
int launch_map(size_t n, qsim_stream_t s, F f)
{
    if (n == 0) return QSIM_OK;
    hipLaunchKernelGGL(k_map<F>, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), n, f);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

// ------------------------------------------------------------ double-double
struct dd {
    double hi, lo;
};

__device__ __forceinline__ dd dd_add(dd a, double b)
{
    double s = a.hi + b;
    double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (b - bb) + a.lo;
    double hi = s + e;
    return dd{hi, e - (hi - s)};
}

__device__ __forceinline__ dd dd_add(dd a, dd b)
{
    double s = a.hi + b.hi;
    double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (b.hi - bb) + (a.lo + b.lo);
    double hi = s + e;
    return dd{hi, e - (hi - s)};
}

// Sum reduction of K values per element: f(i, v) fills v[0..K).
template <int K, typename F>
__global__ void __launch_bounds__(kBlock) k_reduce_sum(size_t n, F f, dd *partial)
{
    dd acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = dd{0.0, 0.0};
    size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        double v[K];
        f(i, v);
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = dd_add(acc[k], v[k]);
    }
    __shared__ dd sh[K][kBlock];
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < K; ++k)
                sh[k][threadIdx.x] = dd_add(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) partial[blockIdx.x * K + k] = sh[k][0];
    }
}

template <int K>
__global__ void __launch_bounds__(kBlock) k_reduce_sum_final(int nb, const dd *partial,
                                                             double *out)
{
    __shared__ dd sh[K][kBlock];
    dd acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = dd{0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += kBlock) {
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = dd_add(acc[k], partial[b * K + k]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < K; ++k)
                sh[k][threadIdx.x] = dd_add(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = sh[k][0].hi + sh[k][0].lo;
    }
}

template <int K, typename F>
int reduce_sum(size_t n, qsim_stream_t s, double *host_out, F f)
{
    size_t nb = grid_for(n);
    Scratch sc;
    QSIM_TRY(scratch_acquire(nb * K * sizeof(dd) + 64, &sc));
    dd *partial = reinterpret_cast<dd *>(sc.dev);
    double *res = reinterpret_cast<double *>(partial + nb * K);
    hipStream_t st = as_stream(s);
    hipLaunchKernelGGL((k_reduce_sum<K, F>), dim3(nb), dim3(kBlock), 0, st, n, f, partial);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_reduce_sum_final<K>, dim3(1), dim3(kBlock), 0, st, (int)nb,
                           partial, res);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(sc.host, res, K * sizeof(double),
                                            hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) std::memcpy(host_out, sc.host, K * sizeof(double));
    scratch_release();
    if (e != hipSuccess) return hip_error(e, "reduction");
    return QSIM_OK;
}

// max (MAX=true) or min of f(i) over i < n
template <bool MAX, typename F>
__global__ void __launch_bounds__(kBlock) k_reduce_ext(size_t n, F f, double init,
                                                       double *partial)
{
    double acc = init;
    size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        double v = f(i);
        acc = MAX ? (acc >= v ? acc : v) : (acc < v ? acc : v);
    }
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = sh[threadIdx.x], b = sh[threadIdx.x + w];
            sh[threadIdx.x] = MAX ? (a >= b ? a : b) : (a < b ? a : b);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

template <bool MAX>
__global__ void __launch_bounds__(kBlock) k_reduce_ext_final(int nb, const double *partial,
                                                             double init, double *out)
{
    __shared__ double sh[kBlock];
    double acc = init;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
        double v = partial[b];
        acc = MAX ? (acc >= v ? acc : v) : (acc < v ? acc : v);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = sh[threadIdx.x], b = sh[threadIdx.x + w];
            sh[threadIdx.x] = MAX ? (a >= b ? a : b) : (a < b ? a : b);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

template <bool MAX, typename F>
int reduce_ext(size_t n, qsim_stream_t s, double init, double *host_out, F f)
{
    size_t nb = grid_for(n);
    Scratch sc;
    QSIM_TRY(scratch_acquire(nb * sizeof(double) + 64, &sc));
    double *partial = reinterpret_cast<double *>(sc.dev);
    double *res = partial + nb;
    hipStream_t st = as_stream(s);
    hipLaunchKernelGGL((k_reduce_ext<MAX, F>), dim3(nb), dim3(kBlock), 0, st, n, f, init,
                       partial);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_reduce_ext_final<MAX>, dim3(1), dim3(kBlock), 0, st, (int)nb,
                           partial, init, res);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(sc.host, res, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) std::memcpy(host_out, sc.host, sizeof(double));
    scratch_release();
    if (e != hipSuccess) return hip_error(e, "reduction");
    return QSIM_OK;
}


}  // namespace red
}  // namespace qsim
