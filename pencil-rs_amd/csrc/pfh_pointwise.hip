// Device pointwise operations on pencil arrays: the pieces a pseudo-spectral
// step needs between transforms. They follow the reference's pointwise and
// grid helpers (crates/pencil-array/src/pointwise.rs:40-598, a closure over
// co-located elements; grid.rs:20-120, per-axis coordinate arrays seen in the
// pencil's memory order) and its l2 reduction (collectives.rs:519-560), as
// memory-bound element kernels.
#include <hip/hip_runtime.h>

#include "pfh_kernels.hpp"

namespace pfh {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int64_t grid_stride() { return (int64_t)gridDim.x * blockDim.x; }
__device__ __forceinline__ int64_t global_id() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

unsigned blocks_for(int64_t n)
{
    const int64_t b = (n + kBlock - 1) / kBlock;
    return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

// data[e][memory index] *= prod_a f_a[i_a]: the memory index is split into
// per-position indices (slowest first), position p holding logical axis
// perm[p]; f holds the per-axis factors concatenated in logical axis order
// (complex factors: (re, im) pairs, e.g. i k_a for a spectral derivative).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale_axes(ScaleAxesArgs s, T *data, const T *f)
{
    const int64_t total = s.len * s.extra;
    for (int64_t i = global_id(); i < total; i += grid_stride()) {
        int64_t rem = i % s.len;
        T wr = (T)1, wi = (T)0;
        for (int p = s.nd - 1; p >= 0; --p) {
            const int64_t idx = rem % s.mext[p];
            rem /= s.mext[p];
            if (s.complex_factors) {
                const T fr = f[2 * (s.foff[p] + idx)], fi = f[2 * (s.foff[p] + idx) + 1];
                const T r = wr * fr - wi * fi;
                wi = wr * fi + wi * fr;
                wr = r;
            } else {
                wr *= f[s.foff[p] + idx];
            }
        }
        if (s.complex_factors) {
            const T x = data[2 * i], y = data[2 * i + 1];
            data[2 * i] = x * wr - y * wi;
            data[2 * i + 1] = x * wi + y * wr;
        } else if (s.complex_data) {
            data[2 * i] *= wr;
            data[2 * i + 1] *= wr;
        } else {
            data[i] *= wr;
        }
    }
}

// u = w + a w^2 (a pointwise quadratic step)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_quadratic(const T *w, T *u, int64_t n, T a)
{
    for (int64_t i = global_id(); i < n; i += grid_stride()) {
        const T v = w[i];
        u[i] = v + a * v * v;
    }
}

// out = g (1 + 2 a w): the transpose of the quadratic step's derivative
template <typename T>
__global__ __launch_bounds__(kBlock) void k_quadratic_adjoint(const T *w, const T *g, T *out, int64_t n, T a)
{
    for (int64_t i = global_id(); i < n; i += grid_stride()) out[i] = g[i] * ((T)1 + (T)2 * a * w[i]);
}

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r += sh[k];
    return r;
}

// Deterministic sum of squares: a fixed grid of partial sums (each block a
// fixed chunk), then one block adds the partials in order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sumsq_partial(const T *x, int64_t n, double *partial)
{
    __shared__ double sh[kBlock / 64];
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b = (int64_t)blockIdx.x * chunk;
    const int64_t e = b + chunk < n ? b + chunk : n;
    double acc = 0.0;
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const double v = (double)x[i];
        acc += v * v;
    }
    const double r = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(kBlock) void k_sumsq_final(const double *partial, int nparts, double *result)
{
    __shared__ double sh[kBlock / 64];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc += partial[i];
    const double r = block_sum(acc, sh);
    if (threadIdx.x == 0) *result = r;
}

} // namespace

template <typename T>
hipError_t launch_scale_axes(const ScaleAxesArgs &s, void *data, const void *factors, hipStream_t st)
{
    const int64_t total = s.len * s.extra;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale_axes<T>, dim3(blocks_for(total)), dim3(kBlock), 0, st, s, static_cast<T *>(data),
                       static_cast<const T *>(factors));
    return hipGetLastError();
}
template hipError_t launch_scale_axes<float>(const ScaleAxesArgs &, void *, const void *, hipStream_t);
template hipError_t launch_scale_axes<double>(const ScaleAxesArgs &, void *, const void *, hipStream_t);

template <typename T>
hipError_t launch_quadratic(const void *w, void *u, int64_t n, double a, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_quadratic<T>, dim3(blocks_for(n)), dim3(kBlock), 0, st, static_cast<const T *>(w),
                       static_cast<T *>(u), n, (T)a);
    return hipGetLastError();
}
template hipError_t launch_quadratic<float>(const void *, void *, int64_t, double, hipStream_t);
template hipError_t launch_quadratic<double>(const void *, void *, int64_t, double, hipStream_t);

template <typename T>
hipError_t launch_quadratic_adjoint(const void *w, const void *g, void *out, int64_t n, double a, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_quadratic_adjoint<T>, dim3(blocks_for(n)), dim3(kBlock), 0, st, static_cast<const T *>(w),
                       static_cast<const T *>(g), static_cast<T *>(out), n, (T)a);
    return hipGetLastError();
}
template hipError_t launch_quadratic_adjoint<float>(const void *, const void *, void *, int64_t, double, hipStream_t);
template hipError_t launch_quadratic_adjoint<double>(const void *, const void *, void *, int64_t, double,
                                                     hipStream_t);

template <typename T>
hipError_t launch_sum_squares(const void *x, int64_t n, double *partial, int nparts, double *result, hipStream_t st)
{
    if (nparts < 1 || nparts > kSumSquaresParts) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sumsq_partial<T>, dim3((unsigned)nparts), dim3(kBlock), 0, st, static_cast<const T *>(x),
                       n, partial);
    hipLaunchKernelGGL(k_sumsq_final, dim3(1), dim3(kBlock), 0, st, partial, nparts, result);
    return hipGetLastError();
}
template hipError_t launch_sum_squares<float>(const void *, int64_t, double *, int, double *, hipStream_t);
template hipError_t launch_sum_squares<double>(const void *, int64_t, double *, int, double *, hipStream_t);

} // namespace pfh
