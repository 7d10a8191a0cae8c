// Pre- and post-processing of long real lines.
//
// A real-data stage (R2C, C2R, the DCT / DST kinds, DHT) whose line does
// not fit one workgroup runs as: k_long_prep (the line -> the input of a
// complex FFT of length m, line-major in the workspace), the pieces of the
// long complex FFT (four-step split or Bluestein's chirp convolution, run by
// the generic line kernel), and k_long_post (the FFT output -> the stage's
// output layout). The promoted FFTs are the textbook ones: R2C / C2R / DHT
// through the length-n complex FFT of the real line, DCT-II / DCT-III
// through Makhoul's reorder and the length-n FFT (J. Makhoul, "A fast
// cosine transform in one and two dimensions", IEEE TASSP 28(1), 1980),
// DCT-IV through the length-2n FFT of the twiddled, zero-padded line, and
// DCT-I / DST-I through the length-2(n-1) / 2(n+1) FFT of the even / odd
// extension. The same kinds as the single-workgroup stages, with the same
// reversal, alternating-sign, adjoint-weight and Hermitian-weight flags.
#include <hip/hip_runtime.h>

#include "pfh_kernels.hpp"

namespace pfh {

namespace {

// Offset of a line's first input (is) or output (os) element.
__device__ __forceinline__ int64_t line_base(const LineGeom &g, int64_t line, bool out)
{
    int64_t off = 0;
    uint32_t x = (uint32_t)line;
#pragma unroll
    for (int d = 0; d < kMaxDigits; ++d) {
        if (d < g.nd) {
            const uint32_t q = fastdiv(x, g.dmag[d], g.dsh[d]);
            const uint32_t r = x - q * (uint32_t)g.ext[d];
            off += (int64_t)r * (out ? g.os[d] : g.is[d]);
            x = q;
        }
    }
    return off;
}

// Input value i of a real line after the reversal, alternating sign and
// adjoint edge weights the stage asks for.
template <typename T>
__device__ __forceinline__ T real_in(const LongArgs<T> &a, const T *s, int i)
{
    const int idx = (a.flags & F_PRE_REV) ? a.n - 1 - i : i;
    T v = s[(int64_t)idx * a.ips];
    if ((a.flags & F_PRE_ALT) && (idx & 1)) v = -v;
    if (a.flags & F_WEIGHTS) {
        if (idx == 0) v = v * a.prew0;
        if (idx == a.n - 1) v = v * a.prew1;
    }
    return v;
}

// Half-spectrum value X_k of a C2R line: Im of DC (and Nyquist) dropped,
// interior bins halved for the adjoint of R2C.
template <typename T>
__device__ __forceinline__ cpx<T> half_in(const LongArgs<T> &a, const cpx<T> *s, int k, int last)
{
    cpx<T> v = s[(int64_t)k * a.ips];
    if (k == 0 || k == last) v.y = (T)0;
    if ((a.flags & F_HERM_HALF) && k > 0 && k < last) v = scal(v, (T)0.5);
    return v;
}

template <typename T>
__device__ __forceinline__ cpx<T> prep_value(const LongArgs<T> &a, const void *src, int64_t io, int j)
{
    const int n = a.n;
    const T *rs = reinterpret_cast<const T *>(src) + io;
    switch (a.kind) {
    case LK_C2R: {
        const cpx<T> *cs = reinterpret_cast<const cpx<T> *>(src) + io;
        const int h = n / 2, last = (n % 2 == 0) ? h : h + 1;
        const cpx<T> v = (j <= h) ? half_in(a, cs, j, last) : conjg(half_in(a, cs, n - j, last));
        return conjg(v);
    }
    case LK_DCT2: {
        const int hh = (n + 1) / 2;
        return mk(j < hh ? real_in(a, rs, 2 * j) : real_in(a, rs, 2 * (n - 1 - j) + 1), (T)0);
    }
    case LK_DCT3: {
        T y = real_in(a, rs, j);
        if (j == 0) y = y * (T)0.5;
        return scal(a.tab[a.pre0 + j], y);
    }
    case LK_DCT4:
        return j < n ? scal(a.tab[a.pre0 + j], real_in(a, rs, j)) : mk((T)0, (T)0);
    case LK_DCT1:
        return mk(j <= n - 1 ? real_in(a, rs, j) : real_in(a, rs, 2 * (n - 1) - j), (T)0);
    case LK_DST1: {
        if (j == 0 || j == n + 1) return mk((T)0, (T)0);
        return mk(j <= n ? real_in(a, rs, j - 1) : -real_in(a, rs, 2 * n + 1 - j), (T)0);
    }
    default: // LK_R2C, LK_DHT
        return mk(real_in(a, rs, j), (T)0);
    }
}

template <typename T>
__device__ __forceinline__ T post_real_value(const LongArgs<T> &a, const cpx<T> *Z, int k)
{
    const int n = a.n;
    const int kk = (a.flags & F_POST_REV) ? n - 1 - k : k;
    T y;
    switch (a.kind) {
    case LK_DCT2:
    case LK_DCT4:
        y = (T)2 * cmul(a.tab[a.post0 + kk], Z[kk]).x;
        break;
    case LK_DCT3: {
        const int i = (kk & 1) ? n - 1 - (kk >> 1) : (kk >> 1);
        y = (T)2 * Z[i].x;
        break;
    }
    case LK_DST1:
        y = -Z[kk + 1].y;
        break;
    case LK_DHT:
        y = Z[kk].x - Z[kk].y;
        break;
    default: // LK_C2R, LK_DCT1
        y = Z[kk].x;
        break;
    }
    if ((a.flags & F_POST_ALT) && (k & 1)) y = -y;
    if (a.flags & F_WEIGHTS) {
        if (k == 0) y = y * a.postw0;
        if (k == n - 1) y = y * a.postw1;
    }
    return y * a.scale;
}

} // namespace

// Blocks (x, y): x covers the points of a line, y the lines (strided by the
// grid); every thread writes whole elements, no LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_long_prep(const LongArgs<T> a)
{
    if (a.gate != nullptr && *a.gate != 0) return;
    cpx<T> *dst = reinterpret_cast<cpx<T> *>(a.dst);
    for (int64_t line = blockIdx.y; line < a.nlines; line += gridDim.y) {
        const int64_t io = line_base(a.g, line, false);
        cpx<T> *z = dst + line * (int64_t)a.m;
        for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < a.m; j += (int64_t)gridDim.x * blockDim.x)
            z[j] = prep_value(a, a.src, io, (int)j);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_long_post(const LongArgs<T> a)
{
    if (a.gate != nullptr && *a.gate != 0) return;
    const cpx<T> *src = reinterpret_cast<const cpx<T> *>(a.src);
    for (int64_t line = blockIdx.y; line < a.nlines; line += gridDim.y) {
        const int64_t oo = line_base(a.g, line, true);
        const cpx<T> *Z = src + line * (int64_t)a.m;
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.nout;
             k += (int64_t)gridDim.x * blockDim.x) {
            if (a.kind == LK_R2C) {
                // crop of the length-n spectrum: DC (and Nyquist) exactly
                // real; adjoint of C2R: interior bins doubled
                const int n = a.n, last = (n % 2 == 0) ? n / 2 : -1;
                cpx<T> v = Z[k];
                if (k == 0 || k == last) v.y = (T)0;
                if (a.flags & F_HERM_DOUBLE) {
                    if (k > 0 && k != last) v = scal(v, (T)2);
                    else v.y = (T)0;
                }
                reinterpret_cast<cpx<T> *>(a.dst)[oo + k * a.ops] = scal(v, a.scale);
            } else {
                reinterpret_cast<T *>(a.dst)[oo + k * a.ops] = post_real_value(a, Z, (int)k);
            }
        }
    }
}

namespace {
dim3 long_grid(int64_t points, int64_t nlines)
{
    int64_t gx = (points + 255) / 256;
    if (gx > 4096) gx = 4096;
    if (gx < 1) gx = 1;
    int64_t gy = 32768 / gx;
    if (gy < 1) gy = 1;
    if (gy > nlines) gy = nlines;
    if (gy > 65535) gy = 65535;
    if (gy < 1) gy = 1;
    return dim3((unsigned)gx, (unsigned)gy, 1);
}
} // namespace

template <typename T>
hipError_t launch_long_prep(const LongArgs<T> &a, hipStream_t s)
{
    if (a.nlines <= 0 || a.m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_long_prep<T>, long_grid(a.m, a.nlines), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_long_post(const LongArgs<T> &a, hipStream_t s)
{
    if (a.nlines <= 0 || a.nout <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_long_post<T>, long_grid(a.nout, a.nlines), dim3(256), 0, s, a);
    return hipGetLastError();
}

template hipError_t launch_long_prep<float>(const LongArgs<float> &, hipStream_t);
template hipError_t launch_long_prep<double>(const LongArgs<double> &, hipStream_t);
template hipError_t launch_long_post<float>(const LongArgs<float> &, hipStream_t);
template hipError_t launch_long_post<double>(const LongArgs<double> &, hipStream_t);

} // namespace pfh
