// Device reductions and broadcast pointwise combinations over pencil arrays
// at one rank: the local parts of the reference's checked global reductions
// (crates/pencil-array/src/collectives.rs:519-620, fold rules :200-520,
// NormState :2150-2208, norm_values :2210-2250, extremes :2066-2094) and of
// its broadcast pointwise operations (pointwise.rs:40-598).
//
// Every reduction is two kernels: a fixed grid of partial states (each block
// a fixed contiguous chunk of the array) and one block that combines the
// partials in chunk order, so results do not depend on scheduling. The
// checked integer sum is order sensitive (an intermediate overflow of the
// storage-order fold is an error even when the total fits): its state is the
// (total, lowest prefix, highest prefix) triple of a contiguous run, combined
// left to right, and each thread owns a contiguous run. The other reductions
// are order free; their threads stride the chunk for coalesced loads.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>

#include "pfh_collectives.hpp"

namespace pfh {

namespace {

constexpr int kRBlock = 256;

// ---------------------------------------------------------------------------
// Element access
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ double as_double(T v)
{
    return (double)v;
}

// bits of the non-finite flags, per component c: NaN 1 << 3c, +inf 2 << 3c,
// -inf 4 << 3c (the reference's six-word flag vector, collectives.rs:300-330)
__device__ __forceinline__ uint32_t nonfinite_bits(double v, int c)
{
    if (isnan(v)) return 1u << (3 * c);
    if (isinf(v)) return (v > 0 ? 2u : 4u) << (3 * c);
    return 0u;
}

// ---------------------------------------------------------------------------
// Reducers. State init(), push(state, x, i), combine(a, b) = a then b,
// finish(state, n, result).
// ---------------------------------------------------------------------------

// Floating sum of C components (1 real, 2 complex). Each component is a
// double-double (TwoSum) accumulator over the finite inputs; non-finite inputs
// only set flags, which decide the result as the reference's collective_sum
// does: NaN wins, opposite infinities give NaN, a lone infinity sign is kept.
template <typename T, int C>
struct FSum {
    static constexpr bool ordered = false;
    struct State {
        double hi[C], lo[C];
        uint32_t flags;
    };
    __device__ static State init()
    {
        State s;
        for (int c = 0; c < C; ++c) s.hi[c] = s.lo[c] = 0.0;
        s.flags = 0;
        return s;
    }
    __device__ static void add(double &hi, double &lo, double v)
    {
        const double s = hi + v;
        const double bp = s - hi;
        const double err = (hi - (s - bp)) + (v - bp);
        hi = s;
        lo += err;
    }
    __device__ static void push(State &s, const T *x, int64_t i)
    {
        for (int c = 0; c < C; ++c) {
            const double v = as_double(x[C * i + c]);
            const uint32_t f = nonfinite_bits(v, c);
            if (f) s.flags |= f;
            else add(s.hi[c], s.lo[c], v);
        }
    }
    __device__ static void combine(State &a, const State &b)
    {
        for (int c = 0; c < C; ++c) {
            add(a.hi[c], a.lo[c], b.hi[c]);
            a.lo[c] += b.lo[c];
        }
        a.flags |= b.flags;
    }
    __device__ static void finish(const State &s, int64_t n, ReduceResult *r)
    {
        (void)n;
        T out[C];
        for (int c = 0; c < C; ++c) {
            const uint32_t f = (s.flags >> (3 * c)) & 7u;
            double v;
            if ((f & 1u) || ((f & 6u) == 6u)) v = NAN;
            else if (f & 2u) v = INFINITY;
            else if (f & 4u) v = -INFINITY;
            else v = s.hi[c] + s.lo[c];
            out[c] = (T)v;
        }
        T *dst = reinterpret_cast<T *>(r->value);
        for (int c = 0; c < C; ++c) dst[c] = out[c];
        r->outcome = kReduceValue;
    }
};

// Checked integer sum (collectives.rs:206-214): every step of the storage-order
// fold must stay inside [MIN, MAX]. State of a contiguous run: its total and
// the lowest and highest of its prefix sums (empty runs have none).
template <typename I>
struct ISum {
    static constexpr bool ordered = true;
    struct State {
        __int128 total, low, high;
        int nonempty;
    };
    __device__ static State init()
    {
        State s;
        s.total = s.low = s.high = 0;
        s.nonempty = 0;
        return s;
    }
    __device__ static void push(State &s, const I *x, int64_t i)
    {
        s.total += (__int128)x[i];
        if (!s.nonempty) {
            s.low = s.high = s.total;
            s.nonempty = 1;
        } else {
            s.low = s.total < s.low ? s.total : s.low;
            s.high = s.total > s.high ? s.total : s.high;
        }
    }
    __device__ static void combine(State &a, const State &b)
    {
        if (!b.nonempty) return;
        const __int128 bl = a.total + b.low, bh = a.total + b.high;
        if (!a.nonempty) {
            a.low = bl;
            a.high = bh;
        } else {
            a.low = bl < a.low ? bl : a.low;
            a.high = bh > a.high ? bh : a.high;
        }
        a.total += b.total;
        a.nonempty = 1;
    }
    __device__ static void finish(const State &s, int64_t n, ReduceResult *r)
    {
        (void)n;
        const __int128 lo = (__int128)IntLimits<I>::lowest(), hi = (__int128)IntLimits<I>::highest();
        if (s.nonempty && (s.low < lo || s.high > hi)) {
            r->outcome = kReduceIntegerOverflow;
            return;
        }
        *reinterpret_cast<I *>(r->value) = (I)s.total;
        r->outcome = kReduceValue;
    }
};

// Minimum / maximum (mapped_extreme_layout, collectives.rs:2066-2094): NaN
// inputs are skipped by the fold but any NaN makes the result NaN; an empty
// array has no extreme.
template <typename T, bool MIN, bool FLOAT>
struct Extreme {
    static constexpr bool ordered = false;
    struct State {
        T v;
        uint32_t nan;
    };
    __device__ static T start()
    {
        if constexpr (FLOAT) return MIN ? (T)INFINITY : (T)-INFINITY;
        else return MIN ? IntLimits<T>::highest() : IntLimits<T>::lowest();
    }
    __device__ static State init() { return State{start(), 0u}; }
    __device__ static void push(State &s, const T *x, int64_t i)
    {
        const T v = x[i];
        if constexpr (FLOAT) {
            if (v != v) {
                s.nan = 1u;
                return;
            }
        }
        s.v = MIN ? (v < s.v ? v : s.v) : (v > s.v ? v : s.v);
    }
    __device__ static void combine(State &a, const State &b)
    {
        a.v = MIN ? (b.v < a.v ? b.v : a.v) : (b.v > a.v ? b.v : a.v);
        a.nan |= b.nan;
    }
    __device__ static void finish(const State &s, int64_t n, ReduceResult *r)
    {
        if (n == 0) {
            r->outcome = kReduceNone;
            return;
        }
        T v = s.v;
        if constexpr (FLOAT) {
            if (s.nan) v = (T)NAN;
        }
        *reinterpret_cast<T *>(r->value) = v;
        r->outcome = kReduceValue;
    }
};

// Scaled L2 norm (NormState, collectives.rs:2150-2208): a running scale (the
// largest magnitude so far) and the sum of squared ratios to it, so no square
// of an input overflows or underflows. Magnitudes are |x| in f64 (complex:
// hypot of the parts). NaN in any component gives NaN; an infinity gives
// +inf; an all-zero or empty array gives 0. NORM is the result type.
template <typename T, int C, typename NORM>
struct Norm {
    static constexpr bool ordered = false;
    struct State {
        double scale, ssq;
        uint32_t flags;
    };
    __device__ static State init() { return State{0.0, 0.0, 0u}; }
    __device__ static void push(State &s, const T *x, int64_t i)
    {
        double a;
        if (C == 1) {
            const double v = as_double(x[i]);
            s.flags |= nonfinite_bits(v, 0);
            a = fabs(v);
        } else {
            const double re = as_double(x[2 * i]), im = as_double(x[2 * i + 1]);
            s.flags |= nonfinite_bits(re, 0) | nonfinite_bits(im, 1);
            a = hypot(re, im);
        }
        if (isinf(a)) {
            s.scale = INFINITY;
            return;
        }
        if (a == 0.0 || isnan(a) || isinf(s.scale)) return;
        if (a > s.scale) {
            const double ratio = s.scale / a;
            s.ssq = 1.0 + (s.ssq * ratio) * ratio;
            s.scale = a;
        } else {
            const double ratio = a / s.scale;
            s.ssq += ratio * ratio;
        }
    }
    __device__ static void combine(State &a, const State &b)
    {
        a.flags |= b.flags;
        if (isinf(a.scale) || isinf(b.scale)) {
            a.scale = INFINITY;
            return;
        }
        if (b.scale == 0.0) return;
        if (a.scale == 0.0) {
            a.scale = b.scale;
            a.ssq = b.ssq;
            return;
        }
        if (b.scale > a.scale) {
            const double ratio = a.scale / b.scale;
            a.ssq = b.ssq + (a.ssq * ratio) * ratio;
            a.scale = b.scale;
        } else {
            const double ratio = b.scale / a.scale;
            a.ssq += (b.ssq * ratio) * ratio;
        }
    }
    __device__ static void finish(const State &s, int64_t n, ReduceResult *r)
    {
        const uint32_t nan_bits = 1u | (1u << 3);
        double v;
        if (s.flags & nan_bits) v = NAN;
        else if ((s.flags & ~nan_bits) || isinf(s.scale)) v = INFINITY;
        else if (s.scale == 0.0 || n == 0) v = 0.0;
        else v = s.scale * sqrt(s.ssq);
        *reinterpret_cast<NORM *>(r->value) = (NORM)v;
        r->outcome = kReduceValue;
    }
};

// any / all of the truth values (x != 0; complex: either part != 0; NaN is
// true), collectives.rs:567-590. any of nothing is false, all of nothing true.
template <typename T, int C, bool ANY>
struct Truth {
    static constexpr bool ordered = false;
    struct State {
        uint32_t v;
    };
    __device__ static State init() { return State{ANY ? 0u : 1u}; }
    __device__ static void push(State &s, const T *x, int64_t i)
    {
        bool t = false;
        for (int c = 0; c < C; ++c) t = t || (x[C * i + c] != (T)0);
        s.v = ANY ? (s.v | (uint32_t)t) : (s.v & (uint32_t)t);
    }
    __device__ static void combine(State &a, const State &b) { a.v = ANY ? (a.v | b.v) : (a.v & b.v); }
    __device__ static void finish(const State &s, int64_t n, ReduceResult *r)
    {
        (void)n;
        r->value[0] = (uint8_t)s.v;
        r->outcome = kReduceValue;
    }
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Ordered tree over the block's states: after it sh[0] holds
// sh[0] . sh[1] . ... . sh[B-1] (pairs always combine earlier with later).
template <class R>
__device__ void block_combine(typename R::State *sh)
{
    const int t = threadIdx.x;
    for (int s = 1; s < kRBlock; s <<= 1) {
        __syncthreads();
        if ((t & (2 * s - 1)) == 0 && t + s < kRBlock) R::combine(sh[t], sh[t + s]);
    }
    __syncthreads();
}

template <class R, typename T>
__global__ __launch_bounds__(kRBlock) void k_reduce_partial(const T *x, int64_t n, int64_t chunk,
                                                            typename R::State *partials)
{
    __shared__ typename R::State sh[kRBlock];
    const int t = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * chunk;
    const int64_t c1 = c0 + chunk < n ? c0 + chunk : n;
    typename R::State s = R::init();
    if (R::ordered) {
        const int64_t per = (chunk + kRBlock - 1) / kRBlock;
        const int64_t a = c0 + (int64_t)t * per;
        const int64_t b = a + per < c1 ? a + per : c1;
        for (int64_t i = a; i < b; ++i) R::push(s, x, i);
    } else {
        for (int64_t i = c0 + t; i < c1; i += kRBlock) R::push(s, x, i);
    }
    sh[t] = s;
    block_combine<R>(sh);
    if (t == 0) partials[blockIdx.x] = sh[0];
}

template <class R>
__global__ __launch_bounds__(kRBlock) void k_reduce_final(const typename R::State *partials, int nparts, int64_t n,
                                                          ReduceResult *result)
{
    __shared__ typename R::State sh[kRBlock];
    const int t = threadIdx.x;
    const int per = (nparts + kRBlock - 1) / kRBlock;
    typename R::State s = R::init();
    for (int k = t * per; k < (t + 1) * per && k < nparts; ++k) R::combine(s, partials[k]);
    sh[t] = s;
    block_combine<R>(sh);
    if (t == 0) {
        ReduceResult r;
        for (int k = 0; k < 16; ++k) r.value[k] = 0;
        r.outcome = kReduceValue;
        r.reserved[0] = r.reserved[1] = r.reserved[2] = 0;
        R::finish(sh[0], n, &r);
        *result = r;
    }
}

template <class R, typename T>
hipError_t run_reduce(const void *x, int64_t n, void *workspace, ReduceResult *result, hipStream_t st)
{
    static_assert(sizeof(typename R::State) <= kReduceStateBytes, "reduction state slot");
    int64_t parts = n / (kRBlock * 16);
    if (parts < 1) parts = 1;
    if (parts > kReduceParts) parts = kReduceParts;
    const int64_t chunk = n > 0 ? (n + parts - 1) / parts : 1;
    auto *partials = static_cast<typename R::State *>(workspace);
    hipLaunchKernelGGL((k_reduce_partial<R, T>), dim3((unsigned)parts), dim3(kRBlock), 0, st,
                       static_cast<const T *>(x), n, chunk, partials);
    hipLaunchKernelGGL((k_reduce_final<R>), dim3(1), dim3(kRBlock), 0, st, partials, (int)parts, n, result);
    return hipGetLastError();
}

template <typename T, bool FLOAT, int C, typename NORM>
hipError_t reduce_typed(int op, const void *x, int64_t n, void *ws, ReduceResult *r, hipStream_t st)
{
    switch (op) {
    case kReduceSum:
        if constexpr (FLOAT) return run_reduce<FSum<T, C>, T>(x, n, ws, r, st);
        else return run_reduce<ISum<T>, T>(x, n, ws, r, st);
    case kReduceMin:
        if constexpr (C == 1) return run_reduce<Extreme<T, true, FLOAT>, T>(x, n, ws, r, st);
        else return hipErrorInvalidValue;
    case kReduceMax:
        if constexpr (C == 1) return run_reduce<Extreme<T, false, FLOAT>, T>(x, n, ws, r, st);
        else return hipErrorInvalidValue;
    case kReduceL2Norm: return run_reduce<Norm<T, C, NORM>, T>(x, n, ws, r, st);
    case kReduceAny: return run_reduce<Truth<T, C, true>, T>(x, n, ws, r, st);
    case kReduceAll: return run_reduce<Truth<T, C, false>, T>(x, n, ws, r, st);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// Broadcast pointwise combinations
// ---------------------------------------------------------------------------

template <typename T>
struct Cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ Cx<T> load_cx(const void *p, int cplx, int64_t i)
{
    const T *q = static_cast<const T *>(p);
    if (cplx) return Cx<T>{q[2 * i], q[2 * i + 1]};
    return Cx<T>{q[i], (T)0};
}

template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return Cx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// out[e][k] = f(in_0[b_0(e)][k], ...): e runs over the output's extra index
// (grid y, strided), k over the spatial elements (grid x, strided); b_i(e) is
// input i's extra index with its extent-1 axes pinned to 0.
template <typename T>
__global__ __launch_bounds__(kRBlock) void k_pointwise(PointwiseArgs a, T *out)
{
    for (int64_t e = blockIdx.y; e < a.extra_count; e += gridDim.y) {
        int64_t base[kPointwiseMaxInputs];
        for (int i = 0; i < a.nin; ++i) {
            int64_t rem = e, lin = 0;
            for (int d = a.extra_ndim - 1; d >= 0; --d) {
                const int64_t idx = rem % a.extra_dims[d];
                rem /= a.extra_dims[d];
                lin += idx * a.extra_stride[i][d];
            }
            base[i] = lin * a.spatial;
        }
        const int64_t obase = e * a.spatial;
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.spatial;
             k += (int64_t)gridDim.x * blockDim.x) {
            Cx<T> r{(T)0, (T)0};
            if (a.op == kPointwiseLinear) {
                for (int i = 0; i < a.nin; ++i) {
                    const Cx<T> c{(T)a.coef[2 * i], (T)a.coef[2 * i + 1]};
                    const Cx<T> v = load_cx<T>(a.in[i], a.in_complex[i], base[i] + k);
                    const Cx<T> p = cmul(c, v);
                    r.re += p.re;
                    r.im += p.im;
                }
            } else {
                const Cx<T> c0{(T)a.coef[0], (T)a.coef[1]};
                const Cx<T> x0 = load_cx<T>(a.in[0], a.in_complex[0], base[0] + k);
                const Cx<T> x1 = load_cx<T>(a.in[1], a.in_complex[1], base[1] + k);
                r = cmul(cmul(c0, x0), x1);
                if (a.op == kPointwiseMulAdd) {
                    const Cx<T> c1{(T)a.coef[2], (T)a.coef[3]};
                    const Cx<T> p = cmul(c1, load_cx<T>(a.in[2], a.in_complex[2], base[2] + k));
                    r.re += p.re;
                    r.im += p.im;
                }
            }
            if (a.out_complex) {
                out[2 * (obase + k)] = r.re;
                out[2 * (obase + k) + 1] = r.im;
            } else {
                out[obase + k] = r.re;
            }
        }
    }
}

} // namespace

hipError_t launch_reduce(int op, int scalar_type, const void *x, int64_t n, void *workspace, ReduceResult *result,
                         hipStream_t st)
{
    switch (scalar_type) {
    case kScalarI8: return reduce_typed<int8_t, false, 1, double>(op, x, n, workspace, result, st);
    case kScalarU8: return reduce_typed<uint8_t, false, 1, double>(op, x, n, workspace, result, st);
    case kScalarI16: return reduce_typed<int16_t, false, 1, double>(op, x, n, workspace, result, st);
    case kScalarU16: return reduce_typed<uint16_t, false, 1, double>(op, x, n, workspace, result, st);
    case kScalarI32: return reduce_typed<int32_t, false, 1, double>(op, x, n, workspace, result, st);
    case kScalarU32: return reduce_typed<uint32_t, false, 1, double>(op, x, n, workspace, result, st);
    case kScalarI64: return reduce_typed<int64_t, false, 1, double>(op, x, n, workspace, result, st);
    case kScalarU64: return reduce_typed<uint64_t, false, 1, double>(op, x, n, workspace, result, st);
    case kScalarF32: return reduce_typed<float, true, 1, float>(op, x, n, workspace, result, st);
    case kScalarF64: return reduce_typed<double, true, 1, double>(op, x, n, workspace, result, st);
    case kScalarC64: return reduce_typed<float, true, 2, float>(op, x, n, workspace, result, st);
    case kScalarC128: return reduce_typed<double, true, 2, double>(op, x, n, workspace, result, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pointwise(const PointwiseArgs &a, void *out, bool f64, hipStream_t st)
{
    if (a.spatial <= 0 || a.extra_count <= 0) return hipSuccess;
    int64_t bx = (a.spatial + kRBlock - 1) / kRBlock;
    if (bx > 4096) bx = 4096;
    int64_t by = a.extra_count;
    if (by > 65535) by = 65535;
    // keep the grid near a few thousand blocks when both dimensions are large
    while (bx * by > 65536 && bx > 1) bx = (bx + 1) / 2;
    const dim3 grid((unsigned)bx, (unsigned)by);
    if (f64) hipLaunchKernelGGL(k_pointwise<double>, grid, dim3(kRBlock), 0, st, a, static_cast<double *>(out));
    else hipLaunchKernelGGL(k_pointwise<float>, grid, dim3(kRBlock), 0, st, a, static_cast<float *>(out));
    return hipGetLastError();
}

} // namespace pfh
