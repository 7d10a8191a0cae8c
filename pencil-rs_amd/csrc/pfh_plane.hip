// gfx950 plane stage: two route stages of a small 3D problem in one pass over
// HBM. A workgroup loads one whole plane (nc rows of nr contiguous complex
// values) into LDS, transforms every row (the first axis, compile-time radix
// plan), then every column (the second axis, read from LDS with the row
// stride), and writes each column as a contiguous run. Three line stages
// read and write the working array three times; a line stage plus a plane
// stage does it twice.
//
// The passes are the Stockham passes of the line kernels (natural-order
// output, decimation in time, pass twiddles from the staged tables); rows
// keep one LDS pad slot per 16 values, rows themselves start at an odd
// stride so that the column passes' lanes (consecutive points of a column)
// fall on distinct banks.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "pfh_kernels.hpp"

namespace pfh {

namespace {

__device__ __forceinline__ int rp(int i) { return i + (i >> 4); }

template <int EM, int R>
constexpr int plane_mb()
{
    return (EM / R) > 0 ? EM / R : 1;
}

// LDS accessors of one line of the plane: a row (padded, contiguous) or a
// column (stride rs between consecutive points).
template <typename T>
struct RowLine {
    cpx<T> *p;
    __device__ __forceinline__ cpx<T> get(int i) const { return p[rp(i)]; }
    __device__ __forceinline__ void set(int i, cpx<T> v) const { p[rp(i)] = v; }
};
template <typename T>
struct ColLine {
    cpx<T> *p;
    int rs;
    __device__ __forceinline__ cpx<T> get(int i) const { return p[i * rs]; }
    __device__ __forceinline__ void set(int i, cpx<T> v) const { p[i * rs] = v; }
};

// Radix-R pass j = tl + tpl * b over a line of M points in LDS: reads
// x[j + r M/R], twiddles by the pass table (Ns > 1), writes the Stockham
// position (j - j mod Ns) R + j mod Ns + r Ns. Every thread of the block
// takes part in both barriers; threads with tl >= M/R do no work.
template <typename T, int R, int EM, int M, typename Line>
__device__ __forceinline__ void plane_pass(const Line &ln, const cpx<T> *twp, int Ns, int tl, int tpl)
{
    constexpr int MB = plane_mb<EM, R>();
    constexpr int NB = M / R;
    cpx<T> v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int j = tl + tpl * b;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = ln.get(j + r * NB);
            if (Ns > 1) {
                const int k = j % Ns;
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], twp[k * (R - 1) + (r - 1)]);
            }
            Dft<T, R>::run(v[b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int j = tl + tpl * b;
        if (j < NB) {
            const int k = j % Ns;
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) ln.set(base + r * Ns, v[b][r]);
        }
    }
    __syncthreads();
}

// Row passes over all rows (rounds of blockDim / tpl_r rows per pass, so
// that the first pass finishes every row before the second starts): the
// first pass reads the raw rows, the prefetched ones from the raw LDS region
// (row < pre_rows, unpadded), the others from global memory.
template <typename T, int R0, int R1, int R2, int EM>
__device__ __forceinline__ void plane_rows(const PlaneArgs<T> &a, const cpx<T> *src, const cpx<T> *raw,
                                           int pre_rows, const cpx<T> *tw, cpx<T> *plane)
{
    constexpr int M = R0 * R1 * (R2 ? R2 : 1);
    constexpr int RL = R2 ? R2 : R1;
    constexpr int MB = plane_mb<EM, R0>();
    constexpr int NB = M / R0;
    const int tpl = a.tpl_r;
    const int per_round = blockDim.x / tpl;
    const int l = threadIdx.x / tpl;
    const bool in_shape = l < per_round;
    const bool cj = (a.flags & F_CONJ_IN) != 0;
    for (int row0 = 0; row0 < a.nc; row0 += per_round) {
        const int row = row0 + l;
        const bool valid = in_shape && row < a.nc;
        const int tl = valid ? threadIdx.x % tpl : (1 << 28);
        const RowLine<T> ln{plane + (valid ? row : 0) * a.rs};
        cpx<T> v[MB][R0];
        if (row < pre_rows) {
            const cpx<T> *s = raw + (valid ? row : 0) * M;
#pragma unroll
            for (int b = 0; b < MB; ++b) {
                const int j = tl + tpl * b;
                if (j < NB) {
#pragma unroll
                    for (int r = 0; r < R0; ++r) v[b][r] = s[j + r * NB];
                }
            }
        } else {
            const cpx<T> *s = src + (int64_t)(valid ? row : 0) * M;
#pragma unroll
            for (int b = 0; b < MB; ++b) {
                const int j = tl + tpl * b;
                if (j < NB) {
#pragma unroll
                    for (int r = 0; r < R0; ++r) v[b][r] = s[j + r * NB];
                }
            }
        }
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            const int j = tl + tpl * b;
            if (j < NB) {
                if (cj) {
#pragma unroll
                    for (int r = 0; r < R0; ++r) v[b][r].y = -v[b][r].y;
                }
                Dft<T, R0>::run(v[b]);
#pragma unroll
                for (int r = 0; r < R0; ++r) ln.set(j * R0 + r, v[b][r]);
            }
        }
    }
    __syncthreads();
}

template <typename T, int R0, int R1, int R2, int EM>
__device__ __forceinline__ void plane_rows_rest(const PlaneArgs<T> &a, const cpx<T> *tw, cpx<T> *plane)
{
    constexpr int M = R0 * R1 * (R2 ? R2 : 1);
    constexpr int RL = R2 ? R2 : R1;
    const int tpl = a.tpl_r;
    const int per_round = blockDim.x / tpl;
    const int l = threadIdx.x / tpl;
    const bool in_shape = l < per_round;
    if constexpr (R2 != 0) {
        for (int row0 = 0; row0 < a.nc; row0 += per_round) {
            const int row = row0 + l;
            const bool valid = in_shape && row < a.nc;
            const int tl = valid ? threadIdx.x % tpl : (1 << 28);
            plane_pass<T, R1, EM, M>(RowLine<T>{plane + (valid ? row : 0) * a.rs}, tw + a.twoff_row[1], R0, tl, tpl);
        }
    }
    for (int row0 = 0; row0 < a.nc; row0 += per_round) {
        const int row = row0 + l;
        const bool valid = in_shape && row < a.nc;
        const int tl = valid ? threadIdx.x % tpl : (1 << 28);
        plane_pass<T, RL, EM, M>(RowLine<T>{plane + (valid ? row : 0) * a.rs}, tw + a.twoff_row[R2 ? 2 : 1], M / RL,
                                 tl, tpl);
    }
}

// Column phase: columns in rounds of blockDim / tpl_c. The passes in LDS
// put consecutive lanes on consecutive columns (the same point of each, so
// the strided column reads hit consecutive LDS slots); the last pass puts
// consecutive lanes on consecutive points of one column and writes the
// column's natural-order output as a contiguous run of global memory
// (conjugation and scale applied).
template <typename T, int R0, int R1, int R2, int EM>
__device__ __forceinline__ void plane_cols(const PlaneArgs<T> &a, cpx<T> *dst, const cpx<T> *tw, cpx<T> *plane)
{
    constexpr int M = R0 * R1 * (R2 ? R2 : 1);
    constexpr int RL = R2 ? R2 : R1;
    constexpr int MB = plane_mb<EM, RL>();
    constexpr int NB = M / RL;
    const int tpl = a.tpl_c;
    const int per_round = blockDim.x / tpl;
    const bool in_shape = (int)threadIdx.x < per_round * tpl;
    // across-columns mapping for the passes in LDS, along-column for the last
    const int lx = threadIdx.x % per_round, tx = threadIdx.x / per_round;
    const int la = threadIdx.x / tpl, ta = threadIdx.x % tpl;
    const bool cj = (a.flags & F_CONJ_OUT) != 0;
    const T sc = a.scale;
    for (int col0 = 0; col0 < a.nr; col0 += per_round) {
        {
            const int col = col0 + lx;
            const bool valid = in_shape && col < a.nr;
            const int tl = valid ? tx : (1 << 28);
            const ColLine<T> ln{plane + rp(valid ? col : 0), a.rs};
            plane_pass<T, R0, EM, M>(ln, tw, 1, tl, tpl);
            if constexpr (R2 != 0) plane_pass<T, R1, EM, M>(ln, tw + a.twoff_col[1], R0, tl, tpl);
        }
        const int col = col0 + la;
        const bool valid = in_shape && col < a.nr;
        const int tl = valid ? ta : (1 << 28);
        const ColLine<T> ln{plane + rp(valid ? col : 0), a.rs};
        const cpx<T> *twp = tw + a.twoff_col[R2 ? 2 : 1];
        cpx<T> v[MB][RL];
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            const int j = tl + tpl * b;
            if (j < NB) {
#pragma unroll
                for (int r = 0; r < RL; ++r) v[b][r] = ln.get(j + r * NB);
#pragma unroll
                for (int r = 1; r < RL; ++r) v[b][r] = cmul(v[b][r], twp[j * (RL - 1) + (r - 1)]);
                Dft<T, RL>::run(v[b]);
            }
        }
        cpx<T> *d = dst + (int64_t)(valid ? col : 0) * M;
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            const int j = tl + tpl * b;
            if (j < NB) {
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    cpx<T> w = v[b][r];
                    if (cj) w.y = -w.y;
                    d[j + r * NB] = scal(w, sc);
                }
            }
        }
    }
}

// Prefetch of the first `rows` rows of a plane into the raw LDS region,
// global -> LDS without registers: wave k moves 1 KB chunks k, k + waves,
// ... (64 lanes x 16 bytes = 128 complex values), lanes past the rows'
// end stay inactive. The caller waits for the loads (s_waitcnt + barrier)
// before reading the region.
template <typename T>
__device__ __forceinline__ void plane_prefetch(const cpx<T> *src, int64_t elems, cpx<T> *raw)
{
    static_assert(sizeof(cpx<T>) == 8, "the prefetch moves pairs of 8-byte values");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    const int64_t chunks = (elems + 127) / 128;
    for (int64_t ch = wave; ch < chunks; ch += waves) {
        const int64_t e = ch * 128 + 2 * lane;
        if (e < elems)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + e),
                                             (__attribute__((address_space(3))) void *)(raw + ch * 128), 16, 0, 0);
    }
}

} // namespace

// Persistent plane stage over a work list: item q covers planes
// prefix[q] .. prefix[q + 1] - 1 of the launch; a block walks work indexes
// w = blockIdx.x, blockIdx.x + gridDim.x, ... and dispatches each plane on
// its item's row and column plans (uniform per block).
template <typename T>
__device__ __forceinline__ int plane_item(const int64_t *__restrict__ prefix, int nitems, int64_t w, int q)
{
    while (q + 1 < nitems && prefix[q + 1] <= w) ++q;
    return q;
}

// Whole rows of an item's plane that fit in the raw region.
template <typename T>
__device__ __forceinline__ int plane_pre_rows(const PlaneArgs<T> &a, int64_t raw_cap)
{
    const int64_t r = raw_cap / a.nr;
    return (int)(r < a.nc ? r : a.nc);
}

// Persistent plane stage over a work list: item q covers planes
// prefix[q] .. prefix[q + 1] - 1 of the launch; a block walks work indexes
// w = blockIdx.x, blockIdx.x + gridDim.x, ... and dispatches each plane on
// its item's row and column plans (uniform per block). LDS: the staged
// tables (tw_cap complex values), the working plane (work_cap) and the raw
// region (raw_cap, 16-byte aligned), into which the next plane's first rows
// are prefetched while the current plane is transformed.
template <typename T>
__global__ __launch_bounds__(kPlaneThreads) void k_plane_batch(const PlaneArgs<T> *__restrict__ items,
                                                              const int64_t *__restrict__ prefix, int nitems,
                                                              int64_t total, int tw_cap, int work_cap, int raw_cap)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
    cpx<T> *plane = lds + tw_cap;
    cpx<T> *raw = plane + work_cap;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    int q = 0;
    {
        const int64_t w0 = blockIdx.x;
        int lo = 0, hi = nitems;
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (prefix[mid] <= w0) lo = mid;
            else hi = mid;
        }
        q = lo;
    }
    if ((int64_t)blockIdx.x < total) {
        const PlaneArgs<T> &a0 = items[q];
        const int64_t p0 = (int64_t)blockIdx.x - prefix[q];
        plane_prefetch<T>(reinterpret_cast<const cpx<T> *>(a0.src) + p0 * a0.nr * a0.nc,
                          (int64_t)plane_pre_rows<T>(a0, raw_cap) * a0.nr, raw);
    }
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        q = plane_item<T>(prefix, nitems, w, q);
        const PlaneArgs<T> &a = items[q];
        const int64_t p = w - prefix[q];
        const int64_t pe = (int64_t)a.nr * a.nc;
        const cpx<T> *src = reinterpret_cast<const cpx<T> *>(a.src) + p * pe;
        cpx<T> *dst = reinterpret_cast<cpx<T> *>(a.dst) + p * pe;
        cpx<T> *tw_r = lds;
        cpx<T> *tw_c = lds + a.twn_row;
        __syncthreads(); // the previous plane is done with the tables and the working plane
        for (int i = tid; i < a.twn_row; i += nthr) tw_r[i] = a.tw_row[i];
        for (int i = tid; i < a.twn_col; i += nthr) tw_c[i] = a.tw_col[i];
        __builtin_amdgcn_s_waitcnt(0); // this wave's prefetch of this plane has landed
        __syncthreads();                // and every wave's
        const int pre = plane_pre_rows<T>(a, raw_cap);
        const int r0 = a.rr[0], r1 = a.rr[1], r2 = a.rr[2], re = a.rem;
#define PFH_PLANE_ROWS(TT, A, B, C, E)                                                                \
    if constexpr (std::is_same<T, TT>::value) {                                                       \
        if (r0 == A && r1 == B && r2 == C && re == E) plane_rows<T, A, B, C, E>(a, src, raw, pre, tw_r, plane); \
    }
        PFH_PLANE_PLANS(PFH_PLANE_ROWS)
#undef PFH_PLANE_ROWS
        // the raw region is free: prefetch the next plane of this block
        const int64_t wn = w + gridDim.x;
        if (wn < total) {
            const int qn = plane_item<T>(prefix, nitems, wn, q);
            const PlaneArgs<T> &an = items[qn];
            const int64_t pn = wn - prefix[qn];
            plane_prefetch<T>(reinterpret_cast<const cpx<T> *>(an.src) + pn * an.nr * an.nc,
                              (int64_t)plane_pre_rows<T>(an, raw_cap) * an.nr, raw);
        }
#define PFH_PLANE_REST(TT, A, B, C, E)                                                       \
    if constexpr (std::is_same<T, TT>::value) {                                              \
        if (r0 == A && r1 == B && r2 == C && re == E) plane_rows_rest<T, A, B, C, E>(a, tw_r, plane); \
    }
        PFH_PLANE_PLANS(PFH_PLANE_REST)
#undef PFH_PLANE_REST
        const int c0 = a.cr[0], c1 = a.cr[1], c2 = a.cr[2], ce = a.cem;
#define PFH_PLANE_COLS(TT, A, B, C, E)                                                       \
    if constexpr (std::is_same<T, TT>::value) {                                              \
        if (c0 == A && c1 == B && c2 == C && ce == E) plane_cols<T, A, B, C, E>(a, dst, tw_c, plane); \
    }
        PFH_PLANE_PLANS(PFH_PLANE_COLS)
#undef PFH_PLANE_COLS
    }
    __builtin_amdgcn_s_waitcnt(0);
}

template <typename T>
hipError_t launch_plane_batch(const PlaneArgs<T> *items, const int64_t *prefix, int nitems, int64_t total,
                              int threads, int tw_cap, int work_cap, int raw_cap, int64_t grid, hipStream_t s)
{
    if (total <= 0 || grid <= 0) return hipSuccess;
    const size_t lds_bytes = (size_t)(tw_cap + work_cap + raw_cap) * sizeof(cpx<T>);
    hipLaunchKernelGGL(k_plane_batch<T>, dim3((unsigned)grid), dim3(threads), lds_bytes, s, items, prefix, nitems,
                       total, tw_cap, work_cap, raw_cap);
    return hipGetLastError();
}
template hipError_t launch_plane_batch<float>(const PlaneArgs<float> *, const int64_t *, int, int64_t, int, int, int,
                                              int, int64_t, hipStream_t);

template <typename T>
int plane_batch_blocks_per_cu(int threads, size_t lds_bytes)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_plane_batch<T>, threads, lds_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}
template int plane_batch_blocks_per_cu<float>(int, size_t);

// Raises the plane kernel's dynamic LDS limit to `bytes` minus its static LDS.
hipError_t set_plane_lds_limit(size_t bytes)
{
    const void *f = reinterpret_cast<const void *>(&k_plane_batch<float>);
    hipFuncAttributes at{};
    hipError_t e = hipFuncGetAttributes(&at, f);
    if (e != hipSuccess) return e;
    const int v = (int)bytes - (int)at.sharedSizeBytes;
    if (v <= 0) return hipErrorInvalidValue;
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, v);
}

bool plane_plan_available(int tbytes, int r0, int r1, int r2, int em)
{
#define PFH_PLANE_HAS(TT, A, B, C, E) \
    if ((int)sizeof(TT) == tbytes && r0 == A && r1 == B && r2 == C && em == E) return true;
    PFH_PLANE_PLANS(PFH_PLANE_HAS)
#undef PFH_PLANE_HAS
    return false;
}

} // namespace pfh
