// Launch description of one line stage: a batch of 1D lines, each read from
// one layout, transformed, and written to another layout. The layout change
// is how a route transition (the reference's pack / exchange / unpack,
// crates/pencil-array/src/transpose.rs:1107-1315, or a LocalTransposePlan
// permutation, local_transpose.rs:87-132) is fused into a stage.
#pragma once

#include <cstdint>

#include "pfh_math.hpp"

namespace pfh {

constexpr int kMaxPasses = 12;
constexpr int kMaxDigits = 5;
// Complex values a thread holds in registers during one radix pass: 16 for
// f32, 8 for f64 (same register footprint).
template <typename T>
struct Regs {
    // generic kernel: 8 values per thread and pass (f32 and f64)
    static constexpr int emax = 8;
    static constexpr int unroll = sizeof(T) == 4 ? 16 : 8; // fast kernels: loads / stores in flight per chunk
#ifndef PFH_GUNROLL_F32
#define PFH_GUNROLL_F32 4
#endif
#ifndef PFH_GUNROLL_F64
#define PFH_GUNROLL_F64 4
#endif
    // generic kernel: loads / stores per unrolled chunk
    static constexpr int gunroll = sizeof(T) == 4 ? PFH_GUNROLL_F32 : PFH_GUNROLL_F64;
    static constexpr int max_radix = sizeof(T) == 4 ? 32 : 8;
};
constexpr int kEmax = 8;         // generic kernel values per thread and pass
// Generic kernel budgets. The instantiation with the odd-prime passes gives
// the radix 5 and 7 passes 3 and 2 butterflies per thread and, in f32, the
// radix 11 / 13 passes 2 and the pre-processing 16 values per thread, so a
// 385-point line (5 x 7 x 11) takes 32 threads (f32) or 64 (f64) instead of
// 128; f64 keeps the smaller budgets, whose registers would spill.
constexpr int generic_pre_e(bool wide, bool f32) { return (wide && f32) ? 16 : kEmax; }

constexpr int generic_mb(int R, bool wide, bool f32)
{
    if (wide && R == 5) return 3;
    if (wide && R == 7) return 2;
    if (wide && f32 && (R == 11 || R == 13)) return 2;
    return kEmax / R > 0 ? kEmax / R : 1;
}
constexpr int kMaxThreadsWide = 512; // generic kernel with the odd-prime passes
constexpr int kMaxThreads = 1024;

// Per-line transform applied by a stage (see pfh_kernels.hip for the math).
enum StageMode : int {
    M_COPY_C = 0,  // identity on complex values (layout change only)
    M_COPY_R,      // identity on real values
    M_C2C,         // complex FFT of length n
    M_R2C,         // real n (even) -> n/2+1 complex via the n/2 complex FFT
    M_R2C_ODD,     // real n (odd) -> (n+1)/2 complex via the n complex FFT
    M_C2R,         // n/2+1 complex -> real n (even), positive sign
    M_C2R_ODD,     // (n+1)/2 complex -> real n (odd)
    M_DCT2,        // REDFT10, n even (Makhoul reorder + n/2 complex FFT)
    M_DCT2_ODD,    // REDFT10, n odd (n complex FFT)
    M_DCT3,        // REDFT01, n even (transpose of M_DCT2)
    M_DCT3_ODD,    // REDFT01, n odd
    M_DCT4,        // REDFT11, n even (n/2 complex FFT, pre/post twiddle)
    M_DCT4_ODD,    // REDFT11, n odd (2n complex FFT of the twiddled line)
    M_DCT1,        // REDFT00 (even extension, n-1 complex FFT)
    M_DST1,        // RODFT00 (odd extension, n+1 complex FFT)
    M_DHT,         // Hartley, n even (n/2 complex FFT)
    M_DHT_ODD,     // Hartley, n odd (n complex FFT)
    M_COUNT
};

enum StageFlags : int {
    F_CONJ_IN = 1,      // conjugate on load (positive-sign C2C)
    F_CONJ_OUT = 2,     // conjugate on store
    F_PRE_REV = 4,      // read x[n-1-j] (DST-III / DST-IV from DCT-III / DCT-IV)
    F_PRE_ALT = 8,      // read (-1)^j x[j] (DST-II from DCT-II)
    F_POST_REV = 16,    // write y[k] = t[n-1-k]
    F_POST_ALT = 32,    // write y[k] = (-1)^k t[k]
    F_HERM_DOUBLE = 64, // R2C adjoint weights: x2 on interior bins, Im := 0 at DC/Nyquist
    F_HERM_HALF = 128,  // C2R adjoint weights: x1/2 on interior bins
    F_WEIGHTS = 256,    // apply pre/post edge weights (R2R adjoints)
    F_XCD_REMAP = 512,  // map consecutive tiles onto one XCD
    F_DIRECT_FIRST = 1024, // first radix pass reads global memory into registers
    F_DIRECT_LAST = 2048,  // last radix pass writes global memory from registers (C2C)
    F_NT = 4096,           // stage arrays exceed the last-level cache: non-temporal loads and stores
    // Pieces of a long line (generic kernel, complex lines; see StageArgs::x*):
    F_XIN = 8192,   // points whose full index reaches xvalid_in load as zero; xtab_in factor on load
    F_XOUT = 16384, // points whose full index reaches xvalid_out are not stored; xtab_out factor on store
    F_XTW = 32768,  // twiddle exp(-2 pi i d k / xtw_m) on the store of point k
    // real column side whose neighbouring lines are neighbouring elements:
    // a lane moves lines 2i and 2i + 1 of a point as one 2-element access
    F_CPAIR_IN = 65536,
    F_CPAIR_OUT = 131072,
    // real row side of even length with even line offsets: a lane moves
    // points 2i and 2i + 1 of a line as one 2-element access
    F_RPAIR_IN = 262144,
    F_RPAIR_OUT = 524288,
};

struct LineGeom {
    int nd;                       // number of line digits
    int64_t ext[kMaxDigits];      // digit extents, digit 0 fastest
    int64_t is[kMaxDigits];       // input element stride of each digit
    int64_t os[kMaxDigits];       // output element stride of each digit
    uint32_t dmag[kMaxDigits];    // invariant-divisor constants of ext (make_fastdiv)
    int dsh[kMaxDigits];
};

// Exact quotient x / d for 0 <= x < 2^31 and an invariant 1 <= d < 2^31
// without a division: q = (x + mulhi(x, mag)) >> sh with sh = ceil(log2 d)
// and mag = floor(2^32 (2^sh - d) / d) + 1, the round-up method of
// Granlund and Montgomery (PLDI 1994, section 4). x + mulhi(x, mag) < 2^32
// because mag < 2^32 and x < 2^31.
struct FastDiv {
    uint32_t mag;
    int sh;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    int sh = 0;
    while ((1ull << sh) < d) ++sh;
    FastDiv f;
    f.mag = (uint32_t)(((1ull << 32) * ((1ull << sh) - d)) / d + 1);
    f.sh = sh;
    return f;
}

__host__ __device__ __forceinline__ uint32_t fastdiv(uint32_t x, uint32_t mag, int sh)
{
    return (x + (uint32_t)(((uint64_t)x * mag) >> 32)) >> sh;
}

template <typename T>
struct StageArgs {
    const void *src;
    void *dst;
    const cpx<T> *tw;       // this stage's twiddle / pre / post tables
    const int *gate;        // if non-null and *gate != 0 the stage writes nothing
    int64_t nlines;
    int64_t ntiles;
    LineGeom g;
    int64_t ips, ops;       // point strides of the transform axis (elements)
    int mode, flags;
    int n;                  // logical line length
    int nin, nout;          // elements per line read / written
    int m;                  // complex FFT length (0: no FFT)
    int npass;
    int radix[kMaxPasses];
    int twoff[kMaxPasses];  // pass twiddle offsets (complex units) into tw
    uint32_t nsmag[kMaxPasses]; // make_fastdiv constants of each pass's Ns (product of the earlier radices)
    int nssh[kMaxPasses];
    int pre0, pre1, post0, post1; // table offsets into tw
    int twn;                // complex entries of tw staged into LDS
    int fem;                // values per thread and pass of the compile-time plan (0: generic kernel)
    int lt, tpl;            // lines per tile, threads per line
    int lc;                 // lines computed at a time (lc * tpl threads; lt is a multiple of lc)
    int ls;                 // LDS stride of one line (complex units)
    int in_col, out_col;    // column (lines adjacent in memory) access on load / store
    int xcd_per;            // tiles per XCD when F_XCD_REMAP
    T scale;
    T prew0, prew1, postw0, postw1;
    // Pieces of a long line (F_XIN / F_XOUT / F_XTW): a line too long for one
    // workgroup runs as several launches. d is the coordinate of line digit
    // xdig (0 when xdig < 0), and point p of a line stands for the index
    // p * xmul + d of the long line.
    const cpx<T> *xtab_in;  // factor of full index f on load (null: none)
    const cpx<T> *xtab_out; // factor of full index f on store (null: none)
    const cpx<T> *xtw;      // exp(-2 pi i q / xtw_m) = xtw[q mod 2^xtw_bsh] * xtw[2^xtw_bsh + (q >> xtw_bsh)]
    int xdig;
    int xmul_in, xmul_out;
    int xvalid_in, xvalid_out;
    int xtw_m, xtw_bsh;
};

// Pre- and post-processing of a long real line (k_long_prep / k_long_post,
// pfh_long.hip): the transform runs as a complex FFT of length m of the
// promoted line (real part, zero imaginary part, or the even / odd
// extension), whose input the prep kernel writes and whose output the post
// kernel reads, line-major (line l at l * m).
enum LongKind : int { LK_R2C = 0, LK_C2R, LK_DCT2, LK_DCT3, LK_DCT4, LK_DCT1, LK_DST1, LK_DHT };

template <typename T>
struct LongArgs {
    const void *src;
    void *dst;
    const cpx<T> *tab;   // plan tables (pre0 / post0 offsets)
    const int *gate;     // if non-null and *gate != 0 nothing is written
    int64_t nlines;
    LineGeom g;          // the stage's lines: is = input strides, os = output strides
    int64_t ips, ops;    // point strides of the stage's input / output
    int kind, flags;     // LongKind, stage flags (F_PRE_* / F_POST_* / F_WEIGHTS / F_HERM_*)
    int n;               // logical line length
    int m;               // promoted complex FFT length
    int nout;            // output points per line (post)
    int pre0, post0;     // table offsets
    T scale, prew0, prew1, postw0, postw1;
};

// Two route stages fused over whole planes (k_plane_batch): a plane of nc
// rows of nr contiguous elements is loaded into LDS, every row is
// transformed (first FFT axis, length nr), then every column (second axis,
// length nc), and the plane is written with the second axis contiguous
// (column c at c * nc). Planes of one item lie back to back (plane p at
// p * nr * nc in both arrays).
template <typename T>
struct PlaneArgs {
    const void *src;
    void *dst;
    const cpx<T> *tw_row;   // pass twiddles of the row FFT's plan (offsets twoff_row)
    const cpx<T> *tw_col;   // and of the column FFT's plan
    int64_t nplanes;
    int nr, nc;             // row length (first FFT), rows (second FFT length)
    int twn_row, twn_col;   // table entries staged into LDS
    int twoff_row[3], twoff_col[3];
    int rr[3], rem;         // compile-time plan of the rows (radices, values per thread)
    int cr[3], cem;         // and of the columns
    int tpl_r, tpl_c;       // threads per row / column line
    int rs;                 // LDS stride of a row (complex units, odd)
    int flags;              // F_CONJ_IN (rows' load), F_CONJ_OUT (columns' store)
    T scale;                // applied at the store
};

// Instantiated fast plans: {R0, R1, R2, EM}.
// Compile-time plans for route stages only (the batched kernels, which
// carry C2C lines, do not instantiate them): the 385-point FFT of a DST-I
// line of length 384 (5 x 7 x 11, 11 values per thread: 64 threads a line;
// with 22 values and 32 threads the 256x384x240 R2R forward + inverse took
// 2% (f32) and 16% (f64) longer on MI355X)
// and the f64 half-length FFTs of 240- and 384-point R2C / C2R / DHT lines.
#define PFH_FAST_ROUTE_PLANS(X) \
    X(float, 11, 7, 5, 11)      \
    X(double, 11, 7, 5, 11)     \
    X(double, 8, 3, 5, 8)       \
    X(double, 8, 8, 3, 8)

// Real-input R2R stages with the stage mode also compiled in: the 128- and
// 120-point plans of the 256- and 240-point DCT-II / DCT-III / DCT-IV lines
// and the 385-point plan of the 384-point DST-I lines.
#define PFH_R2R_MODE_PLANS(X)         \
    X(float, 8, 16, 0, 16, M_DCT2)    \
    X(float, 8, 16, 0, 16, M_DCT3)    \
    X(float, 8, 3, 5, 16, M_DCT4)     \
    X(float, 11, 7, 5, 11, M_DST1)    \
    X(double, 4, 4, 8, 8, M_DCT2)     \
    X(double, 4, 4, 8, 8, M_DCT3)     \
    X(double, 8, 3, 5, 8, M_DCT4)     \
    X(double, 11, 7, 5, 11, M_DST1)

#define PFH_FAST_PLANS(X)      \
    X(float, 8, 8, 0, 16)      \
    X(float, 8, 16, 0, 16)     \
    X(float, 16, 16, 0, 16)    \
    X(float, 16, 32, 0, 16)    \
    X(float, 16, 32, 0, 32)    \
    X(float, 8, 8, 8, 16)      \
    X(float, 32, 32, 0, 32)    \
    X(float, 16, 3, 0, 16)     \
    X(float, 4, 3, 5, 16)      \
    X(float, 16, 5, 0, 16)     \
    X(float, 4, 3, 7, 16)      \
    X(float, 8, 4, 3, 16)      \
    X(float, 16, 7, 0, 16)     \
    X(float, 8, 3, 5, 16)      \
    X(float, 8, 4, 5, 16)      \
    X(float, 16, 4, 3, 16)     \
    X(float, 16, 3, 5, 16)     \
    X(float, 16, 4, 5, 16)     \
    X(float, 16, 8, 3, 16)     \
    X(double, 8, 8, 0, 8)      \
    X(double, 4, 4, 8, 8)      \
    X(double, 4, 8, 8, 8)      \
    X(double, 8, 8, 8, 8)

// Threads of a plane-stage workgroup (the kernel's launch bound).
#ifndef PFH_PLANE_THREADS
#define PFH_PLANE_THREADS 1024
#endif
constexpr int kPlaneThreads = PFH_PLANE_THREADS;

// Compile-time plans the plane stage dispatches on (f32, 16 values per
// thread; lengths up to 384, longer lines do not leave room for a plane).
#define PFH_PLANE_PLANS(X)     \
    X(float, 8, 8, 0, 16)      \
    X(float, 8, 16, 0, 16)     \
    X(float, 16, 16, 0, 16)    \
    X(float, 16, 3, 0, 16)     \
    X(float, 4, 3, 5, 16)      \
    X(float, 16, 5, 0, 16)     \
    X(float, 4, 3, 7, 16)      \
    X(float, 8, 4, 3, 16)      \
    X(float, 16, 7, 0, 16)     \
    X(float, 8, 3, 5, 16)      \
    X(float, 8, 4, 5, 16)      \
    X(float, 16, 4, 3, 16)     \
    X(float, 16, 3, 5, 16)     \
    X(float, 16, 4, 5, 16)     \
    X(float, 16, 8, 3, 16)

} // namespace pfh
