// gfx950 kernels of the pencil-rs single-GPU FFT path.
//
// k_line_stage runs one route stage: it reads a tile of `lt` lines of one
// memory axis (row access when the axis is contiguous, column access with the
// tile's lines adjacent in memory otherwise), stages them in LDS, applies the
// per-mode pre-processing, runs the Stockham radix passes of the complex FFT
// of length m in LDS, and writes the post-processed line into the output
// layout. Writing into a different layout than the one read is how the
// reference's transitions (pack -> exchange -> unpack,
// crates/pencil-array/src/transpose.rs:1107-1315, alltoallv_transpose.rs:
// 283-308) and LocalTransposePlan (local_transpose.rs:87-132) are fused into
// the stage, so every stage is exactly one HBM read and one HBM write of the
// working array.
//
// The real transforms do not embed a line into a 2x..8x longer complex FFT
// (crates/pencil-fft/src/r2r.rs:697-743); they run pre/post twiddle kernels
// around a complex FFT of length n/2, n-1, n+1 or n.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "pfh_kernels.hpp"

namespace pfh {

namespace {

// Offsets of a line's first input and output element (line < 2^31, digit
// quotients by multiply-high with the planner's constants). The digit loop
// is unrolled over the compile-time maximum so that the argument struct is
// never indexed dynamically (which would force a private copy of it).
__device__ __forceinline__ void line_offsets(const LineGeom &g, int64_t line, int64_t &io, int64_t &oo)
{
    io = 0;
    oo = 0;
    uint32_t x = (uint32_t)line;
#pragma unroll
    for (int d = 0; d < kMaxDigits; ++d) {
        if (d < g.nd) {
            const uint32_t q = fastdiv(x, g.dmag[d], g.dsh[d]);
            const uint32_t r = x - q * (uint32_t)g.ext[d];
            io += (int64_t)r * g.is[d];
            oo += (int64_t)r * g.os[d];
            x = q;
        }
    }
}

// Coordinate of line digit xdig of a line (pieces of a long line).
__device__ __forceinline__ int digit_coord(const LineGeom &g, int64_t line, int xdig)
{
    uint32_t x = (uint32_t)line;
    int c = 0;
#pragma unroll
    for (int d = 0; d < kMaxDigits; ++d) {
        if (d < g.nd) {
            const uint32_t q = fastdiv(x, g.dmag[d], g.dsh[d]);
            if (d == xdig) c = (int)(x - q * (uint32_t)g.ext[d]);
            x = q;
        }
    }
    return c;
}

__device__ __forceinline__ bool complex_input(int mode)
{
    return mode == M_COPY_C || mode == M_C2C || mode == M_C2R || mode == M_C2R_ODD;
}

__device__ __forceinline__ bool complex_output(int mode)
{
    return mode == M_COPY_C || mode == M_C2C || mode == M_R2C || mode == M_R2C_ODD;
}

__device__ __forceinline__ bool needs_pre(int mode)
{
    return !(mode == M_COPY_C || mode == M_COPY_R || mode == M_C2C || mode == M_R2C || mode == M_DHT);
}

// An LDS line keeps one pad slot after every 16 complex values, so that the
// stride-R writes of a Stockham pass (R = 8, 16, 32) hit distinct banks.
__device__ __forceinline__ int lp(int i) { return i + (i >> 4); }

template <typename T>
__device__ __forceinline__ cpx<T> L(const cpx<T> *ln, int i)
{
    return ln[lp(i)];
}
template <typename T>
__device__ __forceinline__ void Lset(cpx<T> *ln, int i, cpx<T> v)
{
    ln[lp(i)] = v;
}
// real view of a line: real p lives in complex slot p / 2
template <typename T>
__device__ __forceinline__ T LR(const cpx<T> *ln, int p)
{
    return reinterpret_cast<const T *>(ln)[2 * lp(p >> 1) + (p & 1)];
}
template <typename T>
__device__ __forceinline__ void LRset(cpx<T> *ln, int p, T v)
{
    reinterpret_cast<T *>(ln)[2 * lp(p >> 1) + (p & 1)] = v;
}

// One input value of a real line after the reversal / alternating sign /
// adjoint edge weights requested by the stage.
// Raw real line values for the pre-processing, read from the tile's LDS line.
template <typename T>
struct LdsReal {
    const cpx<T> *ln;
    __device__ __forceinline__ T operator()(int p) const { return LR(ln, p); }
};

template <typename T, typename Rd>
__device__ __forceinline__ T xin(const StageArgs<T> &a, const Rd &rd, int i)
{
    const int idx = (a.flags & F_PRE_REV) ? a.n - 1 - i : i;
    T v = rd(idx);
    if ((a.flags & F_PRE_ALT) && (idx & 1)) v = -v;
    if (a.flags & F_WEIGHTS) {
        if (idx == 0) v = v * a.prew0;
        if (idx == a.n - 1) v = v * a.prew1;
    }
    return v;
}

// Half-spectrum input value X_k of a C2R line (Im of DC / Nyquist dropped,
// adjoint halving of interior bins).
template <typename T>
__device__ __forceinline__ cpx<T> hin(const StageArgs<T> &a, const cpx<T> *ln, int k, int last)
{
    cpx<T> v = L(ln, k);
    if (k == 0 || k == last) v.y = (T)0;
    if ((a.flags & F_HERM_HALF) && k > 0 && k < last) v = scal(v, (T)0.5);
    return v;
}

// FFT input value j of a real-input R2R mode, raw values read through rd
// (MODE >= 0: the stage mode as a compile-time constant).
template <typename T, typename Rd, int MODE = -1>
__device__ __forceinline__ cpx<T> pre_real(const StageArgs<T> &a, const Rd &rd, const cpx<T> *tws, int j)
{
    const int n = a.n;
    switch (MODE >= 0 ? MODE : a.mode) {
    case M_DHT_ODD:
        return mk(xin(a, rd, j), (T)0);
    case M_DCT2: {
        const int h = n / 2;
        const int i0 = 2 * j, i1 = 2 * j + 1;
        T v0 = i0 < h ? xin(a, rd, 2 * i0) : xin(a, rd, 2 * (n - 1 - i0) + 1);
        T v1 = i1 < h ? xin(a, rd, 2 * i1) : xin(a, rd, 2 * (n - 1 - i1) + 1);
        return mk(v0, v1);
    }
    case M_DCT2_ODD: {
        const int h = (n + 1) / 2;
        T v = j < h ? xin(a, rd, 2 * j) : xin(a, rd, 2 * (n - 1 - j) + 1);
        return mk(v, (T)0);
    }
    case M_DCT3: {
        const int m = a.m;
        // V_k = exp(i pi k / 2n) (y_k - i y_{n-k}), y_n = 0
        const int ka = j, kb = m - j;
        cpx<T> Va = cmul(tws[a.pre0 + ka], mk(xin(a, rd, ka), ka == 0 ? (T)0 : -xin(a, rd, n - ka)));
        cpx<T> Vb = cmul(tws[a.pre0 + kb], mk(xin(a, rd, kb), kb == 0 ? (T)0 : -xin(a, rd, n - kb)));
        if (ka == 0) Va.y = (T)0;
        if (kb == m) Vb.y = (T)0;
        cpx<T> B = conjg(Vb);
        cpx<T> E = Va + B, D = Va - B;
        cpx<T> O = cmul(tws[a.pre1 + j], D);
        return conjg(E + mul_pi(O));
    }
    case M_DCT3_ODD: {
        T y = xin(a, rd, j);
        if (j == 0) y = y * (T)0.5;
        return scal(tws[a.pre0 + j], y);
    }
    case M_DCT4:
        return cmul(mk(xin(a, rd, 2 * j), xin(a, rd, n - 1 - 2 * j)), tws[a.pre0 + j]);
    case M_DCT4_ODD:
        return j < n ? scal(tws[a.pre0 + j], xin(a, rd, j)) : mk((T)0, (T)0);
    case M_DCT1: {
        const int Lx = 2 * (n - 1);
        const int i0 = 2 * j, i1 = 2 * j + 1;
        T v0 = i0 <= n - 1 ? xin(a, rd, i0) : xin(a, rd, Lx - i0);
        T v1 = i1 <= n - 1 ? xin(a, rd, i1) : xin(a, rd, Lx - i1);
        return mk(v0, v1);
    }
    case M_DST1: {
        const int i0 = 2 * j, i1 = 2 * j + 1;
        auto o = [&](int i) -> T {
            if (i == 0 || i == n + 1) return (T)0;
            return i <= n ? xin(a, rd, i - 1) : -xin(a, rd, 2 * n + 1 - i);
        };
        return mk(o(i0), o(i1));
    }
    default:
        return mk(xin(a, rd, j), (T)0);
    }
}

template <typename T, int MG, int MODE = -1>
__device__ __forceinline__ cpx<T> pre_value(const StageArgs<T> &a, const cpx<T> *ln, const cpx<T> *tws, int j)
{
    const int n = a.n;
    if constexpr (MG == 1) {
    switch (MODE >= 0 ? MODE : a.mode) {
    case M_R2C_ODD:
        return mk(xin(a, LdsReal<T>{ln}, j), (T)0);
    case M_C2R: {
        const int m = a.m; // n / 2; Nyquist index m
        cpx<T> A = hin(a, ln, j, m);
        cpx<T> B = conjg(hin(a, ln, m - j, m));
        cpx<T> E = A + B, D = A - B;
        cpx<T> O = cmul(tws[a.pre0 + j], D);
        return conjg(E + mul_pi(O));
    }
    case M_C2R_ODD: {
        const int h = (n - 1) / 2;
        cpx<T> v = (j <= h) ? hin(a, ln, j, h + 1) : conjg(hin(a, ln, n - j, h + 1));
        if (n == 1) v.y = (T)0;
        return conjg(v);
    }
    default:
        return L(ln, j);
    }
    } else {
        return pre_real<T, LdsReal<T>, MODE>(a, LdsReal<T>{ln}, tws, j);
    }
}

// X_k of the real sequence whose half-length FFT is Z (k = 0..m).
template <typename T>
__device__ __forceinline__ cpx<T> split(const cpx<T> *Z, const cpx<T> *w, int m, int k)
{
    const int ka = (k == m) ? 0 : k;
    const int kb = (k == 0) ? 0 : m - k;
    cpx<T> A = L(Z, ka), B = conjg(L(Z, kb));
    cpx<T> E = scal(A + B, (T)0.5), D = scal(A - B, (T)0.5);
    return E + mul_mi(cmul(w[k], D));
}

// Real value i of the (optionally conjugated) complex result viewed as reals.
template <typename T>
__device__ __forceinline__ T t_real_view(const cpx<T> *Z, int i, bool conj_result)
{
    cpx<T> v = L(Z, i >> 1);
    if (conj_result) v.y = -v.y;
    return (i & 1) ? v.y : v.x;
}

// Real output y_k of a real-output mode (MODE >= 0: compile-time mode).
template <typename T, int MG, int MODE = -1>
__device__ __forceinline__ T post_real(const StageArgs<T> &a, const cpx<T> *Z, const cpx<T> *tws, int k)
{
    const int n = a.n, m = a.m;
    const int kk = (a.flags & F_POST_REV) ? n - 1 - k : k;
    T y = (T)0;
    if constexpr (MG == 1) {
        if (a.mode == M_C2R) y = t_real_view(Z, kk, true);
        else y = L(Z, kk).x; // M_C2R_ODD
    } else {
    switch (MODE >= 0 ? MODE : a.mode) {
    case M_DCT2: {
        cpx<T> V = kk <= m ? split(Z, tws + a.post0, m, kk) : conjg(split(Z, tws + a.post0, m, n - kk));
        y = (T)2 * cmul(tws[a.post1 + kk], V).x;
        break;
    }
    case M_DCT2_ODD:
        y = (T)2 * cmul(tws[a.post0 + kk], L(Z, kk)).x;
        break;
    case M_DCT3: {
        const int i = (kk & 1) ? n - 1 - (kk >> 1) : (kk >> 1);
        y = t_real_view(Z, i, true);
        break;
    }
    case M_DCT3_ODD: {
        const int i = (kk & 1) ? n - 1 - (kk >> 1) : (kk >> 1);
        y = (T)2 * L(Z, i).x;
        break;
    }
    case M_DCT4: {
        if ((kk & 1) == 0) {
            const int q = kk >> 1;
            y = (T)2 * cmul(L(Z, q), tws[a.post0 + q]).x;
        } else {
            const int q = (n - 1 - kk) >> 1;
            y = (T)-2 * cmul(L(Z, q), tws[a.post0 + q]).y;
        }
        break;
    }
    case M_DCT4_ODD:
        y = (T)2 * cmul(tws[a.post0 + kk], L(Z, kk)).x;
        break;
    case M_DCT1:
        y = split(Z, tws + a.post0, m, kk).x;
        break;
    case M_DST1:
        y = -split(Z, tws + a.post0, m, kk + 1).y;
        break;
    case M_DHT: {
        if (kk <= m) {
            cpx<T> X = split(Z, tws + a.post0, m, kk);
            y = X.x - X.y;
        } else {
            cpx<T> X = split(Z, tws + a.post0, m, n - kk);
            y = X.x + X.y;
        }
        break;
    }
    case M_DHT_ODD: {
        const cpx<T> z = L(Z, kk);
        y = z.x - z.y;
        break;
    }
    default:
        y = (T)0;
    }
    }
    if ((a.flags & F_POST_ALT) && (k & 1)) y = -y;
    if (a.flags & F_WEIGHTS) {
        if (k == 0) y = y * a.postw0;
        if (k == n - 1) y = y * a.postw1;
    }
    return y * a.scale;
}

// Complex output value of a complex-output mode (MODE >= 0: compile-time mode).
template <typename T, int MODE = -1>
__device__ __forceinline__ cpx<T> post_complex(const StageArgs<T> &a, const cpx<T> *Z, const cpx<T> *tws, int k)
{
    cpx<T> v;
    switch (MODE >= 0 ? MODE : a.mode) {
    case M_R2C: {
        v = split(Z, tws + a.post0, a.m, k);
        if (k == 0 || k == a.m) v.y = (T)0; // real-input DC / Nyquist are exactly real
        if (a.flags & F_HERM_DOUBLE) {
            if (k > 0 && k < a.m) v = scal(v, (T)2);
            else v.y = (T)0;
        }
        break;
    }
    case M_R2C_ODD: {
        v = L(Z, k);
        if (k == 0) v.y = (T)0;
        if (a.flags & F_HERM_DOUBLE) {
            if (k > 0) v = scal(v, (T)2);
            else v.y = (T)0;
        }
        break;
    }
    default:
        v = L(Z, k);
        if (a.flags & F_CONJ_OUT) v.y = -v.y;
        break;
    }
    return scal(v, a.scale);
}


// One Stockham radix-R pass in LDS: butterflies j = tl + tpl * b read
// x[j + r m/R], twiddle by exp(-2 pi i r (j mod Ns) / (Ns R)) and write
// y[(j - j mod Ns) R + j mod Ns + r Ns].
template <typename T, int R, bool WIDE>
__device__ __forceinline__ void radix_pass(cpx<T> *ln, const cpx<T> *twp, int m, int Ns, FastDiv nsd, int tl,
                                           int tpl)
{
    constexpr int MB = generic_mb(R, WIDE, sizeof(T) == 4);
    const int nb = m / R;
    cpx<T> v[MB][R];
    int kk[MB];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int j = tl + tpl * b;
        kk[b] = j - (int)fastdiv((uint32_t)j, nsd.mag, nsd.sh) * Ns; // j mod Ns
        if (j < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = L(ln, j + r * nb);
            if (Ns > 1) {
                const int k = kk[b];
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
        if (j < nb) {
            const int k = kk[b];
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) Lset(ln, base + r * Ns, v[b][r]);
        }
    }
    __syncthreads();
}

} // namespace

// Generic radix-R pass for any R (used for prime factors above 13): each
// thread computes whole outputs as O(R) sums from LDS. roots = exp(-2 pi i
// e / R), e < R, followed by the usual Ns (R - 1) pass twiddles.
template <typename T, int E>
__device__ __forceinline__ void radix_pass_generic(cpx<T> *ln, const cpx<T> *roots, int m, int R, int Ns, int tl, int tpl)
{
    const cpx<T> *twp = roots + R;
    const int nb = m / R;
    cpx<T> v[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int o = tl + tpl * i;
        if (o < m) {
            const int r = o / nb, j = o - r * nb;
            const int k = j % Ns;
            cpx<T> acc = mk((T)0, (T)0);
            int e = 0;
            for (int q = 0; q < R; ++q) {
                cpx<T> x = L(ln, j + q * nb);
                if (Ns > 1 && q > 0) x = cmul(x, twp[k * (R - 1) + (q - 1)]);
                acc = acc + cmul(x, roots[e]);
                e += r;
                if (e >= R) e -= R;
            }
            v[i] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int o = tl + tpl * i;
        if (o < m) {
            const int r = o / nb, j = o - r * nb;
            const int k = j % Ns;
            Lset(ln, (j - k) * R + k + r * Ns, v[i]);
        }
    }
    __syncthreads();
}


// ---- load: global -> LDS (raw line values, all loads of a chunk in flight)
// Thread mapping of a tile's load or store: lines ll + i * lstep, i < cnt,
// points p0, p0 + step, ... A column side spreads the tile over all lc *
// tpl threads, consecutive threads on consecutive lines of one point (runs
// of lt elements); a row side goes line by line, tpl threads per line, lc
// lines per round. Threads beyond lc * tpl get nothing.
template <typename T>
__device__ __forceinline__ void tile_mapping(const StageArgs<T> &a, int col, int tid, int l, int tl, int &ll, int &p0,
                                             int &step, int &lstep, int &cnt, bool pair = false)
{
    const int lt = a.lt, lc = a.lc, tpl = a.tpl;
    if (col) {
        const int rows = lc * tpl / lt;
        if (pair) {
            // lines 2i and 2i + 1 per lane (paired column accesses)
            const int hl = lt >> 1;
            ll = 2 * (tid % hl);
            p0 = tid / hl;
            step = 2 * rows;
        } else {
            ll = tid % lt;
            p0 = tid / lt;
            step = rows;
        }
        lstep = 0;
        cnt = tid < rows * lt ? 1 : 0;
    } else {
        ll = l;
        p0 = tl;
        step = tpl;
        lstep = lc;
        cnt = tid < lc * tpl ? lt / lc : 0;
    }
}

// Real-input R2R modes whose FFT input is a signed permutation of the line
// (plus zeros): DST-I (odd extension of length 2n + 2), DCT-II (even
// points forward, odd points backward) and DCT-IV (pairs x_2j, x_n-1-2j,
// the pre-twiddle applied by the first pass). Their compile-time-mode
// kernels put each loaded value straight into its FFT-input slot(s).
template <int MODE>
constexpr bool scatter_mode()
{
    return MODE == M_DST1 || MODE == M_DCT2 || MODE == M_DCT4;
}

// Point p of a raw real line into the FFT input of a scatter mode, after the
// stage's input reversal / alternating sign / edge weights (the values the
// gathered pre-processing reads through xin).
template <typename T, int MODE>
__device__ __forceinline__ void scatter_real(const StageArgs<T> &a, cpx<T> *lln, int p, T v)
{
    const int n = a.n;
    int q = p;
    if (a.flags & (F_PRE_REV | F_PRE_ALT | F_WEIGHTS)) {
        if ((a.flags & F_PRE_ALT) && (p & 1)) v = -v;
        if (a.flags & F_WEIGHTS) {
            if (p == 0) v = v * a.prew0;
            if (p == n - 1) v = v * a.prew1;
        }
        if (a.flags & F_PRE_REV) q = n - 1 - p;
    }
    if constexpr (MODE == M_DST1) {
        // o_{q+1} = x_q, o_{2n+1-q} = -x_q (o_0 = o_{n+1} = 0)
        LRset(lln, q + 1, v);
        LRset(lln, 2 * n + 1 - q, -v);
    } else if constexpr (MODE == M_DCT2) {
        // v_i = x_2i (i < n / 2), v_i = x_{2(n-1-i)+1} (i >= n / 2)
        LRset(lln, (q & 1) ? n - 1 - (q >> 1) : (q >> 1), v);
    } else {
        // z_j = (x_2j, x_{n-1-2j}): real slot 2j or 2j + 1
        LRset(lln, (q & 1) ? n - q : q, v);
    }
}

// Line ll of the tile, points p0, p0 + step, ... : global -> LDS (the FFT
// input of a scatter mode when MODE is one).
template <typename T, bool NT, int MODE = -1>
__device__ __forceinline__ void load_line(const StageArgs<T> &a, cpx<T> *data, int64_t tile, int ll, int p0, int step)
{
    const bool cin = complex_input(a.mode);
    const int64_t line = tile * a.lt + ll;
    if (line < a.nlines) {
        int64_t io, oo;
        line_offsets(a.g, line, io, oo);
        cpx<T> *lln = data + ll * a.ls;
        if (cin && (a.flags & F_XIN)) {
            // piece of a long line: point p is index f = p * xmul_in + d of
            // the long line; indexes from xvalid_in on are zero padding
            const cpx<T> *s = reinterpret_cast<const cpx<T> *>(a.src) + io;
            const bool cj = (a.flags & F_CONJ_IN) != 0;
            const int dco = a.xdig >= 0 ? digit_coord(a.g, line, a.xdig) : 0;
            for (int base = p0; base < a.nin; base += step * Regs<T>::gunroll) {
                cpx<T> v[Regs<T>::gunroll];
#pragma unroll
                for (int u = 0; u < Regs<T>::gunroll; ++u) {
                    const int p = base + u * step;
                    v[u] = mk((T)0, (T)0);
                    if (p < a.nin && p * a.xmul_in + dco < a.xvalid_in) v[u] = gld<false>(s + (int64_t)p * a.ips);
                }
#pragma unroll
                for (int u = 0; u < Regs<T>::gunroll; ++u) {
                    const int p = base + u * step;
                    if (p < a.nin) {
                        cpx<T> w = v[u];
                        const int f = p * a.xmul_in + dco;
                        if (f < a.xvalid_in) {
                            if (cj) w.y = -w.y;
                            if (a.xtab_in) w = cmul(w, a.xtab_in[f]);
                        }
                        Lset(lln, p, w);
                    }
                }
            }
        } else if (cin) {
            const cpx<T> *s = reinterpret_cast<const cpx<T> *>(a.src) + io;
            const bool cj = (a.flags & F_CONJ_IN) != 0;
            for (int base = p0; base < a.nin; base += step * Regs<T>::gunroll) {
                cpx<T> v[Regs<T>::gunroll];
#pragma unroll
                for (int u = 0; u < Regs<T>::gunroll; ++u) {
                    const int p = base + u * step;
                    if (p < a.nin) v[u] = gld<NT>(s + (int64_t)p * a.ips);
                }
#pragma unroll
                for (int u = 0; u < Regs<T>::gunroll; ++u) {
                    const int p = base + u * step;
                    if (p < a.nin) {
                        cpx<T> w = v[u];
                        if (cj) w.y = -w.y;
                        Lset(lln, p, w);
                    }
                }
            }
        } else if (a.in_col && (a.flags & F_CPAIR_IN)) {
            // lines ll and ll + 1: neighbouring elements, one 2-element load
            const cpx<T> *s2 = reinterpret_cast<const cpx<T> *>(reinterpret_cast<const T *>(a.src) + io);
            const int64_t ips2 = a.ips >> 1;
            cpx<T> *lln1 = lln + a.ls;
            if constexpr (MODE == M_DST1) {
                if (p0 == 0) {
                    LRset(lln, 0, (T)0);
                    LRset(lln, a.n + 1, (T)0);
                    LRset(lln1, 0, (T)0);
                    LRset(lln1, a.n + 1, (T)0);
                }
            }
            for (int base = p0; base < a.nin; base += step * Regs<T>::gunroll) {
                cpx<T> v[Regs<T>::gunroll];
#pragma unroll
                for (int u = 0; u < Regs<T>::gunroll; ++u) {
                    const int p = base + u * step;
                    if (p < a.nin) v[u] = gld<NT>(s2 + (int64_t)p * ips2);
                }
#pragma unroll
                for (int u = 0; u < Regs<T>::gunroll; ++u) {
                    const int p = base + u * step;
                    if (p < a.nin) {
                        if constexpr (scatter_mode<MODE>()) {
                            scatter_real<T, MODE>(a, lln, p, v[u].x);
                            scatter_real<T, MODE>(a, lln1, p, v[u].y);
                        } else {
                            LRset(lln, p, v[u].x);
                            LRset(lln1, p, v[u].y);
                        }
                    }
                }
            }
        } else if (!a.in_col && (a.flags & F_RPAIR_IN)) {
            // points 2i and 2i + 1 of the row in one 2-element load
            const cpx<T> *s2 = reinterpret_cast<const cpx<T> *>(reinterpret_cast<const T *>(a.src) + io);
            const int np = a.nin >> 1;
            if constexpr (MODE == M_DST1) {
                if (p0 == 0) {
                    LRset(lln, 0, (T)0);
                    LRset(lln, a.n + 1, (T)0);
                }
            }
            for (int base = p0; base < np; base += step * Regs<T>::gunroll) {
                cpx<T> v[Regs<T>::gunroll];
#pragma unroll
                for (int u = 0; u < Regs<T>::gunroll; ++u) {
                    const int i = base + u * step;
                    if (i < np) v[u] = gld<NT>(s2 + i);
                }
#pragma unroll
                for (int u = 0; u < Regs<T>::gunroll; ++u) {
                    const int i = base + u * step;
                    if (i < np) {
                        if constexpr (scatter_mode<MODE>()) {
                            scatter_real<T, MODE>(a, lln, 2 * i, v[u].x);
                            scatter_real<T, MODE>(a, lln, 2 * i + 1, v[u].y);
                        } else {
                            Lset(lln, i, v[u]);
                        }
                    }
                }
            }
        } else {
            const T *s = reinterpret_cast<const T *>(a.src) + io;
            if constexpr (MODE == M_DST1) {
                if (p0 == 0) {
                    LRset(lln, 0, (T)0);
                    LRset(lln, a.n + 1, (T)0);
                }
            }
            for (int base = p0; base < a.nin; base += step * (2 * Regs<T>::gunroll)) {
                T v[2 * Regs<T>::gunroll];
#pragma unroll
                for (int u = 0; u < 2 * Regs<T>::gunroll; ++u) {
                    const int p = base + u * step;
                    if (p < a.nin) v[u] = gld<NT>(s + (int64_t)p * a.ips);
                }
#pragma unroll
                for (int u = 0; u < 2 * Regs<T>::gunroll; ++u) {
                    const int p = base + u * step;
                    if (p < a.nin) {
                        if constexpr (scatter_mode<MODE>()) scatter_real<T, MODE>(a, lln, p, v[u]);
                        else LRset(lln, p, v[u]);
                    }
                }
            }
        }
    }
}

// Tile load: global -> LDS. lc lines are computed at a time (lc * tpl
// threads) and the tile holds lt = lc * rounds lines. Threads beyond
// lc * tpl only join the barrier.
template <typename T, bool NT, int MODE = -1>
__device__ __forceinline__ void stage_load(const StageArgs<T> &a, cpx<T> *data, int64_t tile, int tid, int l, int tl)
{
    int ll, p0, step, lstep, cnt;
    tile_mapping(a, a.in_col, tid, l, tl, ll, p0, step, lstep, cnt, a.in_col && (a.flags & F_CPAIR_IN));
#pragma unroll 1
    for (int i = 0; i < cnt; ++i) {
        // an opaque copy of the first point per line keeps the compiler from
        // hoisting per-point work out of the loop (and spilling it)
        int p = p0;
        asm volatile("" : "+v"(p));
        load_line<T, NT, MODE>(a, data, tile, ll + i * lstep, p, step);
    }
    __syncthreads();
}

// ---- post-processing + store: LDS -> global (output layout)
// Line lo of the tile, points k0, k0 + tpl, ... : post-processing, LDS -> global.
template <typename T, int MG, bool NT, int MODE = -1>
__device__ __forceinline__ void store_line(const StageArgs<T> &a, const cpx<T> *data, const cpx<T> *tws,
                                           int64_t tile, int lo, int k0, int tpl)
{
    {
        const int64_t line = tile * a.lt + lo;
        if (line < a.nlines) {
            int64_t io, oo;
            line_offsets(a.g, line, io, oo);
            const cpx<T> *Z = data + lo * a.ls;
            constexpr int U = Regs<T>::gunroll;
            if constexpr (MG == 0) {
            if (a.mode == M_COPY_C) {
                // identity stage: move the bits, no arithmetic (bit-exact permutation)
                cpx<T> *d = reinterpret_cast<cpx<T> *>(a.dst) + oo;
                for (int base = k0; base < a.nout; base += tpl * U) {
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = base + u * tpl;
                        if (k < a.nout) gst<NT>(d + (int64_t)k * a.ops, L(Z, k));
                    }
                }
            } else if (a.mode == M_COPY_R) {
                T *d = reinterpret_cast<T *>(a.dst) + oo;
                for (int base = k0; base < a.nout; base += tpl * U) {
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = base + u * tpl;
                        if (k < a.nout) gst<NT>(d + (int64_t)k * a.ops, LR(Z, k));
                    }
                }
            }
            } else {
            if (MG == 1 && complex_output(a.mode) && (a.flags & (F_XOUT | F_XTW))) {
                // piece of a long line: twiddle by d * k, conjugate, scale,
                // then the factor of index f = k * xmul_out + d; indexes from
                // xvalid_out on are not stored
                cpx<T> *d = reinterpret_cast<cpx<T> *>(a.dst) + oo;
                const int dco = a.xdig >= 0 ? digit_coord(a.g, line, a.xdig) : 0;
                const int lo_mask = (1 << a.xtw_bsh) - 1;
                for (int base = k0; base < a.nout; base += tpl * U) {
                    cpx<T> v[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = base + u * tpl;
                        if (k < a.nout) {
                            cpx<T> w = L(Z, k);
                            if (a.flags & F_XTW) {
                                const int q = dco * k; // < xtw_m: d < N2 and k < N1
                                w = cmul(w, cmul(a.xtw[q & lo_mask], a.xtw[(1 << a.xtw_bsh) + (q >> a.xtw_bsh)]));
                            }
                            if (a.flags & F_CONJ_OUT) w.y = -w.y;
                            w = scal(w, a.scale);
                            const int f = k * a.xmul_out + dco;
                            if ((a.flags & F_XOUT) && a.xtab_out && f < a.xvalid_out) w = cmul(w, a.xtab_out[f]);
                            v[u] = w;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = base + u * tpl;
                        if (k < a.nout && (!(a.flags & F_XOUT) || k * a.xmul_out + dco < a.xvalid_out))
                            gst<false>(d + (int64_t)k * a.ops, v[u]);
                    }
                }
            } else if (MG == 1 && complex_output(a.mode)) {
                cpx<T> *d = reinterpret_cast<cpx<T> *>(a.dst) + oo;
                for (int base = k0; base < a.nout; base += tpl * U) {
                    cpx<T> v[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = base + u * tpl;
                        if (k < a.nout) v[u] = post_complex<T, MODE>(a, Z, tws, k);
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = base + u * tpl;
                        if (k < a.nout) gst<NT>(d + (int64_t)k * a.ops, v[u]);
                    }
                }
            } else if (MG == 2 && !a.out_col && (a.flags & F_RPAIR_OUT)) {
                // outputs 2i and 2i + 1 of the row in one 2-element store
                cpx<T> *d2 = reinterpret_cast<cpx<T> *>(reinterpret_cast<T *>(a.dst) + oo);
                const int np = a.nout >> 1;
                for (int base = k0; base < np; base += tpl * U) {
                    cpx<T> v[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int i = base + u * tpl;
                        if (i < np)
                            v[u] = mk(post_real<T, MG, MODE>(a, Z, tws, 2 * i), post_real<T, MG, MODE>(a, Z, tws, 2 * i + 1));
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int i = base + u * tpl;
                        if (i < np) gst<NT>(d2 + i, v[u]);
                    }
                }
            } else {
                T *d = reinterpret_cast<T *>(a.dst) + oo;
                for (int base = k0; base < a.nout; base += tpl * U) {
                    T v[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = base + u * tpl;
                        if (k < a.nout) v[u] = post_real<T, MG, MODE>(a, Z, tws, k);
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = base + u * tpl;
                        if (k < a.nout) gst<NT>(d + (int64_t)k * a.ops, v[u]);
                    }
                }
            }
            }
        }
    }
}

// Tile store, LDS -> global: the thread mapping of stage_load on the
// output side.
template <typename T, int MG, bool NT, int MODE = -1>
__device__ __forceinline__ void stage_store(const StageArgs<T> &a, const cpx<T> *data, const cpx<T> *tws,
                                            int64_t tile, int tid, int l, int tl)
{
    int lo, k0, step, lstep, cnt;
    tile_mapping(a, a.out_col, tid, l, tl, lo, k0, step, lstep, cnt);
#pragma unroll 1
    for (int i = 0; i < cnt; ++i) {
        int k = k0;
        asm volatile("" : "+v"(k));
        store_line<T, MG, NT, MODE>(a, data, tws, tile, lo + i * lstep, k, step);
    }
}

// Lines lo and lo + 1 of the tile (neighbouring output elements), points
// k0, k0 + step, ...: real post-processing of both, one 2-element store.
template <typename T, int MG, bool NT, int MODE = -1>
__device__ __forceinline__ void store_line_pair(const StageArgs<T> &a, const cpx<T> *data, const cpx<T> *tws,
                                                int64_t tile, int lo, int k0, int step)
{
    const int64_t line = tile * a.lt + lo;
    if (line >= a.nlines) return;
    int64_t io, oo;
    line_offsets(a.g, line, io, oo);
    const cpx<T> *Z0 = data + lo * a.ls;
    const cpx<T> *Z1 = Z0 + a.ls;
    cpx<T> *d = reinterpret_cast<cpx<T> *>(reinterpret_cast<T *>(a.dst) + oo);
    const int64_t ops2 = a.ops >> 1;
    constexpr int U = Regs<T>::gunroll;
    for (int base = k0; base < a.nout; base += step * U) {
        cpx<T> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = base + u * step;
            if (k < a.nout) v[u] = mk(post_real<T, MG, MODE>(a, Z0, tws, k), post_real<T, MG, MODE>(a, Z1, tws, k));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = base + u * step;
            if (k < a.nout) gst<NT>(d + (int64_t)k * ops2, v[u]);
        }
    }
}

template <typename T, bool WIDE>
__device__ __forceinline__ void run_pass(int R, cpx<T> *ln, const cpx<T> *twp, int m, int Ns, FastDiv nsd, int tl,
                                         int tpl)
{
    switch (R) {
    case 2: radix_pass<T, 2, WIDE>(ln, twp, m, Ns, nsd, tl, tpl); break;
    case 3: radix_pass<T, 3, WIDE>(ln, twp, m, Ns, nsd, tl, tpl); break;
    case 4: radix_pass<T, 4, WIDE>(ln, twp, m, Ns, nsd, tl, tpl); break;
    case 5: radix_pass<T, 5, WIDE>(ln, twp, m, Ns, nsd, tl, tpl); break;
    case 8: radix_pass<T, 8, WIDE>(ln, twp, m, Ns, nsd, tl, tpl); break;
    case 16:
        if constexpr (Regs<T>::max_radix >= 16) radix_pass<T, 16, WIDE>(ln, twp, m, Ns, nsd, tl, tpl);
        break;
    default:
        if constexpr (WIDE) {
            if (R == 7) radix_pass<T, 7, WIDE>(ln, twp, m, Ns, nsd, tl, tpl);
            else if (R == 11) radix_pass<T, 11, WIDE>(ln, twp, m, Ns, nsd, tl, tpl);
            else if (R == 13) radix_pass<T, 13, WIDE>(ln, twp, m, Ns, nsd, tl, tpl);
            else radix_pass_generic<T, Regs<T>::emax>(ln, twp, m, R, Ns, tl, tpl);
        }
        break;
    }
}

// WIDE instantiations also carry the radix 7 / 11 / 13 butterflies and the
// generic pass; the narrow one keeps register pressure down for the
// power-of-two and 3 / 5 lengths (and adds radix 32 for f32).
// One tile of a generic stage: lt lines, computed lc at a time. Threads
// beyond lc * tpl (a batched launch sizes its blocks for the largest item)
// only take part in the barriers.
template <typename T, bool WIDE, int MG, int MODE = -1>
__device__ __forceinline__ void stage_body(const StageArgs<T> &a, int64_t tile, const cpx<T> *tws, cpx<T> *data,
                                           const int *radix, const int *twoff, const uint32_t *nsmag,
                                           const int *nssh)
{
    const int tid = threadIdx.x;
    const int lc = a.lc, tpl = a.tpl;
    const bool act = tid < lc * tpl;
    // pass mapping: lc lines x tpl threads per round; inactive threads get
    // an index beyond every line so that each pass skips them
    const int l = act ? tid / tpl : 0;
    const int tl = act ? tid % tpl : (1 << 28);
    stage_load<T, false>(a, data, tile, tid, l, tl);
    if constexpr (MG == 0) {
        stage_store<T, 0, false>(a, data, tws, tile, tid, l, tl);
        return;
    }
#pragma unroll 1
    for (int r = 0; r < a.lt; r += lc) {
        // Opaque per round: without it the compiler hoists the rounds'
        // common tl-derived table reads and indexes out of the loop and keeps
        // them live through every pass (hundreds of bytes of spills a lane).
        int tlr = tl;
        asm volatile("" : "+v"(tlr));
        const int tl = tlr;
        cpx<T> *ln = data + (l + r) * a.ls;
        // ---- pre-processing: raw line -> FFT input (in place through registers)
        if (needs_pre(a.mode)) {
            constexpr int E = generic_pre_e(WIDE, sizeof(T) == 4);
            cpx<T> z[E];
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const int j = tl + tpl * i;
                if (j < a.m) z[i] = pre_value<T, MG, MODE>(a, ln, tws, j);
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const int j = tl + tpl * i;
                if (j < a.m) Lset(ln, j, z[i]);
            }
            __syncthreads();
        }
        // ---- Stockham passes in LDS (decimation in time, natural-order output)
        int Ns = 1;
        for (int p = 0; p < a.npass; ++p) {
            const int R = radix[p];
            run_pass<T, WIDE>(R, ln, tws + twoff[p], a.m, Ns, FastDiv{nsmag[p], nssh[p]}, tl, tpl);
            Ns *= R;
        }
    }
    stage_store<T, MG, false, MODE>(a, data, tws, tile, tid, l, tl);
}

template <typename T, bool WIDE, int MG, int MODE = -1>
__global__ __launch_bounds__(WIDE ? kMaxThreadsWide : kMaxThreads) void k_line_stage(const StageArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *tws = reinterpret_cast<cpx<T> *>(smem_raw);
    cpx<T> *data = tws + a.twn;
    if (a.gate != nullptr && *a.gate != 0) return;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    int64_t tile = blockIdx.x;
    if (a.flags & F_XCD_REMAP) {
        tile = (int64_t)(blockIdx.x % 8u) * a.xcd_per + blockIdx.x / 8u;
        if (tile >= a.ntiles) return;
    }
    for (int i = tid; i < a.twn; i += nthr) tws[i] = a.tw[i];
    // pass plan in LDS (read with a runtime pass index; indexing the argument
    // struct dynamically would force a private copy of it)
    __shared__ int s_radix[kMaxPasses], s_twoff[kMaxPasses], s_nssh[kMaxPasses];
    __shared__ uint32_t s_nsmag[kMaxPasses];
    if (tid == 0) {
#pragma unroll
        for (int p = 0; p < kMaxPasses; ++p) {
            s_radix[p] = a.radix[p];
            s_twoff[p] = a.twoff[p];
            s_nsmag[p] = a.nsmag[p];
            s_nssh[p] = a.nssh[p];
        }
    }
    stage_body<T, WIDE, MG, MODE>(a, tile, tws, data, s_radix, s_twoff, s_nsmag, s_nssh);
}

// Persistent batched stage: the blocks of one launch walk a work list of
// (item, tile) pairs, item = one route stage of one problem (its arguments
// live in device memory), prefix[q] = first work index of item q. Each block
// keeps its item index and only moves forward.
template <typename T, bool WIDE>
__global__ __launch_bounds__(WIDE ? kMaxThreadsWide : kMaxThreads) void k_line_batch(const StageArgs<T> *__restrict__ items,
                                                                const int64_t *__restrict__ prefix, int nitems,
                                                                int64_t total)
{
    // LDS per tile: the item's twn table entries, then its lt lines
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *tws = reinterpret_cast<cpx<T> *>(smem_raw);
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    int q = 0;
    {
        // first item of this block: binary search
        const int64_t w0 = blockIdx.x;
        int lo = 0, hi = nitems;
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (prefix[mid] <= w0) lo = mid;
            else hi = mid;
        }
        q = lo;
    }
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        while (q + 1 < nitems && prefix[q + 1] <= w) ++q;
        const StageArgs<T> &a = items[q];
        const int64_t tile = w - prefix[q];
        __syncthreads(); // the previous tile is done with LDS
        for (int i = tid; i < a.twn; i += nthr) tws[i] = a.tw[i];
        stage_body<T, WIDE, 1>(a, tile, tws, tws + a.twn, a.radix, a.twoff, a.nsmag, a.nssh);
    }
}

// ---------------------------------------------------------------------------
// Fast path: power-of-two C2C / R2C / C2R stages with a compile-time radix
// sequence (R0, R1[, R2]) and EM complex registers per thread and pass. The
// first pass of a C2C / R2C stage reads global memory straight into
// registers and the last pass of a C2C / C2R stage writes global memory
// straight from registers, so a two-pass line goes through LDS once.
// ---------------------------------------------------------------------------

template <int EM, int R>
constexpr int fast_mb()
{
    return (EM / R) > 0 ? EM / R : 1;
}

// first pass (Ns = 1) from global memory into LDS
template <typename T, int R, int EM, int M, bool NT>
__device__ __forceinline__ void fast_first(const StageArgs<T> &a, cpx<T> *ln, bool valid, int64_t io, int tl, int tpl,
                                           bool cplx, bool cj)
{
    constexpr int MB = fast_mb<EM, R>();
    constexpr int NB = M / R;
    cpx<T> v[MB][R];
    if (valid) {
        if (cplx) {
            const cpx<T> *s = reinterpret_cast<const cpx<T> *>(a.src) + io;
#pragma unroll
            for (int b = 0; b < MB; ++b) {
                const int j = tl + tpl * b;
                if (j < NB) {
#pragma unroll
                    for (int r = 0; r < R; ++r) v[b][r] = gld<NT>(s + (int64_t)(j + r * NB) * a.ips);
                }
            }
        } else if (a.ips == 1) {
            const cpx<T> *s = reinterpret_cast<const cpx<T> *>(reinterpret_cast<const T *>(a.src) + io);
#pragma unroll
            for (int b = 0; b < MB; ++b) {
                const int j = tl + tpl * b;
                if (j < NB) {
#pragma unroll
                    for (int r = 0; r < R; ++r) v[b][r] = gld<NT>(s + j + r * NB);
                }
            }
        } else {
            const T *s = reinterpret_cast<const T *>(a.src) + io;
#pragma unroll
            for (int b = 0; b < MB; ++b) {
                const int j = tl + tpl * b;
                if (j < NB) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int64_t q = 2 * (int64_t)(j + r * NB);
                        v[b][r] = mk(gld<NT>(s + q * a.ips), gld<NT>(s + (q + 1) * a.ips));
                    }
                }
            }
        }
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            const int j = tl + tpl * b;
            if (j < NB) {
                if (cj) {
#pragma unroll
                    for (int r = 0; r < R; ++r) v[b][r].y = -v[b][r].y;
                }
                Dft<T, R>::run(v[b]);
#pragma unroll
                for (int r = 0; r < R; ++r) Lset(ln, j * R + r, v[b][r]);
            }
        }
    }
    __syncthreads();
}

// First pass of a real-input R2R line whose raw values are already in the
// LDS line (stage_load): the pre-processing gathers each FFT input value from
// LDS (reordering, odd / even extension, pre-twiddles), then, after every
// thread of the tile has read its values, the radix-R0 butterflies write the
// pass output over the raw line.
template <typename T, int R, int EM, int M, int MODE = -1>
__device__ __forceinline__ void fast_first_pre_lds(const StageArgs<T> &a, cpx<T> *ln, const cpx<T> *tws, int tl,
                                                   int tpl)
{
    constexpr int MB = fast_mb<EM, R>();
    constexpr int NB = M / R;
    cpx<T> v[MB][R];
    const LdsReal<T> rd{ln};
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int j = tl + tpl * b;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = pre_real<T, LdsReal<T>, MODE>(a, rd, tws, j + r * NB);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int j = tl + tpl * b;
        if (j < NB) {
            Dft<T, R>::run(v[b]);
#pragma unroll
            for (int r = 0; r < R; ++r) Lset(ln, j * R + r, v[b][r]);
        }
    }
    __syncthreads();
}

// a pass in LDS with compile-time geometry
template <typename T, int R, int EM, int M>
__device__ __forceinline__ void fast_mid(cpx<T> *ln, const cpx<T> *twp, int Ns, int tl, int tpl)
{
    constexpr int MB = fast_mb<EM, R>();
    constexpr int NB = M / R;
    cpx<T> v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int j = tl + tpl * b;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = L(ln, j + r * NB);
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
            for (int r = 0; r < R; ++r) Lset(ln, base + r * Ns, v[b][r]);
        }
    }
    __syncthreads();
}

// first pass (Ns = 1) in LDS, each input value j times pre[j] (the DCT-IV
// pre-twiddle of a scattered line)
template <typename T, int R, int EM, int M>
__device__ __forceinline__ void fast_first_tw(cpx<T> *ln, const cpx<T> *pre, int tl, int tpl)
{
    constexpr int MB = fast_mb<EM, R>();
    constexpr int NB = M / R;
    cpx<T> v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int j = tl + tpl * b;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = cmul(L(ln, j + r * NB), pre[j + r * NB]);
            Dft<T, R>::run(v[b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int j = tl + tpl * b;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) Lset(ln, j * R + r, v[b][r]);
        }
    }
    __syncthreads();
}

// last pass (Ns = M / R) from LDS straight to global memory: complex output
// (C2C, conjugation and scale applied) or the real pairs of a C2R line
// (y_2k = Re conj Z_k, y_2k+1 = Im conj Z_k).
template <typename T, int R, int EM, int M, bool NT>
__device__ __forceinline__ void fast_last(const StageArgs<T> &a, const cpx<T> *ln, const cpx<T> *twp, bool valid,
                                          int64_t oo, int tl, int tpl, bool c2r)
{
    constexpr int MB = fast_mb<EM, R>();
    constexpr int NB = M / R;
    if (!valid) return;
    cpx<T> v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int j = tl + tpl * b;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = L(ln, j + r * NB);
            if (NB > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], twp[j * (R - 1) + (r - 1)]);
            }
            Dft<T, R>::run(v[b]);
        }
    }
    const T sc = a.scale;
    if (!c2r) {
        cpx<T> *d = reinterpret_cast<cpx<T> *>(a.dst) + oo;
        const bool cj = (a.flags & F_CONJ_OUT) != 0;
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            const int j = tl + tpl * b;
            if (j < NB) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    cpx<T> w = v[b][r];
                    if (cj) w.y = -w.y;
                    gst<NT>(d + (int64_t)(j + r * NB) * a.ops, scal(w, sc));
                }
            }
        }
    } else if (a.ops == 1) {
        cpx<T> *d = reinterpret_cast<cpx<T> *>(reinterpret_cast<T *>(a.dst) + oo);
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            const int j = tl + tpl * b;
            if (j < NB) {
#pragma unroll
                for (int r = 0; r < R; ++r) gst<NT>(d + j + r * NB, mk(v[b][r].x * sc, -v[b][r].y * sc));
            }
        }
    } else {
        T *d = reinterpret_cast<T *>(a.dst) + oo;
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            const int j = tl + tpl * b;
            if (j < NB) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int64_t q = 2 * (int64_t)(j + r * NB);
                    gst<NT>(d + q * a.ops, v[b][r].x * sc);
                    gst<NT>(d + (q + 1) * a.ops, -v[b][r].y * sc);
                }
            }
        }
    }
}

// C2R load + merge in one step. The half spectrum X_0..X_M of a 2M-point
// real line becomes the M-point complex input conj(z_j), z_j = E_j + i O_j
// with E_j = A + B, O_j = w_j (A - B), A = X_j, B = conj(X_{M-j}) (z_j =
// x_2j + i x_2j+1 after the positive-sign transform). Each thread loads the
// pair X_q, X_{M-q} (q = 0 .. M/2; every input point is loaded once, the
// tile's lanes still reading 256-byte runs) and writes conj(z_q) and
// conj(z_{M-q}) straight to LDS, so the raw half spectrum never goes
// through LDS.
template <typename T, int M, bool NT>
__device__ __forceinline__ void fast_c2r_load(const StageArgs<T> &a, cpx<T> *data, const cpx<T> *tws, int64_t tile,
                                              int tid, int l, int tl)
{
    const int lt = a.lt, tpl = a.tpl;
    if (tid < lt * tpl) {
        int ll, p0;
        if (a.in_col) {
            ll = tid % lt;
            p0 = tid / lt;
        } else {
            ll = l;
            p0 = tl;
        }
        const int64_t line = tile * lt + ll;
        if (line < a.nlines) {
            int64_t io, oo;
            line_offsets(a.g, line, io, oo);
            cpx<T> *lln = data + ll * a.ls;
            const cpx<T> *s = reinterpret_cast<const cpx<T> *>(a.src) + io;
            const bool cj = (a.flags & F_CONJ_IN) != 0;
            const bool half = (a.flags & F_HERM_HALF) != 0;
            constexpr int U = Regs<T>::unroll / 2;
            constexpr int Q = M / 2; // last pair index (self-paired)
            for (int base = p0; base <= Q; base += tpl * U) {
                cpx<T> x[U], y[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int q = base + u * tpl;
                    if (q <= Q) {
                        x[u] = gld<NT>(s + (int64_t)q * a.ips);
                        y[u] = gld<NT>(s + (int64_t)(M - q) * a.ips);
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int q = base + u * tpl;
                    if (q <= Q) {
                        cpx<T> xq = x[u], yq = y[u];
                        if (cj) {
                            xq.y = -xq.y;
                            yq.y = -yq.y;
                        }
                        if (q == 0) { // X_0 and X_M: imaginary parts dropped
                            xq.y = (T)0;
                            yq.y = (T)0;
                        } else if (half) { // adjoint weights on interior bins
                            xq = scal(xq, (T)0.5);
                            yq = scal(yq, (T)0.5);
                        }
                        {
                            const cpx<T> A = xq, B = conjg(yq);
                            const cpx<T> E = A + B, D = A - B;
                            Lset(lln, q, conjg(E + mul_pi(cmul(tws[a.pre0 + q], D))));
                        }
                        if (q > 0 && q < Q) {
                            const cpx<T> A = yq, B = conjg(xq);
                            const cpx<T> E = A + B, D = A - B;
                            Lset(lln, M - q, conjg(E + mul_pi(cmul(tws[a.pre0 + M - q], D))));
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
}

// R2C split + store of X_0..X_M from the half-length transform in LDS.
template <typename T, bool NT>
__device__ __forceinline__ void fast_r2c_store(const StageArgs<T> &a, const cpx<T> *data, const cpx<T> *tws,
                                               int64_t tile, int tid, int l, int tl)
{
    const int lt = a.lt, tpl = a.tpl;
    int lo, k0;
    if (a.out_col) {
        lo = tid % lt;
        k0 = tid / lt;
    } else {
        lo = l;
        k0 = tl;
    }
    const int64_t line = tile * lt + lo;
    if (line >= a.nlines) return;
    int64_t io, oo;
    line_offsets(a.g, line, io, oo);
    const cpx<T> *Z = data + lo * a.ls;
    cpx<T> *d = reinterpret_cast<cpx<T> *>(a.dst) + oo;
    const cpx<T> *w = tws + a.post0;
    constexpr int U = Regs<T>::unroll;
    const bool dbl = (a.flags & F_HERM_DOUBLE) != 0;
    for (int base = k0; base < a.nout; base += tpl * U) {
        cpx<T> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = base + u * tpl;
            if (k < a.nout) {
                cpx<T> x = split(Z, w, a.m, k);
                if (k == 0 || k == a.m) x.y = (T)0;
                if (dbl && k > 0 && k < a.m) x = scal(x, (T)2);
                v[u] = scal(x, a.scale);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = base + u * tpl;
            if (k < a.nout) gst<NT>(d + (int64_t)k * a.ops, v[u]);
        }
    }
}

// DHT store of an even line from the half-length transform of its real
// pairs in LDS: H_k = Re X_k - Im X_k, with X_k = conj(X_{n-k}) above the
// Nyquist index m (the generic kernel's post-processing of M_DHT).
template <typename T, bool NT>
__device__ __forceinline__ void fast_dht_store(const StageArgs<T> &a, const cpx<T> *data, const cpx<T> *tws,
                                               int64_t tile, int tid, int l, int tl)
{
    const int lt = a.lt, tpl = a.tpl;
    const cpx<T> *w = tws + a.post0;
    const int m = a.m;
    constexpr int U = Regs<T>::unroll;
    if (a.out_col && (a.flags & F_CPAIR_OUT)) {
        // lines lo and lo + 1: neighbouring elements, one 2-element store
        const int hl = lt >> 1;
        const int lo = 2 * (tid % hl), k0 = tid / hl, step = 2 * tpl;
        const int64_t line = tile * lt + lo;
        if (line >= a.nlines) return;
        int64_t io, oo;
        line_offsets(a.g, line, io, oo);
        const cpx<T> *Z0 = data + lo * a.ls;
        const cpx<T> *Z1 = Z0 + a.ls;
        cpx<T> *d = reinterpret_cast<cpx<T> *>(reinterpret_cast<T *>(a.dst) + oo);
        const int64_t ops2 = a.ops >> 1;
        for (int base = k0; base < a.nout; base += step * U) {
            cpx<T> v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = base + u * step;
                if (k < a.nout) {
                    const bool low = k <= m;
                    const int kk = low ? k : 2 * m - k;
                    const cpx<T> x0 = split(Z0, w, m, kk), x1 = split(Z1, w, m, kk);
                    v[u] = scal(low ? mk(x0.x - x0.y, x1.x - x1.y) : mk(x0.x + x0.y, x1.x + x1.y), a.scale);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = base + u * step;
                if (k < a.nout) gst<NT>(d + (int64_t)k * ops2, v[u]);
            }
        }
        return;
    }
    int lo, k0;
    if (a.out_col) {
        lo = tid % lt;
        k0 = tid / lt;
    } else {
        lo = l;
        k0 = tl;
    }
    const int64_t line = tile * lt + lo;
    if (line >= a.nlines) return;
    int64_t io, oo;
    line_offsets(a.g, line, io, oo);
    const cpx<T> *Z = data + lo * a.ls;
    T *d = reinterpret_cast<T *>(a.dst) + oo;
    for (int base = k0; base < a.nout; base += tpl * U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = base + u * tpl;
            if (k < a.nout) {
                const bool low = k <= m;
                const cpx<T> x = split(Z, w, m, low ? k : 2 * m - k);
                v[u] = (low ? x.x - x.y : x.x + x.y) * a.scale;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = base + u * tpl;
            if (k < a.nout) gst<NT>(d + (int64_t)k * a.ops, v[u]);
        }
    }
}

// One tile of a fast stage: tws holds the staged tables, data the tile's
// lines. Threads beyond the stage's lt x tpl shape (a batched launch sized
// for its largest item) take indexes past every line and only join the
// barriers. NT: non-temporal global loads and stores. KIND: 0 = C2C lines
// only (the batched kernels), 1 = route stages of the C2C / R2C / C2R / DHT
// modes, 2 = route stages of the real-input R2R modes (their gathers and
// post-processing live in an instantiation of their own, so the other
// route kernels keep their smaller code).
template <typename T, int R0, int R1, int R2, int EM, bool NT, int KIND = 1, int MODE = -1>
__device__ __forceinline__ void fast_body(const StageArgs<T> &a, int64_t tile, const cpx<T> *tws, cpx<T> *data)
{
    constexpr int M = R0 * R1 * (R2 ? R2 : 1);
    constexpr int RL = R2 ? R2 : R1;
    const int tid = threadIdx.x;
    const int lt = a.lt, tpl = a.tpl;
    const bool act = tid < lt * tpl;
    const int l = act ? tid / tpl : 0;
    const int tl = act ? tid % tpl : (1 << 28);
    cpx<T> *ln = data + l * a.ls;
    const int mode = a.mode;
    if constexpr (KIND == 2) {
        // real-input R2R modes: the raw lines go to LDS with coalesced loads
        // (the generic kernel's tile load), the first pass gathers its
        // pre-processed inputs from there; generic post-processing at the
        // store. (Gathering straight from global memory, 4-byte loads at
        // the reorder's strides, made the 256x384x240 f32 R2R stages run at
        // 0.8-2.0 TB/s.) The compile-time scatter modes place the values
        // at their FFT-input slots as they load, so the first pass reads
        // complex values without the per-value index arithmetic.
        __syncthreads(); // the pre-processing reads the staged tables
        stage_load<T, NT, MODE>(a, data, tile, tid, l, act ? tl : (1 << 28));
        if constexpr (MODE == M_DCT4) fast_first_tw<T, R0, EM, M>(ln, tws + a.pre0, tl, tpl);
        else if constexpr (scatter_mode<MODE>()) fast_mid<T, R0, EM, M>(ln, tws, 1, tl, tpl);
        else fast_first_pre_lds<T, R0, EM, M, MODE>(a, ln, tws, tl, tpl);
        if constexpr (R2 != 0) fast_mid<T, R1, EM, M>(ln, tws + a.twoff[1], R0, tl, tpl);
        fast_mid<T, RL, EM, M>(ln, tws + a.twoff[R2 ? 2 : 1], M / RL, tl, tpl);
        if (act) {
            // the generic kernel's store (column side: consecutive threads
            // on consecutive lines, or on consecutive line pairs)
            if (a.out_col && (a.flags & F_CPAIR_OUT)) {
                const int hl = lt >> 1;
                store_line_pair<T, 2, NT, MODE>(a, data, tws, tile, 2 * (tid % hl), tid / hl, 2 * tpl);
            } else {
                const int lo = a.out_col ? tid % lt : l, k0 = a.out_col ? tid / lt : tl;
                store_line<T, 2, NT, MODE>(a, data, tws, tile, lo, k0, tpl);
            }
        }
        return;
    } else {
        const bool dht = KIND == 1 && mode == M_DHT;
        if (mode == M_C2R) {
            fast_c2r_load<T, M, NT>(a, data, tws, tile, tid, l, tl);
            fast_mid<T, R0, EM, M>(ln, tws, 1, tl, tpl);
        } else {
            // column access: lanes across the tile's lines; row access: along a line
            const int fl = !act ? 0 : (a.in_col ? tid % lt : l);
            const int ft = !act ? (1 << 28) : (a.in_col ? tid / lt : tl);
            const int64_t line = tile * lt + fl;
            int64_t io = 0, oo = 0;
            const bool valid = act && line < a.nlines;
            if (valid) line_offsets(a.g, line, io, oo);
            fast_first<T, R0, EM, M, NT>(a, data + fl * a.ls, valid, io, ft, tpl, mode == M_C2C,
                                         (a.flags & F_CONJ_IN) != 0);
        }
        if constexpr (R2 != 0) fast_mid<T, R1, EM, M>(ln, tws + a.twoff[1], R0, tl, tpl);
        if (mode == M_R2C || dht) {
            fast_mid<T, RL, EM, M>(ln, tws + a.twoff[R2 ? 2 : 1], M / RL, tl, tpl);
            if (act) {
                if (mode == M_R2C) fast_r2c_store<T, NT>(a, data, tws, tile, tid, l, tl);
                else fast_dht_store<T, NT>(a, data, tws, tile, tid, l, tl);
            }
            return;
        }
        const int fl = !act ? 0 : (a.out_col ? tid % lt : l);
        const int ft = !act ? (1 << 28) : (a.out_col ? tid / lt : tl);
        const int64_t line = tile * lt + fl;
        int64_t io = 0, oo = 0;
        const bool valid = act && line < a.nlines;
        if (valid) line_offsets(a.g, line, io, oo);
        fast_last<T, RL, EM, M, NT>(a, data + fl * a.ls, tws + a.twoff[R2 ? 2 : 1], valid, oo, ft, tpl,
                                    mode == M_C2R);
    }
}

template <typename T, int R0, int R1, int R2, int EM, bool NT, int KIND = 1, int MODE = -1>
__global__ __launch_bounds__(512) void k_line_fast(const StageArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *tws = reinterpret_cast<cpx<T> *>(smem_raw);
    if (a.gate != nullptr && *a.gate != 0) return;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    int64_t tile = blockIdx.x;
    if (a.flags & F_XCD_REMAP) {
        tile = (int64_t)(blockIdx.x % 8u) * a.xcd_per + blockIdx.x / 8u;
        if (tile >= a.ntiles) return;
    }
    for (int i = tid; i < a.twn; i += nthr) tws[i] = a.tw[i];
    fast_body<T, R0, R1, R2, EM, NT, KIND, MODE>(a, tile, tws, tws + a.twn);
}

// Persistent batched fast stage: the work list of k_line_batch, every item
// using this compile-time radix plan.
template <typename T, int R0, int R1, int R2, int EM>
__global__ __launch_bounds__(512) void k_line_fast_batch(const StageArgs<T> *__restrict__ items,
                                                        const int64_t *__restrict__ prefix, int nitems, int64_t total)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *tws = reinterpret_cast<cpx<T> *>(smem_raw);
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
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        while (q + 1 < nitems && prefix[q + 1] <= w) ++q;
        const StageArgs<T> &a = items[q];
        const int64_t tile = w - prefix[q];
        __syncthreads(); // the previous tile is done with LDS
        for (int i = tid; i < a.twn; i += nthr) tws[i] = a.tw[i];
        fast_body<T, R0, R1, R2, EM, false, 0>(a, tile, tws, tws + a.twn);
    }
}

// Persistent batched stage over items with different compile-time radix
// plans: each tile dispatches on its item's plan (uniform per workgroup).
// Plans with 32 values per thread are left out (they would raise the
// kernel's register count for every item).
template <typename T>
__global__ __launch_bounds__(512) void k_line_multi_batch(const StageArgs<T> *__restrict__ items,
                                                         const int64_t *__restrict__ prefix, int nitems,
                                                         int64_t total)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *tws = reinterpret_cast<cpx<T> *>(smem_raw);
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
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        while (q + 1 < nitems && prefix[q + 1] <= w) ++q;
        const StageArgs<T> &a = items[q];
        const int64_t tile = w - prefix[q];
        __syncthreads(); // the previous tile is done with LDS
        for (int i = tid; i < a.twn; i += nthr) tws[i] = a.tw[i];
        const int r0 = a.radix[0], r1 = a.radix[1], r2 = a.npass > 2 ? a.radix[2] : 0, em = a.fem;
#define PFH_MULTI_CASE(TT, A, B, C, E)                                                                        \
    if constexpr (std::is_same<T, TT>::value && E != 32) {                                                   \
        if (r0 == A && r1 == B && r2 == C && em == E) {                                                      \
            fast_body<T, A, B, C, E, false, 0>(a, tile, tws, tws + a.twn);                                  \
            continue;                                                                                        \
        }                                                                                                    \
    }
        PFH_FAST_PLANS(PFH_MULTI_CASE)
#undef PFH_MULTI_CASE
    }
}

template <typename T>
hipError_t launch_line_fast(const StageArgs<T> &a, int r0, int r1, int r2, int em, int threads, size_t lds_bytes,
                            int64_t grid, hipStream_t s)
{
    if (grid <= 0) return hipSuccess;
    // real-input R2R modes: their own instantiation, ordinary accesses; the
    // plans and modes of PFH_R2R_MODE_PLANS also with the mode as a
    // compile-time constant
    const bool r2r = !(a.mode == M_C2C || a.mode == M_R2C || a.mode == M_C2R || a.mode == M_DHT);
    static const bool mode_kernels = [] {
        const char *e = std::getenv("PFH_R2R_MODE_KERNELS");
        return e == nullptr || std::atoi(e) != 0;
    }();
    if (r2r && mode_kernels) {
#define PFH_MODE_CASE(TT, A, B, C, E, MD)                                                                     \
        if constexpr (std::is_same<T, TT>::value)                                                             \
            if (r0 == A && r1 == B && r2 == C && em == E && a.mode == MD) {                                   \
                hipLaunchKernelGGL((k_line_fast<T, A, B, C, E, false, 2, MD>), dim3((unsigned)grid),          \
                                   dim3(threads), lds_bytes, s, a);                                           \
                return hipGetLastError();                                                                     \
            }
        PFH_R2R_MODE_PLANS(PFH_MODE_CASE)
#undef PFH_MODE_CASE
    }
#define PFH_FAST_CASE(TT, A, B, C, E)                                                                           \
    if constexpr (std::is_same<T, TT>::value) if (r0 == A && r1 == B && r2 == C && em == E) {                               \
        if (r2r) {                                                                                              \
            hipLaunchKernelGGL((k_line_fast<T, A, B, C, E, false, 2>), dim3((unsigned)grid), dim3(threads),     \
                               lds_bytes, s, a);                                                                \
            return hipGetLastError();                                                                           \
        }                                                                                                       \
        if (a.flags & F_NT)                                                                                     \
            hipLaunchKernelGGL((k_line_fast<T, A, B, C, E, true>), dim3((unsigned)grid), dim3(threads), lds_bytes, \
                               s, a);                                                                           \
        else                                                                                                    \
            hipLaunchKernelGGL((k_line_fast<T, A, B, C, E, false>), dim3((unsigned)grid), dim3(threads),         \
                               lds_bytes, s, a);                                                                \
        return hipGetLastError();                                                                               \
    }
    PFH_FAST_PLANS(PFH_FAST_CASE)
    PFH_FAST_ROUTE_PLANS(PFH_FAST_CASE)
#undef PFH_FAST_CASE
    return hipErrorInvalidValue;
}

template hipError_t launch_line_fast<float>(const StageArgs<float> &, int, int, int, int, int, size_t, int64_t,
                                            hipStream_t);
template hipError_t launch_line_fast<double>(const StageArgs<double> &, int, int, int, int, int, size_t, int64_t,
                                             hipStream_t);

// The plans k_line_multi_batch dispatches on (its PFH_MULTI_CASE list): an
// item of any other plan must not go to it, or its tiles would be skipped.
bool multi_batch_plan_available(int tbytes, int r0, int r1, int r2, int em)
{
#define PFH_MULTI_HAS(TT, A, B, C, E) \
    if (E != 32 && (int)sizeof(TT) == tbytes && r0 == A && r1 == B && r2 == C && em == E) return true;
    PFH_FAST_PLANS(PFH_MULTI_HAS)
#undef PFH_MULTI_HAS
    return false;
}

bool fast_plan_available(int tbytes, int r0, int r1, int r2, int em)
{
#define PFH_FAST_HAS(TT, A, B, C, E) \
    if ((int)sizeof(TT) == tbytes && r0 == A && r1 == B && r2 == C && em == E) return true;
    PFH_FAST_PLANS(PFH_FAST_HAS)
    PFH_FAST_ROUTE_PLANS(PFH_FAST_HAS)
#undef PFH_FAST_HAS
    return false;
}

// Raises a kernel's dynamic LDS limit to `bytes` minus its static LDS.
static hipError_t raise_lds_limit(const void *f, size_t bytes)
{
    hipFuncAttributes at{};
    hipError_t e = hipFuncGetAttributes(&at, f);
    if (e != hipSuccess) return e;
    const int v = (int)bytes - (int)at.sharedSizeBytes;
    if (v <= 0) return hipErrorInvalidValue;
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, v);
}

hipError_t set_fast_lds_limit(size_t bytes)
{
#define PFH_FAST_ATTR(TT, A, B, C, E)                                                                    \
    {                                                                                                     \
        hipError_t e = raise_lds_limit(reinterpret_cast<const void *>(&k_line_fast<TT, A, B, C, E, false>), bytes); \
        if (e == hipSuccess)                                                                              \
            e = raise_lds_limit(reinterpret_cast<const void *>(&k_line_fast<TT, A, B, C, E, true>), bytes); \
        if (e == hipSuccess)                                                                              \
            e = raise_lds_limit(reinterpret_cast<const void *>(&k_line_fast_batch<TT, A, B, C, E>), bytes); \
        if (e != hipSuccess) return e;                                                                    \
    }
    PFH_FAST_PLANS(PFH_FAST_ATTR)
#undef PFH_FAST_ATTR
#define PFH_FAST_ATTR_ROUTE(TT, A, B, C, E)                                                              \
    {                                                                                                     \
        hipError_t e = raise_lds_limit(reinterpret_cast<const void *>(&k_line_fast<TT, A, B, C, E, false>), bytes); \
        if (e == hipSuccess)                                                                              \
            e = raise_lds_limit(reinterpret_cast<const void *>(&k_line_fast<TT, A, B, C, E, true>), bytes); \
        if (e != hipSuccess) return e;                                                                    \
    }
    PFH_FAST_ROUTE_PLANS(PFH_FAST_ATTR_ROUTE)
#undef PFH_FAST_ATTR_ROUTE
#define PFH_FAST_ATTR_R2R(TT, A, B, C, E)                                                                \
    {                                                                                                     \
        hipError_t e = raise_lds_limit(reinterpret_cast<const void *>(&k_line_fast<TT, A, B, C, E, false, 2>), bytes); \
        if (e != hipSuccess) return e;                                                                    \
    }
    PFH_FAST_PLANS(PFH_FAST_ATTR_R2R)
    PFH_FAST_ROUTE_PLANS(PFH_FAST_ATTR_R2R)
#undef PFH_FAST_ATTR_R2R
#define PFH_FAST_ATTR_MODE(TT, A, B, C, E, MD)                                                           \
    {                                                                                                     \
        hipError_t e =                                                                                    \
            raise_lds_limit(reinterpret_cast<const void *>(&k_line_fast<TT, A, B, C, E, false, 2, MD>), bytes); \
        if (e != hipSuccess) return e;                                                                    \
    }
    PFH_R2R_MODE_PLANS(PFH_FAST_ATTR_MODE)
#undef PFH_FAST_ATTR_MODE
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// R2C boundary endpoint statistics (distributed.rs:4205-4363): per extra
// batch and constrained plane, slot 0 = max(|re|,|im|) then sum (re/max)^2,
// slot 1 = max |im| then sum (im/max)^2. Values are accumulated in f64.
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ const cpx<T> *plane_ptr(const EndpointArgs &e, const void *data, int64_t batch, int plane)
{
    const int64_t k = plane == 0 ? 0 : e.clen - 1;
    return reinterpret_cast<const cpx<T> *>(data) + batch * e.len + k * e.stride;
}

// Three block maxima in one pass (one pair of barriers); sh holds 3 x 4.
__device__ __forceinline__ void block_reduce_max3(double &a, double &b, double &c, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) {
        a = fmax(a, __shfl_xor(a, o, 64));
        b = fmax(b, __shfl_xor(b, o, 64));
        c = fmax(c, __shfl_xor(c, o, 64));
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    __syncthreads();
    if (lane == 0) {
        sh[w] = a;
        sh[4 + w] = b;
        sh[8 + w] = c;
    }
    __syncthreads();
    a = sh[0];
    b = sh[4];
    c = sh[8];
    for (int i = 1; i < nw; ++i) {
        a = fmax(a, sh[i]);
        b = fmax(b, sh[4 + i]);
        c = fmax(c, sh[8 + i]);
    }
}

// Two block sums in one pass, each in a fixed order (xor butterflies across
// the lanes, then the waves in index order).
__device__ __forceinline__ void block_reduce_sum2(double &a, double &b, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    __syncthreads();
    if (lane == 0) {
        sh[w] = a;
        sh[4 + w] = b;
    }
    __syncthreads();
    a = sh[0];
    b = sh[4];
    for (int i = 1; i < nw; ++i) {
        a += sh[i];
        b += sh[4 + i];
    }
}

// One chunk of one constrained plane: its max(|re|,|im|), max|im|, and the
// sums of (re/m)^2, (im/m)^2 normalized by the chunk maximum m, plus a
// non-finite flag. grid: x = chunks (e.nblk), y = batch * planes + plane.
template <typename T>
__global__ __launch_bounds__(256) void k_endpoint_partial(EndpointArgs e, const void *data, EndpointPartial *part,
                                                          int *gate)
{
    __shared__ double sh[12];
    if (!e.keep_gate && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *gate = 0;
    const int64_t bp = blockIdx.y;
    const int64_t batch = bp / e.planes;
    const int plane = (int)(bp % e.planes);
    const cpx<T> *base = plane_ptr<T>(e, data, batch, plane);
    const int64_t count = e.outer * e.stride;
    const int64_t chunk = (count + gridDim.x - 1) / gridDim.x;
    const int64_t begin = (int64_t)blockIdx.x * chunk;
    const int64_t end = begin + chunk < count ? begin + chunk : count;
    // a chunk is at most K elements per thread (the planner sizes the grid
    // for it): they stay in registers for the two passes (maximum, then sums
    // scaled by it)
    constexpr int K = kEndpointPerThread;
    cpx<T> v[K];
    bool have[K];
    double ma = 0.0, mi = 0.0;
    int bad = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int64_t i = begin + threadIdx.x + (int64_t)k * blockDim.x;
        have[k] = i < end;
        if (have[k]) {
            const int64_t o = e.stride == 1 ? i : i / e.stride, q = i - o * e.stride;
            v[k] = base[o * e.clen * e.stride + q];
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (!have[k]) continue;
        const double re = (double)v[k].x, im = (double)v[k].y;
        if (!isfinite(re) || !isfinite(im)) {
            bad = 1;
            have[k] = false;
            continue;
        }
        ma = fmax(ma, fmax(fabs(re), fabs(im)));
        mi = fmax(mi, fabs(im));
    }
    double anybad = (double)bad;
    block_reduce_max3(ma, mi, anybad, sh);
    double sr = 0.0, si = 0.0;
    if (ma != 0.0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (!have[k]) continue;
            const double nr = (double)v[k].x / ma, ni = (double)v[k].y / ma;
            sr += nr * nr;
            si += ni * ni;
        }
    }
    block_reduce_sum2(sr, si, sh);
    if (threadIdx.x == 0) {
        EndpointPartial p;
        p.maxa = ma;
        p.maxi = mi;
        p.sre = sr;
        p.sim = si;
        p.bad = anybad != 0.0 ? 1u : 0u;
        part[bp * e.nblk + blockIdx.x] = p;
    }
}

// Combines the chunks of one batch (one block per batch) with a fixed-order
// tree reduction: global max M, sums rescaled by (m_chunk / M)^2 (equal to the
// reference's sums of (x/M)^2), then the acceptance rule
// (distributed.rs:4330-4349). Block (0) of k_endpoint_partial cleared *gate
// (in-place calls: k_inplace_begin set it from the poison word). A rejected
// spectrum of an in-place call also poisons the workspace: the stages before
// the check have already overwritten the caller's array
// (run_in_place_transaction, distributed.rs:3776-3794).
__global__ __launch_bounds__(256) void k_endpoint_combine(EndpointArgs e, const EndpointPartial *part, int *gate,
                                                          int *status)
{
    __shared__ double sh[12];
    const int64_t b = blockIdx.x;
    bool invalid = !isfinite(e.relative) || !isfinite(e.absolute);
    for (int p = 0; p < e.planes; ++p) {
        const EndpointPartial *q = part + (b * e.planes + p) * e.nblk;
        double ma = 0.0, mi = 0.0, bad = 0.0;
        for (int c = threadIdx.x; c < e.nblk; c += blockDim.x) {
            ma = fmax(ma, q[c].maxa);
            mi = fmax(mi, q[c].maxi);
            if (q[c].bad) bad = 1.0;
        }
        block_reduce_max3(ma, mi, bad, sh);
        const double M = ma, MI = mi;
        if (bad != 0.0) invalid = true;
        double s0 = 0.0, s1 = 0.0;
        if (M != 0.0) {
            for (int c = threadIdx.x; c < e.nblk; c += blockDim.x) {
                const double f = q[c].maxa / M;
                s0 += q[c].sre * f * f;
                s1 += q[c].sim * f * f;
            }
        }
        block_reduce_sum2(s0, s1, sh);
        if (!isfinite(M) || !isfinite(MI) || !isfinite(s0) || !isfinite(s1)) invalid = true;
        const bool abs_ok = MI <= e.absolute;
        const bool rel_ok = sqrt(s1) <= e.relative * sqrt(s0);
        if (!(abs_ok || rel_ok)) invalid = true;
    }
    if (threadIdx.x == 0 && invalid) {
        atomicOr(gate, 1);
        atomicOr(status, PFH_DEVICE_INVALID_SPECTRUM);
        if (e.poison != nullptr) atomicOr(e.poison, 1);
    }
}

// First kernel of an in-place call: a workspace poisoned by an earlier
// in-place failure gates every stage of this call (it writes nothing) and
// flags the call in the status word.
__global__ void k_inplace_begin(int *gate, const int *poison, int *status)
{
    if (threadIdx.x == 0) {
        const int p = *poison;
        *gate = p != 0 ? 1 : 0;
        if (p != 0) atomicOr(status, PFH_DEVICE_POISONED);
    }
}

// Local C2R strict endpoint preflight (r2c.rs:798-826): the first failing
// line, DC before Nyquist, as code = batch * 2 + (Nyquist ? 1 : 0).
template <typename T>
__global__ __launch_bounds__(256) void k_local_endpoints(const void *data, int64_t batch, int64_t clen, int even,
                                                         unsigned long long *first_code)
{
    for (int64_t b = threadIdx.x + (int64_t)blockIdx.x * blockDim.x; b < batch; b += (int64_t)gridDim.x * blockDim.x) {
        const cpx<T> *ln = reinterpret_cast<const cpx<T> *>(data) + b * clen;
        const bool dc = ln[0].y != (T)0;
        const bool ny = even && ln[clen - 1].y != (T)0;
        if (dc || ny) atomicMin(first_code, (unsigned long long)(b * 2 + (dc ? 0 : 1)));
    }
}

__global__ void k_local_endpoint_gate(const unsigned long long *first_code, int *gate, int *status)
{
    const int bad = (*first_code != ~0ull) ? 1 : 0;
    *gate = bad;
    if (bad) atomicOr(status, PFH_DEVICE_INVALID_ENDPOINT);
}

__global__ void k_fill_u64(unsigned long long *p, int64_t n, unsigned long long v)
{
    for (int64_t i = threadIdx.x + (int64_t)blockIdx.x * blockDim.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_i32(int *p, int64_t n, int v)
{
    for (int64_t i = threadIdx.x + (int64_t)blockIdx.x * blockDim.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// Generic element permutation for element sizes the line kernel does not
// cover (1 and 2 bytes): one thread per destination element.
template <typename W>
__global__ __launch_bounds__(256) void k_permute_words(PermuteArgs p, const W *src, W *dst)
{
    const int64_t total = p.len * p.extra;
    for (int64_t i = threadIdx.x + (int64_t)blockIdx.x * blockDim.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i / p.len;
        int64_t rem = i - e * p.len;
        int64_t so = 0;
        // i enumerates the destination memory order
        for (int d = p.nd - 1; d >= 0; --d) {
            const int64_t idx = rem % p.dext[d];
            rem /= p.dext[d];
            so += idx * p.sstride[d];
        }
        dst[i] = src[e * p.len + so];
    }
}

// ---------------------------------------------------------------------------
// host-side launch wrappers
// ---------------------------------------------------------------------------

template <typename T>
hipError_t launch_line_stage(const StageArgs<T> &a, bool wide, int threads, size_t lds_bytes, int64_t grid,
                             hipStream_t s)
{
    if (grid <= 0) return hipSuccess;
    const int mg = (a.mode == M_COPY_C || a.mode == M_COPY_R)
                       ? 0
                       : ((a.mode == M_C2C || a.mode == M_R2C || a.mode == M_R2C_ODD || a.mode == M_C2R ||
                           a.mode == M_C2R_ODD)
                              ? 1
                              : 2);
    // the odd-length extension modes (DST-I / DCT-I lines, whose FFT lengths
    // n + 1 / n - 1 rarely have a compile-time plan) with the mode compiled in
    static const bool mode_kernels = [] {
        const char *e = std::getenv("PFH_R2R_MODE_KERNELS");
        return e == nullptr || std::atoi(e) != 0;
    }();
    if (mg == 2 && mode_kernels) {
#define PFH_STAGE_MODE_CASE(W, MD)                                                                               \
        if (wide == W && a.mode == MD) {                                                                         \
            hipLaunchKernelGGL((k_line_stage<T, W, 2, MD>), dim3((unsigned)grid), dim3(threads), lds_bytes, s, a); \
            return hipGetLastError();                                                                           \
        }
        PFH_STAGE_MODE_CASE(true, M_DST1)
        PFH_STAGE_MODE_CASE(false, M_DST1)
        PFH_STAGE_MODE_CASE(true, M_DCT1)
        PFH_STAGE_MODE_CASE(false, M_DCT1)
#undef PFH_STAGE_MODE_CASE
    }
    // C2C lines (lengths without a compile-time plan, the pieces of long
    // lines) with the mode compiled in
    if (mg == 1 && a.mode == M_C2C && mode_kernels) {
        if (wide)
            hipLaunchKernelGGL((k_line_stage<T, true, 1, M_C2C>), dim3((unsigned)grid), dim3(threads), lds_bytes, s, a);
        else
            hipLaunchKernelGGL((k_line_stage<T, false, 1, M_C2C>), dim3((unsigned)grid), dim3(threads), lds_bytes, s,
                               a);
        return hipGetLastError();
    }
#define PFH_STAGE_CASE(W, G)                                                                                     \
    if (wide == W && mg == G) {                                                                                  \
        hipLaunchKernelGGL((k_line_stage<T, W, G>), dim3((unsigned)grid), dim3(threads), lds_bytes, s, a);       \
        return hipGetLastError();                                                                               \
    }
    PFH_STAGE_CASE(false, 0)
    PFH_STAGE_CASE(false, 1)
    PFH_STAGE_CASE(false, 2)
    PFH_STAGE_CASE(true, 1)
    PFH_STAGE_CASE(true, 2)
#undef PFH_STAGE_CASE
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_line_batch(const StageArgs<T> *items, const int64_t *prefix, int nitems, int64_t total, bool wide,
                             int threads, size_t lds_bytes, int64_t grid, hipStream_t s)
{
    if (total <= 0 || grid <= 0) return hipSuccess;
    if (wide)
        hipLaunchKernelGGL((k_line_batch<T, true>), dim3((unsigned)grid), dim3(threads), lds_bytes, s, items, prefix,
                           nitems, total);
    else
        hipLaunchKernelGGL((k_line_batch<T, false>), dim3((unsigned)grid), dim3(threads), lds_bytes, s, items, prefix,
                           nitems, total);
    return hipGetLastError();
}

template hipError_t launch_line_batch<float>(const StageArgs<float> *, const int64_t *, int, int64_t, bool, int,
                                             size_t, int64_t, hipStream_t);
template hipError_t launch_line_batch<double>(const StageArgs<double> *, const int64_t *, int, int64_t, bool, int,
                                              size_t, int64_t, hipStream_t);

template <typename T>
int line_batch_blocks_per_cu(bool wide, int threads, size_t lds_bytes)
{
    int n = 0;
    const hipError_t e =
        wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_line_batch<T, true>, threads, lds_bytes)
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_line_batch<T, false>, threads, lds_bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}
template int line_batch_blocks_per_cu<float>(bool, int, size_t);
template int line_batch_blocks_per_cu<double>(bool, int, size_t);

template <typename T>
hipError_t launch_line_fast_batch(const StageArgs<T> *items, const int64_t *prefix, int nitems, int64_t total, int r0,
                                  int r1, int r2, int em, int threads, size_t lds_bytes, int64_t grid, hipStream_t s)
{
    if (total <= 0 || grid <= 0) return hipSuccess;
#define PFH_FAST_BATCH_CASE(TT, A, B, C, E)                                                                      \
    if constexpr (std::is_same<T, TT>::value) if (r0 == A && r1 == B && r2 == C && em == E) {                  \
        hipLaunchKernelGGL((k_line_fast_batch<T, A, B, C, E>), dim3((unsigned)grid), dim3(threads), lds_bytes, s, \
                           items, prefix, nitems, total);                                                        \
        return hipGetLastError();                                                                                \
    }
    PFH_FAST_PLANS(PFH_FAST_BATCH_CASE)
#undef PFH_FAST_BATCH_CASE
    return hipErrorInvalidValue;
}
template hipError_t launch_line_fast_batch<float>(const StageArgs<float> *, const int64_t *, int, int64_t, int, int,
                                                  int, int, int, size_t, int64_t, hipStream_t);
template hipError_t launch_line_fast_batch<double>(const StageArgs<double> *, const int64_t *, int, int64_t, int, int,
                                                   int, int, int, size_t, int64_t, hipStream_t);

template <typename T>
hipError_t launch_line_multi_batch(const StageArgs<T> *items, const int64_t *prefix, int nitems, int64_t total,
                                   int threads, size_t lds_bytes, int64_t grid, hipStream_t s)
{
    if (total <= 0 || grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_line_multi_batch<T>, dim3((unsigned)grid), dim3(threads), lds_bytes, s, items, prefix, nitems,
                       total);
    return hipGetLastError();
}
template hipError_t launch_line_multi_batch<float>(const StageArgs<float> *, const int64_t *, int, int64_t, int,
                                                   size_t, int64_t, hipStream_t);
template hipError_t launch_line_multi_batch<double>(const StageArgs<double> *, const int64_t *, int, int64_t, int,
                                                    size_t, int64_t, hipStream_t);

template <typename T>
int line_multi_batch_blocks_per_cu(int threads, size_t lds_bytes)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_line_multi_batch<T>, threads, lds_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}
template int line_multi_batch_blocks_per_cu<float>(int, size_t);
template int line_multi_batch_blocks_per_cu<double>(int, size_t);

template <typename T>
int line_fast_batch_blocks_per_cu(int r0, int r1, int r2, int em, int threads, size_t lds_bytes)
{
    int n = 0;
#define PFH_FAST_OCC(TT, A, B, C, E)                                                                             \
    if constexpr (std::is_same<T, TT>::value) if (r0 == A && r1 == B && r2 == C && em == E) {                  \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_line_fast_batch<T, A, B, C, E>, threads,          \
                                                         lds_bytes) != hipSuccess) {                             \
            (void)hipGetLastError();                                                                             \
            return 1;                                                                                            \
        }                                                                                                        \
        return n > 0 ? n : 1;                                                                                    \
    }
    PFH_FAST_PLANS(PFH_FAST_OCC)
#undef PFH_FAST_OCC
    return 1;
}
template int line_fast_batch_blocks_per_cu<float>(int, int, int, int, int, size_t);
template int line_fast_batch_blocks_per_cu<double>(int, int, int, int, int, size_t);

template hipError_t launch_line_stage<float>(const StageArgs<float> &, bool, int, size_t, int64_t, hipStream_t);
template hipError_t launch_line_stage<double>(const StageArgs<double> &, bool, int, size_t, int64_t, hipStream_t);

hipError_t set_line_stage_lds_limit(size_t bytes)
{
    const void *fns[] = {reinterpret_cast<const void *>(&k_line_stage<float, false, 0>),
                         reinterpret_cast<const void *>(&k_line_stage<float, false, 1>),
                         reinterpret_cast<const void *>(&k_line_stage<float, false, 2>),
                         reinterpret_cast<const void *>(&k_line_stage<float, true, 1>),
                         reinterpret_cast<const void *>(&k_line_stage<float, true, 2>),
                         reinterpret_cast<const void *>(&k_line_stage<double, false, 0>),
                         reinterpret_cast<const void *>(&k_line_stage<double, false, 1>),
                         reinterpret_cast<const void *>(&k_line_stage<double, false, 2>),
                         reinterpret_cast<const void *>(&k_line_stage<double, true, 1>),
                         reinterpret_cast<const void *>(&k_line_stage<double, true, 2>),
                         reinterpret_cast<const void *>(&k_line_stage<float, true, 1, M_C2C>),
                         reinterpret_cast<const void *>(&k_line_stage<float, false, 1, M_C2C>),
                         reinterpret_cast<const void *>(&k_line_stage<double, true, 1, M_C2C>),
                         reinterpret_cast<const void *>(&k_line_stage<double, false, 1, M_C2C>),
                         reinterpret_cast<const void *>(&k_line_stage<float, true, 2, M_DST1>),
                         reinterpret_cast<const void *>(&k_line_stage<float, false, 2, M_DST1>),
                         reinterpret_cast<const void *>(&k_line_stage<float, true, 2, M_DCT1>),
                         reinterpret_cast<const void *>(&k_line_stage<float, false, 2, M_DCT1>),
                         reinterpret_cast<const void *>(&k_line_stage<double, true, 2, M_DST1>),
                         reinterpret_cast<const void *>(&k_line_stage<double, false, 2, M_DST1>),
                         reinterpret_cast<const void *>(&k_line_stage<double, true, 2, M_DCT1>),
                         reinterpret_cast<const void *>(&k_line_stage<double, false, 2, M_DCT1>),
                         reinterpret_cast<const void *>(&k_line_batch<float, true>),
                         reinterpret_cast<const void *>(&k_line_batch<float, false>),
                         reinterpret_cast<const void *>(&k_line_batch<double, true>),
                         reinterpret_cast<const void *>(&k_line_batch<double, false>),
                         reinterpret_cast<const void *>(&k_line_multi_batch<float>),
                         reinterpret_cast<const void *>(&k_line_multi_batch<double>)};
    for (const void *f : fns) {
        hipError_t e = raise_lds_limit(f, bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename T>
hipError_t launch_endpoint_check(const EndpointArgs &e, const void *data, EndpointPartial *part, int *gate,
                                 int *status, hipStream_t s)
{
    const int64_t by = e.extra * e.planes;
    if (by > 0)
        hipLaunchKernelGGL(k_endpoint_partial<T>, dim3((unsigned)e.nblk, (unsigned)by), dim3(256), 0, s, e, data,
                           part, gate);
    else if (!e.keep_gate)
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, gate, (int64_t)1, 0);
    if (e.extra > 0)
        hipLaunchKernelGGL(k_endpoint_combine, dim3((unsigned)e.extra), dim3(256), 0, s, e, part, gate, status);
    return hipGetLastError();
}

template hipError_t launch_endpoint_check<float>(const EndpointArgs &, const void *, EndpointPartial *, int *, int *,
                                                 hipStream_t);
template hipError_t launch_endpoint_check<double>(const EndpointArgs &, const void *, EndpointPartial *, int *, int *,
                                                  hipStream_t);

template <typename T>
hipError_t launch_local_endpoint_check(const void *data, int64_t batch, int64_t clen, int even,
                                       unsigned long long *first_code, int *gate, int *status, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, first_code, (int64_t)1, ~0ull);
    if (batch > 0) {
        int64_t g = (batch + 255) / 256;
        if (g > 4096) g = 4096;
        hipLaunchKernelGGL(k_local_endpoints<T>, dim3((unsigned)g), dim3(256), 0, s, data, batch, clen, even, first_code);
    }
    hipLaunchKernelGGL(k_local_endpoint_gate, dim3(1), dim3(1), 0, s, first_code, gate, status);
    return hipGetLastError();
}

template hipError_t launch_local_endpoint_check<float>(const void *, int64_t, int64_t, int, unsigned long long *, int *,
                                                       int *, hipStream_t);
template hipError_t launch_local_endpoint_check<double>(const void *, int64_t, int64_t, int, unsigned long long *,
                                                        int *, int *, hipStream_t);

hipError_t launch_inplace_begin(int *gate, const int *poison, int *status, hipStream_t s)
{
    hipLaunchKernelGGL(k_inplace_begin, dim3(1), dim3(64), 0, s, gate, poison, status);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, int64_t n, unsigned long long v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)g), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int *p, int64_t n, int v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)g), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_permute_words(const PermuteArgs &p, const void *src, void *dst, int word, hipStream_t s)
{
    const int64_t total = p.len * p.extra;
    if (total <= 0) return hipSuccess;
    int64_t g = (total + 255) / 256;
    if (g > 65536) g = 65536;
    switch (word) {
    case 1:
        hipLaunchKernelGGL(k_permute_words<uint8_t>, dim3((unsigned)g), dim3(256), 0, s, p,
                           reinterpret_cast<const uint8_t *>(src), reinterpret_cast<uint8_t *>(dst));
        break;
    case 2:
        hipLaunchKernelGGL(k_permute_words<uint16_t>, dim3((unsigned)g), dim3(256), 0, s, p,
                           reinterpret_cast<const uint16_t *>(src), reinterpret_cast<uint16_t *>(dst));
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace pfh
