// Host emulation of the HIP runtime and of the library's kernel launchers,
// linked with pencil-rs_amd/csrc/pfh_api.cpp (the planner, compiled by g++)
// to check the planner on a machine without a GPU:
//  - every launch's shape against the kernel it targets (threads, LDS,
//    grid, tiles, threads per line against each radix pass);
//  - every global address a launch would read or write lies inside one
//    allocation, and no element is written twice by one launch;
//  - the layout semantics of a whole operation: with each line transform
//    replaced by the identity on its points (R2C keeps the first n/2+1
//    points, C2R zero-fills), the output array must hold the input value at
//    the same logical coordinates, so routes, fused transitions and slab
//    pipelining are checked end to end.
// The FFT arithmetic is not emulated there (the GPU tests check it), except
// in the `long` scenario: the pieces of long lines (split and Bluestein
// line stages, the prep / post kernels of real lines) are computed in long
// double, and whole transforms are compared with naive sums.
#include <hip/hip_runtime_api.h>

#include <cinttypes>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/pencil_fft_hip.h"
#include "../../pencil-rs_amd/csrc/pfh_collectives.hpp"
#include "../../pencil-rs_amd/csrc/pfh_kernels.hpp"

namespace {

int g_fail = 0;

void fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::fprintf(stdout, "FAIL: ");
    std::vfprintf(stdout, fmt, ap);
    std::fprintf(stdout, "\n");
    va_end(ap);
    ++g_fail;
    if (g_fail > 20) std::exit(1);
}

struct Alloc {
    size_t bytes;
    std::vector<uint8_t> writes; // writes per byte in the current launch
};
std::map<uintptr_t, Alloc> g_allocs;

const size_t kLdsLimit = 163840;

// allocation containing [p, p + n)
Alloc *find_alloc(const void *p, size_t n, uintptr_t *base)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    auto it = g_allocs.upper_bound(a);
    if (it == g_allocs.begin()) return nullptr;
    --it;
    if (a + n > it->first + it->second.bytes) return nullptr;
    *base = it->first;
    return &it->second;
}

void begin_launch()
{
    for (auto &kv : g_allocs) std::fill(kv.second.writes.begin(), kv.second.writes.end(), 0);
}

bool check_read(const char *what, const void *p, size_t n)
{
    uintptr_t base;
    if (!find_alloc(p, n, &base)) {
        fail("%s: read of %zu bytes at %p outside every allocation", what, n, p);
        return false;
    }
    return true;
}

bool check_write(const char *what, void *p, size_t n)
{
    uintptr_t base;
    Alloc *a = find_alloc(p, n, &base);
    if (!a) {
        fail("%s: write of %zu bytes at %p outside every allocation", what, n, p);
        return false;
    }
    const size_t off = reinterpret_cast<uintptr_t>(p) - base;
    for (size_t i = 0; i < n; ++i) {
        if (a->writes[off + i]) {
            fail("%s: byte written twice in one launch (offset %zu)", what, off + i);
            return false;
        }
        a->writes[off + i] = 1;
    }
    return true;
}

} // namespace

// ---------------------------------------------------------------------------
// HIP runtime stubs (host memory stands in for device memory)
// ---------------------------------------------------------------------------
extern "C" {

hipError_t hipMalloc(void **p, size_t n)
{
    const size_t sz = n ? n : 1;
    void *q = std::aligned_alloc(256, (sz + 255) / 256 * 256);
    if (!q) return hipErrorOutOfMemory;
    std::memset(q, 0xA5, sz);
    Alloc a;
    a.bytes = sz;
    a.writes.assign(sz, 0);
    g_allocs[reinterpret_cast<uintptr_t>(q)] = std::move(a);
    *p = q;
    return hipSuccess;
}
hipError_t hipFree(void *p)
{
    if (!p) return hipSuccess;
    g_allocs.erase(reinterpret_cast<uintptr_t>(p));
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind)
{
    std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t)
{
    std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemset(void *d, int v, size_t n)
{
    std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipGetDevice(int *d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600 *p, int)
{
    std::memset(p, 0, sizeof(*p));
    p->sharedMemPerBlock = kLdsLimit;
    p->multiProcessorCount = 256;
    std::snprintf(p->gcnArchName, sizeof p->gcnArchName, "gfx950:sramecc+:xnack-");
    return hipSuccess;
}
const char *hipGetErrorString(hipError_t) { return "emulated error"; }
hipError_t hipGetLastError(void) { return hipSuccess; }
static int g_dummy[64];
hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = reinterpret_cast<hipEvent_t>(&g_dummy[0]);
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
// Elapsed times follow a fixed pseudo-random sequence, so that a measured
// choice (the batch autotuner's) is not simply the first candidate.
unsigned g_elapsed_calls = 0;
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t)
{
    const unsigned c = g_elapsed_calls++;
    *ms = 1.0f + (float)((c * 2654435761u) >> 24) / 256.0f;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t *s)
{
    *s = reinterpret_cast<hipStream_t>(&g_dummy[1]);
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
// Graphs: launches run when issued; a captured graph replays nothing, so
// the emulation only checks the first (capturing) execution.
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipSuccess; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g)
{
    *g = reinterpret_cast<hipGraph_t>(&g_dummy[2]);
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t, hipGraphNode_t *, char *, size_t)
{
    *e = reinterpret_cast<hipGraphExec_t>(&g_dummy[3]);
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }

} // extern "C"

// ---------------------------------------------------------------------------
// kernel launcher emulation
// ---------------------------------------------------------------------------
namespace pfh {

static bool mode_cin(int m) { return m == M_COPY_C || m == M_C2C || m == M_C2R || m == M_C2R_ODD; }
static bool mode_cout(int m) { return m == M_COPY_C || m == M_C2C || m == M_R2C || m == M_R2C_ODD; }
static bool mode_pre(int m)
{
    return !(m == M_COPY_C || m == M_COPY_R || m == M_C2C || m == M_R2C || m == M_DHT);
}

static void offsets(const LineGeom &g, int64_t line, int64_t &io, int64_t &oo)
{
    io = oo = 0;
    for (int d = 0; d < g.nd; ++d) {
        const int64_t q = line / g.ext[d], r = line - q * g.ext[d];
        // the kernels' quotient (multiply-high with the planner's constants)
        if (line >= (1LL << 31) || (int64_t)fastdiv((uint32_t)line, g.dmag[d], g.dsh[d]) != q)
            fail("line %" PRId64 " digit %d: fast quotient differs", line, d);
        io += r * g.is[d];
        oo += r * g.os[d];
        line = q;
    }
}

// make_fastdiv / fastdiv against the exact quotient: every divisor below
// 2^12 with numerators around its multiples, then random pairs and the
// extremes of the 31-bit range.
static void fastdiv_selftest()
{
    const int before = g_fail;
    auto check = [&](uint32_t x, uint32_t d) {
        const FastDiv f = make_fastdiv(d);
        if (fastdiv(x, f.mag, f.sh) != x / d) fail("fastdiv %u / %u", x, d);
    };
    for (uint32_t d = 1; d < 4096 && g_fail == before; ++d)
        for (uint32_t q : {0u, 1u, 2u, 3u, 1000u, (0x7fffffffu / d) - 1, 0x7fffffffu / d})
            for (int e = -1; e <= 1; ++e) {
                const int64_t x = (int64_t)q * d + e;
                if (x >= 0 && x < (1LL << 31)) check((uint32_t)x, d);
            }
    uint64_t st = 0x9e3779b97f4a7c15ull;
    auto rnd = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    for (int i = 0; i < 2000000 && g_fail == before; ++i) {
        const uint32_t d = (uint32_t)(rnd() % 0x7fffffffu) + 1;
        const uint32_t dd = (i & 1) ? d : (d >> (rnd() % 31)) + 1;
        check((uint32_t)(rnd() & 0x7fffffffu), dd);
        check(0x7fffffffu, dd);
        check(0x7fffffffu - (0x7fffffffu % dd), dd);
    }
    std::printf("fastdiv: %s\n", g_fail == before ? "ok" : "FAILED");
}


// shape checks shared by the single-stage kernels and the batch items
template <typename T>
static void check_shape(const char *what, const StageArgs<T> &a, int threads, size_t lds, int emax, bool fast,
                        int fem, bool wide = false)
{
    const int cb = 2 * (int)sizeof(T);
    if (a.lc < 1 || a.lt % a.lc != 0) fail("%s: lt %d not a multiple of lc %d", what, a.lt, a.lc);
    if (fast && a.lc != a.lt) fail("%s: fast stage computes in rounds", what);
    if (a.lc * a.tpl > threads) fail("%s: lc*tpl %d > threads %d", what, a.lc * a.tpl, threads);
    if ((a.in_col || a.out_col) && a.lt > a.lc * a.tpl) fail("%s: column tile %d > threads", what, a.lt);
    if (!fast) {
        // stage_load / stage_store thread mapping: each (line, point) of a
        // tile exactly once on both sides
        const int nthr = a.lc * a.tpl;
        for (int side = 0; side < 2; ++side) {
            const bool col = side == 0 ? a.in_col : a.out_col;
            const int np = side == 0 ? a.nin : a.nout;
            std::vector<int> hits((size_t)a.lt * np, 0);
            auto visit = [&](int ll, int p0, int step) {
                for (int p = p0; p < np; p += step) ++hits[(size_t)ll * np + p];
            };
            for (int tid = 0; tid < nthr; ++tid) {
                if (col) {
                    const int rows = nthr / a.lt;
                    if (tid < rows * a.lt) visit(tid % a.lt, tid / a.lt, rows);
                } else {
                    for (int r = 0; r < a.lt; r += a.lc) visit(tid / a.tpl + r, tid % a.tpl, a.tpl);
                }
            }
            for (size_t i = 0; i < hits.size(); ++i)
                if (hits[i] != 1) {
                    fail("%s: %s mapping touches (%zu, %zu) %d times", what, side ? "store" : "load", i / np, i % np,
                         hits[i]);
                    break;
                }
        }
    }
    if ((size_t)(a.twn + (int64_t)a.lt * a.ls) * cb > lds) fail("%s: LDS %zu too small", what, lds);
    if (lds > kLdsLimit - 256) fail("%s: LDS %zu above the budget", what, lds);
    if (a.ntiles * a.lt < a.nlines) fail("%s: tiles do not cover the lines", what);
    int need_c = a.m > 0 ? a.m : 1;
    need_c = std::max(need_c, mode_cin(a.mode) ? a.nin : (a.nin + 1) / 2);
    need_c = std::max(need_c, mode_cout(a.mode) ? a.nout : (a.nout + 1) / 2);
    if (a.ls < need_c) fail("%s: line stride %d < %d", what, a.ls, need_c);
    int prod = 1;
    for (int p = 0; p < a.npass; ++p) {
        const int R = a.radix[p];
        prod *= R;
        if (a.twoff[p] < 0 || a.twoff[p] > a.twn) fail("%s: pass %d table offset", what, p);
        const int em = fast ? fem : emax;
        const bool generic = !(R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 16 || R == 32 || R == 7 ||
                               R == 11 || R == 13);
        if (generic) {
            if ((int64_t)a.tpl * em < a.m) fail("%s: pass %d (generic radix %d) not covered", what, p, R);
        } else {
            const int MB = fast ? std::max(1, em / R) : generic_mb(R, wide, sizeof(T) == 4);
            if ((int64_t)a.tpl * MB < a.m / R) fail("%s: pass %d (radix %d) not covered", what, p, R);
        }
    }
    if (a.m > 0 && prod != a.m) fail("%s: radix product %d != m %d", what, prod, a.m);
    if (!fast) {
        // each pass's j mod Ns through the planner's constants, every j of the pass
        int Ns = 1;
        for (int p = 0; p < a.npass; ++p) {
            const int nb = a.m / a.radix[p];
            for (int j = 0; j < nb; ++j)
                if ((int)fastdiv((uint32_t)j, a.nsmag[p], a.nssh[p]) != j / Ns) {
                    fail("%s: pass %d Ns %d constants", what, p, Ns);
                    break;
                }
            Ns *= a.radix[p];
        }
    }
    if (mode_pre(a.mode) && (int64_t)a.tpl * (fast ? fem : generic_pre_e(wide, sizeof(T) == 4)) < a.m)
        fail("%s: pre-processing not covered", what);
}

bool g_math = false; // the `long` scenario: line stages compute their transform
bool g_math_dft = true; // false: bounds and shapes only (the DFTs of real-size lines take too long here)

using ld = long double;
struct cld {
    ld re, im;
};
static cld cmul_l(cld a, cld b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// Arithmetic emulation of a C2C line stage, pieces of long lines included:
// load (zero padding, conjugation, load factors), DFT of length m, store
// (twiddle d k, conjugation, scale, store factors, crop).
template <typename T>
static void emulate_math(const char *what, const StageArgs<T> &a)
{
    if (a.mode != M_C2C) {
        fail("%s: arithmetic emulation of mode %d", what, a.mode);
        return;
    }
    const int m = a.m;
    if (a.nin != m || a.nout != m) fail("%s: nin %d / nout %d for m %d", what, a.nin, a.nout, m);
    std::vector<cld> z(m), Z(m);
    const ld pi = 3.141592653589793238462643383279502884L;
    for (int64_t line = 0; line < a.nlines; ++line) {
        int64_t io, oo;
        offsets(a.g, line, io, oo);
        int dco = 0;
        {
            int64_t x = line;
            for (int d = 0; d < a.g.nd; ++d) {
                if (d == a.xdig) dco = (int)(x % a.g.ext[d]);
                x /= a.g.ext[d];
            }
        }
        for (int p = 0; p < m; ++p) {
            cld v{0, 0};
            const int64_t f = (int64_t)p * a.xmul_in + dco;
            if (!(a.flags & F_XIN) || f < a.xvalid_in) {
                const cpx<T> *q = reinterpret_cast<const cpx<T> *>(a.src) + io + (int64_t)p * a.ips;
                if (!check_read(what, q, sizeof(*q))) return;
                v = {(ld)q->x, (ld)q->y};
                if (a.flags & F_CONJ_IN) v.im = -v.im;
                if ((a.flags & F_XIN) && a.xtab_in) {
                    if (!check_read("xtab_in", a.xtab_in + f, sizeof(cpx<T>))) return;
                    v = cmul_l(v, {(ld)a.xtab_in[f].x, (ld)a.xtab_in[f].y});
                }
            }
            z[p] = v;
        }
        for (int k = 0; k < m && !g_math_dft; ++k) Z[k] = z[k];
        for (int k = 0; k < m && g_math_dft; ++k) {
            cld acc{0, 0};
            for (int p = 0; p < m; ++p) {
                const ld ang = -2 * pi * (ld)(((int64_t)p * k) % m) / (ld)m;
                acc = {acc.re + z[p].re * cosl(ang) - z[p].im * sinl(ang),
                       acc.im + z[p].re * sinl(ang) + z[p].im * cosl(ang)};
            }
            Z[k] = acc;
        }
        for (int k = 0; k < m; ++k) {
            cld w = Z[k];
            const int64_t f = (int64_t)k * a.xmul_out + dco;
            if ((a.flags & F_XOUT) && f >= a.xvalid_out) continue;
            if (a.flags & F_XTW) {
                const int64_t q = (int64_t)dco * k;
                if (q >= a.xtw_m) fail("%s: twiddle index %" PRId64 " >= %d", what, q, a.xtw_m);
                const int64_t lo = q & ((1 << a.xtw_bsh) - 1), hi = (1 << a.xtw_bsh) + (q >> a.xtw_bsh);
                if (!check_read("xtw", a.xtw + lo, sizeof(cpx<T>)) || !check_read("xtw", a.xtw + hi, sizeof(cpx<T>)))
                    return;
                w = cmul_l(w, cmul_l({(ld)a.xtw[lo].x, (ld)a.xtw[lo].y}, {(ld)a.xtw[hi].x, (ld)a.xtw[hi].y}));
            }
            if (a.flags & F_CONJ_OUT) w.im = -w.im;
            w = {w.re * (ld)a.scale, w.im * (ld)a.scale};
            if ((a.flags & F_XOUT) && a.xtab_out) {
                if (!check_read("xtab_out", a.xtab_out + f, sizeof(cpx<T>))) return;
                w = cmul_l(w, {(ld)a.xtab_out[f].x, (ld)a.xtab_out[f].y});
            }
            cpx<T> *q = reinterpret_cast<cpx<T> *>(a.dst) + oo + (int64_t)k * a.ops;
            if (!check_write(what, q, sizeof(*q))) return;
            q->x = (T)w.re;
            q->y = (T)w.im;
        }
    }
}

// identity-on-points emulation of one line stage
template <typename T>
static void emulate(const char *what, const StageArgs<T> &a)
{
    if (a.gate && *a.gate) return;
    if (g_math) {
        emulate_math(what, a);
        return;
    }
    if (a.flags & (F_XIN | F_XOUT | F_XTW)) fail("%s: long-line piece outside the arithmetic scenario", what);
    const bool ci = mode_cin(a.mode), co = mode_cout(a.mode);
    const size_t ei = ci ? 2 * sizeof(T) : sizeof(T), eo = co ? 2 * sizeof(T) : sizeof(T);
    std::vector<T> re(a.nin), im(a.nin);
    for (int64_t line = 0; line < a.nlines; ++line) {
        int64_t io, oo;
        offsets(a.g, line, io, oo);
        for (int j = 0; j < a.nin; ++j) {
            const char *p = static_cast<const char *>(a.src) + (io + (int64_t)j * a.ips) * (int64_t)ei;
            if (!check_read(what, p, ei)) return;
            const T *v = reinterpret_cast<const T *>(p);
            re[j] = v[0];
            im[j] = ci ? v[1] : (T)0;
        }
        for (int k = 0; k < a.nout; ++k) {
            char *p = static_cast<char *>(a.dst) + (oo + (int64_t)k * a.ops) * (int64_t)eo;
            if (!check_write(what, p, eo)) return;
            T *v = reinterpret_cast<T *>(p);
            const T r = k < a.nin ? re[k] : (T)0, i = k < a.nin ? im[k] : (T)0;
            v[0] = r;
            if (co) v[1] = (a.mode == M_R2C || a.mode == M_R2C_ODD) ? (T)0 : i;
        }
    }
}

template <typename T>
static void check_grid(const char *what, const StageArgs<T> &a, int64_t grid)
{
    if (a.flags & F_XCD_REMAP) {
        if (grid != 8LL * a.xcd_per || 8LL * a.xcd_per < a.ntiles) fail("%s: XCD grid", what);
    } else if (grid < a.ntiles) {
        fail("%s: grid %" PRId64 " < tiles %" PRId64, what, grid, a.ntiles);
    }
}

int g_launches = 0;
int g_all_launches = 0;         // every device launch, fills included
bool g_reject_endpoints = false; // the emulated C2R boundary check rejects every spectrum
static const bool g_trace = std::getenv("PFH_EMU_TRACE") != nullptr;

template <typename T>
static void trace(const char *what, const StageArgs<T> &a, int threads, int64_t grid)
{
    if (!g_trace) return;
    std::printf("  %s mode %d n %d m %d nlines %" PRId64 " lt %d tpl %d threads %d tiles %" PRId64 " grid %" PRId64
                " ips %" PRId64 " ops %" PRId64 " col %d/%d flags %d lc %d digits",
                what, a.mode, a.n, a.m, a.nlines, a.lt, a.tpl, threads, a.ntiles, grid, a.ips, a.ops, a.in_col,
                a.out_col, a.flags, a.lc);
    for (int d = 0; d < a.g.nd; ++d)
        std::printf(" [%" PRId64 ":%" PRId64 "/%" PRId64 "]", a.g.ext[d], a.g.is[d], a.g.os[d]);
    std::printf("\n");
}

template <typename T>
hipError_t launch_line_stage(const StageArgs<T> &a, bool wide, int threads, size_t lds_bytes, int64_t grid,
                             hipStream_t)
{
    ++g_launches;
    ++g_all_launches;
    begin_launch();
    if (threads > (wide ? kMaxThreadsWide : kMaxThreads)) fail("line_stage: %d threads", threads);
    trace("line_stage", a, threads, grid);
    check_shape("line_stage", a, threads, lds_bytes, Regs<T>::emax, false, 0, wide);
    check_grid("line_stage", a, grid);
    if (a.twn > 0) check_read("line_stage tables", a.tw, (size_t)a.twn * 2 * sizeof(T));
    emulate("line_stage", a);
    return hipSuccess;
}
template hipError_t launch_line_stage<float>(const StageArgs<float> &, bool, int, size_t, int64_t, hipStream_t);
template hipError_t launch_line_stage<double>(const StageArgs<double> &, bool, int, size_t, int64_t, hipStream_t);

// The prep / post kernels of long real lines, restated on the host.
template <typename T>
static bool long_checks(const char *what, const LongArgs<T> &a)
{
    ++g_launches;
    ++g_all_launches;
    begin_launch();
    if (!g_math) {
        fail("%s outside the arithmetic scenario", what);
        return false;
    }
    if (a.nlines < 1 || a.m < 1 || a.n < 1) {
        fail("%s: empty shape", what);
        return false;
    }
    return !(a.gate && *a.gate);
}

template <typename T>
static int64_t long_base(const LongArgs<T> &a, int64_t line, bool out)
{
    int64_t io, oo;
    offsets(a.g, line, io, oo);
    return out ? oo : io;
}

template <typename T>
hipError_t launch_long_prep(const LongArgs<T> &a, hipStream_t)
{
    if (!long_checks("long_prep", a)) return hipSuccess;
    const int n = a.n;
    for (int64_t line = 0; line < a.nlines; ++line) {
        const int64_t io = long_base(a, line, false);
        auto X = [&](int i) -> ld {
            const int idx = (a.flags & F_PRE_REV) ? n - 1 - i : i;
            const T *q = reinterpret_cast<const T *>(a.src) + io + (int64_t)idx * a.ips;
            if (idx < 0 || idx >= n) fail("long_prep: point %d of %d", idx, n);
            check_read("long_prep", q, sizeof(T));
            ld v = (ld)*q;
            if ((a.flags & F_PRE_ALT) && (idx & 1)) v = -v;
            if (a.flags & F_WEIGHTS) {
                if (idx == 0) v *= (ld)a.prew0;
                if (idx == n - 1) v *= (ld)a.prew1;
            }
            return v;
        };
        auto H = [&](int k, int last) -> cld {
            const cpx<T> *q = reinterpret_cast<const cpx<T> *>(a.src) + io + (int64_t)k * a.ips;
            check_read("long_prep", q, sizeof(*q));
            cld v{(ld)q->x, (ld)q->y};
            if (k == 0 || k == last) v.im = 0;
            if ((a.flags & F_HERM_HALF) && k > 0 && k < last) v = {v.re / 2, v.im / 2};
            return v;
        };
        auto tab = [&](int off) -> cld {
            check_read("long_prep table", a.tab + off, sizeof(cpx<T>));
            return {(ld)a.tab[off].x, (ld)a.tab[off].y};
        };
        for (int j = 0; j < a.m; ++j) {
            cld z{0, 0};
            switch (a.kind) {
            case LK_C2R: {
                const int h = n / 2, last = (n % 2 == 0) ? h : h + 1;
                cld v = j <= h ? H(j, last) : H(n - j, last);
                if (j > h) v.im = -v.im;
                z = {v.re, -v.im};
                break;
            }
            case LK_DCT2: z = {j < (n + 1) / 2 ? X(2 * j) : X(2 * (n - 1 - j) + 1), 0}; break;
            case LK_DCT3: {
                ld y = X(j);
                if (j == 0) y /= 2;
                const cld t = tab(a.pre0 + j);
                z = {t.re * y, t.im * y};
                break;
            }
            case LK_DCT4:
                if (j < n) {
                    const cld t = tab(a.pre0 + j);
                    const ld y = X(j);
                    z = {t.re * y, t.im * y};
                }
                break;
            case LK_DCT1: z = {j <= n - 1 ? X(j) : X(2 * (n - 1) - j), 0}; break;
            case LK_DST1:
                if (j != 0 && j != n + 1) z = {j <= n ? X(j - 1) : -X(2 * n + 1 - j), 0};
                break;
            default: z = {X(j), 0}; break;
            }
            cpx<T> *q = reinterpret_cast<cpx<T> *>(a.dst) + line * (int64_t)a.m + j;
            if (!check_write("long_prep", q, sizeof(*q))) return hipSuccess;
            q->x = (T)z.re;
            q->y = (T)z.im;
        }
    }
    return hipSuccess;
}

template <typename T>
hipError_t launch_long_post(const LongArgs<T> &a, hipStream_t)
{
    if (!long_checks("long_post", a)) return hipSuccess;
    const int n = a.n;
    for (int64_t line = 0; line < a.nlines; ++line) {
        const int64_t oo = long_base(a, line, true);
        auto Z = [&](int i) -> cld {
            if (i < 0 || i >= a.m) fail("long_post: index %d of %d", i, a.m);
            const cpx<T> *q = reinterpret_cast<const cpx<T> *>(a.src) + line * (int64_t)a.m + i;
            check_read("long_post", q, sizeof(*q));
            return {(ld)q->x, (ld)q->y};
        };
        for (int k = 0; k < a.nout; ++k) {
            if (a.kind == LK_R2C) {
                const int last = (n % 2 == 0) ? n / 2 : -1;
                cld v = Z(k);
                if (k == 0 || k == last) v.im = 0;
                if (a.flags & F_HERM_DOUBLE) {
                    if (k > 0 && k != last) v = {2 * v.re, 2 * v.im};
                    else v.im = 0;
                }
                cpx<T> *q = reinterpret_cast<cpx<T> *>(a.dst) + oo + (int64_t)k * a.ops;
                if (!check_write("long_post", q, sizeof(*q))) return hipSuccess;
                q->x = (T)(v.re * (ld)a.scale);
                q->y = (T)(v.im * (ld)a.scale);
                continue;
            }
            const int kk = (a.flags & F_POST_REV) ? n - 1 - k : k;
            ld y = 0;
            switch (a.kind) {
            case LK_DCT2:
            case LK_DCT4: {
                check_read("long_post table", a.tab + a.post0 + kk, sizeof(cpx<T>));
                const cld t{(ld)a.tab[a.post0 + kk].x, (ld)a.tab[a.post0 + kk].y};
                y = 2 * cmul_l(t, Z(kk)).re;
                break;
            }
            case LK_DCT3: y = 2 * Z((kk & 1) ? n - 1 - (kk >> 1) : (kk >> 1)).re; break;
            case LK_DST1: y = -Z(kk + 1).im; break;
            case LK_DHT: y = Z(kk).re - Z(kk).im; break;
            default: y = Z(kk).re; break;
            }
            if ((a.flags & F_POST_ALT) && (k & 1)) y = -y;
            if (a.flags & F_WEIGHTS) {
                if (k == 0) y *= (ld)a.postw0;
                if (k == n - 1) y *= (ld)a.postw1;
            }
            T *q = reinterpret_cast<T *>(a.dst) + oo + (int64_t)k * a.ops;
            if (!check_write("long_post", q, sizeof(*q))) return hipSuccess;
            *q = (T)(y * (ld)a.scale);
        }
    }
    return hipSuccess;
}
template hipError_t launch_long_prep<float>(const LongArgs<float> &, hipStream_t);
template hipError_t launch_long_prep<double>(const LongArgs<double> &, hipStream_t);
template hipError_t launch_long_post<float>(const LongArgs<float> &, hipStream_t);
template hipError_t launch_long_post<double>(const LongArgs<double> &, hipStream_t);

template <typename T>
hipError_t launch_line_fast(const StageArgs<T> &a, int r0, int r1, int r2, int em, int threads, size_t lds_bytes,
                            int64_t grid, hipStream_t)
{
    ++g_launches;
    ++g_all_launches;
    begin_launch();
    if (threads > 512) fail("line_fast: %d threads", threads);
    if (!fast_plan_available((int)sizeof(T), r0, r1, r2, em)) fail("line_fast: plan not instantiated");
    if (a.radix[0] != r0 || a.radix[1] != r1 || (r2 && a.radix[2] != r2)) fail("line_fast: radices differ");
    if (!(a.mode == M_C2C || a.mode == M_R2C || a.mode == M_C2R || a.mode == M_DHT || a.mode == M_DCT1 ||
          a.mode == M_DST1 || a.mode == M_DCT2 || a.mode == M_DCT2_ODD || a.mode == M_DCT3 || a.mode == M_DCT3_ODD ||
          a.mode == M_DCT4 || a.mode == M_DCT4_ODD || a.mode == M_DHT_ODD))
        fail("line_fast: mode %d", a.mode);
    if (a.mode == M_DHT && (a.flags & (F_PRE_REV | F_PRE_ALT | F_POST_REV | F_POST_ALT | F_WEIGHTS)))
        fail("line_fast: DHT with pre / post flags %d", a.flags);
    trace("line_fast", a, threads, grid);
    check_shape("line_fast", a, threads, lds_bytes, Regs<T>::emax, true, em);
    if (a.twn > 0) check_read("line_fast tables", a.tw, (size_t)a.twn * 2 * sizeof(T));
    check_grid("line_fast", a, grid);
    emulate("line_fast", a);
    return hipSuccess;
}
template hipError_t launch_line_fast<float>(const StageArgs<float> &, int, int, int, int, int, size_t, int64_t,
                                            hipStream_t);
template hipError_t launch_line_fast<double>(const StageArgs<double> &, int, int, int, int, int, size_t, int64_t,
                                             hipStream_t);

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

template <typename T>
static void emulate_batch(const char *what, const StageArgs<T> *items, const int64_t *prefix, int nitems,
                          int64_t total, int threads, size_t lds_bytes, int64_t grid, bool fast, int em,
                          const int *fr, bool wide = false)
{
    ++g_launches;
    ++g_all_launches;
    begin_launch();
    if (threads > kMaxThreadsWide) fail("%s: %d threads", what, threads);
    if (grid <= 0 || grid > total) fail("%s: grid %" PRId64 " for %" PRId64 " tiles", what, grid, total);
    check_read("batch items", items, sizeof(StageArgs<T>) * nitems);
    check_read("batch prefix", prefix, sizeof(int64_t) * (nitems + 1));
    if (prefix[0] != 0 || prefix[nitems] != total) fail("%s: prefix ends", what);
    if (g_trace) {
        double bytes = 0;
        for (int q = 0; q < nitems; ++q)
            bytes += (double)items[q].nlines * (items[q].nin + items[q].nout) * 2.0 * sizeof(T);
        std::printf("  %s items %d tiles %" PRId64 " threads %d grid %" PRId64 " lds %zu bytes %.0f\n", what, nitems,
                    total, threads, grid, lds_bytes, bytes);
    }
    for (int q = 0; q < nitems; ++q) {
        const StageArgs<T> &a = items[q];
        if (prefix[q + 1] - prefix[q] != a.ntiles) fail("%s: item %d tiles", what, q);
        if (a.mode != M_C2C) fail("%s: item %d mode", what, q);
        if (fast && (a.radix[0] != fr[0] || a.radix[1] != fr[1] || (fr[2] && a.radix[2] != fr[2])))
            fail("%s: item %d radices differ from the launch's plan", what, q);
        for (int p = 0; p < a.npass && !fast; ++p)
            if (!wide && !(a.radix[p] == 2 || a.radix[p] == 3 || a.radix[p] == 4 || a.radix[p] == 5 ||
                           a.radix[p] == 8 || a.radix[p] == 16 || a.radix[p] == 32))
                fail("%s: item %d radix %d in a narrow launch", what, q, a.radix[p]);
        check_shape(what, a, threads, lds_bytes, Regs<T>::emax, fast, em, wide);
        emulate(what, a);
    }
}

template <typename T>
hipError_t launch_line_batch(const StageArgs<T> *items, const int64_t *prefix, int nitems, int64_t total, bool wide,
                             int threads, size_t lds_bytes, int64_t grid, hipStream_t)
{
    const int none[3] = {0, 0, 0};
    emulate_batch("line_batch", items, prefix, nitems, total, threads, lds_bytes, grid, false, 0, none, wide);
    return hipSuccess;
}
template hipError_t launch_line_batch<float>(const StageArgs<float> *, const int64_t *, int, int64_t, bool, int,
                                             size_t, int64_t, hipStream_t);
template hipError_t launch_line_batch<double>(const StageArgs<double> *, const int64_t *, int, int64_t, bool, int,
                                              size_t, int64_t, hipStream_t);

template <typename T>
hipError_t launch_line_fast_batch(const StageArgs<T> *items, const int64_t *prefix, int nitems, int64_t total, int r0,
                                  int r1, int r2, int em, int threads, size_t lds_bytes, int64_t grid, hipStream_t)
{
    if (!fast_plan_available((int)sizeof(T), r0, r1, r2, em)) fail("line_fast_batch: plan not instantiated");
    const int fr[3] = {r0, r1, r2};
    emulate_batch("line_fast_batch", items, prefix, nitems, total, threads, lds_bytes, grid, true, em, fr);
    return hipSuccess;
}
template hipError_t launch_line_fast_batch<float>(const StageArgs<float> *, const int64_t *, int, int64_t, int, int,
                                                  int, int, int, size_t, int64_t, hipStream_t);
template hipError_t launch_line_fast_batch<double>(const StageArgs<double> *, const int64_t *, int, int64_t, int, int,
                                                   int, int, int, size_t, int64_t, hipStream_t);

template <typename T>
hipError_t launch_line_multi_batch(const StageArgs<T> *items, const int64_t *prefix, int nitems, int64_t total,
                                   int threads, size_t lds_bytes, int64_t grid, hipStream_t)
{
    // every item must name an instantiated plan of the standard register budget
    for (int q = 0; q < nitems; ++q) {
        const StageArgs<T> &a = items[q];
        const int r2 = a.npass > 2 ? a.radix[2] : 0;
        if (!fast_plan_available((int)sizeof(T), a.radix[0], a.radix[1], r2, a.fem) || a.fem == 32)
            fail("line_multi_batch: item %d has no plan in the kernel", q);
    }
    ++g_launches;
    ++g_all_launches;
    begin_launch();
    if (threads > 512) fail("line_multi_batch: %d threads", threads);
    if (grid <= 0 || grid > total) fail("line_multi_batch: grid");
    if (prefix[0] != 0 || prefix[nitems] != total) fail("line_multi_batch: prefix ends");
    if (g_trace) {
        double bytes = 0;
        for (int q = 0; q < nitems; ++q)
            bytes += (double)items[q].nlines * (items[q].nin + items[q].nout) * 2.0 * sizeof(T);
        std::printf("  line_multi_batch items %d tiles %" PRId64 " threads %d grid %" PRId64 " lds %zu bytes %.0f\n",
                    nitems, total, threads, grid, lds_bytes, bytes);
    }
    for (int q = 0; q < nitems; ++q) {
        const StageArgs<T> &a = items[q];
        if (prefix[q + 1] - prefix[q] != a.ntiles) fail("line_multi_batch: item %d tiles", q);
        check_shape("line_multi_batch", a, threads, lds_bytes, Regs<T>::emax, true, a.fem);
        emulate("line_multi_batch", a);
    }
    return hipSuccess;
}
template hipError_t launch_line_multi_batch<float>(const StageArgs<float> *, const int64_t *, int, int64_t, int,
                                                   size_t, int64_t, hipStream_t);
template hipError_t launch_line_multi_batch<double>(const StageArgs<double> *, const int64_t *, int, int64_t, int,
                                                    size_t, int64_t, hipStream_t);

template <typename T>
int line_multi_batch_blocks_per_cu(int, size_t)
{
    return 2;
}
template int line_multi_batch_blocks_per_cu<float>(int, size_t);
template int line_multi_batch_blocks_per_cu<double>(int, size_t);

template <typename T>
int line_batch_blocks_per_cu(bool, int, size_t)
{
    return 2;
}
template int line_batch_blocks_per_cu<float>(bool, int, size_t);
template int line_batch_blocks_per_cu<double>(bool, int, size_t);

template <typename T>
int line_fast_batch_blocks_per_cu(int, int, int, int, int, size_t)
{
    return 2;
}
template int line_fast_batch_blocks_per_cu<float>(int, int, int, int, int, size_t);
template int line_fast_batch_blocks_per_cu<double>(int, int, int, int, int, size_t);

hipError_t set_line_stage_lds_limit(size_t) { return hipSuccess; }
hipError_t set_plane_lds_limit(size_t) { return hipSuccess; }

bool plane_plan_available(int tbytes, int r0, int r1, int r2, int em)
{
#define PFH_PLANE_HAS(TT, A, B, C, E) \
    if ((int)sizeof(TT) == tbytes && r0 == A && r1 == B && r2 == C && em == E) return true;
    PFH_PLANE_PLANS(PFH_PLANE_HAS)
#undef PFH_PLANE_HAS
    return false;
}

template <typename T>
int plane_batch_blocks_per_cu(int, size_t)
{
    return 1;
}
template int plane_batch_blocks_per_cu<float>(int, size_t);

// Plane stage: every plane's element (row r, column c) is read from
// src + r nr + c and written to dst + c nc + r (the two transforms replaced by
// the identity); shape checks against the kernel's assumptions.
template <typename T>
hipError_t launch_plane_batch(const PlaneArgs<T> *items, const int64_t *prefix, int nitems, int64_t total,
                              int threads, int tw_cap, int work_cap, int raw_cap, int64_t grid, hipStream_t)
{
    ++g_launches;
    ++g_all_launches;
    begin_launch();
    const size_t lds_bytes = (size_t)(tw_cap + work_cap + raw_cap) * sizeof(cpx<T>);
    if ((tw_cap + work_cap) % 2 != 0 || raw_cap < 0) fail("plane_batch: LDS regions");
    if (threads != kPlaneThreads && threads != kPlaneThreads / 2 && threads != kPlaneThreads / 4)
        fail("plane_batch: %d threads", threads);
    if (lds_bytes > kLdsLimit) fail("plane_batch: %zu bytes of LDS", lds_bytes);
    if (grid <= 0 || grid > total) fail("plane_batch: grid");
    if (prefix[0] != 0 || prefix[nitems] != total)
        fail("plane_batch: prefix ends %lld %lld %lld (nitems %d)", (long long)prefix[0], (long long)prefix[nitems], (long long)total, nitems);
    if (g_trace) {
        double bytes = 0;
        for (int q = 0; q < nitems; ++q) bytes += (double)items[q].nplanes * items[q].nr * items[q].nc * 4.0 * sizeof(T);
        std::printf("  plane_batch items %d planes %" PRId64 " threads %d grid %" PRId64 " lds %zu bytes %.0f\n", nitems,
                    total, threads, grid, lds_bytes, bytes);
    }
    for (int q = 0; q < nitems; ++q) {
        const PlaneArgs<T> &a = items[q];
        if (prefix[q + 1] - prefix[q] != a.nplanes) fail("plane_batch: item %d planes", q);
        if (!plane_plan_available((int)sizeof(T), a.rr[0], a.rr[1], a.rr[2], a.rem) ||
            !plane_plan_available((int)sizeof(T), a.cr[0], a.cr[1], a.cr[2], a.cem))
            fail("plane_batch: item %d has no plane plan", q);
        const int mr = a.rr[0] * a.rr[1] * (a.rr[2] ? a.rr[2] : 1), mc = a.cr[0] * a.cr[1] * (a.cr[2] ? a.cr[2] : 1);
        if (mr != a.nr || mc != a.nc) fail("plane_batch: item %d plan lengths %d/%d vs %d/%d", q, mr, mc, a.nr, a.nc);
        // every pass's butterflies are covered by the threads of a line
        auto covered = [](int M, const int r[3], int em, int tpl) {
            for (int p = 0; p < 3; ++p)
                if (r[p] && tpl * std::max(1, em / r[p]) < M / r[p]) return false;
            return tpl <= 256;
        };
        if (!covered(mr, a.rr, a.rem, a.tpl_r) || !covered(mc, a.cr, a.cem, a.tpl_c))
            fail("plane_batch: item %d threads per line", q);
        if (a.rs < (a.nr - 1) + ((a.nr - 1) >> 4) + 1 || a.rs % 2 == 0) fail("plane_batch: item %d row stride", q);
        // the load phase moves pairs of values: even rows, at most one pair per thread per row
        if (a.nr % 2 != 0 || a.nr / 2 > threads) fail("plane_batch: item %d row length %d", q, a.nr);
        if (a.twn_row + a.twn_col > tw_cap || (int64_t)a.nc * a.rs > work_cap) fail("plane_batch: item %d LDS", q);
        check_read("plane_batch tables", a.tw_row, (size_t)a.twn_row * sizeof(cpx<T>));
        check_read("plane_batch tables", a.tw_col, (size_t)a.twn_col * sizeof(cpx<T>));
        const int64_t pe = (int64_t)a.nr * a.nc;
        for (int64_t p = 0; p < a.nplanes; ++p) {
            const cpx<T> *s = static_cast<const cpx<T> *>(a.src) + p * pe;
            cpx<T> *d = static_cast<cpx<T> *>(a.dst) + p * pe;
            if (!check_read("plane_batch", s, (size_t)pe * sizeof(cpx<T>))) return hipSuccess;
            for (int r = 0; r < a.nc; ++r)
                for (int c = 0; c < a.nr; ++c) {
                    if (!check_write("plane_batch", d + (int64_t)c * a.nc + r, sizeof(cpx<T>))) return hipSuccess;
                    d[(int64_t)c * a.nc + r] = s[(int64_t)r * a.nr + c];
                }
        }
    }
    return hipSuccess;
}
template hipError_t launch_plane_batch<float>(const PlaneArgs<float> *, const int64_t *, int, int64_t, int, int, int,
                                              int, int64_t, hipStream_t);
hipError_t set_fast_lds_limit(size_t) { return hipSuccess; }

template <typename T>
hipError_t launch_endpoint_check(const EndpointArgs &e, const void *data, EndpointPartial *part, int *gate,
                                 int *status, hipStream_t)
{
    ++g_launches;
    ++g_all_launches;
    begin_launch();
    const int64_t count = e.outer * e.stride;
    // the partial kernel covers at most kEndpointPerThread x 256 elements of a chunk
    if (e.nblk < 1 || (count + e.nblk - 1) / e.nblk > 256 * kEndpointPerThread)
        fail("endpoint chunk of more than %d elements", 256 * kEndpointPerThread);
    for (int64_t b = 0; b < e.extra; ++b)
        for (int p = 0; p < e.planes; ++p) {
            const int64_t k = p == 0 ? 0 : e.clen - 1;
            for (int64_t i = 0; i < count; ++i) {
                const int64_t o = i / e.stride, q = i - o * e.stride;
                const char *v = static_cast<const char *>(data) +
                                (b * e.len + k * e.stride + o * e.clen * e.stride + q) * 2 * (int64_t)sizeof(T);
                if (!check_read("endpoint_check", v, 2 * sizeof(T))) return hipSuccess;
            }
        }
    check_write("endpoint_check partials", part, sizeof(EndpointPartial) * e.extra * e.planes * e.nblk);
    check_read("endpoint_check status", status, sizeof(int));
    if (g_reject_endpoints) {
        *gate = 1;
        *status |= PFH_DEVICE_INVALID_SPECTRUM;
        if (e.poison) *e.poison = 1;
    } else if (!e.keep_gate) {
        *gate = 0; // accept
    }
    return hipSuccess;
}
template hipError_t launch_endpoint_check<float>(const EndpointArgs &, const void *, EndpointPartial *, int *, int *,
                                                 hipStream_t);
template hipError_t launch_endpoint_check<double>(const EndpointArgs &, const void *, EndpointPartial *, int *,
                                                  int *, hipStream_t);

template <typename T>
hipError_t launch_local_endpoint_check(const void *data, int64_t lines, int64_t clen, int nyq,
                                       unsigned long long *first, int *gate, int *status, hipStream_t)
{
    ++g_launches;
    ++g_all_launches;
    (void)nyq;
    check_read("local_endpoint_check", data, (size_t)(lines * clen) * 2 * sizeof(T));
    check_read("local_endpoint_check first", first, sizeof(*first));
    check_read("local_endpoint_check status", status, sizeof(int));
    *gate = 0;
    return hipSuccess;
}
template hipError_t launch_local_endpoint_check<float>(const void *, int64_t, int64_t, int, unsigned long long *,
                                                       int *, int *, hipStream_t);
template hipError_t launch_local_endpoint_check<double>(const void *, int64_t, int64_t, int, unsigned long long *,
                                                        int *, int *, hipStream_t);

hipError_t launch_fill_i32(int *p, int64_t n, int v, hipStream_t)
{
    ++g_all_launches;
    if (n > 0 && !check_read("fill_i32", p, (size_t)n * sizeof(int))) return hipSuccess;
    for (int64_t i = 0; i < n; ++i) p[i] = v;
    return hipSuccess;
}

hipError_t launch_fill_u64(unsigned long long *p, int64_t n, unsigned long long v, hipStream_t)
{
    ++g_all_launches;
    if (n > 0 && !check_read("fill_u64", p, (size_t)n * sizeof(*p))) return hipSuccess;
    for (int64_t i = 0; i < n; ++i) p[i] = v;
    return hipSuccess;
}

hipError_t launch_inplace_begin(int *gate, const int *poison, int *status, hipStream_t)
{
    ++g_all_launches;
    *gate = *poison != 0 ? 1 : 0;
    if (*poison) *status |= PFH_DEVICE_POISONED;
    return hipSuccess;
}

hipError_t launch_permute_words(const PermuteArgs &p, const void *src, void *dst, int word, hipStream_t)
{
    ++g_launches;
    ++g_all_launches;
    begin_launch();
    const int64_t total = p.len * p.extra;
    for (int64_t i = 0; i < total; ++i) {
        const int64_t e = i / p.len;
        int64_t rem = i - e * p.len, so = 0;
        for (int d = p.nd - 1; d >= 0; --d) {
            so += (rem % p.dext[d]) * p.sstride[d];
            rem /= p.dext[d];
        }
        const char *s = static_cast<const char *>(src) + (e * p.len + so) * word;
        char *d = static_cast<char *>(dst) + i * word;
        if (!check_read("permute", s, word) || !check_write("permute", d, word)) return hipSuccess;
        std::memcpy(d, s, word);
    }
    return hipSuccess;
}

template <typename T>
hipError_t launch_scale_axes(const ScaleAxesArgs &s, void *data, const void *factors, hipStream_t)
{
    ++g_launches;
    ++g_all_launches;
    const int64_t total = s.len * s.extra;
    const size_t eb = (s.complex_data ? 2 : 1) * sizeof(T);
    check_read("scale_axes", data, (size_t)total * eb);
    int64_t nf = 0;
    for (int p = 0; p < s.nd; ++p) nf = std::max(nf, s.foff[p] + s.mext[p]);
    check_read("scale_axes factors", factors, (size_t)nf * sizeof(T) * (s.complex_factors ? 2 : 1));
    if (s.complex_factors) {
        if (!s.complex_data) fail("scale_axes: complex factors on real data");
        return hipSuccess; // bounds only
    }
    T *d = static_cast<T *>(data);
    const T *f = static_cast<const T *>(factors);
    for (int64_t i = 0; i < total; ++i) {
        int64_t rem = i % s.len;
        T w = (T)1;
        for (int p = s.nd - 1; p >= 0; --p) {
            w *= f[s.foff[p] + rem % s.mext[p]];
            rem /= s.mext[p];
        }
        if (s.complex_data) {
            d[2 * i] *= w;
            d[2 * i + 1] *= w;
        } else {
            d[i] *= w;
        }
    }
    return hipSuccess;
}
template hipError_t launch_scale_axes<float>(const ScaleAxesArgs &, void *, const void *, hipStream_t);
template hipError_t launch_scale_axes<double>(const ScaleAxesArgs &, void *, const void *, hipStream_t);

template <typename T>
hipError_t launch_quadratic(const void *w, void *u, int64_t n, double a, hipStream_t)
{
    check_read("quadratic", w, n * sizeof(T));
    check_read("quadratic out", u, n * sizeof(T));
    for (int64_t i = 0; i < n; ++i) {
        const T v = static_cast<const T *>(w)[i];
        static_cast<T *>(u)[i] = v + (T)a * v * v;
    }
    return hipSuccess;
}
template hipError_t launch_quadratic<float>(const void *, void *, int64_t, double, hipStream_t);
template hipError_t launch_quadratic<double>(const void *, void *, int64_t, double, hipStream_t);

template <typename T>
hipError_t launch_quadratic_adjoint(const void *w, const void *g, void *out, int64_t n, double a, hipStream_t)
{
    check_read("quadratic_adjoint", w, n * sizeof(T));
    check_read("quadratic_adjoint g", g, n * sizeof(T));
    check_read("quadratic_adjoint out", out, n * sizeof(T));
    for (int64_t i = 0; i < n; ++i)
        static_cast<T *>(out)[i] =
            static_cast<const T *>(g)[i] * ((T)1 + (T)2 * (T)a * static_cast<const T *>(w)[i]);
    return hipSuccess;
}
template hipError_t launch_quadratic_adjoint<float>(const void *, const void *, void *, int64_t, double, hipStream_t);
template hipError_t launch_quadratic_adjoint<double>(const void *, const void *, void *, int64_t, double,
                                                     hipStream_t);

template <typename T>
hipError_t launch_sum_squares(const void *x, int64_t n, double *partial, int nparts, double *result, hipStream_t)
{
    check_read("sum_squares", x, n * sizeof(T));
    check_read("sum_squares partials", partial, nparts * sizeof(double));
    double acc = 0;
    for (int64_t i = 0; i < n; ++i) acc += (double)static_cast<const T *>(x)[i] * (double)static_cast<const T *>(x)[i];
    *result = acc;
    return hipSuccess;
}
template hipError_t launch_sum_squares<float>(const void *, int64_t, double *, int, double *, hipStream_t);
template hipError_t launch_sum_squares<double>(const void *, int64_t, double *, int, double *, hipStream_t);

// Reductions: bounds of the input, the partial-state slots and the result.
hipError_t launch_reduce(int op, int scalar_type, const void *x, int64_t n, void *workspace, ReduceResult *result,
                         hipStream_t)
{
    static const int width[13] = {0, 1, 1, 2, 2, 4, 4, 8, 8, 4, 8, 8, 16};
    if (op < kReduceSum || op > kReduceAll || scalar_type < 1 || scalar_type > 12) return hipErrorInvalidValue;
    begin_launch();
    if (n > 0) check_read("reduce input", x, (size_t)n * width[scalar_type]);
    int64_t parts = n / (256 * 16);
    parts = parts < 1 ? 1 : parts > kReduceParts ? kReduceParts : parts;
    check_write("reduce partials", workspace, (size_t)parts * kReduceStateBytes);
    begin_launch();
    check_write("reduce result", result, sizeof(ReduceResult));
    return hipSuccess;
}

// Broadcast pointwise: every (extra, spatial) element of every input lies in
// one allocation and each output element is written once.
hipError_t launch_pointwise(const PointwiseArgs &a, void *out, bool f64, hipStream_t)
{
    begin_launch();
    const size_t real = f64 ? 8 : 4;
    const size_t oeb = real * (a.out_complex ? 2 : 1);
    for (int64_t e = 0; e < a.extra_count; ++e) {
        int64_t rem = e;
        int64_t idx[4] = {0, 0, 0, 0};
        for (int d = a.extra_ndim - 1; d >= 0; --d) {
            idx[d] = rem % a.extra_dims[d];
            rem /= a.extra_dims[d];
        }
        for (int i = 0; i < a.nin; ++i) {
            int64_t lin = 0;
            for (int d = 0; d < a.extra_ndim; ++d) lin += idx[d] * a.extra_stride[i][d];
            const size_t ieb = real * (a.in_complex[i] ? 2 : 1);
            if (!check_read("pointwise input", static_cast<const char *>(a.in[i]) + (size_t)(lin * a.spatial) * ieb,
                            (size_t)a.spatial * ieb))
                return hipSuccess;
        }
        if (!check_write("pointwise output", static_cast<char *>(out) + (size_t)(e * a.spatial) * oeb,
                         (size_t)a.spatial * oeb))
            return hipSuccess;
    }
    return hipSuccess;
}

} // namespace pfh

// ---------------------------------------------------------------------------
// scenarios
// ---------------------------------------------------------------------------

namespace {

struct Arr {
    std::vector<int> perm;
    std::vector<int64_t> shape;
    int64_t elems = 0, eb = 0;
};

Arr layout(pfh_plan p, int side, int nd)
{
    int perm[PFH_MAX_DIMS];
    size_t shape[PFH_MAX_DIMS], el = 0, eb = 0;
    if (pfh_plan_layout(p, side, perm, shape, &el, &eb) != PFH_OK) fail("layout query");
    Arr a;
    for (int i = 0; i < nd; ++i) {
        a.perm.push_back(perm[i]);
        a.shape.push_back((int64_t)shape[i]);
    }
    a.elems = (int64_t)el;
    a.eb = (int64_t)eb;
    return a;
}

// memory offset (elements) of logical coordinates c in extra block e
int64_t offset_of(const Arr &a, const std::vector<int64_t> &c, int64_t e)
{
    int64_t len = 1;
    for (int64_t s : a.shape) len *= s;
    int64_t off = 0;
    for (size_t p = 0; p < a.perm.size(); ++p) off = off * a.shape[a.perm[p]] + c[a.perm[p]];
    return e * len + off;
}

void *dmalloc(size_t n, bool zero = false)
{
    void *p = nullptr;
    if (hipMalloc(&p, n) != hipSuccess) return nullptr;
    if (zero) std::memset(p, 0, n);
    return p;
}

// Runs op on coded input and checks the identity-on-points semantics.
template <typename T>
void run_case(const char *name, pfh_plan plan, int nd, int op, bool real_in, bool real_out, int red_axis,
              int64_t red_n)
{
    const bool fwd_like = (op == PFH_FORWARD || op == PFH_ADJOINT_INVERSE || op == PFH_ADJOINT_BACKWARD);
    Arr ain = layout(plan, fwd_like ? 0 : 1, nd), aout = layout(plan, fwd_like ? 1 : 0, nd);
    const int64_t ei = real_in ? sizeof(T) : 2 * sizeof(T), eo = real_out ? sizeof(T) : 2 * sizeof(T);
    void *src = dmalloc(ain.elems * ei), *dst = dmalloc(aout.elems * eo);
    T *s = static_cast<T *>(src);
    for (int64_t i = 0; i < ain.elems; ++i) {
        if (real_in) {
            s[i] = (T)(i + 1);
        } else {
            s[2 * i] = (T)(i + 1);
            s[2 * i + 1] = (T)(-(i + 1));
        }
    }
    size_t wsb = 0;
    pfh_plan_workspace_bytes(plan, op, &wsb);
    void *ws = dmalloc(wsb, true);
    const int before = g_fail;
    pfh_status st = pfh_execute(plan, op, src, ain.elems, dst, aout.elems, ws, wsb, nullptr, 0);
    if (st != PFH_OK) {
        pfh_error_info inf;
        pfh_last_error(&inf);
        fail("%s op %d: status %d (%s)", name, op, (int)st, inf.message);
    }
    // semantics: out(c) = in(c) on the common points
    int64_t len_out = 1;
    for (int64_t v : aout.shape) len_out *= v;
    const int64_t extra = aout.elems / len_out;
    const T *d = static_cast<const T *>(dst);
    std::vector<int64_t> c(nd, 0);
    int64_t bad = 0;
    for (int64_t e = 0; e < extra && bad < 3; ++e) {
        for (int64_t lin = 0; lin < len_out && bad < 3; ++lin) {
            int64_t r = lin;
            for (int a = nd - 1; a >= 0; --a) {
                c[a] = r % aout.shape[a];
                r /= aout.shape[a];
            }
            const int64_t oo = offset_of(aout, c, e);
            double er = 0, eim = 0;
            const bool inside = red_axis < 0 || c[red_axis] < ain.shape[red_axis];
            if (inside) {
                const int64_t io = offset_of(ain, c, e);
                er = real_in ? (double)s[io] : (double)s[2 * io];
                eim = real_in ? 0.0 : (double)s[2 * io + 1];
            }
            (void)red_n;
            const double gr = real_out ? (double)d[oo] : (double)d[2 * oo];
            const double gi = real_out ? 0.0 : (double)d[2 * oo + 1];
            const double wi = (real_in || real_out) ? 0.0 : eim;
            if (gr != er || gi != wi) {
                fail("%s op %d: extra %" PRId64 " logical index %" PRId64 ": got (%g, %g) want (%g, %g)", name, op, e,
                     lin, gr, gi, er, wi);
                ++bad;
            }
        }
    }
    std::printf("%s op %d: %s\n", name, op, g_fail == before ? "ok" : "FAILED");
    (void)hipFree(src);
    (void)hipFree(dst);
    (void)hipFree(ws);
}

pfh_desc route_desc(int kind, int prec, std::vector<size_t> shape, std::vector<size_t> extra)
{
    pfh_desc d;
    std::memset(&d, 0, sizeof(d));
    d.kind = kind;
    d.precision = prec;
    d.ndim = (int)shape.size();
    for (size_t i = 0; i < shape.size(); ++i) d.shape[i] = shape[i];
    d.extra_ndim = (int)extra.size();
    for (size_t i = 0; i < extra.size(); ++i) d.extra_shape[i] = extra[i];
    d.permute_dims = 1;
    return d;
}

// Plans whose stages would index more than 2^31 - 1 lines (the kernels'
// 32-bit line division) are refused at creation; 2^30 lines still plan.
static void line_count_guard()
{
    const int before = g_fail;
    for (size_t ex : {256, 512}) {
        pfh_desc d = route_desc(PFH_C2C, PFH_F32, {2048, 2048, 64}, {ex});
        pfh_plan p = nullptr;
        const pfh_status st = pfh_plan_create(&d, &p);
        if (ex == 256 && st != PFH_OK) fail("line-count guard: 2^30 lines refused (%d)", (int)st);
        if (ex == 512) {
            pfh_error_info e;
            pfh_last_error(&e);
            if (st != PFH_ERR_UNSUPPORTED || std::strstr(e.message, "2^31") == nullptr)
                fail("line-count guard: 2^31 lines gave status %d (%s)", (int)st, e.message);
        }
        if (p) pfh_plan_destroy(p);
    }
    std::printf("line-count guard: %s\n", g_fail == before ? "ok" : "FAILED");
}

template <typename T>
void route_cases(int prec)
{
    const char *pn = prec == PFH_F32 ? "f32" : "f64";
    struct C {
        std::vector<size_t> shape, extra;
    };
    std::vector<C> cases = {{{64, 64, 64}, {}},   {{64, 64, 64}, {3}},  {{48, 60, 64}, {}}, {{16, 96, 40}, {2}},
                            {{3, 4, 5}, {}},      {{128, 32, 8}, {}},   {{2, 3, 2, 3}, {}}, {{8, 12, 16, 4}, {}},
                            {{513, 4, 6}, {}},    {{40, 48, 36}, {3}},  {{67, 32, 40}, {}}, {{37, 41, 64}, {2}}};
    for (const C &c : cases) {
        char name[128];
        std::string sh;
        for (size_t v : c.shape) sh += std::to_string(v) + "x";
        sh.pop_back();
        if (!c.extra.empty()) sh += "+extra" + std::to_string(c.extra[0]);
        const int nd = (int)c.shape.size();
        // C2C
        {
            pfh_desc d = route_desc(PFH_C2C, prec, c.shape, c.extra);
            pfh_plan p = nullptr;
            if (pfh_plan_create(&d, &p) != PFH_OK) {
                fail("C2C %s %s: plan", pn, sh.c_str());
                continue;
            }
            std::snprintf(name, sizeof(name), "C2C %s %s", pn, sh.c_str());
            for (int op : {PFH_FORWARD, PFH_INVERSE, PFH_ADJOINT_FORWARD})
                run_case<T>(name, p, nd, op, false, false, -1, 0);
            pfh_plan_destroy(p);
        }
        // R2C
        {
            pfh_desc d = route_desc(PFH_R2C, prec, c.shape, c.extra);
            pfh_plan p = nullptr;
            if (pfh_plan_create(&d, &p) != PFH_OK) {
                fail("R2C %s %s: plan", pn, sh.c_str());
                continue;
            }
            std::snprintf(name, sizeof(name), "R2C %s %s", pn, sh.c_str());
            const int ra = nd - 1;
            run_case<T>(name, p, nd, PFH_FORWARD, true, false, ra, (int64_t)c.shape[ra]);
            run_case<T>(name, p, nd, PFH_ADJOINT_INVERSE, true, false, ra, (int64_t)c.shape[ra]);
            run_case<T>(name, p, nd, PFH_INVERSE, false, true, ra, (int64_t)c.shape[ra]);
            run_case<T>(name, p, nd, PFH_BACKWARD, false, true, ra, (int64_t)c.shape[ra]);
            pfh_plan_destroy(p);
        }
        // C2R orientation (C2rPlan): forward = C2R from the canonical complex side
        {
            pfh_desc d = route_desc(PFH_C2R, prec, c.shape, c.extra);
            pfh_plan p = nullptr;
            if (pfh_plan_create(&d, &p) != PFH_OK) {
                fail("C2R %s %s: plan", pn, sh.c_str());
                continue;
            }
            std::snprintf(name, sizeof(name), "C2Rplan %s %s", pn, sh.c_str());
            run_case<T>(name, p, nd, PFH_FORWARD, false, true, 0, (int64_t)c.shape[0]);
            run_case<T>(name, p, nd, PFH_BACKWARD, true, false, 0, (int64_t)c.shape[0]);
            run_case<T>(name, p, nd, PFH_INVERSE, true, false, 0, (int64_t)c.shape[0]);
            run_case<T>(name, p, nd, PFH_ADJOINT_FORWARD, true, false, 0, (int64_t)c.shape[0]);
            pfh_plan_destroy(p);
        }
        // mixed plans: complex R2R stages run as real lines with a (re, im) digit
        {
            pfh_desc d = route_desc(PFH_MIXED_C2C, prec, c.shape, c.extra);
            const int codes[4] = {PFH_AXIS_FFT, PFH_DCT2, PFH_DHT, PFH_DST1};
            for (int a = 0; a < nd; ++a) d.axis_kind[a] = codes[a % 4];
            pfh_plan p = nullptr;
            if (pfh_plan_create(&d, &p) != PFH_OK) {
                fail("MixedC2C %s %s: plan", pn, sh.c_str());
                continue;
            }
            std::snprintf(name, sizeof(name), "MixedC2C %s %s", pn, sh.c_str());
            for (int op : {PFH_FORWARD, PFH_INVERSE, PFH_ADJOINT_FORWARD})
                run_case<T>(name, p, nd, op, false, false, -1, 0);
            pfh_plan_destroy(p);
        }
        {
            pfh_desc d = route_desc(PFH_MIXED_R2C, prec, c.shape, c.extra);
            // FFT / R2R before the RFFT axis (axis 1), real R2R after it
            const int codes[4] = {PFH_AXIS_FFT, PFH_AXIS_RFFT, PFH_DCT4, PFH_DHT};
            for (int a = 0; a < nd; ++a) d.axis_kind[a] = codes[a];
            pfh_plan p = nullptr;
            if (pfh_plan_create(&d, &p) != PFH_OK) {
                fail("MixedR2C %s %s: plan", pn, sh.c_str());
                continue;
            }
            std::snprintf(name, sizeof(name), "MixedR2C %s %s", pn, sh.c_str());
            run_case<T>(name, p, nd, PFH_FORWARD, true, false, 1, (int64_t)c.shape[1]);
            run_case<T>(name, p, nd, PFH_INVERSE, false, true, 1, (int64_t)c.shape[1]);
            run_case<T>(name, p, nd, PFH_ADJOINT_INVERSE, true, false, 1, (int64_t)c.shape[1]);
            pfh_plan_destroy(p);
        }
        // R2R (DCT-II, DST-I, DCT-IV, DHT per axis)
        {
            pfh_desc d = route_desc(PFH_R2R, prec, c.shape, c.extra);
            const int kinds[4] = {PFH_DCT2, PFH_DST1, PFH_DCT4, PFH_DHT};
            for (int a = 0; a < nd; ++a) d.axis_kind[a] = kinds[a % 4];
            pfh_plan p = nullptr;
            if (pfh_plan_create(&d, &p) != PFH_OK) {
                fail("R2R %s %s: plan", pn, sh.c_str());
                continue;
            }
            std::snprintf(name, sizeof(name), "R2R %s %s", pn, sh.c_str());
            for (int op : {PFH_FORWARD, PFH_INVERSE})
                run_case<T>(name, p, nd, op, true, true, -1, 0);
            pfh_plan_destroy(p);
        }
    }
    // R2R kinds whose FFT length has a compile-time plan (fast first pass
    // gathers the pre-processed values)
    struct K {
        std::vector<size_t> shape, extra;
        std::vector<int> kinds;
    };
    const std::vector<K> fast_r2r = {{{128, 120, 63}, {}, {PFH_DCT2, PFH_DCT4, PFH_DST1}},
                                     {{120, 64, 128}, {2}, {PFH_DST2, PFH_DCT3, PFH_DST4}},
                                     {{65, 6, 128}, {}, {PFH_DCT1, 0, PFH_DCT3}},
                                     {{128, 127, 2}, {}, {PFH_DST3, PFH_DST1, 0}},
                                     // 384-point DST-I lines (rows, a column axis, an odd line count)
                                     {{384, 3, 5}, {}, {PFH_DST1, 0, PFH_DCT2}},
                                     {{5, 384, 7}, {2}, {0, PFH_DST1, 0}},
                                     {{3, 384}, {}, {0, PFH_DST1}}};
    for (const K &c : fast_r2r) {
        char name[128];
        std::string sh;
        for (size_t v : c.shape) sh += std::to_string(v) + "x";
        pfh_desc d = route_desc(PFH_R2R, prec, c.shape, c.extra);
        for (size_t a = 0; a < c.shape.size(); ++a) d.axis_kind[a] = c.kinds[a];
        pfh_plan p = nullptr;
        if (pfh_plan_create(&d, &p) != PFH_OK) {
            fail("fast R2R %s %s: plan", pn, sh.c_str());
            continue;
        }
        std::snprintf(name, sizeof(name), "fast R2R %s %s", pn, sh.c_str());
        for (int op : {PFH_FORWARD, PFH_INVERSE, PFH_ADJOINT_FORWARD, PFH_ADJOINT_INVERSE})
            run_case<T>(name, p, (int)c.shape.size(), op, true, true, -1, 0);
        pfh_plan_destroy(p);
    }
    // DHT on every axis, lengths whose half has a compile-time plan
    const std::vector<C> dht_cases = {{{128, 96, 40}, {}}, {{6, 128}, {3}}, {{128, 5, 96}, {2}}};
    for (const C &c : dht_cases) {
        char name[128];
        std::string sh;
        for (size_t v : c.shape) sh += std::to_string(v) + "x";
        pfh_desc d = route_desc(PFH_R2R, prec, c.shape, c.extra);
        for (size_t a = 0; a < c.shape.size(); ++a) d.axis_kind[a] = PFH_DHT;
        pfh_plan p = nullptr;
        if (pfh_plan_create(&d, &p) != PFH_OK) {
            fail("DHT %s %s: plan", pn, sh.c_str());
            continue;
        }
        std::snprintf(name, sizeof(name), "DHT %s %s", pn, sh.c_str());
        for (int op : {PFH_FORWARD, PFH_INVERSE, PFH_ADJOINT_FORWARD})
            run_case<T>(name, p, (int)c.shape.size(), op, true, true, -1, 0);
        pfh_plan_destroy(p);
    }
}

void batch_cases()
{
    // the first, {128, 112, 20}, {96, 60, 7} and {160, 84, 5} fuse the inner
    // stage pair into plane stages ((x, y) planes), {160, 128, 48} the outer
    // pair ((y, z) planes: its (x, y) plane does not fit in LDS)
    std::vector<std::vector<size_t>> shapes = {{48, 60, 64},   {12, 7, 9},   {84, 16, 20}, {5, 5, 5},
                                               {64, 32, 16},   {128, 112, 20}, {96, 60, 7}, {160, 84, 5},
                                               {160, 128, 48}};
    std::vector<pfh_desc> ds;
    for (auto &s : shapes) ds.push_back(route_desc(PFH_C2C, PFH_F32, s, {3}));
    // mode 2: problems in groups of at most 1 MiB that share one workspace region
    for (int mode : {0, 1, 2}) {
        const int tune = mode == 1;
        if (mode == 2) setenv("PFH_BATCH_GROUP_MB", "1", 1);
        pfh_batch b = nullptr;
        if (pfh_batch_create(ds.data(), (int)ds.size(), tune, &b) != PFH_OK) {
            fail("batch create (autotune %d)", tune);
            continue;
        }
        size_t wsb = 0;
        pfh_batch_workspace_bytes(b, &wsb);
        void *ws = dmalloc(wsb, true);
        std::vector<const void *> srcs;
        std::vector<void *> dsts;
        std::vector<int64_t> n;
        for (auto &s : shapes) {
            const int64_t e = (int64_t)(s[0] * s[1] * s[2] * 3);
            n.push_back(e);
            float *x = static_cast<float *>(dmalloc(e * 8));
            for (int64_t i = 0; i < 2 * e; ++i) x[i] = (float)(i + 1);
            srcs.push_back(x);
            dsts.push_back(dmalloc(e * 8));
        }
        const int before = g_fail;
        for (int op : {PFH_FORWARD, PFH_INVERSE})
            if (pfh_batch_execute(b, op, srcs.data(), dsts.data(), ws, wsb, nullptr, 0) != PFH_OK)
                fail("batch execute op %d", op);
        // the batch must give the single plans' layout
        for (size_t i = 0; i < shapes.size(); ++i) {
            pfh_plan p = nullptr;
            pfh_desc d = ds[i];
            pfh_plan_create(&d, &p);
            void *ref = dmalloc(n[i] * 8);
            size_t w2 = 0;
            pfh_plan_workspace_bytes(p, PFH_INVERSE, &w2);
            void *ws2 = dmalloc(w2, true);
            pfh_execute(p, PFH_INVERSE, srcs[i], n[i], ref, n[i], ws2, w2, nullptr, 0);
            if (std::memcmp(ref, dsts[i], n[i] * 8) != 0) fail("batch problem %zu differs from its plan", i);
            (void)hipFree(ref);
            (void)hipFree(ws2);
            pfh_plan_destroy(p);
        }
        std::printf("batch mode %d: %s\n", mode, g_fail == before ? "ok" : "FAILED");
        if (mode == 2) unsetenv("PFH_BATCH_GROUP_MB");
        for (auto *x : srcs) (void)hipFree(const_cast<void *>(x));
        for (auto *x : dsts) (void)hipFree(x);
        (void)hipFree(ws);
        pfh_batch_destroy(b);
    }
}

} // namespace

// Broadcast pointwise calls over tight allocations: every input and output
// element the launch touches lies inside its own buffer.
void pointwise_cases()
{
    auto layout = [](pfh_array &a, std::vector<size_t> shape, std::vector<int> perm, std::vector<size_t> extra,
                     int type) {
        std::memset(&a, 0, sizeof(a));
        a.scalar_type = type;
        a.extra_ndim = (int)extra.size();
        for (size_t d = 0; d < extra.size(); ++d) a.extra_shape[d] = extra[d];
        a.pencil.ndim = (int)shape.size();
        for (size_t d = 0; d < shape.size(); ++d) {
            a.pencil.shape[d] = shape[d];
            a.pencil.perm[d] = perm[d];
        }
        size_t n = (type == PFH_SCALAR_C64 || type == PFH_SCALAR_F64) ? 8 : type == PFH_SCALAR_C128 ? 16 : 4;
        for (size_t v : shape) n *= v;
        for (size_t v : extra) n *= v;
        void *p = nullptr;
        if (hipMalloc(&p, n) != hipSuccess) fail("pointwise alloc");
        a.data = p;
    };
    struct Case {
        int op, type_in, type_out;
        std::vector<std::vector<size_t>> extras;
        std::vector<size_t> out_extra;
    };
    const std::vector<Case> cases = {
        {PFH_POINTWISE_LINEAR, PFH_SCALAR_F32, PFH_SCALAR_F32, {{2, 1, 3}, {1, 4, 3}, {2, 4, 1}}, {2, 4, 3}},
        {PFH_POINTWISE_MUL, PFH_SCALAR_C128, PFH_SCALAR_C128, {{3}, {1}}, {3}},
        {PFH_POINTWISE_MUL_ADD, PFH_SCALAR_F64, PFH_SCALAR_C128, {{1, 1}, {5, 2}, {5, 1}}, {5, 2}},
        {PFH_POINTWISE_LINEAR, PFH_SCALAR_C64, PFH_SCALAR_C64, {{}, {}}, {}},
    };
    const std::vector<size_t> shape = {6, 5, 4};
    const std::vector<int> perm = {2, 0, 1};
    for (const Case &c : cases) {
        std::vector<pfh_array> in(c.extras.size());
        for (size_t i = 0; i < in.size(); ++i) layout(in[i], shape, perm, c.extras[i], c.type_in);
        pfh_array out;
        layout(out, shape, perm, c.out_extra, c.type_out);
        const double coef[16] = {1.0, c.type_out == c.type_in ? 0.0 : 0.5, 2.0, 0.0, -1.0, 0.0};
        const pfh_status st = pfh_pointwise(c.op, in.data(), (int)in.size(), &out, coef, nullptr);
        if (st != PFH_OK) fail("pointwise case op %d: %s", c.op, pfh_status_string(st));
    }
}

// The boundary contract (no device work before a refusal): workspace
// binding (WorkspaceMismatch), overlapping buffers, and the in-place poisoned
// state after a failure that follows the first write, synchronous and
// deferred.
static void contract_cases()
{
    auto expect = [](const char *what, pfh_status got, pfh_status want) {
        if (got != want)
            fail("contract: %s: got %d (%s), expected %d", what, (int)got, pfh_status_string(got), (int)want);
    };
    auto mk = [](pfh_desc d) {
        pfh_plan p = nullptr;
        if (pfh_plan_create(&d, &p) != PFH_OK) fail("contract: plan");
        return p;
    };
    auto ws_bytes = [](pfh_plan p) {
        size_t m = 0;
        for (int op = 0; op < 6; ++op) {
            size_t b = 0;
            if (pfh_plan_workspace_bytes(p, op, &b) == PFH_OK) m = std::max(m, b);
        }
        return m;
    };
    auto same = [](const void *a, const std::vector<uint8_t> &b) { return std::memcmp(a, b.data(), b.size()) == 0; };
    auto snap = [](const void *a, size_t n) {
        std::vector<uint8_t> v(n);
        std::memcpy(v.data(), a, n);
        return v;
    };
    const int before_fail = g_fail;

    // workspace binding
    const pfh_desc dc = route_desc(PFH_C2C, PFH_F32, {8, 6, 4}, {2});
    pfh_plan A = mk(dc), B = mk(dc);
    const size_t n = 8 * 6 * 4 * 2, nb = n * 8;
    const size_t wa = ws_bytes(A), wb = ws_bytes(B);
    void *wsA = dmalloc(wa, true), *wsB = dmalloc(wb, true);
    float *src = static_cast<float *>(dmalloc(nb));
    for (size_t i = 0; i < 2 * n; ++i) src[i] = (float)i;
    void *dst = dmalloc(nb, true);
    expect("A with its workspace", pfh_execute(A, PFH_FORWARD, src, n, dst, n, wsA, wa, nullptr, 0), PFH_OK);
    expect("B with its workspace", pfh_execute(B, PFH_FORWARD, src, n, dst, n, wsB, wb, nullptr, 0), PFH_OK);
    {
        std::memset(dst, 0x5A, nb);
        const auto d0 = snap(dst, nb), w0 = snap(wsA, wa);
        const int l0 = pfh::g_all_launches;
        expect("B with A's workspace", pfh_execute(B, PFH_FORWARD, src, n, dst, n, wsA, wa, nullptr, 0),
               PFH_ERR_WORKSPACE_MISMATCH);
        expect("B in place with A's workspace", pfh_execute_in_place(B, PFH_FORWARD, dst, nb, wsA, wa, nullptr, 0),
               PFH_ERR_WORKSPACE_MISMATCH);
        expect("B timed with A's workspace",
               [&] {
                   pfh_transform_timing t;
                   return pfh_execute_with_timing(B, PFH_FORWARD, src, n, dst, n, wsA, wa, nullptr, &t);
               }(),
               PFH_ERR_WORKSPACE_MISMATCH);
        if (pfh::g_all_launches != l0) fail("contract: a refused call launched device work");
        if (!same(dst, d0) || !same(wsA, w0)) fail("contract: a refused call wrote memory");
    }
    // pfh_workspace_init rebinds; a zero-filled workspace is unbound
    expect("init A's workspace for B", pfh_workspace_init(B, wsA, wa, nullptr), PFH_OK);
    expect("B with the rebound workspace", pfh_execute(B, PFH_FORWARD, src, n, dst, n, wsA, wa, nullptr, 0), PFH_OK);
    expect("A with the workspace bound to B", pfh_execute(A, PFH_FORWARD, src, n, dst, n, wsA, wa, nullptr, 0),
           PFH_ERR_WORKSPACE_MISMATCH);
    std::memset(wsA, 0, wa);
    expect("A with a zero-filled workspace", pfh_execute(A, PFH_FORWARD, src, n, dst, n, wsA, wa, nullptr, 0), PFH_OK);
    expect("init of a local plan's workspace",
           [&] {
               pfh_desc d;
               std::memset(&d, 0, sizeof(d));
               d.kind = PFH_LOCAL_C2C;
               d.precision = PFH_F32;
               d.line_len = 8;
               pfh_plan L = mk(d);
               const pfh_status st = pfh_workspace_init(L, wsA, wa, nullptr);
               pfh_plan_destroy(L);
               return st;
           }(),
           PFH_ERR_INVALID_ARGUMENT);

    // overlapping buffers: refused before any device work
    {
        char *big = static_cast<char *>(dmalloc(3 * nb + wa, true));
        const int l0 = pfh::g_all_launches;
        expect("src == dst", pfh_execute(A, PFH_FORWARD, big, n, big, n, wsA, wa, nullptr, 0), PFH_ERR_ALIASED_BUFFERS);
        expect("dst = src + 8 bytes", pfh_execute(A, PFH_FORWARD, big, n, big + 8, n, wsA, wa, nullptr, 0),
               PFH_ERR_ALIASED_BUFFERS);
        expect("dst ends where src starts (no overlap)",
               pfh_execute(A, PFH_FORWARD, big + nb, n, big, n, wsA, wa, nullptr, 0), PFH_OK);
        int l1 = pfh::g_all_launches;
        expect("src inside the workspace", pfh_execute(A, PFH_FORWARD, big + nb, n, big, n, big + nb, wa, nullptr, 0),
               PFH_ERR_ALIASED_BUFFERS);
        expect("dst inside the workspace",
               pfh_execute(A, PFH_FORWARD, src, n, big + 2 * nb, n, big + 2 * nb + 16, wa, nullptr, 0),
               PFH_ERR_ALIASED_BUFFERS);
        expect("in-place data inside the workspace",
               pfh_execute_in_place(A, PFH_FORWARD, big + 64, nb, big, wa + nb, nullptr, 0), PFH_ERR_ALIASED_BUFFERS);
        pfh_desc d;
        std::memset(&d, 0, sizeof(d));
        d.kind = PFH_LOCAL_C2C;
        d.precision = PFH_F32;
        d.line_len = 8;
        pfh_plan L = mk(d);
        expect("local src == dst", pfh_local_execute(L, PFH_FORWARD, big, 16, big, 16, nullptr, 0, nullptr, 0),
               PFH_ERR_ALIASED_BUFFERS);
        const int l2 = pfh::g_all_launches;
        expect("local in place", pfh_local_execute_in_place(L, PFH_FORWARD, big, 16, 16 * 8, nullptr, 0, nullptr, 0),
               PFH_OK);
        if (pfh::g_all_launches == l2) fail("contract: local in place launched nothing");
        l1 = pfh::g_all_launches;
        pfh_plan_destroy(L);
        pfh_desc t = route_desc(PFH_LOCAL_TRANSPOSE, PFH_F32, {4, 5, 6}, {});
        t.element_bytes = 8;
        t.src_perm[0] = 0, t.src_perm[1] = 1, t.src_perm[2] = 2;
        t.dst_perm[0] = 2, t.dst_perm[1] = 0, t.dst_perm[2] = 1;
        pfh_plan T = mk(t);
        expect("transpose src == dst", pfh_execute(T, PFH_FORWARD, big, 120, big, 120, nullptr, 0, nullptr, 0),
               PFH_ERR_ALIASED_BUFFERS);
        pfh_plan_destroy(T);
        if (pfh::g_all_launches != l1) fail("contract: an aliased call launched device work");
        (void)l0;
        (void)hipFree(big);
    }

    // in-place poisoned state (run_in_place_transaction)
    pfh_plan R = mk(route_desc(PFH_R2C, PFH_F32, {8, 6, 4}, {}));
    const size_t wr = ws_bytes(R);
    void *wsR = dmalloc(wr, true);
    const size_t cap = 8 * 6 * 3 * 8;
    void *data = dmalloc(cap, true);
    pfh::g_reject_endpoints = true;
    expect("in-place inverse of a rejected spectrum",
           pfh_execute_in_place(R, PFH_INVERSE, data, cap, wsR, wr, nullptr, 0), PFH_ERR_INVALID_SPECTRUM);
    pfh::g_reject_endpoints = false;
    {
        const auto d0 = snap(data, cap);
        const int l0 = pfh::g_all_launches;
        expect("in-place forward on a poisoned workspace",
               pfh_execute_in_place(R, PFH_FORWARD, data, cap, wsR, wr, nullptr, 0), PFH_ERR_POISONED);
        expect("in-place inverse on a poisoned workspace",
               pfh_execute_in_place(R, PFH_INVERSE, data, cap, wsR, wr, nullptr, 0), PFH_ERR_POISONED);
        if (pfh::g_all_launches != l0) fail("contract: a poisoned in-place call launched device work");
        if (!same(data, d0)) fail("contract: a poisoned in-place call wrote its array");
    }
    {
        float *x = static_cast<float *>(dmalloc(8 * 6 * 4 * 4, true));
        void *y = dmalloc(cap, true);
        expect("out of place with a poisoned workspace",
               pfh_execute(R, PFH_FORWARD, x, 8 * 6 * 4, y, 8 * 6 * 3, wsR, wr, nullptr, 0), PFH_OK);
        (void)hipFree(x);
        (void)hipFree(y);
    }
    expect("workspace init resets", pfh_workspace_init(R, wsR, wr, nullptr), PFH_OK);
    expect("in-place forward after the reset", pfh_execute_in_place(R, PFH_FORWARD, data, cap, wsR, wr, nullptr, 0),
           PFH_OK);
    // deferred: the failure and the gated later call show at the status read
    pfh::g_reject_endpoints = true;
    expect("deferred in-place inverse",
           pfh_execute_in_place(R, PFH_INVERSE, data, cap, wsR, wr, nullptr, PFH_EXEC_DEFER_STATUS), PFH_OK);
    pfh::g_reject_endpoints = false;
    {
        const auto d0 = snap(data, cap);
        expect("deferred in-place forward after it",
               pfh_execute_in_place(R, PFH_FORWARD, data, cap, wsR, wr, nullptr, PFH_EXEC_DEFER_STATUS), PFH_OK);
        if (!same(data, d0)) fail("contract: an in-place call after a deferred failure wrote its array");
    }
    expect("take status", pfh_workspace_take_status(R, wsR, nullptr), PFH_ERR_INVALID_SPECTRUM);
    expect("in-place after the status read", pfh_execute_in_place(R, PFH_FORWARD, data, cap, wsR, wr, nullptr, 0),
           PFH_ERR_POISONED);
    // a synchronous in-place call after a deferred failure reads the status itself
    expect("reset again", pfh_workspace_init(R, wsR, wr, nullptr), PFH_OK);
    pfh::g_reject_endpoints = true;
    expect("deferred in-place inverse (2)",
           pfh_execute_in_place(R, PFH_INVERSE, data, cap, wsR, wr, nullptr, PFH_EXEC_DEFER_STATUS), PFH_OK);
    pfh::g_reject_endpoints = false;
    expect("synchronous in-place forward after it",
           pfh_execute_in_place(R, PFH_FORWARD, data, cap, wsR, wr, nullptr, 0), PFH_ERR_INVALID_SPECTRUM);
    expect("and the next one", pfh_execute_in_place(R, PFH_FORWARD, data, cap, wsR, wr, nullptr, 0), PFH_ERR_POISONED);

    for (void *p : {(void *)wsA, wsB, (void *)src, dst, wsR, data}) (void)hipFree(p);
    pfh_plan_destroy(A);
    pfh_plan_destroy(B);
    pfh_plan_destroy(R);
    // batches refuse problems whose lines run as pieces, before any device work
    {
        std::vector<pfh_desc> bd = {route_desc(PFH_C2C, PFH_F32, {12, 10, 8}, {}),
                                    route_desc(PFH_C2C, PFH_F32, {4, 3, 1009}, {})};
        pfh_batch bt = nullptr;
        expect("batch with a long line", pfh_batch_create(bd.data(), 2, 0, &bt), PFH_ERR_UNSUPPORTED);
        if (bt) pfh_batch_destroy(bt);
    }
    std::printf("contract: %s\n", g_fail == before_fail ? "ok" : "FAILED");
}

// ---------------------------------------------------------------------------
// `long` scenario: lines split into pieces (four-step and Bluestein), with
// the arithmetic emulated, against naive sums in long double
// ---------------------------------------------------------------------------

uint64_t g_rng = 0x243f6a8885a308d3ull;
double urand()
{
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return (double)(g_rng >> 11) / (double)(1ull << 53) * 2.0 - 1.0;
}

const long double kPi = 3.141592653589793238462643383279502884L;

// Naive line transforms (FFTW conventions, unnormalized unless noted).
std::vector<pfh::cld> naive_dft(const std::vector<pfh::cld> &x, int sign, long double scale)
{
    const int64_t n = (int64_t)x.size();
    std::vector<pfh::cld> y(n);
    for (int64_t k = 0; k < n; ++k) {
        long double re = 0, im = 0;
        for (int64_t j = 0; j < n; ++j) {
            const long double a = sign * 2 * kPi * (long double)((j * k) % n) / (long double)n;
            re += x[j].re * cosl(a) - x[j].im * sinl(a);
            im += x[j].re * sinl(a) + x[j].im * cosl(a);
        }
        y[k] = {re * scale, im * scale};
    }
    return y;
}

std::vector<long double> naive_r2r(int kind, const std::vector<long double> &x)
{
    const int64_t n = (int64_t)x.size();
    std::vector<long double> y(n, 0);
    for (int64_t k = 0; k < n; ++k) {
        long double acc = 0;
        for (int64_t j = 0; j < n; ++j) {
            const long double v = x[j];
            switch (kind) {
            case PFH_DCT1:
                if (j == 0) acc += v;
                else if (j == n - 1) acc += (k % 2 ? -v : v);
                else acc += 2 * v * cosl(kPi * j * k / (n - 1));
                break;
            case PFH_DCT2: acc += 2 * v * cosl(kPi * (j + 0.5L) * k / n); break;
            case PFH_DCT3: acc += (j == 0) ? v : 2 * v * cosl(kPi * j * (k + 0.5L) / n); break;
            case PFH_DCT4: acc += 2 * v * cosl(kPi * (j + 0.5L) * (k + 0.5L) / n); break;
            case PFH_DST1: acc += 2 * v * sinl(kPi * (j + 1) * (k + 1) / (n + 1)); break;
            case PFH_DST2: acc += 2 * v * sinl(kPi * (j + 0.5L) * (k + 1) / n); break;
            case PFH_DST3:
                acc += (j == n - 1) ? (k % 2 ? -v : v) : 2 * v * sinl(kPi * (j + 1) * (k + 0.5L) / n);
                break;
            case PFH_DST4: acc += 2 * v * sinl(kPi * (j + 0.5L) * (k + 0.5L) / n); break;
            default: { // DHT
                const long double a = 2 * kPi * (long double)((j * k) % n) / n;
                acc += v * (cosl(a) + sinl(a));
            }
            }
        }
        y[k] = acc;
    }
    return y;
}

template <typename T>
void local_long_case(int kind, int r2r, int64_t n, int op, int64_t batch, double tol)
{
    char name[96];
    std::snprintf(name, sizeof name, "long local kind %d r2r %d n %" PRId64 " op %d %s", kind, r2r, n, op,
                  sizeof(T) == 4 ? "f32" : "f64");
    pfh_desc d;
    std::memset(&d, 0, sizeof d);
    d.kind = kind;
    d.precision = sizeof(T) == 4 ? PFH_F32 : PFH_F64;
    d.line_len = (size_t)n;
    d.r2r_kind = r2r;
    pfh_plan p = nullptr;
    if (pfh_plan_create(&d, &p) != PFH_OK) {
        pfh_error_info e;
        pfh_last_error(&e);
        fail("%s: plan creation: %s", name, e.message);
        return;
    }
    const bool r2c = kind == PFH_LOCAL_R2C, c2r = r2c && op == PFH_INVERSE;
    const bool cin = kind == PFH_LOCAL_C2C || c2r, cout = kind == PFH_LOCAL_C2C || (r2c && !c2r);
    const int64_t lin = c2r ? n / 2 + 1 : n, lout = (r2c && !c2r) ? n / 2 + 1 : n;
    size_t wsb = 0;
    pfh_plan_workspace_bytes(p, op, &wsb);
    T *src = static_cast<T *>(dmalloc((size_t)(batch * lin * (cin ? 2 : 1)) * sizeof(T)));
    T *dst = static_cast<T *>(dmalloc((size_t)(batch * lout * (cout ? 2 : 1)) * sizeof(T)));
    void *ws = dmalloc(wsb ? wsb : 1, true);
    for (int64_t i = 0; i < batch * lin * (cin ? 2 : 1); ++i) src[i] = (T)urand();
    const int64_t launches = pfh::g_launches;
    if (pfh_local_execute(p, op, src, (size_t)(batch * lin), dst, (size_t)(batch * lout), ws, wsb, nullptr, 0) !=
        PFH_OK) {
        pfh_error_info e;
        pfh_last_error(&e);
        fail("%s: execute: %s", name, e.message);
    } else if (pfh::g_launches - launches < 2) {
        fail("%s: ran as %" PRId64 " launch(es), not as pieces", name, pfh::g_launches - launches);
    } else if (!pfh::g_math_dft) {
        // shapes and bounds only
    } else {
        long double num = 0, den = 0;
        for (int64_t b = 0; b < batch; ++b) {
            std::vector<long double> got, ref;
            if (kind == PFH_LOCAL_C2C) {
                std::vector<pfh::cld> x(n);
                for (int64_t j = 0; j < n; ++j) x[j] = {(long double)src[2 * (b * n + j)], (long double)src[2 * (b * n + j) + 1]};
                const int sign = op == PFH_FORWARD ? -1 : 1;
                const auto y = naive_dft(x, sign, op == PFH_INVERSE ? 1.0L / n : 1.0L);
                for (int64_t k = 0; k < n; ++k) {
                    ref.push_back(y[k].re);
                    ref.push_back(y[k].im);
                    got.push_back(dst[2 * (b * n + k)]);
                    got.push_back(dst[2 * (b * n + k) + 1]);
                }
            } else if (r2c && !c2r) {
                std::vector<pfh::cld> x(n);
                for (int64_t j = 0; j < n; ++j) x[j] = {(long double)src[b * n + j], 0};
                const auto y = naive_dft(x, -1, 1.0L);
                for (int64_t k = 0; k < lout; ++k) {
                    ref.push_back(y[k].re);
                    ref.push_back(y[k].im);
                    got.push_back(dst[2 * (b * lout + k)]);
                    got.push_back(dst[2 * (b * lout + k) + 1]);
                }
            } else if (c2r) {
                std::vector<pfh::cld> x(n);
                for (int64_t k = 0; k <= n / 2; ++k) {
                    pfh::cld v{(long double)src[2 * (b * lin + k)], (long double)src[2 * (b * lin + k) + 1]};
                    if (k == 0 || (n % 2 == 0 && k == n / 2)) v.im = 0;
                    x[k] = v;
                    if (k > 0) x[n - k] = {v.re, -v.im};
                }
                const auto y = naive_dft(x, 1, 1.0L / n);
                for (int64_t j = 0; j < n; ++j) {
                    ref.push_back(y[j].re);
                    got.push_back(dst[b * n + j]);
                }
            } else {
                std::vector<long double> x(n);
                for (int64_t j = 0; j < n; ++j) x[j] = src[b * n + j];
                ref = naive_r2r(kind == PFH_LOCAL_DHT ? PFH_DHT : r2r, x);
                for (int64_t k = 0; k < n; ++k) got.push_back(dst[b * n + k]);
            }
            for (size_t i = 0; i < ref.size(); ++i) {
                num += (got[i] - ref[i]) * (got[i] - ref[i]);
                den += ref[i] * ref[i];
            }
        }
        const double err = (double)sqrtl(num / (den > 0 ? den : 1));
        if (!(err <= tol)) fail("%s: relative L2 error %.3g > %.3g", name, err, tol);
    }
    for (void *q : {(void *)src, (void *)dst, ws}) (void)hipFree(q);
    pfh_plan_destroy(p);
}

// A 3D C2C route with a long axis against the naive separable transform.
template <typename T>
void route_long_case(std::vector<size_t> shape, size_t extra, int op, double tol, bool permute = true)
{
    char name[112];
    std::snprintf(name, sizeof name, "long route c2c %zux%zux%zu extra %zu op %d %s%s", shape[0], shape[1], shape[2],
                  extra, op, sizeof(T) == 4 ? "f32" : "f64", permute ? "" : " strided");
    pfh_desc d = route_desc(PFH_C2C, sizeof(T) == 4 ? PFH_F32 : PFH_F64, shape,
                            extra > 1 ? std::vector<size_t>{extra} : std::vector<size_t>{});
    d.permute_dims = permute ? 1 : 0;
    pfh_plan p = nullptr;
    if (pfh_plan_create(&d, &p) != PFH_OK) {
        pfh_error_info e;
        pfh_last_error(&e);
        fail("%s: plan creation: %s", name, e.message);
        return;
    }
    int perm[2][PFH_MAX_DIMS];
    size_t shp[2][PFH_MAX_DIMS], elems[2], eb[2];
    for (int side = 0; side < 2; ++side) pfh_plan_layout(p, side, perm[side], shp[side], &elems[side], &eb[side]);
    const int64_t len = (int64_t)(shape[0] * shape[1] * shape[2]);
    auto mem_index = [&](int side, const int64_t c[3]) {
        int64_t idx = 0;
        for (int pos = 0; pos < 3; ++pos) idx = idx * (int64_t)shape[perm[side][pos]] + c[perm[side][pos]];
        return idx;
    };
    // forward: input layout -> output layout; inverse / backward the other way
    const int sin = op == PFH_FORWARD ? 0 : 1, sout = 1 - sin;
    size_t wsb = 0;
    pfh_plan_workspace_bytes(p, op, &wsb);
    T *src = static_cast<T *>(dmalloc(elems[sin] * 2 * sizeof(T)));
    T *dst = static_cast<T *>(dmalloc(elems[sout] * 2 * sizeof(T)));
    void *ws = dmalloc(wsb, true);
    for (size_t i = 0; i < elems[sin] * 2; ++i) src[i] = (T)urand();
    if (pfh_execute(p, op, src, elems[sin], dst, elems[sout], ws, wsb, nullptr, 0) != PFH_OK) {
        pfh_error_info e;
        pfh_last_error(&e);
        fail("%s: execute: %s", name, e.message);
    } else if (!pfh::g_math_dft) {
        // shapes and bounds only
    } else {
        const int sign = op == PFH_FORWARD ? -1 : 1;
        long double num = 0, den = 0;
        for (size_t e = 0; e < extra; ++e) {
            // logical array [a0][a1][a2]
            std::vector<pfh::cld> A((size_t)len);
            int64_t c[3];
            for (c[0] = 0; c[0] < (int64_t)shape[0]; ++c[0])
                for (c[1] = 0; c[1] < (int64_t)shape[1]; ++c[1])
                    for (c[2] = 0; c[2] < (int64_t)shape[2]; ++c[2]) {
                        const int64_t mi = (int64_t)e * len + mem_index(sin, c);
                        A[(c[0] * shape[1] + c[1]) * shape[2] + c[2]] = {(long double)src[2 * mi], (long double)src[2 * mi + 1]};
                    }
            const int64_t st[3] = {(int64_t)(shape[1] * shape[2]), (int64_t)shape[2], 1};
            for (int ax = 0; ax < 3; ++ax) {
                const int64_t na = (int64_t)shape[ax];
                for (int64_t base = 0; base < len; ++base) {
                    if ((base / st[ax]) % na != 0) continue; // first point of each line along ax
                    std::vector<pfh::cld> x(na);
                    for (int64_t j = 0; j < na; ++j) x[j] = A[base + j * st[ax]];
                    const auto y = naive_dft(x, sign, op == PFH_INVERSE ? 1.0L / na : 1.0L);
                    for (int64_t j = 0; j < na; ++j) A[base + j * st[ax]] = y[j];
                }
            }
            for (c[0] = 0; c[0] < (int64_t)shape[0]; ++c[0])
                for (c[1] = 0; c[1] < (int64_t)shape[1]; ++c[1])
                    for (c[2] = 0; c[2] < (int64_t)shape[2]; ++c[2]) {
                        const int64_t mo = (int64_t)e * len + mem_index(sout, c);
                        const pfh::cld r = A[(c[0] * shape[1] + c[1]) * shape[2] + c[2]];
                        const long double gr = dst[2 * mo], gi = dst[2 * mo + 1];
                        num += (gr - r.re) * (gr - r.re) + (gi - r.im) * (gi - r.im);
                        den += r.re * r.re + r.im * r.im;
                    }
        }
        const double err = (double)sqrtl(num / den);
        if (!(err <= tol)) fail("%s: relative L2 error %.3g > %.3g", name, err, tol);
    }
    for (void *q : {(void *)src, (void *)dst, ws}) (void)hipFree(q);
    pfh_plan_destroy(p);
}

void long_cases()
{
    const int before = g_fail;
    pfh::g_math = true;
    // small thresholds, so that small lines exercise every piece shape:
    // lines above 32 points split, prime factors above 7 go through Bluestein
    setenv("PFH_LONG_LINE_MAX", "32", 1);
    setenv("PFH_BLUESTEIN_PRIME", "7", 1);
    for (int op : {PFH_FORWARD, PFH_BACKWARD, PFH_INVERSE})
        for (int64_t n : {48, 64, 11, 37, 97, 160})
            local_long_case<double>(PFH_LOCAL_C2C, 0, n, op, 3, 1e-12);
    for (int64_t n : {96, 22, 75, 74})
        for (int op : {PFH_FORWARD, PFH_INVERSE}) local_long_case<double>(PFH_LOCAL_R2C, 0, n, op, 2, 1e-12);
    for (int k = PFH_DCT1; k <= PFH_DST4; ++k)
        for (int64_t n : {70, 45}) local_long_case<double>(PFH_LOCAL_R2R, k, n, PFH_FORWARD, 2, 1e-12);
    for (int64_t n : {70, 45}) local_long_case<double>(PFH_LOCAL_DHT, 0, n, PFH_FORWARD, 2, 1e-12);
    local_long_case<float>(PFH_LOCAL_C2C, 0, 48, PFH_FORWARD, 2, 2e-6);
    local_long_case<float>(PFH_LOCAL_C2C, 0, 37, PFH_INVERSE, 2, 2e-6);
    local_long_case<float>(PFH_LOCAL_R2C, 0, 22, PFH_FORWARD, 2, 2e-6);
    for (int op : {PFH_FORWARD, PFH_INVERSE}) {
        route_long_case<double>({37, 6, 4}, 1, op, 1e-12);
        route_long_case<double>({5, 48, 3}, 2, op, 1e-12);
        route_long_case<double>({4, 3, 41}, 1, op, 1e-12);
    }
    route_long_case<float>({6, 37, 5}, 1, PFH_FORWARD, 2e-6);
    // permute_dims = false: every stage's lines are strided in place
    for (int op : {PFH_FORWARD, PFH_INVERSE}) {
        route_long_case<double>({37, 6, 4}, 2, op, 1e-12, false);
        route_long_case<double>({5, 48, 3}, 1, op, 1e-12, false);
    }
    unsetenv("PFH_LONG_LINE_MAX");
    unsetenv("PFH_BLUESTEIN_PRIME");
    // real-size lines with the default thresholds: launch shapes, thread
    // mappings, LDS and every address, without the arithmetic
    pfh::g_math_dft = false;
    for (int64_t n : {32768, 65536, 1009, 10007, 1 << 20}) {
        for (int op : {PFH_FORWARD, PFH_INVERSE}) {
            local_long_case<float>(PFH_LOCAL_C2C, 0, n, op, 3, 0);
            local_long_case<float>(PFH_LOCAL_R2C, 0, 2 * n, op, 2, 0);
        }
        local_long_case<double>(PFH_LOCAL_C2C, 0, n, PFH_FORWARD, 2, 0);
        local_long_case<float>(PFH_LOCAL_R2R, PFH_DCT2, 2 * n, PFH_FORWARD, 2, 0);
    }
    for (int64_t n : {10006, 65535}) local_long_case<double>(PFH_LOCAL_R2R, PFH_DST1, n, PFH_FORWARD, 2, 0);
    for (int op : {PFH_FORWARD, PFH_INVERSE}) {
        route_long_case<float>({10007, 3, 4}, 2, op, 0);
        route_long_case<float>({4, 65536, 3}, 1, op, 0);
        route_long_case<double>({3, 4, 1009}, 1, op, 0);
    }
    pfh::g_math_dft = true;
    pfh::g_math = false;
    std::printf("long lines: %s\n", g_fail == before ? "ok" : "FAILED");
}

// The batch tuning cache: a measured tuning is reused within the process and,
// through the cache file, by a later creation after the process map is
// cleared; a record that does not parse is skipped; a cache switched off
// measures every time.
void tune_cache_cases(const char *path)
{
    const int before_fail = g_fail;
    // small problems: the cache logic does not depend on their size, and every
    // autotuned creation emulates each candidate's timed executions
    std::vector<std::vector<size_t>> shapes = {{12, 10, 8}, {12, 7, 9}, {6, 16, 4}, {8, 8, 5}};
    std::vector<pfh_desc> ds;
    for (auto &s : shapes) ds.push_back(route_desc(PFH_C2C, PFH_F32, s, {3}));
    setenv("PFH_TUNE_CACHE", path, 1);
    std::remove(path);
    struct Got {
        int src = -1, t[7] = {};
    };
    auto create = [&](const std::vector<pfh_desc> &d, int tune) {
        Got g;
        pfh_batch b = nullptr;
        if (pfh_batch_create(d.data(), (int)d.size(), tune, &b) != PFH_OK) {
            fail("tune cache: batch create");
            return g;
        }
        pfh_batch_tuning_source(b, &g.src);
        pfh_batch_tuning(b, &g.t[0], &g.t[1], &g.t[2], &g.t[3], &g.t[4], &g.t[5], &g.t[6]);
        pfh_batch_destroy(b);
        return g;
    };
    auto same = [](const Got &a, const Got &b) {
        for (int i = 0; i < 7; ++i)
            if (a.t[i] != b.t[i]) return false;
        return true;
    };
    const int calls0 = (int)g_elapsed_calls;
    Got m = create(ds, 1);
    if (m.src != PFH_TUNED_MEASURED) fail("tune cache: first creation source %d", m.src);
    const int timed = (int)g_elapsed_calls - calls0;
    if (timed < 24 * 5) fail("tune cache: %d timings for 24 candidates", timed);
    Got d0 = create(ds, 0);
    if (d0.src != PFH_TUNED_DEFAULT) fail("tune cache: default source %d", d0.src);
    const int calls1 = (int)g_elapsed_calls;
    Got p = create(ds, 1);
    if (p.src != PFH_TUNED_PROCESS_CACHE || !same(p, m)) fail("tune cache: process cache (source %d)", p.src);
    if ((int)g_elapsed_calls != calls1) fail("tune cache: a cached creation measured");
    // another process: the map is empty, the file holds the record
    pfh_batch_tuning_cache_clear();
    Got f = create(ds, 1);
    if (f.src != PFH_TUNED_FILE_CACHE || !same(f, m)) fail("tune cache: file cache (source %d)", f.src);
    // a different batch (one shape changed) is measured
    std::vector<pfh_desc> ds2 = ds;
    ds2[1].shape[2] = 10;
    Got o = create(ds2, 1);
    if (o.src != PFH_TUNED_MEASURED) fail("tune cache: another batch's source %d", o.src);
    // records that do not parse are skipped; the last good record of a key wins
    {
        FILE *fp = std::fopen(path, "r");
        std::string first;
        char buf[1 << 14];
        if (fp && std::fgets(buf, sizeof buf, fp)) first = buf;
        if (fp) std::fclose(fp);
        const size_t sp = first.find(' ');
        if (sp == std::string::npos) {
            fail("tune cache: no record in the file");
        } else {
            const std::string key = first.substr(0, sp);
            const int alt_block = m.t[0] == 128 ? 256 : 128;
            fp = std::fopen(path, "a");
            std::fprintf(fp, "%s 256 8 0 1 1\n", key.c_str());            // a field missing
            std::fprintf(fp, "%s 300 8 0 1 1 1 2.0\n", key.c_str());      // not a power of two
            std::fprintf(fp, "%s 256 8 0 1 1 1 2.0 junk\n", key.c_str()); // trailing text
            std::fprintf(fp, "garbage line\n");
            std::fprintf(fp, "%s %d 8 32 1 0 0 1 2.5\n", key.c_str(), alt_block);
            std::fclose(fp);
            pfh_batch_tuning_cache_clear();
            Got g = create(ds, 1);
            if (g.src != PFH_TUNED_FILE_CACHE || g.t[0] != alt_block || g.t[2] != 32 || g.t[4] != 0)
                fail("tune cache: last record (source %d, block %d, col %d, persistent %d)", g.src, g.t[0], g.t[2],
                     g.t[4]);
        }
    }
    // cache off: every autotuned creation measures, nothing is written
    setenv("PFH_TUNE_CACHE", "off", 1);
    pfh_batch_tuning_cache_clear();
    std::remove(path);
    Got x = create(ds, 1);
    Got y = create(ds, 1);
    if (x.src != PFH_TUNED_MEASURED || y.src != PFH_TUNED_PROCESS_CACHE)
        fail("tune cache off: sources %d %d", x.src, y.src);
    if (FILE *fp = std::fopen(path, "r")) {
        std::fclose(fp);
        fail("tune cache off: a file was written");
    }
    std::printf("tune cache: %s\n", g_fail == before_fail ? "ok" : "FAILED");
}

int main(int argc, char **argv)
{
    if (argc == 2 && std::strcmp(argv[1], "long") == 0) {
        long_cases();
        std::printf("launches %d, failures %d\n", pfh::g_launches, g_fail);
        return g_fail ? 1 : 0;
    }
    if (argc == 3 && std::strcmp(argv[1], "tunecache") == 0) {
        tune_cache_cases(argv[2]);
        std::printf("launches %d, failures %d\n", pfh::g_launches, g_fail);
        return g_fail ? 1 : 0;
    }
    // the scenarios below keep their tunings in the process
    setenv("PFH_TUNE_CACHE", "off", 1);
    if (argc == 7 && std::strcmp(argv[1], "plan") == 0) {
        // one R2R (DCT-II, DST-I, DCT-IV) or DHT plan: emu plan r2r|dht n0 n1 n2 f32|f64
        // (launch shapes printed with PFH_EMU_TRACE set)
        const bool dht = std::strcmp(argv[2], "dht") == 0, f32 = std::strcmp(argv[6], "f32") == 0;
        std::vector<size_t> shape = {(size_t)std::atol(argv[3]), (size_t)std::atol(argv[4]), (size_t)std::atol(argv[5])};
        pfh_desc d = route_desc(PFH_R2R, f32 ? PFH_F32 : PFH_F64, shape, {});
        const int kinds[3] = {PFH_DCT2, PFH_DST1, PFH_DCT4};
        for (int a = 0; a < 3; ++a) d.axis_kind[a] = dht ? PFH_DHT : kinds[a];
        pfh_plan p = nullptr;
        if (pfh_plan_create(&d, &p) != PFH_OK) return 2;
        for (int op : {PFH_FORWARD, PFH_INVERSE}) {
            std::printf("op %d\n", op);
            if (f32)
                run_case<float>(argv[2], p, 3, op, true, true, -1, 0);
            else
                run_case<double>(argv[2], p, 3, op, true, true, -1, 0);
        }
        pfh_plan_destroy(p);
        std::printf("launches %d, failures %d\n", pfh::g_launches, g_fail);
        return g_fail ? 1 : 0;
    }
    if (argc == 2 && std::strcmp(argv[1], "contract") == 0) {
        contract_cases();
        std::printf("launches %d, failures %d\n", pfh::g_launches, g_fail);
        return g_fail ? 1 : 0;
    }
    if (argc == 6) {
        // one plan: emu c2c|r2c n0 n1 n2 f32|f64 (planner defaults)
        const bool r2c = std::strcmp(argv[1], "r2c") == 0, f32 = std::strcmp(argv[5], "f32") == 0;
        std::vector<size_t> shape = {(size_t)std::atol(argv[2]), (size_t)std::atol(argv[3]), (size_t)std::atol(argv[4])};
        pfh_desc d = route_desc(r2c ? PFH_R2C : PFH_C2C, f32 ? PFH_F32 : PFH_F64, shape, {});
        pfh_plan p = nullptr;
        if (pfh_plan_create(&d, &p) != PFH_OK) return 2;
        for (int op : {PFH_FORWARD, PFH_INVERSE}) {
            const bool fl = op == PFH_FORWARD;
            if (f32)
                run_case<float>(argv[1], p, 3, op, r2c && fl, r2c && !fl, r2c ? 2 : -1, 0);
            else
                run_case<double>(argv[1], p, 3, op, r2c && fl, r2c && !fl, r2c ? 2 : -1, 0);
        }
        pfh_plan_destroy(p);
        std::printf("launches %d, failures %d\n", pfh::g_launches, g_fail);
        return g_fail ? 1 : 0;
    }
    // 1 MiB slab intermediates so that the small shapes here are pipelined
    setenv("PFH_SLAB_MB", "1", 0);
    pfh::fastdiv_selftest();
    line_count_guard();
    route_cases<double>(PFH_F64);
    route_cases<float>(PFH_F32);
    batch_cases();
    pointwise_cases();
    std::printf("launches %d, failures %d\n", pfh::g_launches, g_fail);
    return g_fail ? 1 : 0;
}
