// Host planner and C ABI of the gfx950 FFT path.
//
// A plan operation is a list of Steps; each Step is one k_line_stage launch
// that transforms one logical axis and, in the same pass, moves the array
// into the next layout of the route. The route itself is the reference's
// (build_route, crates/pencil-fft/src/distributed.rs:2306-2373): stage s
// transforms axis N-1-s; with permute_dims each stage's pencil moves its axis
// last. Which side of a stage absorbs each transition, and which buffer
// holds each intermediate, is chosen here to minimise workspace.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "pfh_collectives.hpp"
#include "pfh_plan.hpp"
#include "pfh_tune_cache.hpp"

using namespace pfh;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

static thread_local pfh_error_info g_err;

static pfh_status set_err(pfh_status st, const char *msg, size_t required = 0, size_t actual = 0, size_t batch = 0,
                          size_t index = 0, int hip_error = 0)
{
    g_err.status = st;
    g_err.required = required;
    g_err.actual = actual;
    g_err.batch = batch;
    g_err.index = index;
    g_err.hip_error = hip_error;
    std::snprintf(g_err.message, sizeof(g_err.message), "%s", msg ? msg : "");
    return st;
}

static pfh_status ok()
{
    g_err.status = PFH_OK;
    g_err.message[0] = 0;
    return PFH_OK;
}

namespace pfh {
// error state shared with the other translation units of the library
pfh_status set_error(pfh_status st, const char *msg, size_t required, size_t actual, size_t batch, size_t index,
                     int hip_error)
{
    return set_err(st, msg, required, actual, batch, index, hip_error);
}
pfh_status clear_error() { return ok(); }
} // namespace pfh

#define PFH_HIP_CHECK(expr)                                                                     \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return set_err(PFH_ERR_DEVICE, hipGetErrorString(e_), 0, 0, 0, 0, (int)e_); \
    } while (0)

namespace pfh {
PlanImpl::~PlanImpl()
{
    if (dev_tables) (void)hipFree(dev_tables);
}
} // namespace pfh

struct pfh_plan_s {
    PlanImpl impl;
};

// ---------------------------------------------------------------------------
// device information
// ---------------------------------------------------------------------------

static std::mutex g_dev_mu;
static std::vector<DeviceInfo> g_devs;
static std::vector<int> g_dev_ready;

static pfh_status device_info(DeviceInfo &out)
{
    int dev = 0;
    PFH_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if ((int)g_devs.size() <= dev) {
        g_devs.resize(dev + 1);
        g_dev_ready.resize(dev + 1, 0);
    }
    if (!g_dev_ready[dev]) {
        hipDeviceProp_t prop;
        PFH_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
        DeviceInfo d;
        d.device = dev;
        d.cus = prop.multiProcessorCount;
        std::snprintf(d.arch, sizeof d.arch, "%s", prop.gcnArchName);
        // The planner's budget leaves kStaticLds bytes per workgroup for the
        // kernels' static LDS (the pass plan copy of k_line_stage).
        constexpr size_t kStaticLds = 256;
        size_t lim = prop.sharedMemPerBlock;
        if (lim < 65536) lim = 65536;
        if (lim > 65536) {
            if (set_line_stage_lds_limit(lim) != hipSuccess || set_fast_lds_limit(lim) != hipSuccess ||
                set_plane_lds_limit(lim) != hipSuccess) {
                (void)hipGetLastError(); // keep the failed attribute call out of later error checks
                lim = 65536;
            }
        }
        d.lds_limit = lim - kStaticLds;
        g_devs[dev] = d;
        g_dev_ready[dev] = 1;
    }
    out = g_devs[dev];
    return PFH_OK;
}

// ---------------------------------------------------------------------------
// tables: exp(-2 pi i q / d) computed in long double and rounded
// ---------------------------------------------------------------------------

static const long double kPiL = 3.141592653589793238462643383279502884L;

static void push_cexp(std::vector<double> &t, long long q, long long d, bool positive = false)
{
    long long r = q % d;
    if (r < 0) r += d;
    long double a = -2.0L * kPiL * (long double)r / (long double)d;
    long double c = cosl(a), s = sinl(a);
    if ((4 * r) % d == 0) { // exact quarter turns
        const long long qt = (4 * r) / d;
        c = (qt == 0) ? 1.0L : (qt == 2 ? -1.0L : 0.0L);
        s = (qt == 1) ? -1.0L : (qt == 3 ? 1.0L : 0.0L);
    }
    if (positive) s = -s;
    t.push_back((double)c);
    t.push_back((double)s);
}

// Radix split of the complex FFT length m: powers of two in 16/8/4/2
// passes, then the odd primes. Empty when m has a factor beyond 13.
static bool radix_plan(int m, std::vector<int> &out, int max_radix)
{
    out.clear();
    if (m <= 1) return true;
    int a = 0, r = m;
    while (r % 2 == 0) {
        r /= 2;
        ++a;
    }
    std::vector<int> p2;
    int odd = r;
    while (odd % 3 == 0) odd /= 3;
    while (odd % 5 == 0) odd /= 5;
    if (max_radix >= 32 && odd == 1 && a >= 5 && a <= 10) {
        // radix 32 halves the LDS passes of the long power-of-two lines
        static const std::vector<int> plans[6] = {{32}, {8, 8}, {8, 16}, {16, 16}, {16, 32}, {32, 32}};
        p2 = plans[a - 5];
        a = -1;
    } else if (max_radix < 16) { // 8 / 4 / 2 passes only
        int left = a;
        while (left >= 3) {
            p2.push_back(8);
            left -= 3;
        }
        if (left == 2) p2.push_back(4);
        else if (left == 1) p2.push_back(2);
        a = -1;
    }
    switch (a) {
    case -1: break;
    case 0: break;
    case 1: p2 = {2}; break;
    case 2: p2 = {4}; break;
    case 3: p2 = {8}; break;
    case 4: p2 = {16}; break;
    case 5: p2 = {8, 4}; break;
    case 6: p2 = {8, 8}; break;
    case 7: p2 = {16, 8}; break;
    case 8: p2 = {16, 16}; break;
    case 9: p2 = {8, 8, 8}; break;
    default: {
        int left = a;
        while (left >= 4) {
            p2.push_back(16);
            left -= 4;
        }
        if (left == 3) p2.push_back(8);
        else if (left == 2) p2.push_back(4);
        else if (left == 1) p2.push_back(2);
    }
    }
    for (int x : p2) out.push_back(x);
    for (int p : {3, 5, 7, 11, 13}) {
        while (r % p == 0) {
            out.push_back(p);
            r /= p;
        }
    }
    // remaining prime factors run through the generic O(R) pass
    for (int p = 17; (int64_t)p * p <= r; p += 2) {
        while (r % p == 0) {
            out.push_back(p);
            r /= p;
        }
    }
    if (r > 1) out.push_back(r);
    if ((int)out.size() > kMaxPasses) return false;
    return true;
}

// ---------------------------------------------------------------------------
// stage specifications
// ---------------------------------------------------------------------------

enum StageKind { SK_COPY, SK_C2C, SK_R2C, SK_C2R, SK_R2R };

struct StageSpec {
    int kind = SK_COPY;
    int axis = 0;
    bool in_c = true, out_c = true;
    int64_t n = 1;
    bool sign_pos = false;
    double scale = 1.0;
    int r2r = 0; // applied kind (1..9)
    int flags = 0;
    double prew0 = 1, prew1 = 1, postw0 = 1, postw1 = 1;
    Layout in_natural, out_natural; // shape/elem on each side with the stage's natural perm
    bool cplx_r2r = false;          // an R2R / DHT stage on complex data (parts transformed separately)
};

static int r2r_backward_kind(int k)
{
    switch (k) {
    case PFH_DCT2: return PFH_DCT3;
    case PFH_DCT3: return PFH_DCT2;
    case PFH_DST2: return PFH_DST3;
    case PFH_DST3: return PFH_DST2;
    default: return k;
    }
}

static int64_t r2r_norm(int k, int64_t n)
{
    switch (k) {
    case PFH_DCT1: return 2 * (n - 1);
    case PFH_DST1: return 2 * (n + 1);
    case PFH_DHT: return n;
    default: return 2 * n;
    }
}

// Resolves the R2R stage for a plan operation (see the adjoint weights in
// the header comment of the oracle and pencil_fft_hip.h).
static void r2r_stage(StageSpec &s, int kind, int64_t n, int op)
{
    s.kind = SK_R2R;
    s.in_c = s.out_c = false;
    s.n = n;
    const bool adjoint = (op == PFH_ADJOINT_FORWARD || op == PFH_ADJOINT_BACKWARD || op == PFH_ADJOINT_INVERSE);
    const bool scaled = (op == PFH_INVERSE || op == PFH_ADJOINT_INVERSE);
    int base = (op == PFH_FORWARD || op == PFH_ADJOINT_FORWARD) ? kind : r2r_backward_kind(kind);
    if (scaled) s.scale = 1.0 / (double)r2r_norm(kind, n);
    s.r2r = base;
    if (!adjoint || kind == PFH_DHT) return;
    switch (base) {
    case PFH_DCT1:
        s.prew0 = 2;
        s.prew1 = 2;
        s.postw0 = 0.5;
        s.postw1 = 0.5;
        s.flags |= F_WEIGHTS;
        break;
    case PFH_DCT2:
        s.r2r = PFH_DCT3;
        s.prew0 = 2;
        s.flags |= F_WEIGHTS;
        break;
    case PFH_DCT3:
        s.r2r = PFH_DCT2;
        s.postw0 = 0.5;
        s.flags |= F_WEIGHTS;
        break;
    case PFH_DST2:
        s.r2r = PFH_DST3;
        s.prew1 = 2;
        s.flags |= F_WEIGHTS;
        break;
    case PFH_DST3:
        s.r2r = PFH_DST2;
        s.postw1 = 0.5;
        s.flags |= F_WEIGHTS;
        break;
    default:
        break;
    }
}

// ---------------------------------------------------------------------------
// step construction
// ---------------------------------------------------------------------------

static int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
static int pow2ceil(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

static bool is_generic_radix(int R)
{
    return !(R == 2 || R == 3 || R == 4 || R == 5 || R == 7 || R == 8 || R == 11 || R == 13 || R == 16 || R == 32);
}

static bool is_wide_radix(int R) { return !(R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 16 || R == 32); }

static bool mode_needs_pre(int mode)
{
    return !(mode == M_COPY_C || mode == M_COPY_R || mode == M_C2C || mode == M_R2C || mode == M_DHT);
}

// Picks a compile-time specialized kernel for power-of-two C2C / R2C / C2R
// lines (k_line_fast); PFH_FAST=0 disables it, PFH_FAST512=<0|1|2> selects
// the 512-point f32 variant (16x32 with 16 registers, 8x8x8, 16x32 with 32).
static int env_int(const char *name, int dflt)
{
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

static void choose_fast(Step &st, int tbytes)
{
    st.fast = false;
    static const int enabled = env_int("PFH_FAST", 1);
    static const int v512 = env_int("PFH_FAST512", 0);
    if (!enabled) return;
    // even DHT lines run the R2C plan with a Hartley store (no input /
    // output reversal, sign alternation or end weights)
    const bool dht = st.mode == M_DHT && !(st.flags & (F_PRE_REV | F_PRE_ALT | F_POST_REV | F_POST_ALT | F_WEIGHTS));
    // real-input R2R modes: the first pass gathers the pre-processed values
    // from global memory, the store applies the generic post-processing.
    // f32 only: on MI355X the 256-point f32 DCT-II stage of a 256x384x240
    // field took 101 us against 161 us generic, the f64 one 174 against 164.
    static const int fast_r2r = env_int("PFH_FAST_R2R", 2);
    const bool r2r_mode = st.mode == M_DCT1 || st.mode == M_DST1 || st.mode == M_DCT2 || st.mode == M_DCT2_ODD ||
                          st.mode == M_DCT3 || st.mode == M_DCT3_ODD || st.mode == M_DCT4 || st.mode == M_DCT4_ODD ||
                          st.mode == M_DHT_ODD;
    // PFH_FAST_R2R: 0 off, 1 f32 only, 2 f32 and f64 (f64 256x384x240 R2R
    // forward + inverse on MI355X: 1.39 against 1.49 ms with the generic
    // kernel for the DCT stages)
    const bool r2r = r2r_mode && (tbytes == 4 ? fast_r2r >= 1 : fast_r2r >= 2);
    if (!(st.mode == M_C2C || st.mode == M_R2C || st.mode == M_C2R || dht || r2r)) return;
    const int m = st.m;
    int r0 = 0, r1 = 0, r2 = 0, em = 0;
    if (tbytes == 4) {
        // mixed-radix lengths with a compile-time plan (first radix <= 16)
        switch (m) {
        case 48: r0 = 16; r1 = 3; em = 16; break;
        case 60: r0 = 4; r1 = 3; r2 = 5; em = 16; break;
        case 80: r0 = 16; r1 = 5; em = 16; break;
        case 84: r0 = 4; r1 = 3; r2 = 7; em = 16; break;
        case 96: r0 = 8; r1 = 4; r2 = 3; em = 16; break;
        case 112: r0 = 16; r1 = 7; em = 16; break;
        case 120: r0 = 8; r1 = 3; r2 = 5; em = 16; break;
        case 160: r0 = 8; r1 = 4; r2 = 5; em = 16; break;
        case 192: r0 = 16; r1 = 4; r2 = 3; em = 16; break;
        case 240: r0 = 16; r1 = 3; r2 = 5; em = 16; break;
        case 320: r0 = 16; r1 = 4; r2 = 5; em = 16; break;
        case 384: r0 = 16; r1 = 8; r2 = 3; em = 16; break;
        case 385: // real-input plan: DST-I lines of length 384 and the odd 385-point
                  // DCT-II / DCT-III / DST-II / DST-III / DHT lines (never C2C)
            if (r2r) { r0 = 11; r1 = 7; r2 = 5; em = 11; }
            break;
        default: break;
        }
    } else if (st.mode == M_R2C || st.mode == M_C2R || dht || r2r) {
        // route-only f64 plans (never C2C, so never a batched line)
        if (m == 120) { r0 = 8; r1 = 3; r2 = 5; em = 8; }
        if (m == 192) { r0 = 8; r1 = 8; r2 = 3; em = 8; }
        if (m == 385 && r2r) { r0 = 11; r1 = 7; r2 = 5; em = 11; }
    }
    if (r0 == 0 && (m < 64 || (m & (m - 1)) != 0)) return;
    if (r0 != 0) {
        // mixed radix plan chosen above
    } else if (tbytes == 4) {
        switch (m) {
        case 64: r0 = 8; r1 = 8; em = 16; break;
        case 128: r0 = 8; r1 = 16; em = 16; break;
        case 256: r0 = 16; r1 = 16; em = 16; break;
        case 512:
            if (v512 == 1) { r0 = 8; r1 = 8; r2 = 8; em = 16; }
            else if (v512 == 2) { r0 = 16; r1 = 32; em = 32; }
            else { r0 = 16; r1 = 32; em = 16; }
            break;
        case 1024: r0 = 32; r1 = 32; em = 32; break;
        default: return;
        }
    } else {
        switch (m) {
        case 64: r0 = 8; r1 = 8; em = 8; break;
        case 128: r0 = 4; r1 = 4; r2 = 8; em = 8; break;
        case 256: r0 = 4; r1 = 8; r2 = 8; em = 8; break;
        case 512: r0 = 8; r1 = 8; r2 = 8; em = 8; break;
        default: return;
        }
    }
    if (r0 == 32) return; // the direct first pass covers radix <= 16
    if (!fast_plan_available(tbytes, r0, r1, r2, em)) return;
    st.fast = true;
    st.fr[0] = r0;
    st.fr[1] = r1;
    st.fr[2] = r2;
    st.fem = em;
    st.radices.clear();
    st.radices.push_back(r0);
    st.radices.push_back(r1);
    if (r2) st.radices.push_back(r2);
}

// Fills mode, lengths, radices and this step's tables (appended to tab).
static pfh_status fill_transform(Step &st, const StageSpec &sp, std::vector<double> &tab, int tbytes,
                                 const Tuning &tu)
{
    const int64_t n = sp.n;
    st.n = (int)n;
    st.axis = sp.axis;
    st.scale = sp.scale;
    st.flags = sp.flags;
    st.prew0 = sp.prew0;
    st.prew1 = sp.prew1;
    st.postw0 = sp.postw0;
    st.postw1 = sp.postw1;
    st.in_cplx = sp.in_c;
    st.out_cplx = sp.out_c;
    switch (sp.kind) {
    case SK_COPY:
        st.mode = sp.in_c ? M_COPY_C : M_COPY_R;
        st.m = 0;
        st.nin = st.nout = (int)n;
        break;
    case SK_C2C:
        st.mode = M_C2C;
        st.m = (int)n;
        st.nin = st.nout = (int)n;
        if (sp.sign_pos) st.flags |= F_CONJ_IN | F_CONJ_OUT;
        break;
    case SK_R2C:
        st.mode = (n % 2 == 0) ? M_R2C : M_R2C_ODD;
        st.m = (n % 2 == 0) ? (int)(n / 2) : (int)n;
        st.nin = (int)n;
        st.nout = (int)(n / 2 + 1);
        break;
    case SK_C2R:
        st.mode = (n % 2 == 0) ? M_C2R : M_C2R_ODD;
        st.m = (n % 2 == 0) ? (int)(n / 2) : (int)n;
        st.nin = (int)(n / 2 + 1);
        st.nout = (int)n;
        break;
    case SK_R2R: {
        st.nin = st.nout = (int)n;
        const bool even = (n % 2 == 0);
        switch (sp.r2r) {
        case PFH_DCT1: st.mode = M_DCT1; st.m = (int)(n - 1); break;
        case PFH_DST1: st.mode = M_DST1; st.m = (int)(n + 1); break;
        case PFH_DCT2: case PFH_DST2:
            st.mode = even ? M_DCT2 : M_DCT2_ODD;
            st.m = even ? (int)(n / 2) : (int)n;
            if (sp.r2r == PFH_DST2) st.flags |= F_PRE_ALT | F_POST_REV;
            break;
        case PFH_DCT3: case PFH_DST3:
            st.mode = even ? M_DCT3 : M_DCT3_ODD;
            st.m = even ? (int)(n / 2) : (int)n;
            if (sp.r2r == PFH_DST3) st.flags |= F_PRE_REV | F_POST_ALT;
            break;
        case PFH_DCT4: case PFH_DST4:
            st.mode = even ? M_DCT4 : M_DCT4_ODD;
            st.m = even ? (int)(n / 2) : (int)(2 * n);
            if (sp.r2r == PFH_DST4) st.flags |= F_PRE_REV | F_POST_ALT;
            break;
        case PFH_DHT:
            if (n == 1) { // bit-exact identity (dht.rs:215-224)
                st.mode = M_COPY_R;
                st.m = 0;
                st.scale = 1.0;
            } else {
                st.mode = even ? M_DHT : M_DHT_ODD;
                st.m = even ? (int)(n / 2) : (int)n;
            }
            break;
        default:
            return set_err(PFH_ERR_INVALID_ARGUMENT, "unknown R2R kind");
        }
        break;
    }
    }
    if (sp.kind == SK_R2C && (sp.flags & F_HERM_DOUBLE)) st.flags |= F_HERM_DOUBLE;
    if (sp.kind == SK_C2R && (sp.flags & F_HERM_HALF)) st.flags |= F_HERM_HALF;
    const int max_r = tu.max_radix ? tu.max_radix : (tbytes == 4 ? 16 : 8);
    st.long_only = !radix_plan(st.m, st.radices, max_r); // more passes than a stage holds: pieces only
    if (st.long_only) st.radices.clear();
    if (tu.allow_fast && !st.long_only) choose_fast(st, tbytes);

    // tables
    st.tw_base = (int64_t)(tab.size() / 2);
    int off = 0;
    auto here = [&]() { return off; };
    int Ns = 1;
    for (size_t p = 0; p < st.radices.size(); ++p) {
        const int R = st.radices[p];
        st.twoff[p] = here();
        if (is_generic_radix(R)) { // roots exp(-2 pi i e / R) precede the pass twiddles
            for (int e = 0; e < R; ++e) {
                push_cexp(tab, e, R);
                ++off;
            }
        }
        if (Ns > 1) {
            for (int k = 0; k < Ns; ++k)
                for (int r = 1; r < R; ++r) {
                    push_cexp(tab, (long long)r * k, (long long)Ns * R);
                    ++off;
                }
        }
        Ns *= R;
    }
    const int m = st.m;
    auto split_table = [&]() { // exp(-2 pi i k / 2m), k = 0..m
        int o = here();
        for (int k = 0; k <= m; ++k) {
            push_cexp(tab, k, 2LL * m);
            ++off;
        }
        return o;
    };
    switch (st.mode) {
    case M_R2C: case M_DCT1: case M_DST1: case M_DHT:
        st.post0 = split_table();
        break;
    case M_DCT2:
        st.post0 = split_table();
        st.post1 = here();
        for (int k = 0; k < n; ++k) { push_cexp(tab, k, 4 * n); ++off; }
        break;
    case M_DCT2_ODD:
        st.post0 = here();
        for (int k = 0; k < n; ++k) { push_cexp(tab, k, 4 * n); ++off; }
        break;
    case M_C2R:
        st.pre0 = here();
        for (int j = 0; j < m; ++j) { push_cexp(tab, j, n, true); ++off; }
        break;
    case M_DCT3:
        st.pre0 = here();
        for (int k = 0; k <= m; ++k) { push_cexp(tab, k, 4 * n, true); ++off; }
        st.pre1 = here();
        for (int j = 0; j < m; ++j) { push_cexp(tab, j, n, true); ++off; }
        break;
    case M_DCT3_ODD:
        st.pre0 = here();
        for (int k = 0; k < n; ++k) { push_cexp(tab, k, 4 * n); ++off; }
        break;
    case M_DCT4:
        st.pre0 = here();
        for (int j = 0; j < m; ++j) { push_cexp(tab, j, 2 * n); ++off; }
        st.post0 = here();
        for (int k = 0; k < m; ++k) { push_cexp(tab, 4LL * k + 1, 8 * n); ++off; }
        break;
    case M_DCT4_ODD:
        st.pre0 = here();
        for (int j = 0; j < n; ++j) { push_cexp(tab, j, 4 * n); ++off; }
        st.post0 = here();
        for (int k = 0; k < n; ++k) { push_cexp(tab, 2LL * k + 1, 8 * n); ++off; }
        break;
    default:
        break;
    }
    st.twn = off;
    return PFH_OK;
}

// Tuning override read once from the environment: PFH_COLUMN_LINES sets the
// lines per tile of stages with a column-access side.
static int g_column_lines = []() {
    const char *e = std::getenv("PFH_COLUMN_LINES");
    return e ? std::atoi(e) : 0;
}();

// Tiles and grid of a step whose lines and lines per tile are set.
static void set_tiles(Step &st, const Tuning &tu)
{
    st.tiles = (st.nlines + st.lt - 1) / st.lt;
    static const int xcd_env = env_int("PFH_XCD", 0);
    if (st.tiles >= 64 && tu.allow_xcd && xcd_env) {
        st.flags |= F_XCD_REMAP;
        st.xcd_per = (int)((st.tiles + 7) / 8);
        st.grid = 8LL * st.xcd_per;
    } else {
        st.flags &= ~F_XCD_REMAP;
        st.xcd_per = 0;
        st.grid = st.tiles;
    }
}

// Line grid and access patterns of a step whose in/out layouts are set.
static pfh_status step_geometry(Step &st)
{
    const int a = st.axis;
    const auto sti = st.in.strides(), sto = st.out.strides();
    st.ips = sti[a];
    st.ops = sto[a];
    std::vector<int> others;
    for (int b = 0; b < st.in.N; ++b)
        if (b != a && st.in.shape[b] > 1) others.push_back(b);
    st.in_col = (st.ips != 1 && st.nin > 1) ? 1 : 0;
    st.out_col = (st.ops != 1 && st.nout > 1) ? 1 : 0;
    if (!st.in_col && st.out_col)
        std::sort(others.begin(), others.end(), [&](int x, int y) { return sto[x] < sto[y]; });
    else
        std::sort(others.begin(), others.end(), [&](int x, int y) { return sti[x] < sti[y]; });
    LineGeom g{};
    int nd = 0;
    int64_t nlines = 1;
    for (int d = 0; d < kMaxDigits; ++d) st.digit_axis[d] = -1;
    for (int b : others) {
        if (nd >= kMaxDigits) return set_err(PFH_ERR_UNSUPPORTED, "too many line digits");
        g.ext[nd] = st.in.shape[b];
        g.is[nd] = sti[b];
        g.os[nd] = sto[b];
        st.digit_axis[nd] = b;
        nlines *= st.in.shape[b];
        ++nd;
    }
    if (st.in.extra > 1) {
        if (nd >= kMaxDigits) return set_err(PFH_ERR_UNSUPPORTED, "too many line digits");
        g.ext[nd] = st.in.extra;
        g.is[nd] = st.in.len();
        g.os[nd] = st.out.len();
        nlines *= st.in.extra;
        ++nd;
    }
    g.nd = nd;
    // the kernels divide line indexes in 32 bits (make_fastdiv)
    if (nlines > INT32_MAX) return set_err(PFH_ERR_UNSUPPORTED, "more than 2^31 - 1 lines in one stage");
    st.g = g;
    st.nlines = nlines;
    return PFH_OK;
}

// Threads per line of a step's radix plan: every pass's butterflies (and
// the pre-processing) covered, rounded up to a power of two.
static int threads_per_line(const Step &st, int tbytes)
{
    bool wide_plan = false;
    for (int R : st.radices) wide_plan = wide_plan || is_wide_radix(R);
    const int emax = st.fast ? st.fem : kEmax;
    int need = 1;
    for (int R : st.radices) {
        if (is_generic_radix(R)) {
            need = std::max(need, ceil_div(st.m, emax));
            continue;
        }
        const int MB = st.fast ? std::max(1, emax / R) : generic_mb(R, wide_plan, tbytes == 4);
        need = std::max(need, ceil_div(st.m / R, MB));
    }
    if (mode_needs_pre(st.mode))
        need = std::max(need, ceil_div(st.m, st.fast ? emax : generic_pre_e(wide_plan, tbytes == 4)));
    return pow2ceil(need);
}

// Launch shape (tile, threads per line, LDS) of a step whose line grid is set.
static pfh_status size_step(Step &st, const DeviceInfo &dev, int tbytes, const Tuning &tu)
{
    const int cb = 2 * tbytes;
    const int in_eb = st.in_cplx ? cb : tbytes, out_eb = st.out_cplx ? cb : tbytes;
    // column access: enough lines per tile that each point of the tile is a
    // 256-byte run for the compile-time kernels (measured faster than
    // 128-byte runs at 512^3 f32) and a 128-byte run for the generic kernel
    // (its f64 stages gain 10-25% from the smaller LDS tiles, f32 is even)
    const int run_bytes = st.fast ? 256 : 128;
    int lt_col = 1;
    if (st.in_col) lt_col = std::max(lt_col, run_bytes / in_eb);
    if (st.out_col) lt_col = std::max(lt_col, run_bytes / out_eb);
    if (g_column_lines > 0) lt_col = g_column_lines;
    if (tu.col_lines > 0) lt_col = tu.col_lines;
    bool wide_plan = false;
    for (int R : st.radices) wide_plan = wide_plan || is_wide_radix(R);
    const int max_threads = st.fast ? 512 : (wide_plan ? kMaxThreadsWide : kMaxThreads);
    int tpl = threads_per_line(st, tbytes);
    static const int block_env = env_int("PFH_BLOCK", 256);
    const int target = tu.block > 0 ? tu.block : block_env;
    int lt = (st.in_col || st.out_col) ? lt_col : std::max(1, target / tpl);
    if (tu.fill_block && (st.in_col || st.out_col)) lt = std::max(lt, target / tpl);
    while (lt > 1 && (int64_t)(lt / 2) >= st.nlines) lt /= 2;
    while (lt * tpl < 64) tpl *= 2;
    const int cap = tu.block_cap > 0 ? std::min(tu.block_cap, max_threads) : max_threads;
    // A generic stage with a column side whose tile of lt lines x tpl
    // threads exceeds the workgroup keeps the tile (runs of lt elements on
    // the column side) and computes it in rounds of lc lines. With lc = lt
    // instead, the 385-point DST-I lines of a 384-long axis (128 threads
    // each) got 4-line tiles: 16-byte column runs, 3.5x slower on MI355X.
    static const int rounds_env = env_int("PFH_ROUNDS", 1);
    const bool rounds = rounds_env && !st.fast && (st.in_col || st.out_col) && (lt & (lt - 1)) == 0;
    int lc = lt;
    while (lc * tpl > cap && lc > 1) lc /= 2;
    if (lc * tpl > cap) return set_err(PFH_ERR_UNSUPPORTED, "line too long for one workgroup");
    if (!rounds) lt = lc;
    // the column mapping spreads one point's lt lines over lc * tpl threads
    while (lt > lc * tpl) lt /= 2;
    int need_c = std::max(st.m, 1);
    need_c = std::max(need_c, st.in_cplx ? st.nin : (st.nin + 1) / 2);
    need_c = std::max(need_c, st.out_cplx ? st.nout : (st.nout + 1) / 2);
    // one pad slot per 16 values inside a line (see lp() in the kernels)
    const int padded = need_c + (need_c >> 4) + 1;
    st.ls = (padded % 2 == 0) ? padded + 1 : padded;
    auto lds_of = [&](int l) { return (size_t)(st.twn + (int64_t)l * st.ls) * (size_t)cb; };
    while (lds_of(lt) > dev.lds_limit && lt > 1) {
        lt /= 2;
        if (lc > lt) lc = lt;
        while (lc * tpl < 64) tpl *= 2;
    }
    if (lds_of(lt) > dev.lds_limit) return set_err(PFH_ERR_UNSUPPORTED, "line does not fit in LDS");
    st.lt = lt;
    st.lc = lc;
    st.tpl = tpl;
    st.threads = lc * tpl;
    st.lds_bytes = lds_of(lt);
    set_tiles(st, tu);
    // Real column sides of the compile-time R2R (load and store) and DHT
    // (store) stages: lines 2i and 2i + 1 of a tile are neighbouring
    // elements when the fastest line digit has stride 1 and an even extent
    // and every other line stride and the point stride are even; a lane
    // then moves both with one 2-element access (PFH_CPAIR=0: one element).
    static const int cpair_env = env_int("PFH_CPAIR", 1);
    st.flags &= ~(F_CPAIR_IN | F_CPAIR_OUT | F_RPAIR_IN | F_RPAIR_OUT);
    const bool kind2 = st.mode != M_C2C && st.mode != M_R2C && st.mode != M_C2R && st.mode != M_DHT;
    auto pairs_ok = [&](bool in) {
        const LineGeom &g = st.g;
        if (g.nd < 1 || g.ext[0] % 2 != 0 || (in ? g.is[0] : g.os[0]) != 1 || (in ? st.ips : st.ops) % 2 != 0)
            return false;
        for (int d = 1; d < g.nd; ++d)
            if ((in ? g.is[d] : g.os[d]) % 2 != 0) return false;
        return st.lt % 2 == 0;
    };
    if (cpair_env && st.fast && !st.long_only) {
        if (kind2 && st.in_col && !st.in_cplx && pairs_ok(true)) st.flags |= F_CPAIR_IN;
        if ((kind2 || st.mode == M_DHT) && st.out_col && !st.out_cplx && pairs_ok(false)) st.flags |= F_CPAIR_OUT;
        // rows of the R2R stages: an even number of points and even line
        // offsets keep every point pair 2-element aligned
        auto rows_ok = [&](bool in) {
            const LineGeom &g = st.g;
            if ((in ? st.ips : st.ops) != 1 || (in ? st.nin : st.nout) % 2 != 0) return false;
            for (int d = 0; d < g.nd; ++d)
                if ((in ? g.is[d] : g.os[d]) % 2 != 0) return false;
            return true;
        };
        if (kind2 && !st.in_col && !st.in_cplx && rows_ok(true)) st.flags |= F_RPAIR_IN;
        if (kind2 && !st.out_col && !st.out_cplx && rows_ok(false)) st.flags |= F_RPAIR_OUT;
    }
    // Non-temporal global accesses (the compile-time radix plans only) once
    // the stage's input and output together exceed the 256 MB last-level
    // cache (PFH_NT_MB overrides the threshold, 0 turns them off): measured
    // on MI355X, +5% on 512^3 f32 stages, but -25% on batches of ~20 MB
    // problems, whose tiles share cache lines; batched launches never use them.
    // They also need whole cache lines per tile: rows of both arrays (the
    // fastest memory axis) and column-tile runs (lines per tile x element)
    // in multiples of 128 bytes; otherwise neighbouring tiles share lines and
    // the hint evicts what the neighbour reads (permute_dims = false at
    // 512^3, 257-element complex rows: 450 against 640 round trips/s).
    static const int nt_mb = env_int("PFH_NT_MB", 256);
    auto rows_aligned = [](const Layout &L) {
        return L.N > 0 && (L.shape[L.perm[L.N - 1]] * (int64_t)L.elem) % 128 == 0;
    };
    const bool runs_aligned = (!st.in_col || ((int64_t)st.lt * st.in.elem) % 128 == 0) &&
                              (!st.out_col || ((int64_t)st.lt * st.out.elem) % 128 == 0);
    if (nt_mb > 0 && st.in.bytes() + st.out.bytes() > ((int64_t)nt_mb << 20) && rows_aligned(st.in) &&
        rows_aligned(st.out) && runs_aligned)
        st.flags |= F_NT;
    else
        st.flags &= ~F_NT;
    return PFH_OK;
}

// Line grid, access patterns and launch shape of a step whose in/out
// layouts are set.
static pfh_status finish_step(Step &st, const DeviceInfo &dev, int tbytes, const Tuning &tu)
{
    pfh_status r = step_geometry(st);
    if (r) return r;
    return size_step(st, dev, tbytes, tu);
}

// ---------------------------------------------------------------------------
// long lines: a stage whose line does not fit one workgroup, or whose FFT
// length has a large prime factor, runs as pieces (generic line stages on
// workspace regions, plus the prep / post kernels of real lines)
// ---------------------------------------------------------------------------

static int64_t largest_prime_factor(int64_t m)
{
    int64_t best = 1;
    for (int64_t p = 2; p * p <= m; ++p)
        while (m % p == 0) {
            best = p;
            m /= p;
        }
    return m > 1 ? std::max(best, m) : best;
}

// Prime factors above this run through Bluestein's chirp convolution. The
// generic pass costs about 4p flops per point for a factor p, against 16
// bytes of HBM traffic per point and pass: at p near 100 a stage is as
// compute bound as it is HBM bound on MI355X (157 TFLOP/s f32 against
// 8 TB/s), while Bluestein's pieces move about five times a stage's bytes.
// PFH_BLUESTEIN_PRIME overrides it; read at every plan creation.
static int64_t bluestein_prime()
{
    const char *e = std::getenv("PFH_BLUESTEIN_PRIME");
    const int64_t v = e ? std::atoll(e) : 0;
    return v > 0 ? v : 128;
}

// PFH_LONG_LINE_MAX: complex FFT lengths above it run as pieces even where
// one workgroup would hold the line (0 or unset: only where it does not);
// read at every plan creation.
static int64_t long_line_max()
{
    const char *e = std::getenv("PFH_LONG_LINE_MAX");
    return e ? std::atoll(e) : 0;
}

static bool wants_pieces(const Step &s)
{
    if (s.long_only) return true;
    if (s.m <= 1) return false;
    const int64_t lim = long_line_max();
    if (lim > 0 && s.m > lim) return true;
    return largest_prime_factor(s.m) > bluestein_prime();
}

// Pass tables of a radix split (roots of generic radices, then the pass
// twiddles), offsets relative to the block; appended to tab when non-null.
static int pass_tables(const std::vector<int> &radices, std::vector<double> *tab, std::array<int, kMaxPasses> &twoff)
{
    int off = 0, Ns = 1;
    for (size_t p = 0; p < radices.size(); ++p) {
        const int R = radices[p];
        twoff[p] = off;
        if (is_generic_radix(R)) {
            if (tab)
                for (int e = 0; e < R; ++e) push_cexp(*tab, e, R);
            off += R;
        }
        if (Ns > 1) {
            if (tab)
                for (int k = 0; k < Ns; ++k)
                    for (int r = 1; r < R; ++r) push_cexp(*tab, (long long)r * k, (long long)Ns * R);
            off += Ns * (R - 1);
        }
        Ns *= R;
    }
    return off;
}

// A named table of the plan (built once; long-line pieces of several
// operations share it). fill appends the entries to host_tables.
template <typename F>
static pfh_status plan_table(PlanImpl &P, const std::string &name, int64_t &offset, F fill)
{
    auto it = P.table_index.find(name);
    if (it != P.table_index.end()) {
        offset = it->second;
        return PFH_OK;
    }
    if (P.tables_frozen) return set_err(PFH_ERR_PREPARATION_FAILED, "long-line table missing after upload");
    offset = (int64_t)(P.host_tables.size() / 2);
    fill(P.host_tables);
    P.table_index[name] = offset;
    return PFH_OK;
}

// Forward DFT of a 5-smooth length in long double (decimation in time over
// the factors 2, 3, 5; w = exp(-2 pi i e / n) for the full length n).
using cxl = std::complex<long double>;
static void host_dft_rec(const cxl *in, size_t stride, cxl *out, size_t len, const std::vector<cxl> &w, size_t wstep)
{
    if (len == 1) {
        out[0] = in[0];
        return;
    }
    const size_t p = (len % 2 == 0) ? 2 : (len % 3 == 0) ? 3 : (len % 5 == 0) ? 5 : len;
    const size_t q = len / p;
    if (p == len) {
        for (size_t k = 0; k < len; ++k) {
            cxl acc = 0;
            for (size_t j = 0; j < len; ++j) acc += in[j * stride] * w[((j * k) % len) * wstep];
            out[k] = acc;
        }
        return;
    }
    for (size_t r = 0; r < p; ++r) host_dft_rec(in + r * stride, stride * p, out + r * q, q, w, wstep * p);
    cxl t[5];
    for (size_t k = 0; k < q; ++k) {
        for (size_t r = 0; r < p; ++r) t[r] = out[r * q + k];
        for (size_t s = 0; s < p; ++s) {
            cxl acc = 0;
            for (size_t r = 0; r < p; ++r) acc += t[r] * w[((r * (k + q * s)) % len) * wstep];
            out[k + q * s] = acc;
        }
    }
}

static std::vector<cxl> host_dft(const std::vector<cxl> &x)
{
    const size_t n = x.size();
    std::vector<cxl> w(n), out(n);
    for (size_t e = 0; e < n; ++e) {
        std::vector<double> t;
        push_cexp(t, (long long)e, (long long)n);
        w[e] = cxl(t[0], t[1]);
    }
    host_dft_rec(x.data(), 1, out.data(), n, w, 1);
    return out;
}

// exp(s i pi j^2 / m), the phase reduced exactly (j^2 mod 2m) before the
// trigonometry.
static cxl chirp(int64_t j, int64_t m, int s)
{
    const unsigned __int128 jj = (unsigned __int128)(uint64_t)j * (uint64_t)j;
    const int64_t r = (int64_t)(jj % (unsigned __int128)(2 * m));
    const long double a = (long double)s * kPiL * (long double)r / (long double)m;
    return cxl(cosl(a), sinl(a));
}

static bool five_smooth(int64_t v)
{
    for (int p : {2, 3, 5})
        while (v % p == 0) v /= p;
    return v == 1;
}

// Does a complex FFT of length m run as one generic line stage?
static bool piece_fits(int64_t m, const DeviceInfo &dev, int tb, const Tuning &tu)
{
    if (m < 1 || m > INT32_MAX) return false;
    const int64_t lim = long_line_max();
    if (lim > 0 && m > lim) return false;
    if (m > 1 && largest_prime_factor(m) > bluestein_prime()) return false;
    Step p;
    p.mode = M_C2C;
    p.n = p.m = p.nin = p.nout = (int)m;
    const int max_r = tu.max_radix ? tu.max_radix : (tb == 4 ? 16 : 8);
    if (!radix_plan((int)m, p.radices, max_r)) return false;
    p.twn = pass_tables(p.radices, nullptr, p.twoff);
    p.g.nd = 1;
    p.g.ext[0] = 64;
    p.g.is[0] = p.g.os[0] = m;
    p.nlines = 64;
    p.ips = p.ops = 1;
    const pfh_error_info saved = g_err;
    const bool fits = size_step(p, dev, tb, tu) == PFH_OK;
    g_err = saved;
    return fits;
}

// N1 (the first piece's length) of a split N1 x N2 = m whose pieces both
// fit one workgroup, N1 as close to sqrt(m) as the divisors allow; 0: none.
static int64_t split_factor(int64_t m, const DeviceInfo &dev, int tb, const Tuning &tu)
{
    int64_t best = 0;
    double best_d = 1e300;
    for (int64_t a = 2; a * a <= m; ++a) {
        if (m % a) continue;
        for (int64_t n1 : {a, m / a}) {
            const int64_t n2 = m / n1;
            if (!piece_fits(n1, dev, tb, tu) || !piece_fits(n2, dev, tb, tu)) continue;
            const double d = std::fabs(std::log((double)n1) - 0.5 * std::log((double)m));
            if (d < best_d) {
                best_d = d;
                best = n1;
            }
        }
    }
    return best;
}

// Radix passes of a generic line stage of length len.
static int radix_passes(int64_t len, int tb, const Tuning &tu)
{
    std::vector<int> r;
    const int max_r = tu.max_radix ? tu.max_radix : (tb == 4 ? 16 : 8);
    return radix_plan((int)len, r, max_r) ? (int)r.size() : 64;
}

// Length of Bluestein's cyclic convolution for an FFT of length m: a
// 5-smooth M >= 2m - 1 that runs as one piece or a split, the one with the
// fewest M x (radix passes of its pieces) among those up to 25% above the
// smallest. For m = 10007: 20480 = 128 x 160 (2 + 3 passes) rather than
// 20250 = 135 x 150 (4 + 4 passes).
static int64_t bluestein_length(int64_t m, const DeviceInfo &dev, int tb, const Tuning &tu)
{
    int64_t best = 0;
    double best_cost = 0;
    const int64_t lo = 2 * m - 1;
    for (int64_t M = lo; M <= 4 * m && M <= ((int64_t)1 << 25); ++M) {
        if (best && (double)M > 1.25 * (double)lo) break;
        if (!five_smooth(M)) continue;
        int passes = 0;
        if (piece_fits(M, dev, tb, tu)) {
            passes = radix_passes(M, tb, tu);
        } else {
            const int64_t n1 = split_factor(M, dev, tb, tu);
            if (n1 <= 0) continue;
            passes = radix_passes(n1, tb, tu) + radix_passes(M / n1, tb, tu);
        }
        const double cost = (double)M * passes;
        if (!best || cost < best_cost) {
            best = M;
            best_cost = cost;
        }
    }
    return best;
}

// Strides of one side of the long lines: the stage's line digits (is or os
// of its geometry) and the point stride; line-major regions have point
// stride 1 and line l at l * len.
struct LineSide {
    int buf = BUF_SRC;
    int64_t point = 1;
    int64_t dig[kMaxDigits] = {0, 0, 0, 0, 0};
};

static LineSide region_side(const Step &outer, int buf, int64_t len)
{
    LineSide v;
    v.buf = buf;
    v.point = 1;
    int64_t run = len;
    for (int d = 0; d < outer.g.nd; ++d) {
        v.dig[d] = run;
        run *= outer.g.ext[d];
    }
    return v;
}

struct PieceDigit {
    int64_t ext, is, os;
    bool special;
};

struct PieceX {
    int flags = 0; // F_CONJ_IN / F_CONJ_OUT / F_XIN / F_XOUT / F_XTW
    double scale = 1.0;
    int64_t tab_in = -1, tab_out = -1, tw = -1;
    int mul_in = 1, mul_out = 1, valid_in = 0, valid_out = 0, tw_m = 0, tw_bsh = 0;
};

// Two-level table of exp(-2 pi i q / M): 2^bsh low entries, then the
// powers of exp(-2 pi i 2^bsh / M).
static pfh_status twiddle_table(PlanImpl &P, int64_t M, int64_t &off, int &bsh)
{
    bsh = 0;
    while (((int64_t)1 << (2 * bsh)) < M) ++bsh;
    const int b = bsh;
    return plan_table(P, "tw2:" + std::to_string(M), off, [&](std::vector<double> &t) {
        const int64_t B = (int64_t)1 << b;
        for (int64_t q = 0; q < B; ++q) push_cexp(t, q, M);
        for (int64_t h = 0; h * B < M; ++h) push_cexp(t, h * B, M);
    });
}

// One generic C2C line stage of length N over the given digits.
static pfh_status make_piece(PlanImpl &P, Step &p, int64_t N, std::vector<PieceDigit> digs, int64_t ips,
                             int64_t ops, int in_buf, int out_buf, const PieceX &x, int tb)
{
    p = Step();
    p.mode = M_C2C;
    p.n = p.m = p.nin = p.nout = (int)N;
    p.in_cplx = p.out_cplx = true;
    p.flags = x.flags;
    p.scale = x.scale;
    const int max_r = P.tuning.max_radix ? P.tuning.max_radix : (tb == 4 ? 16 : 8);
    if (!radix_plan((int)N, p.radices, max_r)) return set_err(PFH_ERR_UNSUPPORTED, "long-line piece radix plan");
    int64_t base = 0;
    std::array<int, kMaxPasses> twoff{};
    const int twn = pass_tables(p.radices, nullptr, twoff);
    pfh_status st = plan_table(P, "pass:" + std::to_string(N) + ":" + std::to_string(max_r), base,
                               [&](std::vector<double> &t) { pass_tables(p.radices, &t, twoff); });
    if (st) return st;
    p.tw_base = base;
    p.twn = twn;
    p.twoff = twoff;
    std::vector<PieceDigit> keep;
    for (const PieceDigit &d : digs)
        if (d.ext > 1) keep.push_back(d);
    if ((int)keep.size() > kMaxDigits) return set_err(PFH_ERR_UNSUPPORTED, "too many line digits");
    const bool in_col = ips != 1, out_col = ops != 1;
    std::stable_sort(keep.begin(), keep.end(), [&](const PieceDigit &a, const PieceDigit &b) {
        return (!in_col && out_col) ? a.os < b.os : a.is < b.is;
    });
    int64_t nlines = 1;
    p.g.nd = (int)keep.size();
    for (int d = 0; d < p.g.nd; ++d) {
        p.g.ext[d] = keep[d].ext;
        p.g.is[d] = keep[d].is;
        p.g.os[d] = keep[d].os;
        if (keep[d].special) p.xdig = d;
        nlines *= keep[d].ext;
    }
    if (nlines > INT32_MAX) return set_err(PFH_ERR_UNSUPPORTED, "more than 2^31 - 1 lines in one stage");
    p.nlines = nlines;
    p.ips = ips;
    p.ops = ops;
    p.in_col = in_col ? 1 : 0;
    p.out_col = out_col ? 1 : 0;
    p.in_buf = in_buf;
    p.out_buf = out_buf;
    p.xtab_in = x.tab_in;
    p.xtab_out = x.tab_out;
    p.xtw = x.tw;
    p.xmul_in = x.mul_in;
    p.xmul_out = x.mul_out;
    p.xvalid_in = x.valid_in;
    p.xvalid_out = x.valid_out;
    p.xtw_m = x.tw_m;
    p.xtw_bsh = x.tw_bsh;
    Tuning tu = P.tuning;
    tu.allow_fast = false;
    return size_step(p, P.dev, tb, tu);
}

// A complex FFT of length N from side `in` to side `out` over the outer
// stage's lines: one piece, or the four-step split N = N1 x N2 through
// region rs (line-major [line][k1][n2]): the first piece transforms n1 (index
// N2 n1 + n2, lines n2) and multiplies by exp(-2 pi i n2 k1 / N) at its
// store, the second transforms n2 (lines k1) and writes index k1 + N1 k2.
// xin / xout (load / store factors, padding and crop) go to the first /
// last piece, with full indexes N2 n1 + n2 and k1 + N1 k2.
static pfh_status fft_pieces(PlanImpl &P, const Step &outer, int64_t N, const LineSide &in, const LineSide &out,
                             const PieceX &xin, const PieceX &xout, int rs_buf, int64_t xb[3], int tb,
                             std::vector<Step> &pieces)
{
    const int nd = outer.g.nd;
    if (piece_fits(N, P.dev, tb, P.tuning)) {
        std::vector<PieceDigit> digs;
        for (int d = 0; d < nd; ++d) digs.push_back({outer.g.ext[d], in.dig[d], out.dig[d], false});
        PieceX x = xout;
        x.flags = (xin.flags & (F_CONJ_IN | F_XIN)) | (xout.flags & (F_CONJ_OUT | F_XOUT));
        x.tab_in = xin.tab_in;
        x.mul_in = 1;
        x.valid_in = xin.valid_in;
        x.mul_out = 1;
        Step p;
        pfh_status st = make_piece(P, p, N, digs, in.point, out.point, in.buf, out.buf, x, tb);
        if (st) return st;
        pieces.push_back(p);
        return PFH_OK;
    }
    const int64_t N1 = split_factor(N, P.dev, tb, P.tuning);
    if (N1 <= 0) return set_err(PFH_ERR_UNSUPPORTED, "no split of the long line into workgroup pieces");
    const int64_t N2 = N / N1;
    const LineSide rs = region_side(outer, rs_buf, N);
    xb[rs_buf - BUF_X0] = std::max(xb[rs_buf - BUF_X0], outer.nlines * N * 2 * tb);
    // first piece: lines (n2, outer lines), points n1
    {
        std::vector<PieceDigit> digs;
        digs.push_back({N2, in.point, 1, true});
        for (int d = 0; d < nd; ++d) digs.push_back({outer.g.ext[d], in.dig[d], rs.dig[d], false});
        PieceX x;
        x.flags = (xin.flags & (F_CONJ_IN | F_XIN)) | F_XTW;
        x.tab_in = xin.tab_in;
        x.mul_in = (int)N2;
        x.valid_in = xin.valid_in;
        pfh_status st = twiddle_table(P, N, x.tw, x.tw_bsh);
        if (st) return st;
        x.tw_m = (int)N;
        Step p;
        st = make_piece(P, p, N1, digs, N2 * in.point, N2, in.buf, rs_buf, x, tb);
        if (st) return st;
        pieces.push_back(p);
    }
    // second piece: lines (k1, outer lines), points n2 -> index k1 + N1 k2
    {
        std::vector<PieceDigit> digs;
        digs.push_back({N1, N2, out.point, true});
        for (int d = 0; d < nd; ++d) digs.push_back({outer.g.ext[d], rs.dig[d], out.dig[d], false});
        PieceX x = xout;
        x.flags = xout.flags & (F_CONJ_OUT | F_XOUT);
        x.mul_out = (int)N1;
        Step p;
        pfh_status st = make_piece(P, p, N2, digs, 1, N1 * out.point, rs_buf, out.buf, x, tb);
        if (st) return st;
        pieces.push_back(p);
    }
    return PFH_OK;
}

// The complex FFT of length m (sign from conj: F_CONJ_IN | F_CONJ_OUT =
// positive exponent), scaled, from `in` to `out`: directly in pieces, or
// through Bluestein's convolution when m has a prime factor above the
// threshold. Bluestein (sign s = -1 forward, +1 positive):
//   Z_k = c_k sum_j (x_j c_j) b_(k-j),  c_j = exp(s i pi j^2 / m),
//   b_t = exp(-s i pi t^2 / m),
// a cyclic convolution of length M >= 2m - 1: piece B1 loads x_j c_j (zero
// from m on), runs FFT_M and stores A_k * FFT_M(b)_k / M into region rb;
// piece B2 loads the conjugate, runs FFT_M, stores the conjugate (= the
// convolution) times c_k * scale for k < m.
static pfh_status core_pieces(PlanImpl &P, const Step &outer, int64_t m, const LineSide &in, const LineSide &out,
                              int conj, double scale, int rs_buf, int rb_buf, int64_t xb[3], int tb,
                              std::vector<Step> &pieces)
{
    if (m <= 1 || largest_prime_factor(m) <= bluestein_prime()) {
        PieceX xi, xo;
        xi.flags = conj & F_CONJ_IN;
        xi.valid_in = (int)m;
        xo.flags = conj & F_CONJ_OUT;
        xo.scale = scale;
        xo.valid_out = (int)m;
        return fft_pieces(P, outer, m, in, out, xi, xo, rs_buf, xb, tb, pieces);
    }
    const int64_t M = bluestein_length(m, P.dev, tb, P.tuning);
    if (M <= 0) return set_err(PFH_ERR_UNSUPPORTED, "no Bluestein length for this line");
    const int s = (conj & F_CONJ_IN) ? 1 : -1;
    int64_t c_off = 0, bh_off = 0;
    pfh_status st = plan_table(P, "chirp:" + std::to_string(m) + ":" + std::to_string(s), c_off,
                               [&](std::vector<double> &t) {
                                   for (int64_t j = 0; j < m; ++j) {
                                       const cxl c = chirp(j, m, s);
                                       t.push_back((double)c.real());
                                       t.push_back((double)c.imag());
                                   }
                               });
    if (st) return st;
    st = plan_table(P, "blue:" + std::to_string(m) + ":" + std::to_string(M) + ":" + std::to_string(s), bh_off,
                    [&](std::vector<double> &t) {
                        std::vector<cxl> b((size_t)M, cxl(0, 0));
                        for (int64_t j = 0; j < m; ++j) {
                            const cxl v = chirp(j, m, -s);
                            b[(size_t)j] = v;
                            if (j > 0) b[(size_t)(M - j)] = v;
                        }
                        const std::vector<cxl> B = host_dft(b);
                        for (int64_t k = 0; k < M; ++k) {
                            t.push_back((double)(B[(size_t)k].real() / (long double)M));
                            t.push_back((double)(B[(size_t)k].imag() / (long double)M));
                        }
                    });
    if (st) return st;
    const LineSide rb = region_side(outer, rb_buf, M);
    xb[rb_buf - BUF_X0] = std::max(xb[rb_buf - BUF_X0], outer.nlines * M * 2 * tb);
    {
        PieceX xi, xo;
        xi.flags = F_XIN;
        xi.tab_in = c_off;
        xi.valid_in = (int)m;
        xo.flags = F_XOUT;
        xo.tab_out = bh_off;
        xo.valid_out = (int)M;
        st = fft_pieces(P, outer, M, in, rb, xi, xo, rs_buf, xb, tb, pieces);
        if (st) return st;
    }
    PieceX xi, xo;
    xi.flags = F_CONJ_IN;
    xi.valid_in = (int)M;
    xo.flags = F_CONJ_OUT | F_XOUT;
    xo.scale = scale;
    xo.tab_out = c_off;
    xo.valid_out = (int)m;
    return fft_pieces(P, outer, M, rb, out, xi, xo, rs_buf, xb, tb, pieces);
}

// Expands a stage (line grid set, launch shape not) into its pieces:
// complex lines directly; real lines through prep (the promoted FFT input,
// line-major in region X2), the complex FFT (X2 -> X2) and post. Regions:
// X0 the split intermediate, X1 Bluestein's, X2 the promoted lines.
static pfh_status expand_long(PlanImpl &P, Step &s, int tb, int64_t xb[3])
{
    s.pieces.clear();
    s.fast = false;
    if (s.g.nd + 1 > kMaxDigits) return set_err(PFH_ERR_UNSUPPORTED, "too many line digits for a long line");
    LineSide in, out;
    in.buf = s.in_buf;
    in.point = s.ips;
    out.buf = s.out_buf;
    out.point = s.ops;
    for (int d = 0; d < s.g.nd; ++d) {
        in.dig[d] = s.g.is[d];
        out.dig[d] = s.g.os[d];
    }
    if (s.mode == M_C2C)
        return core_pieces(P, s, s.n, in, out, s.flags & (F_CONJ_IN | F_CONJ_OUT), s.scale, BUF_X0, BUF_X1, xb, tb,
                           s.pieces);
    const int64_t n = s.n;
    int kind;
    int64_t m;
    switch (s.mode) {
    case M_R2C: case M_R2C_ODD: kind = LK_R2C; m = n; break;
    case M_C2R: case M_C2R_ODD: kind = LK_C2R; m = n; break;
    case M_DCT2: case M_DCT2_ODD: kind = LK_DCT2; m = n; break;
    case M_DCT3: case M_DCT3_ODD: kind = LK_DCT3; m = n; break;
    case M_DCT4: case M_DCT4_ODD: kind = LK_DCT4; m = 2 * n; break;
    case M_DCT1: kind = LK_DCT1; m = 2 * (n - 1); break;
    case M_DST1: kind = LK_DST1; m = 2 * (n + 1); break;
    case M_DHT: case M_DHT_ODD: kind = LK_DHT; m = n; break;
    default: return set_err(PFH_ERR_UNSUPPORTED, "long line of this stage kind");
    }
    if (m > INT32_MAX) return set_err(PFH_ERR_UNSUPPORTED, "line too long");
    int64_t pre0 = 0, post0 = 0;
    const std::string nk = std::to_string(n);
    pfh_status st = PFH_OK;
    if (kind == LK_DCT3 || kind == LK_DCT4)
        st = plan_table(P, "lpre4n:" + nk, pre0, [&](std::vector<double> &t) {
            for (int64_t j = 0; j < n; ++j) push_cexp(t, j, 4 * n);
        });
    if (st) return st;
    if (kind == LK_DCT2)
        st = plan_table(P, "lpre4n:" + nk, post0, [&](std::vector<double> &t) {
            for (int64_t j = 0; j < n; ++j) push_cexp(t, j, 4 * n);
        });
    if (kind == LK_DCT4)
        st = plan_table(P, "lpost8n:" + nk, post0, [&](std::vector<double> &t) {
            for (int64_t k = 0; k < n; ++k) push_cexp(t, 2 * k + 1, 8 * n);
        });
    if (st) return st;
    xb[2] = std::max(xb[2], s.nlines * m * 2 * tb);
    const LineSide ra = region_side(s, BUF_X2, m);
    Step pre;
    pre.launch = 1;
    pre.long_kind = kind;
    pre.g = s.g;
    pre.nlines = s.nlines;
    pre.ips = s.ips;
    pre.ops = 1;
    pre.n = (int)n;
    pre.m = (int)m;
    pre.nout = s.nout;
    pre.pre0 = (int)pre0;
    pre.post0 = (int)post0;
    pre.tw_base = 0;
    pre.flags = s.flags;
    pre.scale = s.scale;
    pre.prew0 = s.prew0;
    pre.prew1 = s.prew1;
    pre.postw0 = s.postw0;
    pre.postw1 = s.postw1;
    pre.in_buf = s.in_buf;
    pre.out_buf = BUF_X2;
    s.pieces.push_back(pre);
    st = core_pieces(P, s, m, ra, ra, 0, 1.0, BUF_X0, BUF_X1, xb, tb, s.pieces);
    if (st) return st;
    Step post = pre;
    post.launch = 2;
    post.in_buf = BUF_X2;
    post.out_buf = s.out_buf;
    post.ops = s.ops;
    s.pieces.push_back(post);
    return PFH_OK;
}

// Shape of a stage: its launch shape, or its pieces when its line needs them.
static pfh_status finish_or_expand(PlanImpl &P, Step &s, int tb, int64_t xb[3])
{
    if (!wants_pieces(s)) {
        const pfh_error_info saved = g_err;
        const pfh_status st = finish_step(s, P.dev, tb, P.tuning);
        if (st != PFH_ERR_UNSUPPORTED) return st;
        g_err = saved;
    }
    pfh_status st = step_geometry(s);
    if (st) return st;
    return expand_long(P, s, tb, xb);
}

// ---------------------------------------------------------------------------
// routes
// ---------------------------------------------------------------------------

static void build_route(int N, bool permute, std::vector<std::array<int, 8>> &perm)
{
    perm.assign(N, {});
    for (int p = 0; p < N; ++p) perm[0][p] = p;
    int s = 1;
    for (int axis = N - 2; axis >= 0; --axis, ++s) {
        if (!permute) {
            for (int p = 0; p < N; ++p) perm[s][p] = p;
            continue;
        }
        int q = 0;
        for (int p = 0; p < N; ++p)
            if (perm[s - 1][p] != axis) perm[s][q++] = perm[s - 1][p];
        perm[s][N - 1] = axis;
    }
}

static Layout make_layout(int N, const std::array<int, 8> &perm, const std::array<int64_t, 8> &shape, int64_t extra,
                          int elem)
{
    Layout L;
    L.N = N;
    L.perm = perm;
    L.shape = shape;
    L.extra = extra;
    L.elem = elem;
    return L;
}

// Stage specs of one operation in execution order.
// The real route's operation that a plan operation runs: a C2rPlan runs the
// route in C2R orientation, its FORWARD being the raw C2R (the route's
// BACKWARD), BACKWARD the raw R2C (FORWARD) and INVERSE the R2C normalized
// by the selected real extents (c2r.rs:306-360).
static int route_op(int kind, int op)
{
    if (kind != PFH_C2R) return op;
    switch (op) {
    case PFH_FORWARD: return PFH_BACKWARD;
    case PFH_BACKWARD: return PFH_FORWARD;
    case PFH_INVERSE: return PFH_FORWARD;
    case PFH_ADJOINT_FORWARD: return PFH_ADJOINT_BACKWARD;
    case PFH_ADJOINT_BACKWARD: return PFH_ADJOINT_FORWARD;
    case PFH_ADJOINT_INVERSE: return PFH_ADJOINT_FORWARD;
    default: return op;
    }
}

static pfh_status route_specs(const PlanImpl &P, int op_in, std::vector<StageSpec> &specs)
{
    const pfh_desc &d = P.desc;
    const int N = P.N;
    const int tb = (P.precision == PFH_F32) ? 4 : 8;
    // C2R orientation: stage s transforms axis s in the permutation of route
    // stage N-1-s (stages and transitions reversed, r2c.rs:1340-1348)
    const bool c2r = P.kind == PFH_C2R;
    const bool mixed_c2c = P.kind == PFH_MIXED_C2C, mixed_r2c = P.kind == PFH_MIXED_R2C;
    const bool real_kind = P.kind == PFH_R2C || c2r || mixed_r2c;
    const bool scaled = op_in == PFH_INVERSE || op_in == PFH_ADJOINT_INVERSE;
    const int op = route_op(P.kind, op_in);
    std::vector<std::array<int, 8>> route, perm;
    build_route(N, d.permute_dims != 0, route);
    perm.resize(N);
    for (int s = 0; s < N; ++s) perm[s] = route[c2r ? N - 1 - s : s];
    std::array<int64_t, 8> orig{}, red{};
    for (int a = 0; a < N; ++a) orig[a] = red[a] = (int64_t)d.shape[a];
    unsigned sel = d.selected ? d.selected : ((1u << N) - 1u);
    int boundary = -1, raxis = -1;
    if (mixed_r2c) {
        for (int a = 0; a < N; ++a)
            if (d.axis_kind[a] == PFH_AXIS_RFFT) raxis = a;
        boundary = N - 1 - raxis;
        red[raxis] = orig[raxis] / 2 + 1;
    } else if (real_kind) {
        // R2C: the last selected axis; C2R orientation: the first
        for (int i = 0; i < N; ++i) {
            const int a = c2r ? i : N - 1 - i;
            if (sel & (1u << a)) {
                raxis = a;
                break;
            }
        }
        boundary = c2r ? raxis : N - 1 - raxis;
        red[raxis] = orig[raxis] / 2 + 1;
    }
    const bool fwd = (op == PFH_FORWARD || op == PFH_ADJOINT_INVERSE || op == PFH_ADJOINT_BACKWARD);
    std::vector<StageSpec> by_stage(N);
    for (int s = 0; s < N; ++s) {
        StageSpec sp;
        const int axis = c2r ? s : N - 1 - s;
        sp.axis = axis;
        // forward-orientation shapes / element kinds
        std::array<int64_t, 8> sh_in = orig, sh_out = orig;
        bool c_in = true, c_out = true;
        if (real_kind) {
            sh_in = (s <= boundary) ? orig : red;
            sh_out = (s < boundary) ? orig : red;
            c_in = s > boundary;
            c_out = s >= boundary;
        } else if (P.kind == PFH_R2R) {
            c_in = c_out = false;
        }
        const int code = d.axis_kind[axis];
        if (mixed_c2c || (mixed_r2c && s != boundary)) {
            // mixed stage (mixed.rs:3880-3990): identity, complex FFT, or
            // R2R / DHT on complex data (parts separately) or real data
            const bool cplx = mixed_c2c || s > boundary;
            sp.n = (mixed_r2c && s > boundary) ? red[axis] : orig[axis];
            if (code == PFH_AXIS_NONE) {
                sp.kind = SK_COPY;
            } else if (code == PFH_AXIS_FFT) {
                sp.kind = SK_C2C;
                const bool neg_forward = d.direction[axis] == 0;
                sp.sign_pos = fwd ? !neg_forward : neg_forward;
                if (scaled) sp.scale = 1.0 / (double)sp.n;
            } else {
                r2r_stage(sp, code, sp.n, op);
                sp.cplx_r2r = cplx;
            }
        } else if (P.kind == PFH_C2C) {
            sp.n = orig[axis];
            if (!(sel & (1u << axis))) {
                sp.kind = SK_COPY;
            } else {
                sp.kind = SK_C2C;
                const bool neg_forward = d.direction[axis] == 0;
                sp.sign_pos = fwd ? !neg_forward : neg_forward;
                if (op == PFH_INVERSE || op == PFH_ADJOINT_INVERSE) sp.scale = 1.0 / (double)orig[axis];
            }
        } else if (real_kind) {
            if (s < boundary) {
                sp.kind = SK_COPY;
                sp.n = orig[axis];
            } else if (s > boundary) {
                sp.n = red[axis];
                if (!(sel & (1u << axis))) {
                    sp.kind = SK_COPY;
                } else {
                    sp.kind = SK_C2C;
                    sp.sign_pos = !fwd;
                    if (scaled) sp.scale = 1.0 / (double)red[axis];
                }
            } else {
                sp.n = orig[axis];
                if (fwd) {
                    sp.kind = SK_R2C;
                    if (op != PFH_FORWARD) sp.flags |= F_HERM_DOUBLE;
                } else {
                    sp.kind = SK_C2R;
                    if (op == PFH_ADJOINT_FORWARD) sp.flags |= F_HERM_HALF;
                }
                if (scaled) sp.scale = 1.0 / (double)orig[axis];
            }
        } else { // R2R / DHT
            const int k = d.axis_kind[axis];
            sp.n = orig[axis];
            if (k == 0) sp.kind = SK_COPY;
            else r2r_stage(sp, k, orig[axis], op);
        }
        // element kinds and natural layouts in execution orientation
        if (fwd) {
            sp.in_c = c_in;
            sp.out_c = c_out;
            sp.in_natural = make_layout(N, perm[s], sh_in, P.extra, (c_in ? 2 : 1) * tb);
            sp.out_natural = make_layout(N, perm[s], sh_out, P.extra, (c_out ? 2 : 1) * tb);
        } else {
            sp.in_c = c_out;
            sp.out_c = c_in;
            sp.in_natural = make_layout(N, perm[s], sh_out, P.extra, (c_out ? 2 : 1) * tb);
            sp.out_natural = make_layout(N, perm[s], sh_in, P.extra, (c_in ? 2 : 1) * tb);
        }
        if (sp.kind == SK_COPY || sp.cplx_r2r) {
            sp.in_c = sp.out_c = fwd ? c_in : c_out;
        }
        by_stage[s] = sp;
    }
    specs.clear();
    if (fwd)
        for (int s = 0; s < N; ++s) specs.push_back(by_stage[s]);
    else
        for (int s = N - 1; s >= 0; --s) specs.push_back(by_stage[s]);
    (void)raxis;
    return PFH_OK;
}

// Chooses which side of each stage absorbs each transition and assigns
// buffers (SRC, DST, W0, W1), minimising workspace. in_place: SRC == DST.
// Chooses, for every route transition, whether the stage before it
// absorbs it at its store or the stage after it at its load, and assigns
// every stage's input and output to SRC / DST / W0 / W1. Preference: the
// most slab-pipelined pairs (pair_score), then the fewest workspace arrays.
template <typename Fn>
static bool assign_route(const std::vector<StageSpec> &specs, bool in_place, int64_t dst_capacity, Fn &&pair_score,
                         const std::vector<int> &line_threads, std::vector<Layout> &X, std::vector<Layout> &Y,
                         std::vector<int> &inb, std::vector<int> &outb, int64_t wbytes[2])
{
    const int K = (int)specs.size();
    int best_w = 99, best_pairs = -1;
    int64_t best_cost = INT64_MAX;
    // Column cost of a layout choice: a stage that reads or writes a column
    // side (its axis not the fastest in memory) moves runs of one element
    // per line of its tile, and a tile holds at most 512 / (threads per line)
    // lines (the compile-time kernels' bound). Runs shorter than a 128-byte
    // cache line cost their shortfall: the 385-point DST-I lines (32
    // threads, 16-line tiles) write 64-byte runs of reals, so their stage
    // should read and write rows and leave the transposes to its neighbours.
    auto column_cost = [&](const std::vector<Layout> &x, const std::vector<Layout> &y) {
        int64_t c = 0;
        for (int e = 0; e < K; ++e) {
            const int a = specs[e].axis;
            const int N = x[e].N;
            if (N == 0 || a < 0 || x[e].shape[a] <= 1) continue;
            const int t = e < (int)line_threads.size() ? std::max(1, line_threads[e]) : 1;
            const int64_t lines = std::max(1, 512 / t);
            if (x[e].perm[N - 1] != a) c += std::max<int64_t>(0, 128 - lines * x[e].elem);
            if (y[e].perm[N - 1] != a) c += std::max<int64_t>(0, 128 - lines * y[e].elem);
        }
        return c;
    };
    std::vector<Layout> bX, bY;
    std::vector<int> bin, bout;
    int64_t bw[2] = {0, 0};
    for (int combo = (1 << (std::max(K - 1, 0) + K)) - 1; combo >= 0; --combo) {
        const int choice = combo >> K;     // which transitions are absorbed at a store
        const int keep = combo & ((1 << K) - 1); // which in-place-capable stages run in place
        std::vector<Layout> x(K), y(K);
        for (int e = 0; e < K; ++e) {
            if (e == 0) x[e] = specs[e].in_natural;
            else {
                x[e] = y[e - 1];
            }
            if (e == K - 1) {
                y[e] = specs[e].out_natural;
            } else {
                const bool absorb_store = (choice >> e) & 1;
                y[e] = specs[e].out_natural;
                if (absorb_store) y[e].perm = specs[e + 1].in_natural.perm;
            }
        }
        // the last stage must produce the natural output; the first reads the natural input
        std::vector<int> ib(K), ob(K);
        ob[K - 1] = BUF_DST;
        bool okc = true;
        int64_t wb[2] = {0, 0};
        for (int e = K - 1; e >= 0; --e) {
            const bool same = x[e].same_order(y[e]) && specs[e].in_c == specs[e].out_c &&
                              x[e].bytes() == y[e].bytes() && specs[e].kind != SK_R2C &&
                              specs[e].kind != SK_C2R && ((keep >> e) & 1);
            if (e == 0) {
                ib[e] = BUF_SRC;
                if (ob[e] == BUF_DST && in_place && !same) okc = false;
                if (ob[e] == BUF_SRC) okc = false;
                break;
            }
            if (same) {
                ib[e] = ob[e];
            } else {
                const int64_t need = x[e].bytes();
                int pick = -1;
                if (ob[e] != BUF_DST && need <= dst_capacity) pick = BUF_DST;
                else if (ob[e] != BUF_W0) pick = BUF_W0;
                else pick = BUF_W1;
                ib[e] = pick;
            }
            if (ib[e] == BUF_W0) wb[0] = std::max(wb[0], x[e].bytes());
            if (ib[e] == BUF_W1) wb[1] = std::max(wb[1], x[e].bytes());
            ob[e - 1] = ib[e];
        }
        if (!okc) continue;
        // in place: stage 0 reads the data buffer, which is also DST
        if (in_place) {
            for (int e = 0; e < K; ++e) {
                if (ib[e] == BUF_SRC) ib[e] = BUF_DST;
                if (ob[e] == BUF_SRC) ob[e] = BUF_DST;
            }
            // a non in-place stage must not read and write the same buffer
            bool bad = false;
            for (int e = 0; e < K; ++e) {
                const bool same = x[e].same_order(y[e]) && specs[e].in_c == specs[e].out_c &&
                                  x[e].bytes() == y[e].bytes() && specs[e].kind != SK_R2C &&
                                  specs[e].kind != SK_C2R && ((keep >> e) & 1);
                if (ib[e] == ob[e] && !same) bad = true;
            }
            if (bad) continue;
        }
        const int nw = (wb[0] > 0 ? 1 : 0) + (wb[1] > 0 ? 1 : 0);
        const int pairs = pair_score(x, y);
        const int64_t cost = column_cost(x, y);
        if (pairs > best_pairs || (pairs == best_pairs && nw < best_w) ||
            (pairs == best_pairs && nw == best_w && cost < best_cost)) {
            best_pairs = pairs;
            best_w = nw;
            best_cost = cost;
            bX = x;
            bY = y;
            bin = ib;
            bout = ob;
            bw[0] = wb[0];
            bw[1] = wb[1];
        }
    }
    if (best_w == 99) return false;
    X = bX;
    Y = bY;
    inb = bin;
    outb = bout;
    wbytes[0] = bw[0];
    wbytes[1] = bw[1];
    return true;
}

template <typename T>
static pfh_status upload_tables(PlanImpl &P)
{
    const size_t nent = P.host_tables.size() / 2;
    P.table_entries = (int64_t)nent;
    P.tables_frozen = true;
    if (nent == 0) return PFH_OK;
    std::vector<T> tmp(P.host_tables.size());
    for (size_t i = 0; i < P.host_tables.size(); ++i) tmp[i] = (T)P.host_tables[i];
    PFH_HIP_CHECK(hipMalloc(&P.dev_tables, tmp.size() * sizeof(T)));
    PFH_HIP_CHECK(hipMemcpy(P.dev_tables, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    return PFH_OK;
}

static const int kStatusBytes = 256;

static int64_t align256(int64_t v) { return (v + 255) / 256 * 256; }

static int64_t status_area_bytes(const PlanImpl &P)
{
    return kStatusBytes + align256((int64_t)sizeof(EndpointPartial) * std::max<int64_t>(P.partials, 1));
}

// Builds the steps of one distributed-plan operation.
// Slab intermediate size: PFH_SLAB_MB (MiB), 0 (off) by default. Slabs of
// 16-128 MiB keep the intermediate in the 256 MB last-level cache, but on
// MI355X the slab launches ran 1.6-1.9x slower per byte than one launch per
// stage at 512^3 (390-438 against 713 round trips/s), so the pipelining is
// opt-in.
static int64_t slab_target_bytes(const Tuning &tu)
{
    static const int env_mb = env_int("PFH_SLAB_MB", 0);
    const int mb = tu.slab_mb >= 0 ? tu.slab_mb : env_mb;
    return (int64_t)mb << 20;
}

// The layout of complex data as real values: one more logical axis (the
// real / imaginary part) of extent 2 at the fastest memory position.
static Layout real_view(const Layout &L)
{
    Layout R = L;
    R.N = L.N + 1;
    R.perm[L.N] = L.N;
    R.shape[L.N] = 2;
    R.elem = L.elem / 2;
    return R;
}

static pfh_status build_op(PlanImpl &P, int op, bool in_place, OpPlan &out)
{
    std::vector<StageSpec> specs;
    pfh_status st = route_specs(P, op, specs);
    if (st) return st;
    const int tb = (P.precision == PFH_F32) ? 4 : 8;
    const int K = (int)specs.size();
    int64_t dst_cap = specs[K - 1].out_natural.bytes();
    if (in_place) dst_cap = std::max(dst_cap, specs[0].in_natural.bytes());
    std::vector<Layout> X, Y;
    std::vector<int> ib, ob;
    int64_t wb[2];
    // slab pipelining (out of place only): the pair (e, e+1) runs slab by
    // slab over an axis A that neither stage transforms and that is the
    // slowest memory axis of stage e's input and of the intermediate array
    const int64_t target = in_place ? 0 : slab_target_bytes(P.tuning);
    std::vector<char> checked(K, 0);
    for (int e = 0; e < K; ++e)
        checked[e] = ((P.kind == PFH_R2C || P.kind == PFH_C2R || P.kind == PFH_MIXED_R2C) &&
                      specs[e].kind == SK_C2R &&
                      (route_op(P.kind, op) == PFH_INVERSE || route_op(P.kind, op) == PFH_BACKWARD));
    auto slab_axis_of = [&](int e, const Layout &xin, const Layout &mid) -> int {
        if (target <= 0 || e + 1 >= K || checked[e] || checked[e + 1]) return -1;
        const int A = xin.perm[0];
        if (mid.perm[0] != A || A == specs[e].axis || A == specs[e + 1].axis) return -1;
        const int64_t nA = xin.shape[A];
        if (nA < 2 || mid.bytes() < 2 * target) return -1;
        return A;
    };
    auto pair_score = [&](const std::vector<Layout> &x, const std::vector<Layout> &y) {
        int n = 0;
        for (int e = 0; e + 1 < K;) {
            if (slab_axis_of(e, x[e], y[e]) >= 0) {
                ++n;
                e += 2;
            } else {
                ++e;
            }
        }
        return n;
    };
    // An out-of-place operation with an endpoint check must leave dst
    // untouched when the check rejects the spectrum (c2r.rs:62-64), so no
    // stage before the check may use dst as an intermediate, and every stage
    // from the check on reads the device gate.
    int e_chk = -1;
    for (int e = 0; e < K; ++e)
        if (checked[e] && e_chk < 0) e_chk = e;
    if (e_chk >= 0 && !in_place) dst_cap = 0;
    // threads per line of every stage (its radix plan, as the stage will be
    // planned), for the route's column-side choice
    std::vector<int> line_threads(K, 1);
    for (int e = 0; e < K; ++e) {
        std::vector<double> scratch;
        Step probe;
        const pfh_error_info saved = g_err;
        if (fill_transform(probe, specs[e], scratch, tb, P.tuning) == PFH_OK && !probe.long_only &&
            !wants_pieces(probe))
            line_threads[e] = threads_per_line(probe, tb);
        g_err = saved;
    }
    if (!assign_route(specs, in_place, dst_cap, pair_score, line_threads, X, Y, ib, ob, wb))
        return set_err(PFH_ERR_PREPARATION_FAILED, "no buffer assignment");
    out.steps.clear();
    for (int r = 0; r < 3; ++r) out.x_bytes[r] = 0;
    for (int e = 0; e < K; ++e) {
        Step s;
        st = fill_transform(s, specs[e], P.host_tables, tb, P.tuning);
        if (st) return st;
        s.in = X[e];
        s.out = Y[e];
        if (specs[e].cplx_r2r) {
            // complex data seen as real lines: the (re, im) part is one more,
            // fastest axis of extent 2, so the stage runs 2 real lines per
            // complex line with the same fused layout change
            s.in = real_view(X[e]);
            s.out = real_view(Y[e]);
            s.in_cplx = s.out_cplx = false;
        }
        s.in_buf = ib[e];
        s.out_buf = ob[e];
        st = finish_or_expand(P, s, tb, out.x_bytes);
        if (st) return st;
        if (checked[e]) {
            // acceptance check on the C2R stage's input (distributed/r2c.rs:4709-4807)
            const pfh_desc &d = P.desc;
            unsigned sel = d.selected ? d.selected : ((1u << P.N) - 1u);
            const int ra = specs[e].axis;
            const int64_t real_len = (int64_t)d.shape[ra];
            double depth = 1.0, transverse = 1.0;
            for (int a = 0; a < P.N; ++a) {
                if (a == ra || !(sel & (1u << a))) continue;
                if (d.shape[a] > 1) {
                    int64_t v = (int64_t)d.shape[a] - 1;
                    int bits = 0;
                    while (v) {
                        ++bits;
                        v >>= 1;
                    }
                    depth += bits;
                }
                transverse *= (double)d.shape[a];
            }
            if (P.kind == PFH_MIXED_R2C) {
                // MixedR2cPlan reverse endpoint contract (mixed.rs:409-427):
                // over the non-identity complex-suffix axes, depth adds
                // ceil_log2(E_a) + allowance_a and the raw absolute threshold
                // scales by factor_a
                depth = 1.0;
                transverse = 1.0;
                for (int a = 0; a < ra; ++a) {
                    const int c = d.axis_kind[a];
                    if (c == PFH_AXIS_NONE) continue;
                    const int64_t n = (int64_t)d.shape[a];
                    int64_t Ea, fa;
                    int allow = 2;
                    if (c == PFH_AXIS_FFT) {
                        Ea = n;
                        fa = n;
                        allow = 0;
                    } else if (c == PFH_DCT1) {
                        Ea = fa = 2 * (n - 1);
                    } else if (c == PFH_DST1) {
                        Ea = fa = 2 * (n + 1);
                    } else if (c == PFH_DCT4 || c == PFH_DST4) {
                        Ea = 8 * n;
                        fa = 2 * n;
                    } else if (c == PFH_DHT) {
                        Ea = fa = n;
                    } else {
                        Ea = 4 * n;
                        fa = 2 * n;
                    }
                    int bits = 0;
                    for (int64_t v = Ea - 1; Ea > 1 && v; v >>= 1) ++bits;
                    depth += bits + allow;
                    transverse *= (double)fa;
                }
            }
            const double eps = (tb == 4) ? (double)FLT_EPSILON : DBL_EPSILON;
            const double minsub = (tb == 4) ? (double)1.40129846e-45f : 4.9406564584124654e-324;
            EndpointArgs ep{};
            ep.extra = P.extra;
            ep.len = X[e].len();
            const auto strd = X[e].strides();
            ep.stride = strd[ra];
            ep.clen = X[e].shape[ra];
            ep.outer = ep.len / (ep.clen * ep.stride);
            ep.planes = (real_len % 2 == 0) ? 2 : 1;
            {
                // the plane elements are one per cache line when the reduction
                // axis is the fastest (the C2R input): every thread keeps its
                // kEndpointPerThread loads in flight at once
                const int64_t count = ep.outer * ep.stride;
                const int64_t per = 256 * (int64_t)kEndpointPerThread;
                int64_t nb = (count + per - 1) / per;
                if (nb < 1) nb = 1;
                ep.nblk = (int)nb;
                P.partials = std::max<int64_t>(P.partials, ep.extra * ep.planes * nb);
            }
            ep.relative = 128.0 * eps * depth;
            ep.absolute = 128.0 * minsub * depth * ((route_op(P.kind, op) == PFH_BACKWARD) ? transverse : 1.0);
            s.endpoint_check = true;
            s.ep = ep;
        }
        if (e_chk >= 0 && e > e_chk) s.gated = true;
        out.steps.push_back(s);
    }
    out.pairs.clear();
    out.t_bytes = 0;
    // slab-shaped copies of steps e (input X[e] full, output T) and e + 1
    // (input T, output Y[e + 1] full) for a slab of extent S along A
    auto slab_steps = [&](int e, int A, int64_t S, Step &f, Step &g2) -> pfh_status {
        const int64_t nA = X[e].shape[A];
        Layout T = Y[e];
        T.shape[A] = S;
        f = out.steps[e];
        f.out = specs[e].cplx_r2r ? real_view(T) : T;
        pfh_status r = finish_step(f, P.dev, tb, P.tuning);
        if (r) return r;
        bool found = false;
        for (int d = 0; d < f.g.nd; ++d)
            if (f.digit_axis[d] == A) {
                f.g.ext[d] = S;
                found = true;
            }
        if (!found) return PFH_ERR_UNSUPPORTED;
        f.nlines = f.nlines / nA * S;
        set_tiles(f, P.tuning);
        g2 = out.steps[e + 1];
        g2.in = specs[e + 1].cplx_r2r ? real_view(T) : T;
        return finish_step(g2, P.dev, tb, P.tuning);
    };
    for (int e = 0; e + 1 < K;) {
        const int A = slab_axis_of(e, X[e], Y[e]);
        const Step &a = out.steps[e], &b = out.steps[e + 1];
        if (A < 0 || a.in_buf == b.out_buf || !a.pieces.empty() || !b.pieces.empty()) {
            ++e;
            continue;
        }
        // slabs of about `target` intermediate bytes; when the slab axis
        // groups the lines of a column-access tile, the slab extent is a
        // multiple of the tile's lines so that no tile straddles two slabs
        const int64_t nA = X[e].shape[A];
        const int64_t slice = Y[e].bytes() / nA;
        const int64_t want = std::max<int64_t>(2, (nA * slice + target - 1) / target);
        int64_t q = 1;
        if (a.out_col && a.digit_axis[0] == A) q = std::max<int64_t>(q, a.lt);
        if ((b.in_col || b.out_col) && b.digit_axis[0] == A) q = std::max<int64_t>(q, b.lt);
        const int64_t St = (nA + want - 1) / want;
        const int64_t S = std::max<int64_t>(q, (St + q / 2) / q * q);
        const int64_t nfull = nA / S, tail = nA - nfull * S;
        if (nfull + (tail ? 1 : 0) < 2) {
            ++e;
            continue;
        }
        SlabPair pr;
        pr.e = e;
        pr.axis = A;
        pr.nslab = (int)nfull;
        pr.extent = S;
        pr.tail = tail;
        if (slab_steps(e, A, S, pr.first, pr.second) != PFH_OK ||
            (tail && slab_steps(e, A, tail, pr.first_tail, pr.second_tail) != PFH_OK)) {
            ++e;
            continue;
        }
        pr.in_plane_bytes = X[e].strides()[A] * X[e].elem;
        pr.out_plane_bytes = Y[e + 1].strides()[A] * Y[e + 1].elem;
        out.t_bytes = std::max(out.t_bytes, slice * S);
        out.pairs.push_back(pr);
        e += 2;
    }
    out.in = specs[0].in_natural;
    out.out = specs[K - 1].out_natural;
    out.w_bytes[0] = wb[0];
    out.w_bytes[1] = wb[1];
    out.valid = true;
    return PFH_OK;
}

// Local line plans: one stage on lines of a 1-D array with `batch` extras.
static pfh_status build_local_op(PlanImpl &P, int op, OpPlan &out)
{
    const int tb = (P.precision == PFH_F32) ? 4 : 8;
    StageSpec sp;
    sp.axis = 0;
    const int64_t n = (int64_t)P.line_len;
    sp.n = n;
    const bool fwd = (op == PFH_FORWARD || op == PFH_ADJOINT_INVERSE || op == PFH_ADJOINT_BACKWARD);
    std::array<int64_t, 8> shin{}, shout{};
    shin[0] = shout[0] = n;
    switch (P.kind) {
    case PFH_LOCAL_C2C:
        sp.kind = SK_C2C;
        sp.sign_pos = !fwd;
        if (op == PFH_INVERSE || op == PFH_ADJOINT_INVERSE) sp.scale = 1.0 / (double)n;
        sp.in_c = sp.out_c = true;
        break;
    case PFH_LOCAL_R2C:
        if (fwd) {
            sp.kind = SK_R2C;
            sp.in_c = false;
            sp.out_c = true;
            shout[0] = n / 2 + 1;
            if (op != PFH_FORWARD) sp.flags |= F_HERM_DOUBLE;
        } else {
            sp.kind = SK_C2R;
            sp.in_c = true;
            sp.out_c = false;
            shin[0] = n / 2 + 1;
            if (op == PFH_ADJOINT_FORWARD) sp.flags |= F_HERM_HALF;
        }
        if (op == PFH_INVERSE || op == PFH_ADJOINT_INVERSE) sp.scale = 1.0 / (double)n;
        break;
    case PFH_LOCAL_R2R:
        r2r_stage(sp, P.r2r_kind, n, op);
        break;
    case PFH_LOCAL_DHT:
        r2r_stage(sp, PFH_DHT, n, op);
        break;
    }
    std::array<int, 8> id{};
    Step s;
    pfh_status st = fill_transform(s, sp, P.host_tables, tb, P.tuning);
    if (st) return st;
    s.in = make_layout(1, id, shin, 1, (sp.in_c ? 2 : 1) * tb);
    s.out = make_layout(1, id, shout, 1, (sp.out_c ? 2 : 1) * tb);
    s.in_buf = BUF_SRC;
    s.out_buf = BUF_DST;
    for (int r = 0; r < 3; ++r) out.x_bytes[r] = 0;
    out.long_lines = 0;
    st = finish_or_expand(P, s, tb, out.x_bytes);
    if (st) return st;
    if (!s.pieces.empty()) {
        // long lines run in chunks of lines whose piece regions take about
        // 256 MiB of the workspace (at least one line): each chunk is a
        // handful of launches (838 lines of 10007 points, f32, on MI355X:
        // 0.72 ms in chunks of 828 lines, 0.81 ms in chunks of 207, with
        // M = 20250)
        const int64_t one = out.x_bytes[0] + out.x_bytes[1] + out.x_bytes[2];
        const int64_t chunk = std::max<int64_t>(1, ((int64_t)256 << 20) / std::max<int64_t>(one, 1));
        s.in.extra = s.out.extra = chunk;
        for (int r = 0; r < 3; ++r) out.x_bytes[r] = 0;
        st = step_geometry(s);
        if (!st) st = expand_long(P, s, tb, out.x_bytes);
        if (st) return st;
        s.in.extra = s.out.extra = 1;
        out.long_lines = chunk;
    }
    if (P.kind == PFH_LOCAL_R2C && !fwd && op != PFH_ADJOINT_FORWARD) s.local_endpoint_check = true;
    out.steps.assign(1, s);
    out.in = s.in;
    out.out = s.out;
    out.valid = true;
    return PFH_OK;
}

// ---------------------------------------------------------------------------
// plan creation
// ---------------------------------------------------------------------------

static bool valid_perm(int N, const int *p)
{
    bool seen[8] = {false};
    for (int i = 0; i < N; ++i) {
        if (p[i] < 0 || p[i] >= N || seen[p[i]]) return false;
        seen[p[i]] = true;
    }
    return true;
}

static std::atomic<uint64_t> g_next_tag{1};

static pfh_status create_impl(const pfh_desc *d, PlanImpl &P)
{
    P.tag = g_next_tag.fetch_add(1);
    P.desc = *d;
    P.kind = d->kind;
    P.precision = d->precision;
    if (P.precision != PFH_F32 && P.precision != PFH_F64) return set_err(PFH_ERR_INVALID_ARGUMENT, "precision");
    pfh_status st = device_info(P.dev);
    if (st) return st;
    P.extra = 1;
    if (d->extra_ndim < 0 || d->extra_ndim > PFH_MAX_DIMS) return set_err(PFH_ERR_INVALID_ARGUMENT, "extra_ndim");
    for (int i = 0; i < d->extra_ndim; ++i) P.extra *= (int64_t)d->extra_shape[i];
    switch (P.kind) {
    case PFH_LOCAL_C2C:
    case PFH_LOCAL_R2C:
    case PFH_LOCAL_R2R:
    case PFH_LOCAL_DHT: {
        P.N = 1;
        P.line_len = d->line_len;
        P.r2r_kind = d->r2r_kind;
        if (d->line_len == 0) return set_err(PFH_ERR_INVALID_LENGTH, "line length must be non-zero");
        if (P.kind == PFH_LOCAL_R2R) {
            if (d->r2r_kind < PFH_DCT1 || d->r2r_kind > PFH_DST4) return set_err(PFH_ERR_INVALID_ARGUMENT, "r2r kind");
            if (d->r2r_kind == PFH_DCT1 && d->line_len == 1)
                return set_err(PFH_ERR_INVALID_LENGTH, "DCT-I needs length >= 2");
        }
        if (d->line_len > (size_t)1 << 30) return set_err(PFH_ERR_LENGTH_OVERFLOW, "line length");
        for (int op = 0; op < 6; ++op) {
            st = build_local_op(P, op, P.ops[op]);
            if (st) return st;
        }
        break;
    }
    case PFH_C2C:
    case PFH_R2C:
    case PFH_R2R:
    case PFH_C2R:
    case PFH_MIXED_C2C:
    case PFH_MIXED_R2C: {
        const int N = d->ndim;
        const int M = d->topology_dims ? d->topology_dims : N - 1;
        if (N < 2 || N > PFH_MAX_DIMS || M < 1 || M >= N) return set_err(PFH_ERR_INVALID_DIMENSIONS, "N >= 2, 1 <= M < N");
        P.N = N;
        for (int a = 0; a < N; ++a)
            if (d->shape[a] == 0) return set_err(PFH_ERR_ZERO_EXTENT, "zero global extent", 0, 0, (size_t)a);
        if (P.kind == PFH_R2C || P.kind == PFH_C2R) {
            unsigned sel = d->selected ? d->selected : ((1u << N) - 1u);
            if ((sel & ((1u << N) - 1u)) == 0) return set_err(PFH_ERR_INVALID_DIMENSIONS, "empty selection");
        }
        if (P.kind == PFH_R2R) {
            for (int a = 0; a < N; ++a) {
                const int k = d->axis_kind[a];
                if (k < 0 || k > PFH_DHT) return set_err(PFH_ERR_INVALID_ARGUMENT, "axis kind");
                if (k == PFH_DCT1 && d->shape[a] == 1) return set_err(PFH_ERR_INVALID_LENGTH, "DCT-I needs length >= 2");
            }
        }
        if (P.kind == PFH_MIXED_C2C || P.kind == PFH_MIXED_R2C) {
            // validate_c2c_graph / validate_r2c_graph (mixed.rs:1592-1629) and
            // validate_r2c_directions (:1631-1643)
            int rfft = -1;
            for (int a = 0; a < N; ++a) {
                const int k = d->axis_kind[a];
                if (k < 0 || k > PFH_AXIS_RFFT) return set_err(PFH_ERR_INVALID_ARGUMENT, "axis transform");
                if (k == PFH_AXIS_RFFT) {
                    if (P.kind == PFH_MIXED_C2C || rfft >= 0)
                        return set_err(PFH_ERR_INVALID_GRAPH, "invalid mixed transform axis graph");
                    rfft = a;
                }
                if (k == PFH_DCT1 && d->shape[a] == 1) return set_err(PFH_ERR_INVALID_LENGTH, "DCT-I needs length >= 2");
                if (d->direction[a] != 0 && k != PFH_AXIS_FFT)
                    return set_err(PFH_ERR_PREPARATION_FAILED, "a backward direction needs an FFT axis");
            }
            if (P.kind == PFH_MIXED_R2C) {
                if (rfft < 0) return set_err(PFH_ERR_INVALID_GRAPH, "invalid mixed transform axis graph");
                for (int a = rfft + 1; a < N; ++a)
                    if (d->axis_kind[a] == PFH_AXIS_FFT)
                        return set_err(PFH_ERR_INVALID_GRAPH, "invalid mixed transform axis graph");
            }
        }
        for (int op = 0; op < 6; ++op) {
            st = build_op(P, op, false, P.ops[op]);
            if (st) return st;
            if (P.kind == PFH_C2R) continue; // C2rPlan is out of place only (c2r.rs:69-70)
            st = build_op(P, op, true, P.ops_inplace[op]);
            if (st) return st;
        }
        break;
    }
    case PFH_LOCAL_TRANSPOSE:
    case PFH_ALLTOALLV_TRANSPOSE: {
        const int N = d->ndim;
        if (N < 1 || N > PFH_MAX_DIMS) return set_err(PFH_ERR_INVALID_DIMENSIONS, "ndim");
        if (!valid_perm(N, d->src_perm) || !valid_perm(N, d->dst_perm))
            return set_err(PFH_ERR_INVALID_PERMUTATION, "invalid axis permutation");
        for (int a = 0; a < N; ++a)
            if (d->shape[a] == 0) return set_err(PFH_ERR_ZERO_EXTENT, "zero global extent", 0, 0, (size_t)a);
        if (d->element_bytes == 0) return set_err(PFH_ERR_INVALID_ARGUMENT, "element_bytes");
        P.N = N;
        const size_t eb = d->element_bytes;
        int w = 1;
        for (int c : {16, 8, 4, 2}) {
            if (eb % (size_t)c == 0) {
                w = c;
                break;
            }
        }
        P.word = w;
        P.words_per_elem = (int)(eb / (size_t)w);
        std::array<int, 8> sp{}, dp{};
        std::array<int64_t, 8> sh{};
        for (int a = 0; a < N; ++a) {
            sp[a] = d->src_perm[a];
            dp[a] = d->dst_perm[a];
            sh[a] = (int64_t)d->shape[a];
        }
        P.tr_src = make_layout(N, sp, sh, P.extra, (int)eb);
        P.tr_dst = make_layout(N, dp, sh, P.extra, (int)eb);
        // forward (src -> dst) and backward (dst -> src) as copy steps
        const bool tiled = (P.words_per_elem == 1 && (w == 4 || w == 8 || w == 16));
        for (int dir = 0; dir < 2; ++dir) {
            OpPlan &o = P.ops[dir == 0 ? PFH_FORWARD : PFH_BACKWARD];
            const Layout &a = dir == 0 ? P.tr_src : P.tr_dst;
            const Layout &b = dir == 0 ? P.tr_dst : P.tr_src;
            o.in = a;
            o.out = b;
            o.valid = true;
            o.steps.clear();
            if (tiled) {
                // word type: 4 -> f32 real, 8 -> f32 complex, 16 -> f64 complex
                Step s;
                StageSpec spc;
                spc.kind = SK_COPY;
                spc.axis = b.perm[N - 1];
                spc.n = b.shape[spc.axis];
                spc.in_c = spc.out_c = (w != 4);
                fill_transform(s, spc, P.host_tables, (w == 16) ? 8 : 4, P.tuning);
                s.in = a;
                s.out = b;
                const int tb = (w == 16) ? 8 : 4;
                s.in.elem = s.out.elem = w;
                s.in_buf = BUF_SRC;
                s.out_buf = BUF_DST;
                st = finish_step(s, P.dev, tb, P.tuning);
                if (st) return st;
                o.steps.push_back(s);
            }
        }
        P.precision = (w == 16) ? PFH_F64 : PFH_F32;
        break;
    }
    default:
        return set_err(PFH_ERR_INVALID_ARGUMENT, "unknown plan kind");
    }
    if (P.precision == PFH_F32) st = upload_tables<float>(P);
    else st = upload_tables<double>(P);
    return st;
}

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------

struct WsView {
    int *status = nullptr;
    int *gate = nullptr;
    unsigned long long *first_code = nullptr;
    EndpointPartial *ep = nullptr;
    void *w[2] = {nullptr, nullptr};
    void *t = nullptr; // slab intermediate
    void *x[3] = {nullptr, nullptr, nullptr}; // regions of the pieces of long lines
    // in-place calls: every stage reads the gate (set from the poison word by
    // the call's first kernel), and a rejected spectrum poisons the workspace
    int *poison = nullptr;
    bool gate_all = false;
};

static void status_view(void *ws, WsView &v)
{
    char *b = static_cast<char *>(ws);
    v.status = reinterpret_cast<int *>(b + kStatusOffStatus);
    v.gate = reinterpret_cast<int *>(b + kStatusOffGate);
    v.first_code = reinterpret_cast<unsigned long long *>(b + kStatusOffFirst);
}

// ---------------------------------------------------------------------------
// Workspace binding and in-place failure state.
//
// The reference binds a workspace to the plan that made it and rejects any
// other (Arc::ptr_eq, distributed.rs:2188-2191, :4015-4017), and marks an
// in-place array Poisoned once an operation fails after its first write
// (run_in_place_transaction, distributed.rs:3776-3794; many.rs:25). Here the
// workspace is caller memory: its status area holds the tag of the plan it
// is bound to (written at its first use, or by pfh_workspace_init) and the
// poison word. The host keeps what it last saw per workspace address, so the
// device tag is read (one synchronous 16-byte copy) only when a workspace is
// used with a plan other than the one it was last seen with.
// ---------------------------------------------------------------------------

struct WsEntry {
    uint64_t tag = 0;
    bool poisoned = false;
    bool pending = false; // an in-place call with a device check ran with deferred status
};
static std::mutex g_ws_mu;
static std::unordered_map<uintptr_t, WsEntry> g_ws;

static pfh_status bind_workspace(const PlanImpl &P, void *ws, hipStream_t stream)
{
    const uintptr_t key = reinterpret_cast<uintptr_t>(ws);
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_ws.find(key);
    if (it != g_ws.end() && it->second.tag == P.tag) return PFH_OK;
    PFH_HIP_CHECK(hipStreamSynchronize(stream));
    unsigned char area[32];
    PFH_HIP_CHECK(hipMemcpy(area, ws, sizeof(area), hipMemcpyDeviceToHost));
    uint64_t tag;
    int poison;
    std::memcpy(&tag, area + kStatusOffTag, sizeof(tag));
    std::memcpy(&poison, area + kStatusOffPoison, sizeof(poison));
    if (tag != 0 && tag != P.tag)
        return set_err(PFH_ERR_WORKSPACE_MISMATCH, "workspace is bound to another plan");
    if (tag == 0) {
        PFH_HIP_CHECK(launch_fill_u64(reinterpret_cast<unsigned long long *>(static_cast<char *>(ws) + kStatusOffTag),
                                      1, (unsigned long long)P.tag, stream));
    }
    WsEntry e;
    e.tag = P.tag;
    e.poisoned = poison != 0;
    g_ws[key] = e;
    return PFH_OK;
}

static bool ws_poisoned(const void *ws)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_ws.find(reinterpret_cast<uintptr_t>(ws));
    return it != g_ws.end() && it->second.poisoned;
}

static void ws_set_poisoned(const void *ws, bool v)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    g_ws[reinterpret_cast<uintptr_t>(ws)].poisoned = v;
}

static bool ws_pending(const void *ws)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_ws.find(reinterpret_cast<uintptr_t>(ws));
    return it != g_ws.end() && it->second.pending;
}

static void ws_set_pending(const void *ws, bool v)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_ws.find(reinterpret_cast<uintptr_t>(ws));
    if (it != g_ws.end()) it->second.pending = v;
}

// [a, a + na) and [b, b + nb) share a byte
static bool overlaps(const void *a, size_t na, const void *b, size_t nb)
{
    if (!a || !b || na == 0 || nb == 0) return false;
    const uintptr_t x = reinterpret_cast<uintptr_t>(a), y = reinterpret_cast<uintptr_t>(b);
    return x < y + nb && y < x + na;
}

// Workspace of one operation: status area, W0, W1, slab intermediate T.
static int64_t op_ws_bytes(const PlanImpl &P, const OpPlan &o)
{
    return status_area_bytes(P) + align256(o.w_bytes[0]) + align256(o.w_bytes[1]) + align256(o.t_bytes) +
           align256(o.x_bytes[0]) + align256(o.x_bytes[1]) + align256(o.x_bytes[2]);
}

static pfh_status view_workspace(const PlanImpl &P, const OpPlan &o, void *ws, size_t ws_bytes, WsView &v)
{
    const int64_t need = op_ws_bytes(P, o);
    if (ws == nullptr || (int64_t)ws_bytes < need)
        return set_err(PFH_ERR_WORKSPACE_TOO_SMALL, "workspace too small", (size_t)need, ws ? ws_bytes : 0);
    char *b = static_cast<char *>(ws);
    status_view(ws, v);
    v.ep = reinterpret_cast<EndpointPartial *>(b + kStatusBytes);
    char *w = b + status_area_bytes(P);
    v.w[0] = w;
    v.w[1] = w + align256(o.w_bytes[0]);
    v.t = w + align256(o.w_bytes[0]) + align256(o.w_bytes[1]);
    char *x = static_cast<char *>(v.t) + align256(o.t_bytes);
    for (int r = 0; r < 3; ++r) {
        v.x[r] = o.x_bytes[r] > 0 ? x : nullptr;
        x += align256(o.x_bytes[r]);
    }
    return PFH_OK;
}

template <typename T>
static StageArgs<T> fill_args(const PlanImpl &P, const Step &s, void *const bufs[BUF_COUNT], const int *gate)
{
    StageArgs<T> a{};
    a.src = bufs[s.in_buf];
    a.dst = bufs[s.out_buf];
    a.tw = reinterpret_cast<const cpx<T> *>(P.dev_tables) + s.tw_base;
    a.gate = gate;
    a.nlines = s.nlines;
    a.ntiles = s.tiles;
    a.g = s.g;
    a.ips = s.ips;
    a.ops = s.ops;
    a.mode = s.mode;
    a.flags = s.flags;
    // paired column accesses need 2-element alignment of the array bases
    if ((a.flags & F_CPAIR_IN) && reinterpret_cast<uintptr_t>(a.src) % (2 * sizeof(T)) != 0) a.flags &= ~F_CPAIR_IN;
    if ((a.flags & F_CPAIR_OUT) && reinterpret_cast<uintptr_t>(a.dst) % (2 * sizeof(T)) != 0) a.flags &= ~F_CPAIR_OUT;
    if ((a.flags & F_RPAIR_IN) && reinterpret_cast<uintptr_t>(a.src) % (2 * sizeof(T)) != 0) a.flags &= ~F_RPAIR_IN;
    if ((a.flags & F_RPAIR_OUT) && reinterpret_cast<uintptr_t>(a.dst) % (2 * sizeof(T)) != 0) a.flags &= ~F_RPAIR_OUT;
    a.n = s.n;
    a.nin = s.nin;
    a.nout = s.nout;
    a.m = s.m;
    a.npass = (int)s.radices.size();
    int Ns = 1;
    for (int p = 0; p < a.npass; ++p) {
        a.radix[p] = s.radices[p];
        a.twoff[p] = s.twoff[p];
        const FastDiv f = make_fastdiv((uint32_t)Ns);
        a.nsmag[p] = f.mag;
        a.nssh[p] = f.sh;
        Ns *= s.radices[p];
    }
    for (int d = 0; d < a.g.nd; ++d) {
        const FastDiv f = make_fastdiv((uint32_t)a.g.ext[d]);
        a.g.dmag[d] = f.mag;
        a.g.dsh[d] = f.sh;
    }
    a.pre0 = s.pre0;
    a.pre1 = s.pre1;
    a.post0 = s.post0;
    a.post1 = s.post1;
    a.twn = s.twn;
    a.fem = s.fast ? s.fem : 0;
    a.lt = s.lt;
    a.lc = s.lc;
    a.tpl = s.tpl;
    a.ls = s.ls;
    a.in_col = s.in_col;
    a.out_col = s.out_col;
    a.xcd_per = s.xcd_per;
    a.scale = (T)s.scale;
    a.prew0 = (T)s.prew0;
    a.prew1 = (T)s.prew1;
    a.postw0 = (T)s.postw0;
    a.postw1 = (T)s.postw1;
    const cpx<T> *tabs = reinterpret_cast<const cpx<T> *>(P.dev_tables);
    a.xtab_in = s.xtab_in >= 0 ? tabs + s.xtab_in : nullptr;
    a.xtab_out = s.xtab_out >= 0 ? tabs + s.xtab_out : nullptr;
    a.xtw = s.xtw >= 0 ? tabs + s.xtw : nullptr;
    a.xdig = s.xdig;
    a.xmul_in = s.xmul_in;
    a.xmul_out = s.xmul_out;
    a.xvalid_in = s.xvalid_in;
    a.xvalid_out = s.xvalid_out;
    a.xtw_m = s.xtw_m;
    a.xtw_bsh = s.xtw_bsh;
    return a;
}

// Arguments of a long real line's prep (launch 1) or post (launch 2) piece.
template <typename T>
static LongArgs<T> fill_long_args(const PlanImpl &P, const Step &s, const void *in, void *out, const int *gate)
{
    LongArgs<T> a{};
    a.src = in;
    a.dst = out;
    a.tab = reinterpret_cast<const cpx<T> *>(P.dev_tables);
    a.gate = gate;
    a.nlines = s.nlines;
    a.g = s.g;
    for (int d = 0; d < a.g.nd; ++d) {
        const FastDiv f = make_fastdiv((uint32_t)a.g.ext[d]);
        a.g.dmag[d] = f.mag;
        a.g.dsh[d] = f.sh;
    }
    a.ips = s.ips;
    a.ops = s.ops;
    a.kind = s.long_kind;
    a.flags = s.flags;
    a.n = s.n;
    a.m = s.m;
    a.nout = s.nout;
    a.pre0 = s.pre0;
    a.post0 = s.post0;
    a.scale = (T)s.scale;
    a.prew0 = (T)s.prew0;
    a.prew1 = (T)s.prew1;
    a.postw0 = (T)s.postw0;
    a.postw1 = (T)s.postw1;
    return a;
}

// The C2R boundary check of a step; in-place calls keep the gate that their
// first kernel set and poison the workspace on a rejected spectrum.
template <typename T>
static hipError_t launch_check(const Step &s, const void *in, const WsView &wv, hipStream_t stream)
{
    EndpointArgs e = s.ep;
    e.keep_gate = wv.gate_all ? 1 : 0;
    e.poison = wv.poison;
    return launch_endpoint_check<T>(e, in, wv.ep, wv.gate, wv.status, stream);
}

template <typename T>
static pfh_status launch_step_ptrs(const PlanImpl &P, const Step &s, const void *in, void *out, const WsView &wv,
                                   hipStream_t stream)
{
    if (!s.pieces.empty()) return set_err(PFH_ERR_PREPARATION_FAILED, "a long-line stage runs through its pieces");
    if (s.endpoint_check) {
        PFH_HIP_CHECK(launch_check<T>(s, in, wv, stream));
    }
    if (s.local_endpoint_check) {
        PFH_HIP_CHECK(launch_local_endpoint_check<T>(in, s.nlines, s.nin, (s.n % 2 == 0) ? 1 : 0, wv.first_code,
                                                     wv.gate, wv.status, stream));
    }
    void *io[BUF_COUNT] = {const_cast<void *>(in), out};
    Step sp = s;
    sp.in_buf = 0;
    sp.out_buf = 1;
    const bool gated = s.endpoint_check || s.local_endpoint_check || s.gated || wv.gate_all;
    const StageArgs<T> a = fill_args<T>(P, sp, io, gated ? wv.gate : nullptr);
    if (s.fast) {
        PFH_HIP_CHECK(launch_line_fast<T>(a, s.fr[0], s.fr[1], s.fr[2], s.fem, s.threads, s.lds_bytes, s.grid, stream));
        return PFH_OK;
    }
    bool wide = false;
    for (int R : s.radices) wide = wide || is_wide_radix(R);
    PFH_HIP_CHECK(launch_line_stage<T>(a, wide, s.threads, s.lds_bytes, s.grid, stream));
    return PFH_OK;
}

template <typename T>
static pfh_status launch_step(const PlanImpl &P, const Step &s, void *const bufs[BUF_COUNT], const WsView &wv,
                              hipStream_t stream)
{
    return launch_step_ptrs<T>(P, s, bufs[s.in_buf], bufs[s.out_buf], wv, stream);
}

// A stage that runs as pieces: its boundary check on the stage's input,
// then every piece in order, all reading the gate when the stage does.
template <typename T>
static pfh_status launch_pieces(const PlanImpl &P, const Step &s, void *const bufs[BUF_COUNT], const WsView &wv,
                                hipStream_t stream)
{
    const void *in = bufs[s.in_buf];
    if (s.endpoint_check) {
        PFH_HIP_CHECK(launch_check<T>(s, in, wv, stream));
    }
    if (s.local_endpoint_check) {
        PFH_HIP_CHECK(launch_local_endpoint_check<T>(in, s.nlines, s.nin, (s.n % 2 == 0) ? 1 : 0, wv.first_code,
                                                     wv.gate, wv.status, stream));
    }
    const bool gated = s.endpoint_check || s.local_endpoint_check || s.gated || wv.gate_all;
    const int *gate = gated ? wv.gate : nullptr;
    for (const Step &pc : s.pieces) {
        const void *pin = bufs[pc.in_buf];
        void *pout = bufs[pc.out_buf];
        if (!pin || !pout) return set_err(PFH_ERR_PREPARATION_FAILED, "long-line region missing");
        if (pc.launch == 1) {
            PFH_HIP_CHECK(launch_long_prep<T>(fill_long_args<T>(P, pc, pin, pout, gate), stream));
        } else if (pc.launch == 2) {
            PFH_HIP_CHECK(launch_long_post<T>(fill_long_args<T>(P, pc, pin, pout, gate), stream));
        } else {
            void *io[BUF_COUNT] = {const_cast<void *>(pin), pout};
            Step sp = pc;
            sp.in_buf = 0;
            sp.out_buf = 1;
            const StageArgs<T> a = fill_args<T>(P, sp, io, gate);
            bool wide = false;
            for (int R : pc.radices) wide = wide || is_wide_radix(R);
            PFH_HIP_CHECK(launch_line_stage<T>(a, wide, pc.threads, pc.lds_bytes, pc.grid, stream));
        }
    }
    return PFH_OK;
}

static pfh_status launch_any(const PlanImpl &P, const Step &s, const void *in, void *out, const WsView &wv,
                             hipStream_t stream)
{
    return (P.precision == PFH_F32) ? launch_step_ptrs<float>(P, s, in, out, wv, stream)
                                    : launch_step_ptrs<double>(P, s, in, out, wv, stream);
}

static pfh_status launch_any_pieces(const PlanImpl &P, const Step &s, void *const bufs[BUF_COUNT], const WsView &wv,
                                    hipStream_t stream)
{
    return (P.precision == PFH_F32) ? launch_pieces<float>(P, s, bufs, wv, stream)
                                    : launch_pieces<double>(P, s, bufs, wv, stream);
}

// Device timing of an execution: events recorded between launches; each
// entry (one stage, or one endpoint check) sums its launch intervals.
struct TimingTrace {
    struct Entry {
        int axis = -1;
        double bytes = 0;
        std::vector<int> iv; // interval i spans ev[i - 1] .. ev[i]
    };
    std::vector<hipEvent_t> ev; // ev[0] recorded before the first launch
    std::vector<Entry> entries;
    pfh_status mark(hipStream_t stream, int &idx)
    {
        hipEvent_t e;
        PFH_HIP_CHECK(hipEventCreate(&e));
        ev.push_back(e);
        PFH_HIP_CHECK(hipEventRecord(e, stream));
        idx = (int)ev.size() - 1;
        return PFH_OK;
    }
};

static double step_bytes(const Step &s)
{
    return (double)s.nlines * ((double)s.nin * s.in.elem + (double)s.nout * s.out.elem);
}

static double step_bytes(const Step &s);

static pfh_status run_op(const PlanImpl &P, const OpPlan &o, void *const bufs_in[4], const WsView &wv,
                         hipStream_t stream, TimingTrace *tr)
{
    void *bufs[BUF_COUNT] = {bufs_in[0], bufs_in[1], bufs_in[2], bufs_in[3], wv.x[0], wv.x[1], wv.x[2]};
    size_t pi = 0;
    for (size_t i = 0; i < o.steps.size(); ++i) {
        const Step &s = o.steps[i];
        int idx = 0;
        if (!s.pieces.empty()) {
            pfh_status st = launch_any_pieces(P, s, bufs, wv, stream);
            if (st) return st;
            if (tr) {
                if (tr->mark(stream, idx)) return PFH_ERR_DEVICE;
                TimingTrace::Entry es;
                es.axis = s.axis;
                es.bytes = step_bytes(s);
                es.iv.push_back(idx);
                tr->entries.push_back(es);
            }
            continue;
        }
        if (pi < o.pairs.size() && o.pairs[pi].e == (int)i) {
            // slab-pipelined pair: steps i and i + 1 through the slab region
            const SlabPair &pr = o.pairs[pi++];
            const char *in0 = static_cast<const char *>(bufs[pr.first.in_buf]);
            char *out0 = static_cast<char *>(bufs[pr.second.out_buf]);
            TimingTrace::Entry ea, eb;
            ea.axis = pr.first.axis;
            eb.axis = pr.second.axis;
            const int ns = pr.nslab + (pr.tail ? 1 : 0);
            for (int k = 0; k < ns; ++k) {
                const int64_t a0 = (int64_t)k * pr.extent; // first plane of the slab
                const bool part = k == pr.nslab;
                const Step &f = part ? pr.first_tail : pr.first;
                const Step &g2 = part ? pr.second_tail : pr.second;
                pfh_status st = launch_any(P, f, in0 + a0 * pr.in_plane_bytes, wv.t, wv, stream);
                if (st) return st;
                if (tr) {
                    if (tr->mark(stream, idx)) return PFH_ERR_DEVICE;
                    ea.iv.push_back(idx);
                    ea.bytes += step_bytes(f);
                }
                st = launch_any(P, g2, wv.t, out0 + a0 * pr.out_plane_bytes, wv, stream);
                if (st) return st;
                if (tr) {
                    if (tr->mark(stream, idx)) return PFH_ERR_DEVICE;
                    eb.iv.push_back(idx);
                    eb.bytes += step_bytes(g2);
                }
            }
            if (tr) {
                tr->entries.push_back(ea);
                tr->entries.push_back(eb);
            }
            ++i;
            continue;
        }
        if (tr && (s.endpoint_check || s.local_endpoint_check)) {
            // time the check on its own, then the stage with its gate
            Step chk = s;
            chk.gated = true;
            if (P.precision == PFH_F32) {
                if (s.endpoint_check) PFH_HIP_CHECK(launch_check<float>(s, bufs[s.in_buf], wv, stream));
            } else {
                if (s.endpoint_check) PFH_HIP_CHECK(launch_check<double>(s, bufs[s.in_buf], wv, stream));
            }
            if (tr->mark(stream, idx)) return PFH_ERR_DEVICE;
            TimingTrace::Entry ec;
            ec.axis = -1;
            ec.bytes = (double)(s.ep.extra * s.ep.planes * s.ep.outer * s.ep.stride) * (double)s.in.elem;
            ec.iv.push_back(idx);
            tr->entries.push_back(ec);
            chk.endpoint_check = false;
            pfh_status st = launch_any(P, chk, bufs[s.in_buf], bufs[s.out_buf], wv, stream);
            if (st) return st;
        } else {
            pfh_status st = launch_any(P, s, bufs[s.in_buf], bufs[s.out_buf], wv, stream);
            if (st) return st;
        }
        if (tr) {
            if (tr->mark(stream, idx)) return PFH_ERR_DEVICE;
            TimingTrace::Entry es;
            es.axis = s.axis;
            es.bytes = step_bytes(s);
            es.iv.push_back(idx);
            tr->entries.push_back(es);
        }
    }
    return PFH_OK;
}

static bool op_has_device_check(const OpPlan &o)
{
    for (const Step &s : o.steps)
        if (s.endpoint_check || s.local_endpoint_check) return true;
    return false;
}

// Reads (and clears) the sticky status after a synchronous execution, and
// records a poisoned workspace (the poison word itself stays set).
static pfh_status collect_status(const PlanImpl &P, const WsView &wv, hipStream_t stream)
{
    PFH_HIP_CHECK(hipStreamSynchronize(stream));
    unsigned char area[32];
    unsigned long long code = ~0ull;
    PFH_HIP_CHECK(hipMemcpy(area, wv.status, sizeof(area), hipMemcpyDeviceToHost));
    int status, poison;
    std::memcpy(&status, area + kStatusOffStatus, sizeof(status));
    std::memcpy(&poison, area + kStatusOffPoison, sizeof(poison));
    std::memcpy(&code, area + kStatusOffFirst, sizeof(code));
    if (status) {
        PFH_HIP_CHECK(launch_fill_i32(wv.status, 1, 0, stream));
        PFH_HIP_CHECK(hipStreamSynchronize(stream));
    }
    if (poison) ws_set_poisoned(wv.status, true);
    ws_set_pending(wv.status, false);
    if (status & PFH_DEVICE_INVALID_ENDPOINT) {
        const size_t batch = (size_t)(code / 2);
        const size_t index = (code & 1) ? (size_t)(P.line_len / 2) : 0;
        return set_err(PFH_ERR_INVALID_SPECTRUM_ENDPOINT, "non-zero imaginary endpoint", 0, 0, batch, index);
    }
    if (status & PFH_DEVICE_INVALID_SPECTRUM)
        return set_err(PFH_ERR_INVALID_SPECTRUM, (status & PFH_DEVICE_POISONED)
                                                     ? "inverse spectrum has an invalid constrained boundary plane; "
                                                       "later in-place calls on the poisoned workspace wrote nothing"
                                                     : "inverse spectrum has an invalid constrained boundary plane");
    if (status & PFH_DEVICE_POISONED)
        return set_err(PFH_ERR_POISONED, "an earlier in-place call failed after it started: the call wrote nothing");
    return ok();
}

static pfh_status execute_impl(PlanImpl &P, int op, const void *src, size_t src_elems, void *dst, size_t dst_elems,
                               void *ws, size_t ws_bytes, hipStream_t stream, int flags, pfh_transform_timing *timing,
                               bool bind = true)
{
    if (op < 0 || op > 5) return set_err(PFH_ERR_INVALID_ARGUMENT, "op");
    const bool is_tr = (P.kind == PFH_LOCAL_TRANSPOSE || P.kind == PFH_ALLTOALLV_TRANSPOSE);
    if (is_tr && op != PFH_FORWARD && op != PFH_BACKWARD) return set_err(PFH_ERR_INVALID_ARGUMENT, "transpose op");
    const OpPlan &o = P.ops[op];
    if (!o.valid) return set_err(PFH_ERR_INVALID_ARGUMENT, "op not supported by this plan");
    // an empty extra shape gives empty arrays: nothing to read or write
    if ((!src && o.in.total() > 0) || (!dst && o.out.total() > 0))
        return set_err(PFH_ERR_INVALID_ARGUMENT, "null buffer");
    if ((int64_t)src_elems != o.in.total())
        return set_err(is_tr ? PFH_ERR_SOURCE_LAYOUT_MISMATCH : PFH_ERR_INPUT_LAYOUT_MISMATCH, "source length",
                       (size_t)o.in.total(), src_elems);
    if ((int64_t)dst_elems != o.out.total())
        return set_err(is_tr ? PFH_ERR_DESTINATION_LAYOUT_MISMATCH : PFH_ERR_OUTPUT_LAYOUT_MISMATCH, "destination length",
                       (size_t)o.out.total(), dst_elems);
    if (o.in.total() == 0 && o.out.total() == 0) return ok();
    if (is_tr && overlaps(src, (size_t)o.in.bytes(), dst, (size_t)o.out.bytes()))
        return set_err(PFH_ERR_ALIASED_BUFFERS, "source and destination overlap");
    if (is_tr && o.steps.empty()) {
        // generic word permutation
        PermuteArgs pa{};
        const Layout &a = o.in, &b = o.out;
        const int N = a.N;
        const auto sst = a.strides();
        pa.nd = N + 1;
        pa.len = a.len() * P.words_per_elem;
        pa.extra = a.extra;
        for (int p = 0; p < N; ++p) {
            pa.dext[p] = b.shape[b.perm[p]];
            pa.sstride[p] = sst[b.perm[p]] * P.words_per_elem;
        }
        pa.dext[N] = P.words_per_elem;
        pa.sstride[N] = 1;
        if (P.word == 1 || P.word == 2) {
            PFH_HIP_CHECK(launch_permute_words(pa, src, dst, P.word, stream));
        } else {
            // wider words: permute as 2-byte words with the word split into a trailing axis
            const int k = P.word / 2;
            pa.len *= k;
            for (int p = 0; p <= N; ++p) pa.sstride[p] *= k;
            pa.nd = N + 2;
            pa.dext[N + 1] = k;
            pa.sstride[N + 1] = 1;
            PFH_HIP_CHECK(launch_permute_words(pa, src, dst, 2, stream));
        }
        if (timing) PFH_HIP_CHECK(hipStreamSynchronize(stream));
        return ok();
    }
    WsView wv;
    const bool needs_ws = !is_tr;
    if (needs_ws) {
        pfh_status st = view_workspace(P, o, ws, ws_bytes, wv);
        if (st) return st;
    }
    // Rust's borrow rules keep source, destination and workspace apart; here
    // overlapping buffers are refused before any work
    if (overlaps(src, (size_t)o.in.bytes(), dst, (size_t)o.out.bytes()))
        return set_err(PFH_ERR_ALIASED_BUFFERS, "source and destination overlap");
    if (needs_ws) {
        const size_t wb = (size_t)op_ws_bytes(P, o);
        if (overlaps(src, (size_t)o.in.bytes(), ws, wb))
            return set_err(PFH_ERR_ALIASED_BUFFERS, "source and workspace overlap");
        if (overlaps(dst, (size_t)o.out.bytes(), ws, wb))
            return set_err(PFH_ERR_ALIASED_BUFFERS, "destination and workspace overlap");
        if (bind) {
            pfh_status st = bind_workspace(P, ws, stream);
            if (st) return st;
        }
    }
    void *bufs[4] = {const_cast<void *>(src), dst, wv.w[0], wv.w[1]};
    TimingTrace tr;
    if (timing) {
        int idx = 0;
        pfh_status e0 = tr.mark(stream, idx);
        if (e0) return e0;
    }
    pfh_status st = run_op(P, o, bufs, wv, stream, timing ? &tr : nullptr);
    if (st) {
        for (auto &e : tr.ev) (void)hipEventDestroy(e);
        return st;
    }
    if (timing) {
        PFH_HIP_CHECK(hipEventSynchronize(tr.ev.back()));
        const int ne = (int)tr.entries.size();
        timing->stages = ne < 12 ? ne : 12;
        timing->total_ms = 0;
        for (int i = 0; i < timing->stages; ++i) {
            const TimingTrace::Entry &en = tr.entries[i];
            double ms = 0;
            for (int j : en.iv) {
                float f = 0;
                PFH_HIP_CHECK(hipEventElapsedTime(&f, tr.ev[j - 1], tr.ev[j]));
                ms += f;
            }
            timing->stage_ms[i] = ms;
            timing->stage_axis[i] = en.axis;
            timing->stage_bytes[i] = en.bytes;
            timing->stage_launches[i] = (int)en.iv.size();
            timing->total_ms += ms;
        }
        for (auto &e : tr.ev) (void)hipEventDestroy(e);
    }
    if (needs_ws && op_has_device_check(o) && !(flags & PFH_EXEC_DEFER_STATUS)) return collect_status(P, wv, stream);
    return ok();
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char *pfh_version(void) { return "pencil-fft-hip 0.1.0 (gfx950)"; }

void pfh_last_error(pfh_error_info *out)
{
    if (out) *out = g_err;
}

const char *pfh_status_string(pfh_status s)
{
    switch (s) {
    case PFH_OK: return "ok";
    case PFH_ERR_INVALID_LENGTH: return "FFT line length must be non-zero";
    case PFH_ERR_LENGTH_OVERFLOW: return "FFT length or derived length cannot be represented";
    case PFH_ERR_NON_INTEGRAL_BATCH: return "buffer length is not an integral number of FFT lines";
    case PFH_ERR_BUFFER_LENGTH_MISMATCH: return "source and destination lengths differ";
    case PFH_ERR_BATCH_COUNT_MISMATCH: return "source and destination batch counts differ";
    case PFH_ERR_SCRATCH_TOO_SMALL: return "scratch is too small";
    case PFH_ERR_INVALID_SPECTRUM_ENDPOINT: return "inverse spectrum has a non-zero imaginary endpoint";
    case PFH_ERR_INVALID_DIMENSIONS: return "distributed FFT requires N >= 2 and 1 <= M < N";
    case PFH_ERR_INVALID_INPUT_LAYOUT: return "input pencil must use identity permutation";
    case PFH_ERR_INPUT_LAYOUT_MISMATCH: return "source layout does not match the plan";
    case PFH_ERR_OUTPUT_LAYOUT_MISMATCH: return "destination layout does not match the plan";
    case PFH_ERR_EXTRA_SHAPE_MISMATCH: return "extra shape does not match the plan";
    case PFH_ERR_WORKSPACE_MISMATCH: return "workspace does not belong to this plan";
    case PFH_ERR_WORKSPACE_TOO_SMALL: return "workspace is too small";
    case PFH_ERR_PREPARATION_FAILED: return "preparation failed";
    case PFH_ERR_ALLOCATION_FAILED: return "allocation failed";
    case PFH_ERR_INVALID_SPECTRUM: return "inverse spectrum has an invalid constrained boundary plane";
    case PFH_ERR_INCOMPATIBLE_DISTRIBUTION: return "source and destination distributions are incompatible";
    case PFH_ERR_SOURCE_LAYOUT_MISMATCH: return "source view does not match the plan";
    case PFH_ERR_DESTINATION_LAYOUT_MISMATCH: return "destination view does not match the plan";
    case PFH_ERR_INVALID_PERMUTATION: return "invalid axis permutation";
    case PFH_ERR_ZERO_EXTENT: return "zero global extent";
    case PFH_ERR_INVALID_ARGUMENT: return "invalid argument";
    case PFH_ERR_UNSUPPORTED: return "length not supported by the device kernels";
    case PFH_ERR_DEVICE: return "HIP runtime error";
    case PFH_ERR_INVALID_GRAPH: return "invalid mixed transform axis graph";
    case PFH_ERR_INCOMPATIBLE_PENCILS: return "pointwise arrays must have identical spatial layouts";
    case PFH_ERR_EXTRA_RANK_MISMATCH: return "pointwise extra rank mismatch";
    case PFH_ERR_EXTRA_EXTENT_MISMATCH: return "pointwise extra extent mismatch";
    case PFH_ERR_EMPTY_INPUTS: return "pointwise many operations require at least one input";
    case PFH_ERR_IO_INVALID_INPUT: return "invalid I/O argument";
    case PFH_ERR_IO_INVALID_PATH: return "path cannot be represented as a native path";
    case PFH_ERR_IO_SIZE_LIMIT: return "I/O size or native count limit exceeded";
    case PFH_ERR_IO_METADATA_MISMATCH: return "file metadata mismatch";
    case PFH_ERR_IO_INVALID_FILE: return "invalid I/O file";
    case PFH_ERR_IO_INCOMPLETE_FILE: return "I/O file has no completed commit marker";
    case PFH_ERR_IO_NATIVE: return "native backend operation failed";
    case PFH_ERR_IO_WRITE_INCOMPLETE: return "write is incomplete; the failed file was preserved";
    case PFH_ERR_IO_COMMIT_UNCERTAIN: return "write commit state is uncertain; the failed file was preserved";
    case PFH_ERR_IO_INVALID_NAME: return "invalid dataset name";
    case PFH_ERR_IO_DUPLICATE_NAME: return "dataset name already exists";
    case PFH_ERR_IO_NOT_FOUND: return "dataset name was not found";
    case PFH_ERR_IO_INVALID_TAIL: return "dataset container has an invalid tail";
    case PFH_ERR_POISONED: return "array is poisoned by an incomplete in-place operation";
    case PFH_ERR_ALIASED_BUFFERS: return "source, destination and workspace must not overlap";
    }
    return "unknown status";
}

pfh_status pfh_plan_create(const pfh_desc *desc, pfh_plan *out)
{
    if (!desc || !out) return set_err(PFH_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    pfh_plan p = new (std::nothrow) pfh_plan_s;
    if (!p) return set_err(PFH_ERR_ALLOCATION_FAILED, "plan");
    pfh_status st = create_impl(desc, p->impl);
    if (st) {
        delete p;
        return st;
    }
    *out = p;
    return ok();
}

void pfh_plan_destroy(pfh_plan plan) { delete plan; }

pfh_status pfh_plan_workspace_bytes(pfh_plan plan, int op, size_t *bytes)
{
    if (!plan || !bytes || op < 0 || op > 5) return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    const PlanImpl &P = plan->impl;
    if (P.kind == PFH_LOCAL_TRANSPOSE || P.kind == PFH_ALLTOALLV_TRANSPOSE) {
        *bytes = (size_t)align256(P.tr_src.bytes());
        return ok();
    }
    const OpPlan &o = P.ops[op];
    const OpPlan &oi = P.ops_inplace[op];
    int64_t b = op_ws_bytes(P, o);
    if (oi.valid) b = std::max(b, op_ws_bytes(P, oi));
    if (P.kind == PFH_LOCAL_R2C || P.kind == PFH_LOCAL_C2C || P.kind == PFH_LOCAL_R2R || P.kind == PFH_LOCAL_DHT) {
        // in-place local execution stages the input here
        (void)0;
    }
    *bytes = (size_t)b;
    return ok();
}

pfh_status pfh_plan_layout(pfh_plan plan, int side, int perm[PFH_MAX_DIMS], size_t shape[PFH_MAX_DIMS],
                           size_t *elements, size_t *element_bytes)
{
    if (!plan || (side != 0 && side != 1)) return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    const PlanImpl &P = plan->impl;
    const OpPlan &o = P.ops[PFH_FORWARD];
    const Layout &L = side == 0 ? o.in : o.out;
    for (int a = 0; a < PFH_MAX_DIMS; ++a) {
        if (perm) perm[a] = a < L.N ? L.perm[a] : -1;
        if (shape) shape[a] = a < L.N ? (size_t)L.shape[a] : 0;
    }
    if (elements) *elements = (size_t)L.total();
    if (element_bytes) *element_bytes = (size_t)L.elem;
    return ok();
}

pfh_status pfh_execute(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst, size_t dst_elems,
                       void *workspace, size_t workspace_bytes, pfh_stream_t stream, int flags)
{
    if (!plan) return set_err(PFH_ERR_INVALID_ARGUMENT, "null plan");
    PlanImpl &P = plan->impl;
    if (P.kind <= PFH_LOCAL_DHT)
        return pfh_local_execute(plan, op, src, src_elems, dst, dst_elems, workspace, workspace_bytes, stream, flags);
    return execute_impl(P, op, src, src_elems, dst, dst_elems, workspace, workspace_bytes,
                        reinterpret_cast<hipStream_t>(stream), flags, nullptr);
}

pfh_status pfh_execute_with_timing(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst,
                                   size_t dst_elems, void *workspace, size_t workspace_bytes, pfh_stream_t stream,
                                   pfh_transform_timing *timing)
{
    if (!plan || !timing) return set_err(PFH_ERR_INVALID_ARGUMENT, "null argument");
    PlanImpl &P = plan->impl;
    if (P.kind <= PFH_LOCAL_DHT) return set_err(PFH_ERR_INVALID_ARGUMENT, "timing is for distributed plans");
    return execute_impl(P, op, src, src_elems, dst, dst_elems, workspace, workspace_bytes,
                        reinterpret_cast<hipStream_t>(stream), 0, timing);
}

pfh_status pfh_execute_in_place(pfh_plan plan, int op, void *data, size_t capacity_bytes, void *workspace,
                                size_t workspace_bytes, pfh_stream_t stream_, int flags)
{
    if (!plan) return set_err(PFH_ERR_INVALID_ARGUMENT, "null argument");
    if (op < 0 || op > 5) return set_err(PFH_ERR_INVALID_ARGUMENT, "op");
    PlanImpl &P = plan->impl;
    if (!data && P.extra != 0) return set_err(PFH_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
    if (P.kind == PFH_LOCAL_TRANSPOSE || P.kind == PFH_ALLTOALLV_TRANSPOSE) {
        // stage the source into the workspace, then permute back (local_transpose.rs:159-243)
        const OpPlan &o = P.ops[op == PFH_BACKWARD ? PFH_BACKWARD : PFH_FORWARD];
        const size_t need = (size_t)o.in.bytes();
        if ((int64_t)capacity_bytes < o.in.bytes())
            return set_err(PFH_ERR_SOURCE_LAYOUT_MISMATCH, "data too small", need, capacity_bytes);
        if (!workspace || workspace_bytes < need)
            return set_err(PFH_ERR_SCRATCH_TOO_SMALL, "scratch too small", need, workspace ? workspace_bytes : 0);
        if (overlaps(workspace, need, data, need))
            return set_err(PFH_ERR_ALIASED_BUFFERS, "data and scratch overlap");
        PFH_HIP_CHECK(hipMemcpyAsync(workspace, data, need, hipMemcpyDeviceToDevice, stream));
        return execute_impl(P, op == PFH_BACKWARD ? PFH_BACKWARD : PFH_FORWARD, workspace, (size_t)o.in.total(), data,
                            (size_t)o.out.total(), nullptr, 0, stream, flags, nullptr);
    }
    if (P.kind <= PFH_LOCAL_DHT) return set_err(PFH_ERR_INVALID_ARGUMENT, "local plans: use pfh_local_execute_in_place");
    if (P.kind == PFH_C2R) return set_err(PFH_ERR_INVALID_ARGUMENT, "a C2R plan is out of place only");
    const OpPlan &o = P.ops_inplace[op];
    if (!o.valid) return set_err(PFH_ERR_INVALID_ARGUMENT, "op");
    const int64_t need = std::max(o.in.bytes(), o.out.bytes());
    if ((int64_t)capacity_bytes < need)
        return set_err(PFH_ERR_WORKSPACE_TOO_SMALL, "in-place storage", (size_t)need, capacity_bytes);
    if (need == 0) return ok();
    WsView wv;
    pfh_status st = view_workspace(P, o, workspace, workspace_bytes, wv);
    if (st) return st;
    if (overlaps(data, (size_t)need, workspace, (size_t)op_ws_bytes(P, o)))
        return set_err(PFH_ERR_ALIASED_BUFFERS, "data and workspace overlap");
    st = bind_workspace(P, workspace, stream);
    if (st) return st;
    // ArrayError::Poisoned (many.rs:234): an earlier in-place call on this
    // workspace failed after it started writing its array
    if (ws_poisoned(workspace))
        return set_err(PFH_ERR_POISONED, "an earlier in-place call on this workspace failed after it started");
    // a failure found on the device by an earlier deferred call is not known
    // here yet: the call's first kernel gates every stage on the poison word
    wv.poison = reinterpret_cast<int *>(static_cast<char *>(workspace) + kStatusOffPoison);
    wv.gate_all = true;
    PFH_HIP_CHECK(launch_inplace_begin(wv.gate, wv.poison, wv.status, stream));
    void *bufs[4] = {data, data, wv.w[0], wv.w[1]};
    st = run_op(P, o, bufs, wv, stream, nullptr);
    if (st) {
        // a launch failed after the call started writing the array
        (void)launch_fill_i32(wv.poison, 1, 1, stream);
        ws_set_poisoned(workspace, true);
        return st;
    }
    // the status is read when this call checks the spectrum on the device, or
    // when an earlier deferred in-place call may have poisoned the workspace
    // (this call's stages were then gated off)
    const bool check = op_has_device_check(o);
    if (flags & PFH_EXEC_DEFER_STATUS) {
        if (check) ws_set_pending(workspace, true);
        return ok();
    }
    if (check || ws_pending(workspace)) return collect_status(P, wv, stream);
    return ok();
}

} // extern "C"

namespace pfh {
// Out-of-place execution of a distributed plan on a workspace the library
// owns (the allocation-free entry point's static scratch): no plan binding,
// the status area is cleared by a fill kernel before the operation, device
// errors are read synchronously. ws_bytes is the scratch capacity.
pfh_status execute_on_scratch(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst, size_t dst_elems,
                              void *ws, size_t ws_bytes, hipStream_t stream)
{
    if (!plan) return set_err(PFH_ERR_INVALID_ARGUMENT, "null plan");
    if (op < 0 || op > 5) return set_err(PFH_ERR_INVALID_ARGUMENT, "op");
    PlanImpl &P = plan->impl;
    if (P.kind <= PFH_LOCAL_DHT || P.kind == PFH_LOCAL_TRANSPOSE || P.kind == PFH_ALLTOALLV_TRANSPOSE)
        return set_err(PFH_ERR_INVALID_ARGUMENT, "the static-scratch entry point runs distributed plans only");
    const OpPlan &o = P.ops[op];
    if (!o.valid) return set_err(PFH_ERR_INVALID_ARGUMENT, "op not supported by this plan");
    const int64_t need = op_ws_bytes(P, o);
    if (need > (int64_t)ws_bytes)
        return set_err(PFH_ERR_WORKSPACE_TOO_SMALL,
                       "the operation's workspace exceeds the library's static scratch (required > actual bytes)",
                       (size_t)need, ws_bytes);
    PFH_HIP_CHECK(launch_fill_i32(static_cast<int *>(ws), kStatusBytes / 4, 0, stream));
    return execute_impl(P, op, src, src_elems, dst, dst_elems, ws, ws_bytes, stream, 0, nullptr, false);
}
} // namespace pfh

extern "C" {

pfh_status pfh_workspace_init(pfh_plan plan, void *workspace, size_t workspace_bytes, pfh_stream_t stream_)
{
    if (!plan || !workspace) return set_err(PFH_ERR_INVALID_ARGUMENT, "null argument");
    PlanImpl &P = plan->impl;
    if (P.kind <= PFH_LOCAL_DHT || P.kind == PFH_LOCAL_TRANSPOSE || P.kind == PFH_ALLTOALLV_TRANSPOSE)
        return set_err(PFH_ERR_INVALID_ARGUMENT, "local and transpose plans have no plan-bound workspace");
    if (workspace_bytes < (size_t)kStatusBytes)
        return set_err(PFH_ERR_WORKSPACE_TOO_SMALL, "workspace too small", (size_t)kStatusBytes, workspace_bytes);
    hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
    PFH_HIP_CHECK(launch_fill_i32(static_cast<int *>(workspace), kStatusBytes / 4, 0, stream));
    PFH_HIP_CHECK(launch_fill_u64(reinterpret_cast<unsigned long long *>(static_cast<char *>(workspace) + kStatusOffTag),
                                  1, (unsigned long long)P.tag, stream));
    std::lock_guard<std::mutex> lk(g_ws_mu);
    WsEntry e;
    e.tag = P.tag;
    g_ws[reinterpret_cast<uintptr_t>(workspace)] = e;
    return ok();
}

static pfh_status local_execute_impl(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst,
                                     size_t dst_elems, void *workspace, size_t workspace_bytes,
                                     pfh_stream_t stream_, int flags, bool in_place);

pfh_status pfh_local_execute(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst,
                             size_t dst_elems, void *workspace, size_t workspace_bytes, pfh_stream_t stream_,
                             int flags)
{
    return local_execute_impl(plan, op, src, src_elems, dst, dst_elems, workspace, workspace_bytes, stream_, flags,
                              false);
}

// in_place: src == dst is the caller's in-place array (every workgroup reads
// its lines into LDS before it writes them); any other overlap is refused
static pfh_status local_execute_impl(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst,
                                     size_t dst_elems, void *workspace, size_t workspace_bytes,
                                     pfh_stream_t stream_, int flags, bool in_place)
{
    if (!plan) return set_err(PFH_ERR_INVALID_ARGUMENT, "null plan");
    if (op < 0 || op > 5) return set_err(PFH_ERR_INVALID_ARGUMENT, "op");
    PlanImpl &P = plan->impl;
    if (P.kind > PFH_LOCAL_DHT) return set_err(PFH_ERR_INVALID_ARGUMENT, "not a local plan");
    hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
    const OpPlan &tmpl = P.ops[op];
    const int64_t lin = tmpl.in.shape[0], lout = tmpl.out.shape[0];
    int64_t batch;
    if (P.kind == PFH_LOCAL_R2C) {
        // validate_batches (r2c.rs:571-590): source, then destination, then equality
        if ((int64_t)src_elems % lin) return set_err(PFH_ERR_NON_INTEGRAL_BATCH, "source");
        if ((int64_t)dst_elems % lout) return set_err(PFH_ERR_NON_INTEGRAL_BATCH, "destination");
        const int64_t sc = (int64_t)src_elems / lin, dc = (int64_t)dst_elems / lout;
        if (sc != dc)
            return set_err(PFH_ERR_BATCH_COUNT_MISMATCH, "batch counts differ", 0, 0, (size_t)sc, (size_t)dc);
        batch = sc;
    } else {
        if (src_elems != dst_elems) return set_err(PFH_ERR_BUFFER_LENGTH_MISMATCH, "lengths differ");
        if ((int64_t)src_elems % lin) return set_err(PFH_ERR_NON_INTEGRAL_BATCH, "batch");
        batch = (int64_t)src_elems / lin;
    }
    const bool c2r_check = tmpl.steps[0].local_endpoint_check;
    const bool long_lines = tmpl.long_lines > 0;
    // workspace: the status area (C2R) and the piece regions of long lines
    const int64_t ws_need = long_lines ? op_ws_bytes(P, tmpl) : (c2r_check ? status_area_bytes(P) : 0);
    WsView wv;
    if (ws_need > 0) {
        if (!workspace || (int64_t)workspace_bytes < ws_need)
            return set_err(PFH_ERR_SCRATCH_TOO_SMALL, long_lines ? "long-line workspace" : "status area",
                           (size_t)ws_need, workspace ? workspace_bytes : 0);
        char *b = static_cast<char *>(workspace);
        wv.status = reinterpret_cast<int *>(b);
        wv.gate = reinterpret_cast<int *>(b + 4);
        wv.first_code = reinterpret_cast<unsigned long long *>(b + 8);
        if (long_lines) {
            WsView full;
            pfh_status vs = view_workspace(P, tmpl, workspace, workspace_bytes, full);
            if (vs) return vs;
            for (int r = 0; r < 3; ++r) wv.x[r] = full.x[r];
        }
    }
    if (batch == 0) return ok();
    if (!src || !dst) return set_err(PFH_ERR_INVALID_ARGUMENT, "null buffer");
    {
        const size_t sb = (size_t)(batch * lin * tmpl.in.elem), db = (size_t)(batch * lout * tmpl.out.elem);
        if (!(in_place && src == dst) && overlaps(src, sb, dst, db))
            return set_err(PFH_ERR_ALIASED_BUFFERS, "source and destination overlap");
        if (ws_need > 0 && (overlaps(src, sb, workspace, (size_t)ws_need) || overlaps(dst, db, workspace, (size_t)ws_need)))
            return set_err(PFH_ERR_ALIASED_BUFFERS, "buffers and workspace overlap");
    }
    const int tb = (P.precision == PFH_F32) ? 4 : 8;
    if (long_lines) {
        // the C2R boundary check covers the whole batch before any chunk
        // writes; then chunks of long_lines lines run their pieces
        const Step &s0 = tmpl.steps[0];
        if (c2r_check) {
            if (P.precision == PFH_F32) {
                PFH_HIP_CHECK(launch_local_endpoint_check<float>(src, batch, s0.nin, (s0.n % 2 == 0) ? 1 : 0,
                                                                 wv.first_code, wv.gate, wv.status, stream));
            } else {
                PFH_HIP_CHECK(launch_local_endpoint_check<double>(src, batch, s0.nin, (s0.n % 2 == 0) ? 1 : 0,
                                                                  wv.first_code, wv.gate, wv.status, stream));
            }
        }
        for (int64_t c0 = 0; c0 < batch; c0 += tmpl.long_lines) {
            const int64_t cnt = std::min<int64_t>(tmpl.long_lines, batch - c0);
            // the chunk's pieces are planned once per (op, cnt): planning them
            // (factorizations, the Bluestein length search, piece sizing)
            // took about 0.2 ms of host time per call for 10007-point lines
            std::shared_ptr<const OpPlan> oc;
            {
                std::lock_guard<std::mutex> lk(P.chunks->mu);
                auto it = P.chunks->plans.find({op, cnt});
                if (it != P.chunks->plans.end()) {
                    oc = it->second;
                } else {
                    auto o = std::make_shared<OpPlan>(tmpl);
                    Step &s = o->steps[0];
                    s.in.extra = s.out.extra = cnt;
                    s.local_endpoint_check = false;
                    s.gated = c2r_check;
                    int64_t xb[3] = {0, 0, 0};
                    pfh_status st = step_geometry(s);
                    if (!st) st = expand_long(P, s, tb, xb);
                    if (st) return st;
                    for (int r = 0; r < 3; ++r)
                        if (xb[r] > tmpl.x_bytes[r])
                            return set_err(PFH_ERR_PREPARATION_FAILED, "long-line chunk regions");
                    P.chunks->plans[{op, cnt}] = o;
                    oc = o;
                }
            }
            const char *cs = static_cast<const char *>(src) + c0 * lin * tmpl.in.elem;
            char *cd = static_cast<char *>(dst) + c0 * lout * tmpl.out.elem;
            void *bufs[4] = {const_cast<char *>(cs), cd, nullptr, nullptr};
            pfh_status st = run_op(P, *oc, bufs, wv, stream, nullptr);
            if (st) return st;
        }
        if (c2r_check && !(flags & PFH_EXEC_DEFER_STATUS)) return collect_status(P, wv, stream);
        return ok();
    }
    OpPlan o = tmpl;
    Step &s = o.steps[0];
    s.in.extra = s.out.extra = batch;
    s.flags &= ~F_XCD_REMAP;
    pfh_status st = finish_step(s, P.dev, tb, P.tuning);
    if (st) return st;
    void *bufs[4] = {const_cast<void *>(src), dst, nullptr, nullptr};
    st = run_op(P, o, bufs, wv, stream, nullptr);
    if (st) return st;
    if (c2r_check && !(flags & PFH_EXEC_DEFER_STATUS)) return collect_status(P, wv, stream);
    return ok();
}

pfh_status pfh_local_execute_in_place(pfh_plan plan, int op, void *data, size_t elems, size_t capacity_bytes,
                                      void *workspace, size_t workspace_bytes, pfh_stream_t stream_, int flags)
{
    if (!plan) return set_err(PFH_ERR_INVALID_ARGUMENT, "null plan");
    if (op < 0 || op > 5) return set_err(PFH_ERR_INVALID_ARGUMENT, "op");
    PlanImpl &P = plan->impl;
    if (P.kind > PFH_LOCAL_DHT) return set_err(PFH_ERR_INVALID_ARGUMENT, "not a local plan");
    hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
    const OpPlan &tmpl = P.ops[op];
    const int64_t lin = tmpl.in.shape[0], lout = tmpl.out.shape[0];
    if ((int64_t)elems % lin) return set_err(PFH_ERR_NON_INTEGRAL_BATCH, "batch");
    const int64_t batch = (int64_t)elems / lin;
    const int64_t in_bytes = batch * lin * tmpl.in.elem, out_bytes = batch * lout * tmpl.out.elem;
    const int64_t need = std::max(in_bytes, out_bytes);
    if ((int64_t)capacity_bytes < need)
        return set_err(PFH_ERR_ALLOCATION_FAILED, "in-place storage too small", (size_t)need, capacity_bytes);
    if (P.kind != PFH_LOCAL_R2C) {
        // same layout on both sides: every workgroup reads its lines into LDS before writing them
        return local_execute_impl(plan, op, data, elems, data, elems, workspace, workspace_bytes, stream_, flags,
                                  true);
    }
    // packed real / half-complex storage: stage the input through the
    // workspace, after the status area and the piece regions of long lines
    const int64_t base = tmpl.long_lines > 0 ? op_ws_bytes(P, tmpl) : status_area_bytes(P);
    const int64_t ws_need = base + align256(in_bytes);
    if (!workspace || (int64_t)workspace_bytes < ws_need)
        return set_err(PFH_ERR_SCRATCH_TOO_SMALL, "workspace too small", (size_t)ws_need,
                       workspace ? workspace_bytes : 0);
    if (batch == 0) return ok();
    char *stage = static_cast<char *>(workspace) + base;
    if (overlaps(data, (size_t)need, workspace, (size_t)ws_need))
        return set_err(PFH_ERR_ALIASED_BUFFERS, "data and workspace overlap");
    PFH_HIP_CHECK(hipMemcpyAsync(stage, data, (size_t)in_bytes, hipMemcpyDeviceToDevice, stream));
    return pfh_local_execute(plan, op, stage, (size_t)(batch * lin), data, (size_t)(batch * lout), workspace,
                             workspace_bytes, stream_, flags);
}

pfh_status pfh_workspace_take_status(pfh_plan plan, void *workspace, pfh_stream_t stream)
{
    if (!plan || !workspace) return set_err(PFH_ERR_INVALID_ARGUMENT, "null argument");
    WsView wv;
    status_view(workspace, wv);
    return collect_status(plan->impl, wv, reinterpret_cast<hipStream_t>(stream));
}

} // extern "C"

// ---------------------------------------------------------------------------
// Batched uneven-size transforms. Each problem keeps its own plan (tables,
// route, buffer assignment). Route stage s of every problem becomes one item
// of a work list; one persistent launch per stage walks all (item, tile)
// pairs, so a batch is a handful of launches instead of one per stage and
// problem. The launches can be captured once into a hipGraph and replayed.
// ---------------------------------------------------------------------------

// One persistent launch: the items of one route stage that share a kernel
// (a compile-time radix plan, or the generic kernel's narrow / wide radix set).
struct BatchStage {
    void *d_items = nullptr;     // StageArgs<T>[nitems] (PlaneArgs<T>[nitems] for a plane stage)
    int64_t *d_prefix = nullptr; // [nitems + 1]
    int nitems = 0;
    int64_t total = 0;
    size_t lds = 0; // largest (table + lines) footprint of an item
    int threads = 64;
    int64_t grid = 0;
    bool plane = false;            // k_plane_batch: fused two-axis plane stages
    int tw_cap = 0, work_cap = 0, raw_cap = 0; // plane stages: LDS regions (complex values)
    bool multi = false;            // k_line_multi_batch over items of several compile-time plans
    int chain = 0;                 // 1: problems whose first pass is a plane stage (second stream)
    bool fast = false, wide = false; // else one compile-time plan, or the generic kernel
    int fr[3] = {0, 0, 0};
    int fem = 0;
};

// A fused plane stage of one batched problem (host side of PlaneArgs).
struct PlaneDesc {
    int in_buf = BUF_SRC, out_buf = BUF_DST;
    int64_t nplanes = 0;
    int nr = 0, nc = 0;
    int rr[3] = {0, 0, 0}, rem = 0, cr[3] = {0, 0, 0}, cem = 0;
    int tpl_r = 1, tpl_c = 1, rs = 1;
    int flags = 0;
    double scale = 1.0;
    int64_t tw_base_row = 0, tw_base_col = 0;
    int twn_row = 0, twn_col = 0;
    int twoff_row[3] = {0, 0, 0}, twoff_col[3] = {0, 0, 0};
    size_t lds_bytes = 0;
};

// One launch item of a batched problem's operation: a line step or a fused
// plane stage.
struct BatchItem {
    bool plane = false;
    Step step;
    PlaneDesc pd;
};

struct pfh_batch_s {
    int count = 0;
    int precision = PFH_F32;
    int threads = 256;
    std::vector<std::unique_ptr<PlanImpl>> plans;
    // per problem, per op (FORWARD, INVERSE, BACKWARD): the launch items in order
    std::vector<std::array<std::vector<BatchItem>, 3>> items;
    Tuning tuning;
    std::vector<int64_t> ws_off; // per problem
    int64_t ws_bytes = 0;
    int tune_source = PFH_TUNED_DEFAULT;
    std::vector<BatchStage> stages;
    // pointers the uploaded work list was built for
    int cur_op = -1;
    std::vector<const void *> cur_src;
    std::vector<void *> cur_dst;
    void *cur_ws = nullptr;
    hipGraphExec_t graph = nullptr;
    hipStream_t graph_stream = nullptr;
    // graphs requested on the legacy null stream run on this stream, ordered
    // after / before the caller's work by events
    hipStream_t own_stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // two-chain execution: chain 1 runs on side_stream between ev_fork / ev_join
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    ~pfh_batch_s() { release(); }
    void release_stages()
    {
        for (auto &s : stages) {
            if (s.d_items) (void)hipFree(s.d_items);
            if (s.d_prefix) (void)hipFree(s.d_prefix);
        }
        stages.clear();
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        cur_op = -1;
    }
    void release()
    {
        release_stages();
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (own_stream) (void)hipStreamDestroy(own_stream);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side_stream) (void)hipStreamDestroy(side_stream);
        ev_in = ev_out = ev_fork = ev_join = nullptr;
        own_stream = side_stream = nullptr;
    }
};

// Threads per line of a compile-time plan: every pass's butterflies covered
// by MB = EM / R of them per thread.
static int plan_tpl(int M, const int r[3], int em)
{
    int need = 1;
    for (int p = 0; p < 3; ++p) {
        if (!r[p]) continue;
        const int mb = std::max(1, em / r[p]);
        need = std::max(need, ceil_div(M / r[p], mb));
    }
    return need;
}

// The launch items of one operation of a batched problem. A 3D f32 C2C
// problem runs two consecutive route stages as one plane stage when both
// lengths have a plane plan and the plane fits in LDS: two passes over HBM
// instead of three; every FFT line is computed by the same radix plan as in
// the route. The pair is the inner one (axes 1 and 0, planes (x, y) per z)
// or the outer one (axes 2 and 1, planes (y, z) per x), whichever has the
// smaller plane. The remaining line stage writes (forward inner, inverse
// outer) or reads (forward outer, inverse inner) an intermediate whose
// planes are the plane stage's output or input: memory order (z, x, y) for
// the inner pair, (x, z, y) for the outer one (the plane stage reads rows of
// its first axis and writes columns of its second axis contiguously).
static pfh_status batch_items(const PlanImpl &P, int op, const Tuning &tu, std::vector<BatchItem> &out)
{
    const OpPlan &o = P.ops[op];
    out.clear();
    for (const Step &s : o.steps) {
        BatchItem it;
        it.step = s;
        out.push_back(it);
    }
    if (!tu.batch_plane || P.precision != PFH_F32 || P.N != 3 || o.steps.size() != 3) return PFH_OK;
    static const int plane_env = env_int("PFH_BATCH_PLANE", 1);
    if (!plane_env) return PFH_OK;
    const bool fwd = op == PFH_FORWARD;
    // the plan's outer layouts: [extra][x][y][z] and [extra][z][y][x]
    const Layout &xyz = fwd ? o.in : o.out, &zyx = fwd ? o.out : o.in;
    if (!(xyz.perm[0] == 0 && xyz.perm[1] == 1 && xyz.perm[2] == 2)) return PFH_OK;
    if (!(zyx.perm[0] == 2 && zyx.perm[1] == 1 && zyx.perm[2] == 0)) return PFH_OK;
    for (const Step &s : o.steps)
        if (s.mode != M_C2C) return PFH_OK;
    auto plannable = [](const Step &s) {
        return s.fast && plane_plan_available(4, s.fr[0], s.fr[1], s.fr[2], s.fem);
    };
    const int conj = F_CONJ_IN | F_CONJ_OUT;
    struct Option {
        bool ok = false;
        int a = 0, b = 0, l = 0; // step indexes: plane rows, plane columns, line stage
        bool line_first = false;
        int mid0 = 0, mid1 = 0, mid2 = 0; // intermediate memory order
        int64_t nplanes = 0;
        PlaneDesc pd;
    };
    auto make = [&](int a, int b, int l, bool line_first, int m0, int m1, int m2, int64_t nplanes) {
        Option op_;
        const Step &A = o.steps[a], &Bs = o.steps[b];
        if (!plannable(A) || !plannable(Bs)) return op_;
        const int fa = A.flags & conj, fb = Bs.flags & conj;
        if (fa != fb || (fa != 0 && fa != conj)) return op_; // one sign on both axes
        const int nr = A.m, nc = Bs.m;
        if (nr % 2 != 0) return op_; // the prefetch moves pairs of values
        PlaneDesc &pd = op_.pd;
        pd.nr = nr;
        pd.nc = nc;
        for (int r = 0; r < 3; ++r) {
            pd.rr[r] = A.fr[r];
            pd.cr[r] = Bs.fr[r];
            pd.twoff_row[r] = A.twoff[r];
            pd.twoff_col[r] = Bs.twoff[r];
        }
        pd.rem = A.fem;
        pd.cem = Bs.fem;
        pd.tpl_r = plan_tpl(nr, pd.rr, pd.rem);
        pd.tpl_c = plan_tpl(nc, pd.cr, pd.cem);
        if (pd.tpl_r > kPlaneThreads / 4 || pd.tpl_c > kPlaneThreads / 4) return op_; // any workgroup size
        int rs = (nr - 1) + ((nr - 1) >> 4) + 1; // padded row (one slot per 16 values)
        if (rs % 2 == 0) ++rs;
        pd.rs = rs;
        pd.flags = fa;
        pd.scale = A.scale * Bs.scale;
        pd.tw_base_row = A.tw_base;
        pd.tw_base_col = Bs.tw_base;
        pd.twn_row = A.twn;
        pd.twn_col = Bs.twn;
        pd.lds_bytes = (size_t)(A.twn + Bs.twn + (int64_t)nc * rs) * 8;
        if (pd.lds_bytes + 1024 > P.dev.lds_limit) return op_;
        pd.nplanes = nplanes;
        op_.ok = true;
        op_.a = a;
        op_.b = b;
        op_.l = l;
        op_.line_first = line_first;
        op_.mid0 = m0;
        op_.mid1 = m1;
        op_.mid2 = m2;
        op_.nplanes = nplanes;
        return op_;
    };
    // route steps: forward transforms axes 2, 1, 0; inverse / backward 0, 1, 2
    const int64_t E = xyz.extra, X = xyz.shape[0], Z = xyz.shape[2];
    const Option inner = fwd ? make(1, 2, 0, true, 2, 0, 1, E * Z) : make(0, 1, 2, false, 2, 0, 1, E * Z);
    const Option outer = fwd ? make(0, 1, 2, false, 0, 2, 1, E * X) : make(1, 2, 0, true, 0, 2, 1, E * X);
    const Option *pick = nullptr;
    if (inner.ok) pick = &inner;
    if (outer.ok && (!pick || outer.pd.lds_bytes < inner.pd.lds_bytes)) pick = &outer;
    if (!pick) return PFH_OK;
    Layout mid = xyz;
    mid.perm[0] = pick->mid0;
    mid.perm[1] = pick->mid1;
    mid.perm[2] = pick->mid2;
    if (o.w_bytes[0] < mid.bytes()) return PFH_OK;
    Step line = o.steps[pick->l];
    PlaneDesc pd = pick->pd;
    if (pick->line_first) {
        line.out = mid;
        line.in_buf = BUF_SRC;
        line.out_buf = BUF_W0;
        pd.in_buf = BUF_W0;
        pd.out_buf = BUF_DST;
    } else {
        line.in = mid;
        line.in_buf = BUF_W0;
        line.out_buf = BUF_DST;
        pd.in_buf = BUF_SRC;
        pd.out_buf = BUF_W0;
    }
    pfh_status st = finish_step(line, P.dev, 4, tu);
    if (st) return st;
    BatchItem li, pi;
    li.step = line;
    pi.plane = true;
    pi.pd = pd;
    out.clear();
    if (pick->line_first) {
        out.push_back(li);
        out.push_back(pi);
    } else {
        out.push_back(pi);
        out.push_back(li);
    }
    return PFH_OK;
}

static pfh_status batch_plan(pfh_batch_s &B, const pfh_desc *descs, int count, const Tuning &tu)
{
    B.plans.clear();
    B.items.clear();
    B.release_stages();
    B.tuning = tu;
    B.ws_off.assign(count, 0);
    int64_t off = 0;
    int threads = 64;
    for (int i = 0; i < count; ++i) {
        if (descs[i].kind != PFH_C2C) return set_err(PFH_ERR_INVALID_ARGUMENT, "batch problems must be C2C plans");
        if (descs[i].precision != descs[0].precision) return set_err(PFH_ERR_INVALID_ARGUMENT, "mixed precision");
        std::unique_ptr<PlanImpl> P(new (std::nothrow) PlanImpl);
        if (!P) return set_err(PFH_ERR_ALLOCATION_FAILED, "batch plan");
        P->tuning = tu;
        pfh_status st = create_impl(&descs[i], *P);
        if (st) return st;
        int64_t need = 0;
        for (int op : {PFH_FORWARD, PFH_INVERSE, PFH_BACKWARD}) {
            const OpPlan &o = P->ops[op];
            for (const Step &s : o.steps) {
                if (s.mode != M_C2C || s.endpoint_check || s.local_endpoint_check)
                    return set_err(PFH_ERR_UNSUPPORTED, "batched problems must select every axis");
                if (!s.pieces.empty())
                    return set_err(PFH_ERR_UNSUPPORTED, "batched problems with lines longer than one workgroup");
                threads = std::max(threads, s.threads);
            }
            need = std::max(need, op_ws_bytes(*P, o));
        }
        std::array<std::vector<BatchItem>, 3> its;
        for (int op : {PFH_FORWARD, PFH_INVERSE, PFH_BACKWARD}) {
            st = batch_items(*P, op, tu, its[op]);
            if (st) return st;
        }
        B.items.push_back(std::move(its));
        B.ws_off[i] = off;
        off += align256(need);
        B.plans.push_back(std::move(P));
    }
    B.ws_bytes = off;
    B.count = count;
    B.precision = descs[0].precision;
    B.threads = threads;
    return PFH_OK;
}

template <typename T>
static pfh_status batch_upload(pfh_batch_s &B, int op, const void *const *srcs, void *const *dsts, void *ws,
                               hipStream_t stream)
{
    B.release_stages();
    int nst = 0;
    for (auto &its : B.items) nst = std::max(nst, (int)its[op].size());
    const int cus = B.plans.empty() ? 256 : B.plans[0]->dev.cus;
    // Problem groups: consecutive problems whose arrays add up to at most the
    // group budget run all their stages before the next group starts, and
    // every group's problems take their workspace from the start of the
    // workspace, so a group's intermediates can stay in the last-level cache.
    int group_mb = B.tuning.batch_group_mb;
    if (const char *e = std::getenv("PFH_BATCH_GROUP_MB")) group_mb = std::atoi(e);
    std::vector<std::pair<int, int>> pgroups;
    std::vector<int64_t> off(B.ws_off);
    if (group_mb <= 0) {
        pgroups.push_back({0, B.count});
    } else {
        const int64_t budget = (int64_t)group_mb << 20;
        int a = 0;
        int64_t acc = 0;
        for (int i = 0; i < B.count; ++i) {
            const int64_t b = B.plans[i]->ops[op].in.bytes();
            if (i > a && acc + b > budget) {
                pgroups.push_back({a, i});
                a = i;
                acc = 0;
            }
            acc += b;
        }
        pgroups.push_back({a, B.count});
        for (const auto &g : pgroups) {
            int64_t o = 0;
            for (int i = g.first; i < g.second; ++i) {
                off[i] = o;
                o += (i + 1 < B.count ? B.ws_off[i + 1] : B.ws_bytes) - B.ws_off[i];
            }
        }
    }
    // chain of a problem: 1 when its first pass is a plane stage and the
    // batch runs two chains (the chains share no problem, so their launches
    // need no ordering between them)
    // (problem groups reuse one workspace region group after group, so they
    // keep one chain: a second stream would let two groups overlap in it)
    const bool two_chains = B.tuning.batch_streams > 1 && pgroups.size() == 1;
    auto chain_of = [&](int i) {
        return (two_chains && !B.items[i][op].empty() && B.items[i][op][0].plane) ? 1 : 0;
    };
    for (const auto &pg : pgroups)
    for (int chain = 0; chain < 2; ++chain)
    for (int s = 0; s < nst; ++s) {
        // group this stage's items by kernel
        struct Group {
            BatchStage bs;
            std::vector<StageArgs<T>> items;
            std::vector<PlaneArgs<T>> planes;
            std::vector<int64_t> prefix{0};
        };
        std::vector<Group> groups;
        for (int i = pg.first; i < pg.second; ++i) {
            const PlanImpl &P = *B.plans[i];
            const OpPlan &o = P.ops[op];
            const std::vector<BatchItem> &its = B.items[i][op];
            if (s >= (int)its.size() || chain_of(i) != chain) continue;
            if (its[s].plane) {
                const PlaneDesc &pd = its[s].pd;
                // Plane items go to launches by footprint: a plane that fits
                // in a half / quarter of the LDS runs in 512- / 256-thread
                // workgroups, two / four per CU (the kernel's registers allow
                // 16 waves per CU), so their load, transform and store
                // phases overlap; a launch's workgroups all take the LDS of
                // its largest item.
                const int64_t lim = (int64_t)P.dev.lds_limit;
                const int64_t foot = (int64_t)pd.lds_bytes + 1024; // + one prefetch chunk
                const int share = foot * 4 <= lim ? 4 : foot * 2 <= lim ? 2 : 1;
                Group *g = nullptr;
                for (Group &c : groups)
                    if (c.bs.plane && c.bs.threads == kPlaneThreads / share) g = &c;
                if (!g) {
                    groups.emplace_back();
                    g = &groups.back();
                    g->bs.plane = true;
                    g->bs.threads = kPlaneThreads / share;
                }
                WsView wv;
                pfh_status e = view_workspace(P, o, static_cast<char *>(ws) + off[i], (size_t)(B.ws_bytes), wv);
                if (e) return e;
                void *bufs[4] = {const_cast<void *>(srcs[i]), dsts[i], wv.w[0], wv.w[1]};
                PlaneArgs<T> pa{};
                pa.src = bufs[pd.in_buf];
                pa.dst = bufs[pd.out_buf];
                pa.tw_row = reinterpret_cast<const cpx<T> *>(P.dev_tables) + pd.tw_base_row;
                pa.tw_col = reinterpret_cast<const cpx<T> *>(P.dev_tables) + pd.tw_base_col;
                pa.nplanes = pd.nplanes;
                pa.nr = pd.nr;
                pa.nc = pd.nc;
                pa.twn_row = pd.twn_row;
                pa.twn_col = pd.twn_col;
                for (int r = 0; r < 3; ++r) {
                    pa.twoff_row[r] = pd.twoff_row[r];
                    pa.twoff_col[r] = pd.twoff_col[r];
                    pa.rr[r] = pd.rr[r];
                    pa.cr[r] = pd.cr[r];
                }
                pa.rem = pd.rem;
                pa.cem = pd.cem;
                pa.tpl_r = pd.tpl_r;
                pa.tpl_c = pd.tpl_c;
                pa.rs = pd.rs;
                pa.flags = pd.flags;
                pa.scale = (T)pd.scale;
                g->planes.push_back(pa);
                g->prefix.push_back(g->prefix.back() + pd.nplanes);
                g->bs.tw_cap = std::max(g->bs.tw_cap, pd.twn_row + pd.twn_col);
                g->bs.work_cap = std::max(g->bs.work_cap, pd.nc * pd.rs);
                continue;
            }
            const Step &st = its[s].step;
            bool wide = false;
            for (int R : st.radices) wide = wide || is_wide_radix(R);
            // every compile-time plan of the standard register budget shares
            // one launch; the others get a launch per plan / radix set
            const bool multi = B.tuning.batch_multi && st.fast &&
                               multi_batch_plan_available((int)sizeof(T), st.fr[0], st.fr[1], st.fr[2], st.fem);
            Group *g = nullptr;
            for (Group &c : groups) {
                const BatchStage &k = c.bs;
                if (k.plane || k.multi != multi || k.fast != (st.fast && !multi)) continue;
                if (multi || (st.fast ? (k.fr[0] == st.fr[0] && k.fr[1] == st.fr[1] && k.fr[2] == st.fr[2] &&
                                         k.fem == st.fem)
                                      : (k.wide == wide))) {
                    g = &c;
                    break;
                }
            }
            if (!g) {
                groups.emplace_back();
                g = &groups.back();
                g->bs.multi = multi;
                g->bs.fast = st.fast && !multi;
                g->bs.wide = !st.fast && wide;
                for (int r = 0; r < 3; ++r) g->bs.fr[r] = st.fr[r];
                g->bs.fem = st.fem;
            }
            WsView wv;
            pfh_status e = view_workspace(P, o, static_cast<char *>(ws) + off[i], (size_t)(B.ws_bytes), wv);
            if (e) return e;
            void *bufs[4] = {const_cast<void *>(srcs[i]), dsts[i], wv.w[0], wv.w[1]};
            g->items.push_back(fill_args<T>(P, st, bufs, nullptr));
            g->prefix.push_back(g->prefix.back() + st.tiles);
            g->bs.lds = std::max(g->bs.lds, st.lds_bytes);
            g->bs.threads = std::max(g->bs.threads, st.threads);
        }
        for (Group &g : groups) {
            BatchStage bs = g.bs;
            bs.chain = chain;
            if (bs.plane) {
                // tables and working plane, then the prefetch region (16-byte
                // aligned) with the rest of the workgroup's LDS share
                if ((bs.tw_cap + bs.work_cap) % 2) ++bs.tw_cap;
                const int share = kPlaneThreads / bs.threads;
                const int64_t cap = (int64_t)(B.plans[pg.first]->dev.lds_limit / share / sizeof(cpx<T>));
                bs.raw_cap = (int)std::max<int64_t>(0, cap - bs.tw_cap - bs.work_cap);
                bs.lds = (size_t)(bs.tw_cap + bs.work_cap + bs.raw_cap) * sizeof(cpx<T>);
            }
            const size_t isz = bs.plane ? sizeof(PlaneArgs<T>) : sizeof(StageArgs<T>);
            const size_t nit = bs.plane ? g.planes.size() : g.items.size();
            const void *hdata = bs.plane ? static_cast<const void *>(g.planes.data())
                                         : static_cast<const void *>(g.items.data());
            bs.nitems = (int)nit;
            bs.total = g.prefix.back();
            PFH_HIP_CHECK(hipMalloc(&bs.d_items, std::max<size_t>(1, nit) * isz));
            B.stages.push_back(bs); // owned from here (freed by release_stages)
            BatchStage &own = B.stages.back();
            PFH_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&own.d_prefix), g.prefix.size() * sizeof(int64_t)));
            PFH_HIP_CHECK(hipMemcpy(own.d_items, hdata, nit * isz, hipMemcpyHostToDevice));
            PFH_HIP_CHECK(
                hipMemcpy(own.d_prefix, g.prefix.data(), g.prefix.size() * sizeof(int64_t), hipMemcpyHostToDevice));
            int plane_per_cu = 1;
            if constexpr (std::is_same<T, float>::value)
                if (own.plane) plane_per_cu = plane_batch_blocks_per_cu<T>(own.threads, own.lds);
            const int per_cu = own.plane  ? plane_per_cu
                               : own.multi ? line_multi_batch_blocks_per_cu<T>(own.threads, own.lds)
                               : own.fast ? line_fast_batch_blocks_per_cu<T>(own.fr[0], own.fr[1], own.fr[2],
                                                                             own.fem, own.threads, own.lds)
                                          : line_batch_blocks_per_cu<T>(own.wide, own.threads, own.lds);
            static const int persistent_env = env_int("PFH_BATCH_PERSISTENT", -1);
            const bool persistent = persistent_env < 0 ? B.tuning.batch_persistent : persistent_env != 0;
            own.grid = persistent ? std::min<int64_t>(own.total, (int64_t)per_cu * cus) : own.total;
        }
    }
    (void)stream;
    B.cur_op = op;
    B.cur_src.assign(srcs, srcs + B.count);
    B.cur_dst.assign(dsts, dsts + B.count);
    B.cur_ws = ws;
    return PFH_OK;
}

// One persistent launch of a batch (a kernel class of one route stage).
template <typename T>
static pfh_status batch_launch_one(const BatchStage &bs, hipStream_t stream)
{
    const StageArgs<T> *items = reinterpret_cast<const StageArgs<T> *>(bs.d_items);
    if (bs.plane) {
        if constexpr (std::is_same<T, float>::value)
            PFH_HIP_CHECK(launch_plane_batch<T>(reinterpret_cast<const PlaneArgs<T> *>(bs.d_items), bs.d_prefix,
                                                bs.nitems, bs.total, bs.threads, bs.tw_cap, bs.work_cap, bs.raw_cap,
                                                bs.grid, stream));
        else
            return set_err(PFH_ERR_UNSUPPORTED, "plane stages are f32 only");
    } else if (bs.multi)
        PFH_HIP_CHECK(launch_line_multi_batch<T>(items, bs.d_prefix, bs.nitems, bs.total, bs.threads, bs.lds,
                                                 bs.grid, stream));
    else if (bs.fast)
        PFH_HIP_CHECK(launch_line_fast_batch<T>(items, bs.d_prefix, bs.nitems, bs.total, bs.fr[0], bs.fr[1],
                                                bs.fr[2], bs.fem, bs.threads, bs.lds, bs.grid, stream));
    else
        PFH_HIP_CHECK(launch_line_batch<T>(items, bs.d_prefix, bs.nitems, bs.total, bs.wide, bs.threads, bs.lds,
                                           bs.grid, stream));
    return PFH_OK;
}

template <typename T>
static pfh_status batch_launch(pfh_batch_s &B, hipStream_t stream)
{
    // two chains: chain 1 on the side stream, forked from and joined back to
    // the caller's stream by events (a graph captured on the caller's stream
    // gets the two branches)
    bool two = false;
    for (const BatchStage &bs : B.stages) two = two || bs.chain == 1;
    if (two) {
        if (!B.side_stream) {
            PFH_HIP_CHECK(hipStreamCreateWithFlags(&B.side_stream, hipStreamNonBlocking));
            PFH_HIP_CHECK(hipEventCreateWithFlags(&B.ev_fork, hipEventDisableTiming));
            PFH_HIP_CHECK(hipEventCreateWithFlags(&B.ev_join, hipEventDisableTiming));
        }
        PFH_HIP_CHECK(hipEventRecord(B.ev_fork, stream));
        PFH_HIP_CHECK(hipStreamWaitEvent(B.side_stream, B.ev_fork, 0));
    }
    for (const BatchStage &bs : B.stages) {
        const hipStream_t s = bs.chain == 1 ? B.side_stream : stream;
        pfh_status st = batch_launch_one<T>(bs, s);
        if (st) return st;
    }
    if (two) {
        PFH_HIP_CHECK(hipEventRecord(B.ev_join, B.side_stream));
        PFH_HIP_CHECK(hipStreamWaitEvent(stream, B.ev_join, 0));
    }
    return PFH_OK;
}

static pfh_status batch_run(pfh_batch_s &B, int op, const void *const *srcs, void *const *dsts, void *ws,
                            hipStream_t stream, int use_graph)
{
    bool same = B.cur_op == op && B.cur_ws == ws;
    for (int i = 0; same && i < B.count; ++i) same = B.cur_src[i] == srcs[i] && B.cur_dst[i] == dsts[i];
    if (!same) {
        pfh_status st = B.precision == PFH_F32 ? batch_upload<float>(B, op, srcs, dsts, ws, stream)
                                               : batch_upload<double>(B, op, srcs, dsts, ws, stream);
        if (st) return st;
    }
    if (!use_graph) return B.precision == PFH_F32 ? batch_launch<float>(B, stream) : batch_launch<double>(B, stream);
    // the legacy null stream cannot be captured: hop to a stream of our own
    const bool hop = stream == nullptr;
    hipStream_t caller = stream;
    if (hop) {
        if (!B.own_stream) {
            PFH_HIP_CHECK(hipStreamCreateWithFlags(&B.own_stream, hipStreamNonBlocking));
            PFH_HIP_CHECK(hipEventCreateWithFlags(&B.ev_in, hipEventDisableTiming));
            PFH_HIP_CHECK(hipEventCreateWithFlags(&B.ev_out, hipEventDisableTiming));
        }
        PFH_HIP_CHECK(hipEventRecord(B.ev_in, caller));
        PFH_HIP_CHECK(hipStreamWaitEvent(B.own_stream, B.ev_in, 0));
        stream = B.own_stream;
    }
    if (!B.graph || B.graph_stream != stream) {
        if (B.graph) (void)hipGraphExecDestroy(B.graph);
        B.graph = nullptr;
        hipGraph_t g = nullptr;
        PFH_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        pfh_status st = B.precision == PFH_F32 ? batch_launch<float>(B, stream) : batch_launch<double>(B, stream);
        hipError_t ce = hipStreamEndCapture(stream, &g);
        if (st) return st;
        if (ce != hipSuccess) return set_err(PFH_ERR_DEVICE, hipGetErrorString(ce), 0, 0, 0, 0, (int)ce);
        hipError_t ie = hipGraphInstantiate(&B.graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return set_err(PFH_ERR_DEVICE, hipGetErrorString(ie), 0, 0, 0, 0, (int)ie);
        B.graph_stream = stream;
    }
    PFH_HIP_CHECK(hipGraphLaunch(B.graph, stream));
    if (hop) {
        PFH_HIP_CHECK(hipEventRecord(B.ev_out, stream));
        PFH_HIP_CHECK(hipStreamWaitEvent(caller, B.ev_out, 0));
    }
    return PFH_OK;
}

// The tuning of one autotuner candidate.
static Tuning batch_candidate(int block, int col_lines, bool plane, bool persistent, bool multi, int max_radix,
                              int streams = 1)
{
    Tuning t;
    t.batch_streams = streams;
    t.batch_multi = multi;
    t.batch_persistent = persistent;
    t.batch_plane = plane;
    t.block = block;
    t.max_radix = max_radix;
    t.col_lines = col_lines;
    t.block_cap = kMaxThreadsWide;
    t.allow_fast = true;
    t.allow_xcd = false;
    t.slab_mb = 0;
    t.fill_block = true;
    return t;
}

// Measured autotuning: every candidate tuning re-plans the whole batch and
// times forward executions (device events); the fastest is kept, and
// recorded in the tuning cache under the batch's key. A batch whose key is
// cached is planned with the recorded tuning without measuring.
static pfh_status batch_autotune(pfh_batch_s &B, const pfh_desc *descs, int count)
{
    DeviceInfo dev;
    {
        pfh_status st = device_info(dev);
        if (st) return st;
    }
    const std::string key = batch_tune_key(dev.arch, dev.cus, dev.lds_limit, descs, count);
    {
        TuneRecord rec;
        int source = PFH_TUNED_DEFAULT;
        if (tune_cache_lookup(key, rec, &source)) {
            const Tuning t = batch_candidate(rec.block, rec.col_lines, rec.plane != 0, rec.persistent != 0,
                                             rec.multi != 0, rec.max_radix, rec.streams);
            if (batch_plan(B, descs, count, t) == PFH_OK) {
                B.tune_source = source;
                return ok();
            }
            tune_cache_forget(key); // a record this build cannot plan: measure again
        }
    }
    // Candidates: threads per block of the line stages, column-tile lines,
    // persistent workgroups or one per tile, plane stages on or off (every
    // compile-time plan shares one launch: a launch per plan measured slower
    // on every candidate of the 96-problem workload; the largest generic
    // radix only matters to lengths without a compile-time plan)
    std::vector<Tuning> cands;
    for (int block : {128, 256, 512})
        for (int col : {0, 32})
            for (bool plane : {true, false})
                for (bool pers : {true, false})
                    for (int streams : {1, 2}) {
                        if (streams == 2 && !plane) continue; // only plane stages make a second chain
                        cands.push_back(batch_candidate(block, col, plane, pers, true, 8, streams));
                    }
    // measurement buffers: inputs hold finite nonzero values (every byte
    // 0x3c), so the timed launches move and compute what real data makes
    // them move and compute
    std::vector<void *> bufs;
    std::vector<const void *> srcs(count);
    std::vector<void *> dsts(count);
    auto cleanup = [&]() {
        for (void *p : bufs) (void)hipFree(p);
        bufs.clear();
    };
    {
        pfh_status st = batch_plan(B, descs, count, cands[0]);
        if (st) return st;
    }
    for (int i = 0; i < count; ++i) {
        const Layout &L = B.plans[i]->ops[PFH_FORWARD].in;
        void *a = nullptr, *b = nullptr;
        if (hipMalloc(&a, (size_t)L.bytes()) != hipSuccess || hipMalloc(&b, (size_t)L.bytes()) != hipSuccess) {
            if (a) (void)hipFree(a);
            cleanup();
            return set_err(PFH_ERR_ALLOCATION_FAILED, "autotune buffers");
        }
        (void)hipMemset(a, 0x3c, (size_t)L.bytes());
        bufs.push_back(a);
        bufs.push_back(b);
        srcs[i] = a;
        dsts[i] = b;
    }
    hipStream_t stream = nullptr;
    if (hipStreamCreate(&stream) != hipSuccess) {
        cleanup();
        return set_err(PFH_ERR_DEVICE, "stream");
    }
    constexpr int kMaxReps = 9;
    hipEvent_t ev[kMaxReps + 1];
    for (hipEvent_t &e : ev) (void)hipEventCreate(&e);
    int64_t maxws = 0;
    void *ws = nullptr;
    // The median of `reps` executions of candidate t, each timed on its own
    // after one untimed warm-up; negative if t does not plan or run.
    auto time_candidate = [&](const Tuning &t, int reps) -> double {
        if (batch_plan(B, descs, count, t)) return -1;
        if (B.ws_bytes > maxws) {
            if (ws) (void)hipFree(ws);
            ws = nullptr;
            maxws = 0;
            if (hipMalloc(&ws, (size_t)B.ws_bytes) != hipSuccess) return -1;
            maxws = B.ws_bytes;
        }
        if (batch_run(B, PFH_FORWARD, srcs.data(), dsts.data(), ws, stream, 0)) return -1;
        (void)hipEventRecord(ev[0], stream);
        for (int r = 0; r < reps; ++r) {
            if (batch_run(B, PFH_FORWARD, srcs.data(), dsts.data(), ws, stream, 0)) return -1;
            (void)hipEventRecord(ev[r + 1], stream);
        }
        (void)hipEventSynchronize(ev[reps]);
        std::vector<float> ms(reps);
        for (int r = 0; r < reps; ++r) (void)hipEventElapsedTime(&ms[r], ev[r], ev[r + 1]);
        std::nth_element(ms.begin(), ms.begin() + reps / 2, ms.end());
        return ms[reps / 2];
    };
    // Round 1: every candidate, median of 5. Round 2: the four fastest again,
    // median of 9, each keeping the better of its two medians, so that one
    // disturbed round does not decide.
    std::vector<std::pair<double, int>> score;
    for (int c = 0; c < (int)cands.size(); ++c) {
        const double m = time_candidate(cands[c], 5);
        if (m >= 0) score.push_back({m, c});
    }
    std::sort(score.begin(), score.end());
    for (int i = 0; i < (int)score.size() && i < 4; ++i) {
        const double m = time_candidate(cands[score[i].second], kMaxReps);
        if (m >= 0 && m < score[i].first) score[i].first = m;
    }
    std::sort(score.begin(), score.end());
    if (ws) (void)hipFree(ws);
    for (hipEvent_t &e : ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(stream);
    cleanup();
    if (score.empty()) return batch_plan(B, descs, count, cands[0]);
    const Tuning &best = cands[score[0].second];
    pfh_status st = batch_plan(B, descs, count, best);
    if (st) return st;
    TuneRecord rec;
    rec.block = best.block;
    rec.max_radix = best.max_radix;
    rec.col_lines = best.col_lines;
    rec.multi = best.batch_multi ? 1 : 0;
    rec.persistent = best.batch_persistent ? 1 : 0;
    rec.plane = best.batch_plane ? 1 : 0;
    rec.streams = best.batch_streams;
    rec.ms = score[0].first;
    tune_cache_store(key, rec);
    B.tune_source = PFH_TUNED_MEASURED;
    return ok();
}

extern "C" {

pfh_status pfh_batch_create(const pfh_desc *descs, int count, int autotune, pfh_batch *out)
{
    if (!descs || !out || count <= 0) return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    *out = nullptr;
    pfh_batch b = new (std::nothrow) pfh_batch_s;
    if (!b) return set_err(PFH_ERR_ALLOCATION_FAILED, "batch");
    Tuning t;
    t.allow_fast = true;
    t.allow_xcd = false;
    t.block = 256;
    t.block_cap = kMaxThreadsWide;
    t.slab_mb = 0;
    t.fill_block = true;
    pfh_status st = autotune ? batch_autotune(*b, descs, count) : batch_plan(*b, descs, count, t);
    if (st) {
        delete b;
        return st;
    }
    *out = b;
    return ok();
}

void pfh_batch_destroy(pfh_batch batch) { delete batch; }

pfh_status pfh_batch_workspace_bytes(pfh_batch batch, size_t *bytes)
{
    if (!batch || !bytes) return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    *bytes = (size_t)std::max<int64_t>(batch->ws_bytes, 256);
    return ok();
}

pfh_status pfh_batch_tuning_source(pfh_batch batch, int *source)
{
    if (!batch || !source) return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    *source = batch->tune_source;
    return ok();
}

void pfh_batch_tuning_cache_clear(void) { tune_cache_clear(); }

pfh_status pfh_batch_tuning(pfh_batch batch, int *block, int *max_radix, int *col_lines, int *one_launch,
                            int *persistent, int *plane, int *streams)
{
    if (!batch) return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    if (plane) {
        // fused plane stages on, and used by at least one problem
        int used = 0;
        for (const auto &its : batch->items)
            for (const auto &v : its)
                for (const BatchItem &it : v) used |= it.plane ? 1 : 0;
        *plane = batch->tuning.batch_plane && used ? 1 : 0;
    }
    if (block) *block = batch->tuning.block;
    if (max_radix) *max_radix = batch->tuning.max_radix;
    if (col_lines) *col_lines = batch->tuning.col_lines;
    if (one_launch) *one_launch = batch->tuning.batch_multi ? 1 : 0;
    if (persistent) *persistent = batch->tuning.batch_persistent ? 1 : 0;
    if (streams) *streams = batch->tuning.batch_streams;
    return ok();
}

pfh_status pfh_batch_execute(pfh_batch batch, int op, const void *const *srcs, void *const *dsts, void *workspace,
                             size_t workspace_bytes, pfh_stream_t stream, int use_graph)
{
    if (!batch || !srcs || !dsts) return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    if (op != PFH_FORWARD && op != PFH_INVERSE && op != PFH_BACKWARD)
        return set_err(PFH_ERR_INVALID_ARGUMENT, "batched op must be forward, inverse or backward");
    if (!workspace || (int64_t)workspace_bytes < batch->ws_bytes)
        return set_err(PFH_ERR_WORKSPACE_TOO_SMALL, "batch workspace", (size_t)batch->ws_bytes,
                       workspace ? workspace_bytes : 0);
    for (int i = 0; i < batch->count; ++i)
        if (!srcs[i] || !dsts[i]) return set_err(PFH_ERR_INVALID_ARGUMENT, "null problem buffer");
    pfh_status st = batch_run(*batch, op, srcs, dsts, workspace, reinterpret_cast<hipStream_t>(stream), use_graph);
    if (st) return st;
    return ok();
}

} // extern "C"

// ---------------------------------------------------------------------------
// pointwise operations
// ---------------------------------------------------------------------------

extern "C" {

static pfh_status scale_axes_impl(pfh_plan plan, int side, void *data, size_t elems, const void *factors,
                                  bool complex_factors, pfh_stream_t stream)
{
    if (!plan || !data || !factors || (side != 0 && side != 1)) return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    const PlanImpl &P = plan->impl;
    if (!(P.kind == PFH_C2C || P.kind == PFH_R2C || P.kind == PFH_R2R || P.kind == PFH_C2R ||
          P.kind == PFH_MIXED_C2C || P.kind == PFH_MIXED_R2C))
        return set_err(PFH_ERR_INVALID_ARGUMENT, "needs a distributed transform plan");
    const OpPlan &o = P.ops[PFH_FORWARD];
    const Layout &L = side == 0 ? o.in : o.out;
    if ((int64_t)elems != L.total())
        return set_err(side == 0 ? PFH_ERR_INPUT_LAYOUT_MISMATCH : PFH_ERR_OUTPUT_LAYOUT_MISMATCH, "array length",
                       (size_t)L.total(), elems);
    const int tb = P.precision == PFH_F32 ? 4 : 8;
    ScaleAxesArgs a{};
    a.nd = L.N;
    a.complex_data = L.elem == 2 * tb ? 1 : 0;
    a.complex_factors = complex_factors ? 1 : 0;
    if (complex_factors && !a.complex_data)
        return set_err(PFH_ERR_INVALID_ARGUMENT, "complex factors need the complex side of the plan");
    a.len = L.len();
    a.extra = L.extra;
    for (int p = 0; p < L.N; ++p) {
        a.mext[p] = L.shape[L.perm[p]];
        int64_t off = 0;
        for (int b = 0; b < L.perm[p]; ++b) off += L.shape[b];
        a.foff[p] = off;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    PFH_HIP_CHECK(P.precision == PFH_F32 ? launch_scale_axes<float>(a, data, factors, s)
                                         : launch_scale_axes<double>(a, data, factors, s));
    return ok();
}

pfh_status pfh_pointwise_scale_axes(pfh_plan plan, int side, void *data, size_t elems, const void *factors,
                                    pfh_stream_t stream)
{
    return scale_axes_impl(plan, side, data, elems, factors, false, stream);
}

pfh_status pfh_pointwise_scale_axes_complex(pfh_plan plan, int side, void *data, size_t elems, const void *factors,
                                            pfh_stream_t stream)
{
    return scale_axes_impl(plan, side, data, elems, factors, true, stream);
}

pfh_status pfh_pointwise_quadratic(int precision, const void *w, void *u, size_t n, double a, pfh_stream_t stream)
{
    if (!w || !u || (precision != PFH_F32 && precision != PFH_F64)) return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    PFH_HIP_CHECK(precision == PFH_F32 ? launch_quadratic<float>(w, u, (int64_t)n, a, s)
                                       : launch_quadratic<double>(w, u, (int64_t)n, a, s));
    return ok();
}

pfh_status pfh_pointwise_quadratic_adjoint(int precision, const void *w, const void *g, void *out, size_t n,
                                           double a, pfh_stream_t stream)
{
    if (!w || !g || !out || (precision != PFH_F32 && precision != PFH_F64))
        return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    PFH_HIP_CHECK(precision == PFH_F32 ? launch_quadratic_adjoint<float>(w, g, out, (int64_t)n, a, s)
                                       : launch_quadratic_adjoint<double>(w, g, out, (int64_t)n, a, s));
    return ok();
}

pfh_status pfh_sum_squares(int precision, const void *x, size_t n, void *workspace, size_t workspace_bytes,
                           double *result, pfh_stream_t stream)
{
    if (!x || !result || (precision != PFH_F32 && precision != PFH_F64))
        return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    const size_t need = sizeof(double) * kSumSquaresParts;
    if (!workspace || workspace_bytes < need)
        return set_err(PFH_ERR_WORKSPACE_TOO_SMALL, "sum of squares workspace", need, workspace ? workspace_bytes : 0);
    int64_t parts = (int64_t)n / 16384;
    if (parts < 1) parts = 1;
    if (parts > kSumSquaresParts) parts = kSumSquaresParts;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    double *partial = static_cast<double *>(workspace);
    PFH_HIP_CHECK(precision == PFH_F32 ? launch_sum_squares<float>(x, (int64_t)n, partial, (int)parts, result, s)
                                       : launch_sum_squares<double>(x, (int64_t)n, partial, (int)parts, result, s));
    return ok();
}

// ---------------------------------------------------------------------------
// Reductions and broadcast pointwise combinations
// ---------------------------------------------------------------------------

static_assert(sizeof(pfh_reduce_result) == sizeof(ReduceResult), "pfh_reduce_result layout");
static_assert((int)PFH_SCALAR_C128 == (int)kScalarC128 && (int)PFH_REDUCE_ALL == (int)kReduceAll &&
                  (int)PFH_POINTWISE_MUL_ADD == (int)kPointwiseMulAdd,
              "enum values shared with the kernels");

pfh_status pfh_reduce_workspace_bytes(size_t *bytes)
{
    if (!bytes) return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    *bytes = (size_t)kReduceWorkspaceBytes;
    return ok();
}

pfh_status pfh_reduce(int op, int scalar_type, const void *x, size_t n, void *workspace, size_t workspace_bytes,
                      pfh_reduce_result *result, pfh_stream_t stream)
{
    if (op < PFH_REDUCE_SUM || op > PFH_REDUCE_ALL) return set_err(PFH_ERR_INVALID_ARGUMENT, "reduction");
    if (scalar_type < PFH_SCALAR_I8 || scalar_type > PFH_SCALAR_C128)
        return set_err(PFH_ERR_INVALID_ARGUMENT, "scalar type");
    const bool complex_type = scalar_type == PFH_SCALAR_C64 || scalar_type == PFH_SCALAR_C128;
    if (complex_type && (op == PFH_REDUCE_MIN || op == PFH_REDUCE_MAX))
        return set_err(PFH_ERR_INVALID_ARGUMENT, "complex values have no order");
    if (!result || (n > 0 && !x)) return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    if (n > (size_t)INT64_MAX / 16) return set_err(PFH_ERR_LENGTH_OVERFLOW, "element count");
    if (!workspace || workspace_bytes < (size_t)kReduceWorkspaceBytes)
        return set_err(PFH_ERR_WORKSPACE_TOO_SMALL, "reduction workspace", (size_t)kReduceWorkspaceBytes,
                       workspace ? workspace_bytes : 0);
    PFH_HIP_CHECK(launch_reduce(op, scalar_type, x, (int64_t)n, workspace, reinterpret_cast<ReduceResult *>(result),
                                reinterpret_cast<hipStream_t>(stream)));
    return ok();
}

static bool layout_valid(const pfh_pencil_layout &p)
{
    if (p.ndim < 1 || p.ndim > PFH_MAX_DIMS || p.topology_dims < 0 || p.topology_dims > p.ndim) return false;
    bool seen[PFH_MAX_DIMS] = {false, false, false, false};
    for (int a = 0; a < p.ndim; ++a) {
        if (p.perm[a] < 0 || p.perm[a] >= p.ndim || seen[p.perm[a]]) return false;
        seen[p.perm[a]] = true;
    }
    return true;
}

// Pencil::same_layout (pencil.rs:238-240) at one rank: same global shape,
// decomposition, topology rank and permutation.
static bool same_layout(const pfh_pencil_layout &a, const pfh_pencil_layout &b)
{
    if (a.ndim != b.ndim || a.topology_dims != b.topology_dims) return false;
    for (int d = 0; d < a.ndim; ++d)
        if (a.shape[d] != b.shape[d] || a.perm[d] != b.perm[d]) return false;
    for (int d = 0; d < a.topology_dims; ++d)
        if (a.decomposition[d] != b.decomposition[d]) return false;
    return true;
}

// validate_extents (pointwise.rs:178-197)
static pfh_status check_extents(const pfh_array &in, const pfh_array &out, size_t index)
{
    if (in.extra_ndim != out.extra_ndim)
        return set_err(PFH_ERR_EXTRA_RANK_MISMATCH, "pointwise extra rank mismatch", (size_t)out.extra_ndim,
                       (size_t)in.extra_ndim, index);
    for (int d = 0; d < in.extra_ndim; ++d)
        if (in.extra_shape[d] != out.extra_shape[d] && in.extra_shape[d] != 1)
            return set_err(PFH_ERR_EXTRA_EXTENT_MISMATCH, "pointwise extra extent mismatch", out.extra_shape[d],
                           in.extra_shape[d], index, (size_t)d);
    return PFH_OK;
}

static bool float_type(int t)
{
    return t == PFH_SCALAR_F32 || t == PFH_SCALAR_F64 || t == PFH_SCALAR_C64 || t == PFH_SCALAR_C128;
}
static bool f64_type(int t) { return t == PFH_SCALAR_F64 || t == PFH_SCALAR_C128; }
static bool complex_type(int t) { return t == PFH_SCALAR_C64 || t == PFH_SCALAR_C128; }

pfh_status pfh_pointwise(int op, const pfh_array *inputs, int n_inputs, const pfh_array *out, const double *coeffs,
                         pfh_stream_t stream)
{
    if (op < PFH_POINTWISE_LINEAR || op > PFH_POINTWISE_MUL_ADD || !out || n_inputs < 0 ||
        (n_inputs > 0 && !inputs))
        return set_err(PFH_ERR_INVALID_ARGUMENT, "argument");
    const int want = op == PFH_POINTWISE_MUL ? 2 : op == PFH_POINTWISE_MUL_ADD ? 3 : -1;
    if (op == PFH_POINTWISE_LINEAR) {
        if (n_inputs == 0) return set_err(PFH_ERR_EMPTY_INPUTS, "pointwise many operations require at least one input");
        if (n_inputs > kPointwiseMaxInputs) return set_err(PFH_ERR_UNSUPPORTED, "at most 8 inputs");
    } else if (n_inputs != want) {
        return set_err(PFH_ERR_INVALID_ARGUMENT, "input count of the operation");
    }
    // layouts and extra ranks must be well formed before they are compared
    if (!layout_valid(out->pencil)) return set_err(PFH_ERR_INVALID_PERMUTATION, "output pencil");
    if (out->extra_ndim < 0 || out->extra_ndim > PFH_MAX_DIMS) return set_err(PFH_ERR_INVALID_ARGUMENT, "extra rank");
    for (int i = 0; i < n_inputs; ++i) {
        if (!layout_valid(inputs[i].pencil)) return set_err(PFH_ERR_INVALID_PERMUTATION, "input pencil", 0, 0, (size_t)i);
        if (inputs[i].extra_ndim < 0 || inputs[i].extra_ndim > PFH_MAX_DIMS)
            return set_err(PFH_ERR_INVALID_ARGUMENT, "extra rank", 0, 0, (size_t)i);
    }
    if (op == PFH_POINTWISE_MUL) {
        // validate_two_inputs (pointwise.rs:155-176): both layouts, then extents
        if (!same_layout(inputs[0].pencil, inputs[1].pencil) || !same_layout(inputs[0].pencil, out->pencil))
            return set_err(PFH_ERR_INCOMPATIBLE_PENCILS, "pointwise arrays must have identical spatial layouts");
        for (int i = 0; i < 2; ++i) {
            pfh_status st = check_extents(inputs[i], *out, (size_t)i);
            if (st) return st;
        }
    } else {
        // OutputPlan::new / input (pointwise.rs:262-330), input by input
        for (int i = 0; i < n_inputs; ++i) {
            if (!same_layout(inputs[i].pencil, out->pencil))
                return set_err(PFH_ERR_INCOMPATIBLE_PENCILS, "pointwise arrays must have identical spatial layouts", 0,
                               0, (size_t)i);
            pfh_status st = check_extents(inputs[i], *out, (size_t)i);
            if (st) return st;
        }
    }
    // element types (fixed by the Rust types in the reference)
    const int ncoef = op == PFH_POINTWISE_LINEAR ? n_inputs : op == PFH_POINTWISE_MUL ? 1 : 2;
    if (!coeffs || !float_type(out->scalar_type)) return set_err(PFH_ERR_INVALID_ARGUMENT, "coefficients or type");
    bool any_complex = false;
    for (int i = 0; i < n_inputs; ++i) {
        if (!float_type(inputs[i].scalar_type) || f64_type(inputs[i].scalar_type) != f64_type(out->scalar_type))
            return set_err(PFH_ERR_INVALID_ARGUMENT, "inputs and output must share one floating precision", 0, 0,
                           (size_t)i);
        any_complex = any_complex || complex_type(inputs[i].scalar_type);
    }
    for (int c = 0; c < ncoef; ++c) any_complex = any_complex || coeffs[2 * c + 1] != 0.0;
    if (any_complex && !complex_type(out->scalar_type))
        return set_err(PFH_ERR_INVALID_ARGUMENT, "a complex input or coefficient needs a complex output");
    // sizes
    int64_t spatial = 1, extra = 1;
    for (int d = 0; d < out->pencil.ndim; ++d) {
        if (out->pencil.shape[d] > (size_t)INT64_MAX / (size_t)std::max<int64_t>(spatial, 1))
            return set_err(PFH_ERR_LENGTH_OVERFLOW, "spatial length");
        spatial *= (int64_t)out->pencil.shape[d];
    }
    for (int d = 0; d < out->extra_ndim; ++d) {
        if (out->extra_shape[d] > (size_t)INT64_MAX / (size_t)std::max<int64_t>(extra, 1))
            return set_err(PFH_ERR_LENGTH_OVERFLOW, "extra count");
        extra *= (int64_t)out->extra_shape[d];
    }
    if (spatial > 0 && extra > INT64_MAX / 2 / spatial) return set_err(PFH_ERR_LENGTH_OVERFLOW, "element count");
    if (spatial == 0 || extra == 0) return ok(); // no element is visited
    if (!out->data) return set_err(PFH_ERR_INVALID_ARGUMENT, "output data");
    PointwiseArgs a;
    a.op = op;
    a.nin = n_inputs;
    a.out_complex = complex_type(out->scalar_type) ? 1 : 0;
    a.extra_ndim = out->extra_ndim;
    a.extra_count = extra;
    a.spatial = spatial;
    for (int d = 0; d < out->extra_ndim; ++d) a.extra_dims[d] = (int64_t)out->extra_shape[d];
    for (int i = 0; i < n_inputs; ++i) {
        const pfh_array &in = inputs[i];
        if (!in.data) return set_err(PFH_ERR_INVALID_ARGUMENT, "input data", 0, 0, (size_t)i);
        if (in.data == out->data) {
            bool same = in.scalar_type == out->scalar_type;
            for (int d = 0; d < out->extra_ndim; ++d) same = same && in.extra_shape[d] == out->extra_shape[d];
            if (!same)
                return set_err(PFH_ERR_INVALID_ARGUMENT, "an input that is the output needs its type and extra shape",
                               0, 0, (size_t)i);
        }
        a.in[i] = in.data;
        a.in_complex[i] = complex_type(in.scalar_type) ? 1 : 0;
        // row-major strides of the input's own extra shape; 0 on broadcast axes
        int64_t stride = 1;
        for (int d = in.extra_ndim - 1; d >= 0; --d) {
            a.extra_stride[i][d] = in.extra_shape[d] == 1 ? 0 : stride;
            stride *= (int64_t)in.extra_shape[d];
        }
    }
    for (int c = 0; c < ncoef; ++c) {
        a.coef[2 * c] = coeffs[2 * c];
        a.coef[2 * c + 1] = coeffs[2 * c + 1];
    }
    PFH_HIP_CHECK(launch_pointwise(a, out->data, f64_type(out->scalar_type), reinterpret_cast<hipStream_t>(stream)));
    return ok();
}

} // extern "C"
