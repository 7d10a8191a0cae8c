/*
 * CPU oracle for the single-GPU FFT path of pencil-rs.
 *
 * TEST INFRASTRUCTURE ONLY. liboracle.so is loaded by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py, always as the checker or as
 * the timed CPU baseline, never as the measured or shipped path. The product
 * library (libpencilfft_hip.so) does not link or call it.
 *
 * Pinning (what checks this restatement, see tests/test_oracle.py):
 *  - C2C / R2C / C2R lines: direct O(n^2) long double DFT sums on lengths
 *    1..13 (the reference's dft_oracle, crates/pencil-fft/src/lib.rs:774-799,
 *    r2c.rs:961-1004), numpy FFT goldens, the peak/DC KATs (lib.rs:916-966,
 *    r2c.rs:1195-1245), Parseval and round trips;
 *  - R2R: the reference's direct definitions (r2r.rs:1064-1112) evaluated in
 *    long double, and scipy.fft.dct/dst goldens (FFTW convention);
 *  - DHT: the cas-sum definition (dht.rs:434-449) and Re F - Im F;
 *  - routes: numpy fftn/rfftn/ifftn goldens on the reference's fixture case
 *    list and input formulas (tools/fftw-reference/generate.jl:201-250);
 *  - transposes: integer-coded values (pencil-array/tests/local_transpose.rs
 *    :78-91) against numpy transposes, exhaustively over permutations;
 *  - endpoint statistics: hand-built spectra just inside / outside the
 *    thresholds of distributed/r2c.rs:4841-4878;
 *  - adjoints: dot-product identities <Ax, y> = <x, A*y>.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define PFO_PI_L 3.141592653589793238462643383279502884L

/* ---------------------------------------------------------------------- */
/* Small helpers shared by every scalar type                                 */
/* ---------------------------------------------------------------------- */

typedef void (*pfo_range_fn)(void *ctx, size_t begin, size_t end);

typedef struct {
    pfo_range_fn fn;
    void *ctx;
    size_t begin, end;
} pfo_task;

static void *pfo_thread_main(void *p)
{
    pfo_task *t = (pfo_task *)p;
    if (t->begin < t->end) t->fn(t->ctx, t->begin, t->end);
    return NULL;
}

static int pfo_default_threads(void)
{
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    const char *env = getenv("PFO_THREADS");
    if (env && atoi(env) > 0) n = atoi(env);
    if (n < 1) n = 1;
    if (n > 64) n = 64;
    return (int)n;
}

/* Runs fn over [0, count) split into contiguous ranges on up to `threads`
 * threads (threads <= 0: all cores). */
static void pfo_parallel_for(int threads, size_t count, pfo_range_fn fn, void *ctx)
{
    if (threads <= 0) threads = pfo_default_threads();
    if ((size_t)threads > count) threads = (int)(count ? count : 1);
    if (threads <= 1) {
        fn(ctx, 0, count);
        return;
    }
    pthread_t th[64];
    pfo_task tk[64];
    for (int i = 0; i < threads; ++i) {
        tk[i].fn = fn;
        tk[i].ctx = ctx;
        tk[i].begin = count * (size_t)i / (size_t)threads;
        tk[i].end = count * (size_t)(i + 1) / (size_t)threads;
        pthread_create(&th[i], NULL, pfo_thread_main, &tk[i]);
    }
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
}

static int pfo_r2r_backward_kind(int kind)
{
    switch (kind) {
    case PFO_DCT2: return PFO_DCT3;
    case PFO_DCT3: return PFO_DCT2;
    case PFO_DST2: return PFO_DST3;
    case PFO_DST3: return PFO_DST2;
    default: return kind;
    }
}

/* build_route, distributed.rs:2306-2373: from the identity input pencil, for
 * axis = N-2 down to 0 the next pencil removes `axis` from the current memory
 * order and appends it last (permute_dims), or keeps the identity. */
void pfo_build_route(int ndim, int permute_dims, int perm[PFO_MAXD][PFO_MAXD])
{
    for (int p = 0; p < ndim; ++p) perm[0][p] = p;
    int s = 1;
    for (int axis = ndim - 2; axis >= 0; --axis, ++s) {
        if (!permute_dims) {
            for (int p = 0; p < ndim; ++p) perm[s][p] = p;
            continue;
        }
        int q = 0;
        for (int p = 0; p < ndim; ++p)
            if (perm[s - 1][p] != axis) perm[s][q++] = perm[s - 1][p];
        perm[s][ndim - 1] = axis;
    }
}

/* Memory strides (in elements) of every logical axis for a permutation. */
static void pfo_strides(int N, const size_t *shape, const int *perm, size_t *stride)
{
    size_t run = 1;
    for (int p = N - 1; p >= 0; --p) {
        stride[perm[p]] = run;
        run *= shape[perm[p]];
    }
}

/* Region walk of the whole local region (transpose.rs:952-1067): visits the
 * logical row-major order; the last logical axis is one run. Copies between
 * storage in `perm` and payload order, per extra block. */
static void pfo_region_copy(int N, const size_t *shape, size_t extra, size_t elem, const int *perm,
                            const unsigned char *from, unsigned char *to, int to_payload)
{
    size_t stride[PFO_MAXD];
    pfo_strides(N, shape, perm, stride);
    size_t len = 1;
    for (int a = 0; a < N; ++a) len *= shape[a];
    if (len == 0) return;
    size_t outer = len / shape[N - 1];
    size_t run = shape[N - 1], run_stride = stride[N - 1];
    for (size_t e = 0; e < extra; ++e) {
        size_t counter[PFO_MAXD] = {0, 0, 0, 0};
        size_t offset = 0, pay = 0;
        for (size_t o = 0; o < outer; ++o) {
            for (size_t i = 0; i < run; ++i) {
                size_t so = (e * len + offset + i * run_stride) * elem;
                size_t po = (e * len + pay + i) * elem;
                if (to_payload) memcpy(to + po, from + so, elem);
                else memcpy(to + so, from + po, elem);
            }
            pay += run;
            /* advance_region, transpose.rs:1034-1048 */
            for (int a = N - 2; a >= 0; --a) {
                if (counter[a] + 1 < shape[a]) {
                    counter[a]++;
                    offset += stride[a];
                    break;
                }
                offset -= counter[a] * stride[a];
                counter[a] = 0;
            }
        }
    }
}

static void pfo_pack_region(int N, const size_t *shape, size_t extra, size_t elem, const int *perm,
                            const unsigned char *storage, unsigned char *payload)
{
    pfo_region_copy(N, shape, extra, elem, perm, storage, payload, 1);
}

static void pfo_unpack_region(int N, const size_t *shape, size_t extra, size_t elem, const int *perm,
                              const unsigned char *payload, unsigned char *storage)
{
    pfo_region_copy(N, shape, extra, elem, perm, payload, storage, 0);
}

/* ---------------------------------------------------------------------- */
/* C2R boundary acceptance thresholds                                       */
/* ---------------------------------------------------------------------- */

/* ceil(log2(n)) for n >= 1; 0 for n <= 1 (ceil_log2_len, mixed.rs:4076-4082;
 * usize::BITS - (n - 1).leading_zeros(), r2c.rs:4848-4850). */
static int pfo_ceil_log2(size_t n)
{
    int bits = 0;
    if (n <= 1) return 0;
    for (size_t v = n - 1; v; v >>= 1) ++bits;
    return bits;
}

/* Error-model length E_a, raw composition factor and allowance of one
 * complex-suffix axis of a MixedR2cPlan (mixed.rs:408-417, :4093-4108):
 * FFT (n, n, 0); R2R: E_a = 2(n-1) DCT-I, 2(n+1) DST-I, 4n II/III, 8n IV,
 * factor = the R2R normalization factor (2(n-1), 2(n+1), 2n: r2r.rs:1222-1228),
 * allowance 2; DHT (n, n, 2). Returns 0 for the identity. */
int pfo_mixed_axis_model(int code, size_t n, size_t *E, size_t *factor, int *allowance)
{
    *allowance = 2;
    switch (code) {
    case PFO_AXIS_FFT: *E = n; *factor = n; *allowance = 0; return 1;
    case PFO_DCT1: *E = 2 * (n - 1); *factor = 2 * (n - 1); return 1;
    case PFO_DST1: *E = 2 * (n + 1); *factor = 2 * (n + 1); return 1;
    case PFO_DCT2: case PFO_DCT3: case PFO_DST2: case PFO_DST3: *E = 4 * n; *factor = 2 * n; return 1;
    case PFO_DCT4: case PFO_DST4: *E = 8 * n; *factor = 2 * n; return 1;
    case PFO_DHT: *E = n; *factor = n; return 1;
    default: *E = 0; *factor = 1; *allowance = 0; return 0;
    }
}

/* Thresholds of the C2R boundary check of plan D for route operation op
 * (PFO_OP_INVERSE: normalized; PFO_OP_BACKWARD: raw), precision 4 (f32) or
 * 8 (f64):
 *  - R2C / C2R orientation: depth = 1 + sum ceil_log2(n_a) over the selected
 *    axes other than the reduction axis with n_a > 1 (normalization_depth,
 *    r2c.rs:4841-4853); raw absolute = inverse absolute * prod n_a over the
 *    same axes, extents of 1 included (raw_absolute_threshold_for_shape,
 *    r2c.rs:4855-4878);
 *  - MixedR2cPlan: depth = 1 + sum (ceil_log2(E_a) + allowance_a), raw
 *    absolute scaled by prod factor_a, over the non-identity axes before the
 *    RFFT axis (mixed_depth_normalization, mixed.rs:4085-4117);
 *  - relative = 128 * epsilon * depth, absolute = 128 * min_subnormal * depth
 *    with epsilon and the smallest positive subnormal of the real type taken
 *    to f64 (r2c.rs:4737-4742; lib.rs:255-281: f32::EPSILON, f32::from_bits(1)).
 * Returns PFO_ERR_ARG for a plan without a real boundary. */
int pfo_endpoint_thresholds(const pfo_desc *D, int precision, int op, double *relative, double *absolute,
                            double *depth_out)
{
    const int N = D->ndim;
    int raxis = -1;
    if (N < 1 || N > PFO_MAXD) return PFO_ERR_ARG;
    if (D->kind == PFO_KIND_R2C) {
        for (int a = N - 1; a >= 0; --a) if (D->selected & (1u << a)) { raxis = a; break; }
    } else if (D->kind == PFO_KIND_C2R) {
        for (int a = 0; a < N; ++a) if (D->selected & (1u << a)) { raxis = a; break; }
    } else if (D->kind == PFO_KIND_MIXED_R2C) {
        for (int a = 0; a < N; ++a) if (D->r2r[a] == PFO_AXIS_RFFT) { raxis = a; break; }
    }
    if (raxis < 0) return PFO_ERR_ARG;
    double depth = 1.0, transverse = 1.0;
    if (D->kind == PFO_KIND_MIXED_R2C) {
        for (int a = 0; a < raxis; ++a) {
            size_t E, f;
            int allow;
            if (!pfo_mixed_axis_model(D->r2r[a], D->shape[a], &E, &f, &allow)) continue;
            depth += pfo_ceil_log2(E) + allow;
            transverse *= (double)f;
        }
    } else {
        for (int a = 0; a < N; ++a) {
            if (a == raxis || !(D->selected & (1u << a))) continue;
            depth += pfo_ceil_log2(D->shape[a]);
            transverse *= (double)D->shape[a];
        }
    }
    /* f32::EPSILON = 2^-23, f32::from_bits(1) = 2^-149; f64: 2^-52, 2^-1074 */
    const double eps = precision == 4 ? ldexp(1.0, -23) : ldexp(1.0, -52);
    const double minsub = precision == 4 ? ldexp(1.0, -149) : ldexp(1.0, -1074);
    *relative = 128.0 * eps * depth;
    *absolute = 128.0 * minsub * depth;
    if (op == PFO_OP_BACKWARD) *absolute *= transverse;
    if (depth_out) *depth_out = depth;
    return PFO_OK;
}

/* ---------------------------------------------------------------------- */
/* Typed instantiations                                                      */
/* ---------------------------------------------------------------------- */

#define REAL double
#define SFX f64
#include "pfo_fft.inc"
#include "pfo_route.inc"
#undef REAL
#undef SFX

#define REAL float
#define SFX f32
#include "pfo_fft.inc"
#include "pfo_route.inc"
#undef REAL
#undef SFX

#define REAL long double
#define SFX f80
#include "pfo_fft.inc"
#undef REAL
#undef SFX

/* ---------------------------------------------------------------------- */
/* Public API                                                                */
/* ---------------------------------------------------------------------- */

static int pfo_check_desc(const pfo_desc *d)
{
    if (!d || d->ndim < 1 || d->ndim > PFO_MAXD || d->extra < 1) return PFO_ERR_ARG;
    for (int a = 0; a < d->ndim; ++a)
        if (d->shape[a] == 0) return PFO_ERR_ARG;
    return PFO_OK;
}

int pfo_execute_f64(const pfo_desc *d, int op, const void *src, void *dst, int threads)
{
    int st = pfo_check_desc(d);
    if (st) return st;
    return pfo_execute_impl_f64(d, op, src, dst, threads);
}

int pfo_execute_f32(const pfo_desc *d, int op, const void *src, void *dst, int threads)
{
    int st = pfo_check_desc(d);
    if (st) return st;
    return pfo_execute_impl_f32(d, op, src, dst, threads);
}

int pfo_c2c_lines_f64(const double *src, double *dst, size_t n, size_t batch, int op)
{
    if (n == 0) return PFO_ERR_ARG;
    return pfo_run_lines_f64(PFO_LINE_C2C, 0, op, n, batch, src, dst);
}

int pfo_c2c_lines_f32(const float *src, float *dst, size_t n, size_t batch, int op)
{
    if (n == 0) return PFO_ERR_ARG;
    return pfo_run_lines_f32(PFO_LINE_C2C, 0, op, n, batch, src, dst);
}

/* Local plans validate the endpoints strictly (r2c.rs:798-826); the oracle
 * leaves that to the caller and treats them as zero. */
int pfo_r2c_lines_f64(const double *src, double *dst, size_t n, size_t batch, int op)
{
    if (n == 0) return PFO_ERR_ARG;
    return pfo_run_lines_f64(PFO_LINE_R2C, 0, op, n, batch, src, dst);
}

int pfo_r2c_lines_f32(const float *src, float *dst, size_t n, size_t batch, int op)
{
    if (n == 0) return PFO_ERR_ARG;
    return pfo_run_lines_f32(PFO_LINE_R2C, 0, op, n, batch, src, dst);
}

int pfo_r2r_lines_f64(const double *src, double *dst, size_t n, size_t batch, int kind, int op)
{
    if (kind < PFO_DCT1 || kind > PFO_DHT || n == 0) return PFO_ERR_ARG;
    if (kind == PFO_DCT1 && n < 2) return PFO_ERR_ARG;
    return pfo_run_lines_f64(PFO_LINE_R2R, kind, op, n, batch, src, dst);
}

int pfo_r2r_lines_f32(const float *src, float *dst, size_t n, size_t batch, int kind, int op)
{
    if (kind < PFO_DCT1 || kind > PFO_DHT || n == 0) return PFO_ERR_ARG;
    if (kind == PFO_DCT1 && n < 2) return PFO_ERR_ARG;
    return pfo_run_lines_f32(PFO_LINE_R2R, kind, op, n, batch, src, dst);
}

/* Extended-precision C2C lines (long double arithmetic, double storage),
 * used to pin the double oracle on small lengths. */
int pfo_c2c_lines_f80(const double *src, double *dst, size_t n, size_t batch, int op)
{
    if (n == 0) return PFO_ERR_ARG;
    pfo_plan_f80 p;
    if (pfo_plan_init_f80(&p, n)) return PFO_ERR_ARG;
    long double *a = (long double *)malloc(sizeof(long double) * 2 * n);
    long double *b = (long double *)malloc(sizeof(long double) * 2 * n);
    for (size_t l = 0; l < batch; ++l) {
        for (size_t j = 0; j < 2 * n; ++j) a[j] = src[2 * n * l + j];
        pfo_fft_line_f80(&p, a, 1, b, op != PFO_OP_FORWARD);
        for (size_t j = 0; j < 2 * n; ++j) {
            long double v = b[j];
            if (op == PFO_OP_INVERSE) v = v / (long double)n;
            dst[2 * n * l + j] = (double)v;
        }
    }
    free(a);
    free(b);
    pfo_plan_free_f80(&p);
    return PFO_OK;
}

/* logical_offsets, local_transpose.rs:279-318, applied element by element. */
int pfo_local_transpose(size_t elem, int ndim, const size_t *shape, size_t extra, const int *src_perm,
                        const int *dst_perm, const void *src, void *dst)
{
    if (ndim < 1 || ndim > PFO_MAXD) return PFO_ERR_ARG;
    size_t ss[PFO_MAXD], ds[PFO_MAXD];
    pfo_strides(ndim, shape, src_perm, ss);
    pfo_strides(ndim, shape, dst_perm, ds);
    size_t len = 1;
    for (int a = 0; a < ndim; ++a) len *= shape[a];
    const unsigned char *s = (const unsigned char *)src;
    unsigned char *d = (unsigned char *)dst;
    for (size_t linear = 0; linear < len * extra; ++linear) {
        size_t spatial = linear % len, e = linear / len, rem = spatial;
        size_t so = 0, dof = 0;
        for (int a = ndim - 1; a >= 0; --a) {
            size_t idx = rem % shape[a];
            rem /= shape[a];
            so += idx * ss[a];
            dof += idx * ds[a];
        }
        memcpy(d + (e * len + dof) * elem, s + (e * len + so) * elem, elem);
    }
    return PFO_OK;
}

int pfo_alltoallv_1x1(size_t elem, int ndim, const size_t *shape, size_t extra, const int *src_perm,
                      const int *dst_perm, const void *src, void *dst, void *payload)
{
    if (ndim < 1 || ndim > PFO_MAXD) return PFO_ERR_ARG;
    size_t len = 1;
    for (int a = 0; a < ndim; ++a) len *= shape[a];
    unsigned char *send = (unsigned char *)malloc(len * extra * elem + 1);
    unsigned char *recv = (unsigned char *)malloc(len * extra * elem + 1);
    pfo_pack_region(ndim, shape, extra, elem, src_perm, (const unsigned char *)src, send);
    memcpy(recv, send, len * extra * elem);
    if (payload) memcpy(payload, send, len * extra * elem);
    pfo_unpack_region(ndim, shape, extra, elem, dst_perm, recv, (unsigned char *)dst);
    free(send);
    free(recv);
    return PFO_OK;
}

int pfo_endpoint_accept_f64(int ndim, const int *perm, const size_t *rshape, size_t extra, int axis,
                            size_t real_len, const double *data, double relative, double absolute)
{
    return pfo_endpoint_accept_impl_f64(ndim, perm, rshape, extra, axis, real_len, data, relative, absolute);
}

int pfo_endpoint_accept_f32(int ndim, const int *perm, const size_t *rshape, size_t extra, int axis,
                            size_t real_len, const float *data, double relative, double absolute)
{
    return pfo_endpoint_accept_impl_f32(ndim, perm, rshape, extra, axis, real_len, data, relative, absolute);
}
