/*
 * CPU oracle for the single-GPU FFT path of pencil-rs (TEST INFRASTRUCTURE).
 *
 * Only tests/, the smoke check and the CPU-baseline leg of bench.py load the
 * library built from this directory, and only as the checker (or as the
 * timed CPU baseline). The product library never links it.
 *
 * Conventions follow the reference:
 *  - complex arrays are interleaved (re, im);
 *  - R2R kind codes are the reference's descriptor codes
 *    (crates/pencil-fft/src/r2r.rs:53-64): 1 DCT-I, 2 DCT-II, 3 DCT-III,
 *    4 DCT-IV, 5 DST-I, 6 DST-II, 7 DST-III, 8 DST-IV; 9 is the DHT
 *    (r2r.rs:101-106);
 *  - arrays of a plan are [extra][memory-ordered spatial axes], the memory
 *    order being the pencil permutation (pencil-array/src/lib.rs:12-16).
 *
 * Adjoint operations have no reference counterpart; they are the exact
 * matrix transposes of the forward operations under the real inner product
 * (complex values counted as two reals):
 *  - C2C: adjoint(forward) = raw backward, adjoint(backward) = forward,
 *    adjoint(inverse) = forward with the per-axis 1/n scale;
 *  - R2C: adjoint(forward) = C2R backward of D*y with D = 1/2 on interior
 *    half-spectrum bins (1 elsewhere), imaginary parts of DC/Nyquist
 *    ignored; adjoint(C2R backward) = W*R2C with W = 2 on interior bins and
 *    the imaginary parts of DC/Nyquist set to 0;
 *  - R2R: adjoint(DCT-II) = DCT-III(y with y_0*2); adjoint(DCT-III) =
 *    DCT-II(y) with out_0/2; adjoint(DST-II) = DST-III(y with y_{n-1}*2);
 *    adjoint(DST-III) = DST-II(y) with out_{n-1}/2; adjoint(DCT-I) =
 *    DCT-I(y with y_0, y_{n-1} * 2) with out_0, out_{n-1} / 2; DCT-IV,
 *    DST-I, DST-IV and the DHT are symmetric matrices (self-adjoint).
 */
#ifndef PENCIL_FFT_ORACLE_H
#define PENCIL_FFT_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFO_MAXD 4

enum { PFO_OK = 0, PFO_ERR_ARG = 1, PFO_ERR_INVALID_SPECTRUM = 2 };

enum { PFO_DCT1 = 1, PFO_DCT2, PFO_DCT3, PFO_DCT4, PFO_DST1, PFO_DST2, PFO_DST3, PFO_DST4, PFO_DHT };

/* PFO_KIND_C2R: the C2R orientation of the real route (C2rPlan,
 * distributed/c2r.rs:72-361, construct_oriented distributed/r2c.rs:1204-1448):
 * reduction axis = first selected axis, stage order and transitions
 * reversed; the complex side is in the canonical memory order. Its ops:
 * FORWARD = raw C2R, BACKWARD = raw R2C, INVERSE = R2C normalized by every
 * selected real extent (c2r.rs:340-360), and their adjoints. */
enum { PFO_KIND_C2C = 1, PFO_KIND_R2C = 2, PFO_KIND_R2R = 3, PFO_KIND_C2R = 4 };
/* Heterogeneous plans (distributed/mixed.rs:86-127, 392-441): one transform
 * per logical axis in r2r[a]: 0 none, PFO_DCT1..PFO_DHT, PFO_AXIS_FFT or
 * (mixed R2C only, exactly one axis) PFO_AXIS_RFFT. MIXED_C2C works on
 * complex data (an R2R axis transforms real and imaginary parts
 * separately); MIXED_R2C has the R2C boundary at the RFFT axis, real R2R or
 * identity stages on the axes after it and complex stages on the axes
 * before it (validate_r2c_graph, mixed.rs:1602-1629). */
enum { PFO_KIND_MIXED_C2C = 5, PFO_KIND_MIXED_R2C = 6 };
enum { PFO_AXIS_FFT = 10, PFO_AXIS_RFFT = 11 };

enum {
    PFO_OP_FORWARD = 0,
    PFO_OP_INVERSE = 1,
    PFO_OP_BACKWARD = 2,
    PFO_OP_ADJ_FORWARD = 3,
    PFO_OP_ADJ_INVERSE = 4,
    PFO_OP_ADJ_BACKWARD = 5
};

enum {
    PFO_LINE_COPY_C = 0,
    PFO_LINE_COPY_R,
    PFO_LINE_C2C,
    PFO_LINE_R2C,
    PFO_LINE_R2C_W,
    PFO_LINE_C2R,
    PFO_LINE_C2R_W,
    PFO_LINE_R2R,
    PFO_LINE_DHT
};

typedef struct {
    int kind;                    /* PFO_KIND_* */
    int ndim;                    /* spatial dimensions N (1..4) */
    size_t shape[PFO_MAXD];      /* global logical shape (the real shape for R2C) */
    size_t extra;                /* product of the extra shape (>= 1) */
    unsigned selected;           /* bit a set: axis a transformed (C2C / R2C) */
    int r2r[PFO_MAXD];           /* R2R: per-axis kind code, 0 = identity */
    int backward_dir[PFO_MAXD];  /* C2C: axis uses FourierDirection::Backward */
    int permute_dims;            /* DistributedLayout::permute_dims */
} pfo_desc;

/* Plan route: perm[s][pos] = logical axis at memory position pos of stage s. */
void pfo_build_route(int ndim, int permute_dims, int perm[PFO_MAXD][PFO_MAXD]);

/* Plan operations (whole route, 1x1 process grid). threads <= 0: all cores. */
int pfo_execute_f64(const pfo_desc *d, int op, const void *src, void *dst, int threads);
int pfo_execute_f32(const pfo_desc *d, int op, const void *src, void *dst, int threads);

/* Local batched line transforms on contiguous lines. op: FORWARD (negative
 * sign), BACKWARD (raw positive), INVERSE (backward * 1/n in the element
 * type). For R2C: FORWARD maps n reals to n/2+1 complex; BACKWARD/INVERSE
 * map back (DC / Nyquist imaginary parts ignored). */
int pfo_c2c_lines_f64(const double *src, double *dst, size_t n, size_t batch, int op);
int pfo_c2c_lines_f32(const float *src, float *dst, size_t n, size_t batch, int op);
int pfo_c2c_lines_f80(const double *src, double *dst, size_t n, size_t batch, int op);
int pfo_r2c_lines_f64(const double *src, double *dst, size_t n, size_t batch, int op);
int pfo_r2c_lines_f32(const float *src, float *dst, size_t n, size_t batch, int op);
int pfo_r2r_lines_f64(const double *src, double *dst, size_t n, size_t batch, int kind, int op);
int pfo_r2r_lines_f32(const float *src, float *dst, size_t n, size_t batch, int kind, int op);

/* Memory-axis permutation (LocalTransposePlan, local_transpose.rs:87-132,
 * 279-318): element size elem bytes, logical shape, extra blocks. */
int pfo_local_transpose(size_t elem, int ndim, const size_t *shape, size_t extra, const int *src_perm,
                        const int *dst_perm, const void *src, void *dst);

/* One-peer Alltoallv transpose: pack (payload order = global row-major per
 * extra block), self exchange, unpack. payload may be NULL; when not NULL
 * it receives the packed payload. */
int pfo_alltoallv_1x1(size_t elem, int ndim, const size_t *shape, size_t extra, const int *src_perm,
                      const int *dst_perm, const void *src, void *dst, void *payload);

/* Endpoint acceptance of a complex R2C boundary array (1 accepted, 0 not). */
int pfo_endpoint_accept_f64(int ndim, const int *perm, const size_t *rshape, size_t extra, int axis,
                            size_t real_len, const double *data, double relative, double absolute);
int pfo_endpoint_accept_f32(int ndim, const int *perm, const size_t *rshape, size_t extra, int axis,
                            size_t real_len, const float *data, double relative, double absolute);

/* C2R boundary thresholds of plan d for op (PFO_OP_INVERSE normalized,
 * PFO_OP_BACKWARD raw) in precision 4 (f32) or 8 (f64): relative,
 * absolute and the error-model depth (normalization_depth /
 * raw_absolute_threshold_for_shape, r2c.rs:4841-4878; MixedR2cPlan,
 * mixed.rs:400-427, :4085-4117). */
int pfo_endpoint_thresholds(const pfo_desc *d, int precision, int op, double *relative, double *absolute,
                            double *depth);
/* Error model of one MixedR2cPlan suffix axis: E_a, factor_a, allowance_a
 * (mixed.rs:408-417); 0 for the identity. */
int pfo_mixed_axis_model(int code, size_t n, size_t *E, size_t *factor, int *allowance);

#ifdef __cplusplus
}
#endif

#endif
