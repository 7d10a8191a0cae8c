/*
 * pencil_fft_hip.h - C ABI of the MI355X (gfx950) single-GPU FFT path of
 * pencil-rs.
 *
 * The reference is Rust; its plugin seam for this path is the backend
 * selection of crates/pencil-fft/src/backend.rs:12-131 (BackendKind,
 * LocalBackend, per-line trait objects) plus the stage and transition
 * executors of crates/pencil-fft/src/distributed.rs:3126-3867. A per-line
 * trait object would force a host round trip per line, so this library
 * replaces whole stages and transitions: one call runs a full plan operation
 * on device-resident buffers. Every entry point below names the reference
 * interface it replaces.
 *
 * Conventions
 *  - All data pointers are device pointers (hipMalloc'd or equivalent).
 *  - Complex values are interleaved (re, im), like num_complex::Complex.
 *  - Arrays are laid out [extra...][memory-ordered spatial axes], the memory
 *    order being the pencil permutation (crates/pencil-array/src/lib.rs:12-16).
 *  - The process grid is one rank (a 1x1 / 1 topology); every route
 *    transition is therefore a local permutation fused into a stage kernel.
 *  - Validation completes before any device write: a rejected call leaves
 *    src, dst and workspace untouched (crates/pencil-fft/src/lib.rs:1015-1077).
 *  - Execution is stream ordered. Errors found on the device (an invalid C2R
 *    spectrum) gate the writing kernels on the device, so dst is not written,
 *    and are reported synchronously unless PFH_EXEC_DEFER_STATUS is passed;
 *    deferred status is read with pfh_workspace_take_status.
 */
#ifndef PENCIL_FFT_HIP_H
#define PENCIL_FFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *pfh_stream_t; /* hipStream_t */

typedef struct pfh_plan_s *pfh_plan;

/* Status codes. Each maps to one variant of the reference's error enums:
 * LocalC2cError (lib.rs:337-360), LocalR2cError (r2c.rs:24-109),
 * LocalR2rError (r2r.rs:204-235), FftError (distributed.rs:338-418),
 * R2cError (distributed.rs:440-455), LocalTransposeError
 * (local_transpose.rs:12-44) and TransposeError (transpose.rs:112-175). */
typedef enum pfh_status {
    PFH_OK = 0,
    PFH_ERR_INVALID_LENGTH = 1,           /* Local*Error::InvalidLength */
    PFH_ERR_LENGTH_OVERFLOW = 2,          /* Local*Error::LengthOverflow */
    PFH_ERR_NON_INTEGRAL_BATCH = 3,       /* Local*Error::NonIntegralBatch */
    PFH_ERR_BUFFER_LENGTH_MISMATCH = 4,   /* LocalC2cError/LocalR2rError::BufferLengthMismatch */
    PFH_ERR_BATCH_COUNT_MISMATCH = 5,     /* LocalR2cError::BatchCountMismatch {source_count, destination_count} */
    PFH_ERR_SCRATCH_TOO_SMALL = 6,        /* ScratchTooSmall {required, actual} */
    PFH_ERR_INVALID_SPECTRUM_ENDPOINT = 7,/* LocalR2cError::InvalidSpectrumEndpoint {batch, index} */
    PFH_ERR_INVALID_DIMENSIONS = 8,       /* FftError::InvalidDimensions */
    PFH_ERR_INVALID_INPUT_LAYOUT = 9,     /* FftError::InvalidInputLayout */
    PFH_ERR_INPUT_LAYOUT_MISMATCH = 10,   /* FftError::InputLayoutMismatch */
    PFH_ERR_OUTPUT_LAYOUT_MISMATCH = 11,  /* FftError::OutputLayoutMismatch */
    PFH_ERR_EXTRA_SHAPE_MISMATCH = 12,    /* FftError::ExtraShapeMismatch */
    PFH_ERR_WORKSPACE_MISMATCH = 13,      /* FftError::WorkspaceMismatch */
    PFH_ERR_WORKSPACE_TOO_SMALL = 14,     /* FftError::WorkspaceTooSmall {required, actual} */
    PFH_ERR_PREPARATION_FAILED = 15,      /* FftError::PreparationFailed */
    PFH_ERR_ALLOCATION_FAILED = 16,       /* FftError::AllocationFailed */
    PFH_ERR_INVALID_SPECTRUM = 17,        /* R2cError::InvalidSpectrum */
    PFH_ERR_INCOMPATIBLE_DISTRIBUTION = 18, /* LocalTransposeError::IncompatibleDistribution */
    PFH_ERR_SOURCE_LAYOUT_MISMATCH = 19,  /* LocalTransposeError::SourceLayoutMismatch */
    PFH_ERR_DESTINATION_LAYOUT_MISMATCH = 20, /* LocalTransposeError::DestinationLayoutMismatch */
    PFH_ERR_INVALID_PERMUTATION = 21,     /* AxisError::{OutOfBounds, Duplicate} (axis.rs:40-66) */
    PFH_ERR_ZERO_EXTENT = 22,             /* PencilError::ZeroGlobalExtent */
    PFH_ERR_INVALID_ARGUMENT = 23,        /* null pointer / unknown enum value (no Rust analogue: types forbid it) */
    PFH_ERR_UNSUPPORTED = 24,             /* a shape the device kernels do not cover (more than 2^31 - 1
                                             lines in one stage, more than 5 line digits, Bluestein
                                             lengths above 2^25, batched long lines) */
    PFH_ERR_DEVICE = 25,                  /* a HIP runtime call failed */
    PFH_ERR_INVALID_GRAPH = 26,           /* MixedError::InvalidGraph (mixed.rs:118-120) */
    /* pointwise validation (pointwise.rs:8-33, :216-230); batch = the failing
     * input's index (MultiInputError::Input { index, .. }) */
    PFH_ERR_INCOMPATIBLE_PENCILS = 27,    /* PointwiseError::IncompatiblePencils */
    PFH_ERR_EXTRA_RANK_MISMATCH = 28,     /* PointwiseError::ExtraRankMismatch {expected = required, actual} */
    PFH_ERR_EXTRA_EXTENT_MISMATCH = 29,   /* PointwiseError::ExtraExtentMismatch {axis = index, output = required, input = actual} */
    PFH_ERR_EMPTY_INPUTS = 30,            /* MultiInputError::EmptyInputs */
    /* on-file formats: IoError (crates/pencil-io/src/lib.rs:196-300); the
     * message carries the variant's string field (reason / field / what /
     * stage / operation), actual carries errno for PFH_ERR_IO_NATIVE */
    PFH_ERR_IO_INVALID_INPUT = 31,        /* IoError::InvalidInput */
    PFH_ERR_IO_INVALID_PATH = 32,         /* IoError::InvalidPath */
    PFH_ERR_IO_SIZE_LIMIT = 33,           /* IoError::SizeLimit {what} */
    PFH_ERR_IO_METADATA_MISMATCH = 34,    /* IoError::MetadataMismatch {field} */
    PFH_ERR_IO_INVALID_FILE = 35,         /* IoError::InvalidFile {reason} */
    PFH_ERR_IO_INCOMPLETE_FILE = 36,      /* IoError::IncompleteFile */
    PFH_ERR_IO_NATIVE = 37,               /* IoError::Native {operation, code} (the reference's MPI_File_open failures) */
    PFH_ERR_IO_WRITE_INCOMPLETE = 38,     /* IoError::WriteIncomplete {stage} */
    PFH_ERR_IO_COMMIT_UNCERTAIN = 39,     /* IoError::CommitUncertain {stage} */
    /* named containers: NamedIoError (crates/pencil-io/src/lib.rs:302-322) */
    PFH_ERR_IO_INVALID_NAME = 40,         /* NamedIoError::InvalidName */
    PFH_ERR_IO_DUPLICATE_NAME = 41,       /* NamedIoError::DuplicateName */
    PFH_ERR_IO_NOT_FOUND = 42,            /* NamedIoError::NotFound */
    PFH_ERR_IO_INVALID_TAIL = 43,         /* NamedIoError::InvalidTail */
    PFH_ERR_POISONED = 44,                /* ArrayError::Poisoned (pencil-array error.rs:158, many.rs:234) */
    PFH_ERR_ALIASED_BUFFERS = 45          /* overlapping source / destination / workspace (no Rust analogue:
                                             the borrow rules forbid it) */
} pfh_status;

/* Payload of the last error on the calling thread (Rust error fields). */
typedef struct pfh_error_info {
    pfh_status status;
    size_t required;   /* ScratchTooSmall / WorkspaceTooSmall */
    size_t actual;
    size_t batch;      /* InvalidSpectrumEndpoint / BatchCountMismatch source */
    size_t index;      /* InvalidSpectrumEndpoint / BatchCountMismatch destination */
    int hip_error;     /* hipError_t for PFH_ERR_DEVICE */
    char message[160];
} pfh_error_info;

typedef enum pfh_precision { PFH_F32 = 1, PFH_F64 = 2 } pfh_precision;

typedef enum pfh_kind {
    /* local line plans: contiguous lines, batch = buffer length / line length */
    PFH_LOCAL_C2C = 1,  /* LocalC2cPlan       lib.rs:367-655 */
    PFH_LOCAL_R2C = 2,  /* LocalR2cPlan       r2c.rs:260-813 */
    PFH_LOCAL_R2R = 3,  /* LocalR2rPlan       r2r.rs:287-778 */
    PFH_LOCAL_DHT = 4,  /* LocalDhtPlan       dht.rs:32-303 */
    /* distributed plans on a one-rank process grid */
    PFH_C2C = 5,        /* C2cPlan            distributed.rs:1016-1800 */
    PFH_R2C = 6,        /* R2cPlan            distributed/r2c.rs:255-1130 */
    PFH_R2R = 7,        /* R2rPlan / DhtPlan  distributed/r2r.rs:71-1460 */
    /* layout permutations */
    PFH_LOCAL_TRANSPOSE = 8,    /* LocalTransposePlan       local_transpose.rs:46-277 */
    PFH_ALLTOALLV_TRANSPOSE = 9, /* AllToAllvTransposePlan  alltoallv_transpose.rs:50-311 (one peer) */
    /* C2rPlan (distributed/c2r.rs:72-361): the real route in C2R orientation
     * (construct_oriented, distributed/r2c.rs:1204-1448). shape is the real
     * shape; the reduction axis is the first selected axis; the complex side
     * is in the canonical memory order and the real side in the route's last
     * order. FORWARD = raw C2R (with the endpoint-plane check), BACKWARD = raw
     * R2C, INVERSE = R2C normalized by every selected real extent. Out of
     * place only, like the reference. */
    PFH_C2R = 10,
    /* MixedC2cPlan / MixedR2cPlan (distributed/mixed.rs:86-127, 392-441):
     * one transform per axis in axis_kind[a] (pfh_axis_transform). MIXED_C2C
     * works on complex data, an R2R axis transforming the real and imaginary
     * parts separately; MIXED_R2C has exactly one PFH_AXIS_RFFT axis (the R2C
     * boundary), R2R or identity on the axes after it and any complex
     * transform before it (validate_r2c_graph, mixed.rs:1602-1629); other
     * graphs give PFH_ERR_INVALID_GRAPH. direction[a] applies to FFT axes. */
    PFH_MIXED_C2C = 11,
    PFH_MIXED_R2C = 12
} pfh_kind;

/* Axis transforms of the mixed plans (AxisTransform, mixed.rs:86-101): the
 * R2R kind codes, plus the complex FFT and the real-to-half-complex axis. */
enum { PFH_AXIS_NONE = 0, PFH_AXIS_FFT = 10, PFH_AXIS_RFFT = 11 };

/* R2R kind codes = the reference's descriptor codes (r2r.rs:53-64, :101-106). */
typedef enum pfh_r2r_kind {
    PFH_IDENTITY = 0,
    PFH_DCT1 = 1, PFH_DCT2 = 2, PFH_DCT3 = 3, PFH_DCT4 = 4,
    PFH_DST1 = 5, PFH_DST2 = 6, PFH_DST3 = 7, PFH_DST4 = 8,
    PFH_DHT = 9
} pfh_r2r_kind;

/* Plan operations. FORWARD / INVERSE / BACKWARD are the reference's
 * forward, inverse (normalized) and backward (raw) methods. The adjoint
 * operations have no reference counterpart: each is the exact transpose of
 * the named operation under the real inner product (complex values count as
 * two reals). */
typedef enum pfh_op {
    PFH_FORWARD = 0,
    PFH_INVERSE = 1,
    PFH_BACKWARD = 2,
    PFH_ADJOINT_FORWARD = 3,  /* adjoint of FORWARD  (output-layout input) */
    PFH_ADJOINT_INVERSE = 4,  /* adjoint of INVERSE  (input-layout input) */
    PFH_ADJOINT_BACKWARD = 5  /* adjoint of BACKWARD (input-layout input) */
} pfh_op;

enum {
    PFH_EXEC_DEFER_STATUS = 1 /* do not synchronize for device-found errors */
};

#define PFH_MAX_DIMS 4

/* Plan descriptor. Fields unused by a kind are ignored. */
typedef struct pfh_desc {
    int kind;                       /* pfh_kind */
    int precision;                  /* pfh_precision */
    int ndim;                       /* N: spatial dimensions */
    int topology_dims;              /* M: MpiTopology<M> (all extents 1); 0 = N-1 */
    size_t shape[PFH_MAX_DIMS];     /* global logical shape (the real shape for R2C) */
    int extra_ndim;
    size_t extra_shape[PFH_MAX_DIMS];
    unsigned selected;              /* AxisSelection mask (bit a = axis a); 0 = all */
    int axis_kind[PFH_MAX_DIMS];    /* R2R: pfh_r2r_kind per axis */
    int direction[PFH_MAX_DIMS];    /* C2C: 0 FourierDirection::Forward, 1 Backward */
    int permute_dims;               /* DistributedLayout::permute_dims */
    int transpose_method;           /* TransposeMethod: 0 AllToAllv, 1 PointToPoint */
    int src_perm[PFH_MAX_DIMS];     /* transposes: source memory order */
    int dst_perm[PFH_MAX_DIMS];     /* transposes: destination memory order */
    size_t element_bytes;           /* transposes: bytes per element (any size) */
    size_t line_len;                /* local plans: line length (real length for R2C) */
    int r2r_kind;                   /* PFH_LOCAL_R2R: kind (1..8) */
    int flags;                      /* reserved, 0 */
} pfh_desc;

/* Stage timing (TransformTiming / StageTiming, distributed.rs:290-325). */
typedef struct pfh_transform_timing {
    int stages;              /* number of entries below */
    double stage_ms[12];     /* device time of each entry (sum over its launches) */
    int stage_axis[12];      /* transformed logical axis; -1 for an endpoint check */
    double stage_bytes[12];  /* algorithmic bytes of the entry (one read of its input array + one write of
                                its output array) */
    int stage_launches[12];  /* kernel launches of the entry (a stage run slab by slab has several) */
    double total_ms;
} pfh_transform_timing;

const char *pfh_version(void);
void pfh_last_error(pfh_error_info *out);
const char *pfh_status_string(pfh_status status);

/* Plan construction: LocalC2cPlan::new (lib.rs:391-438), LocalR2cPlan::new
 * (r2c.rs:260-400), LocalR2rPlan::new (r2r.rs:471-495), LocalDhtPlan::new
 * (dht.rs:53-68), C2cPlan::from_shape_with_layout (distributed.rs:1016-1300),
 * R2cPlan::from_shape* (distributed/r2c.rs:255-420, construct_oriented
 * :1204-1448), R2rPlan/DhtPlan::from_shape (distributed/r2r.rs:71-360),
 * LocalTransposePlan::new (local_transpose.rs:59-71),
 * AllToAllvTransposePlan::new (alltoallv_transpose.rs:50-106). Twiddle
 * tables are built here (FP64-generated, rounded to the plan precision) and
 * uploaded to the current device; the plan is immutable afterwards. */
pfh_status pfh_plan_create(const pfh_desc *desc, pfh_plan *out);
void pfh_plan_destroy(pfh_plan plan);

/* Workspace bytes needed by op (TransposeWorkspaceRequirements,
 * transpose.rs:226-254, scratch_len / embedding_len of the local plans). The
 * workspace also holds the plan's small device status area: fill a new
 * workspace with zeros once before its first use, or call
 * pfh_workspace_init (the library keeps the status area consistent
 * afterwards). Besides the route's intermediate arrays it holds the slab
 * intermediate of slab-pipelined stage pairs, which are off unless
 * PFH_SLAB_MB in the environment sets a slab size in MiB.
 *
 * Binding (the reference's workspaces belong to the plan that made them,
 * distributed.rs:2188-2191, :4015-4017): a distributed plan's workspace is
 * bound to the first plan that uses it (or to the plan given to
 * pfh_workspace_init); any other plan then gets PFH_ERR_WORKSPACE_MISMATCH
 * before any device work. The binding is recorded in the workspace itself, so
 * a workspace refilled with zeros (a new allocation) is unbound again. */
pfh_status pfh_plan_workspace_bytes(pfh_plan plan, int op, size_t *bytes);

/* Binds a workspace to plan and resets its status area: clears the sticky
 * status and the in-place poisoned state (the reference's equivalent is a new
 * workspace and a new in-place array from plan.workspace() / the array
 * constructors). workspace_bytes must cover the status area (256 bytes);
 * stream ordered, no synchronization. Distributed plans only. */
pfh_status pfh_workspace_init(pfh_plan plan, void *workspace, size_t workspace_bytes, pfh_stream_t stream);

/* Layout of the plan's input (side 0) or output (side 1) array for the
 * forward direction: input_pencil() / output_pencil()
 * (distributed.rs:1316-1329). perm: memory order; shape: logical shape;
 * elements: elements per array (extras included); element_bytes: bytes per
 * element. */
pfh_status pfh_plan_layout(pfh_plan plan, int side, int perm[PFH_MAX_DIMS], size_t shape[PFH_MAX_DIMS],
                           size_t *elements, size_t *element_bytes);

/* Distributed-plan execution, out of place, source preserved (src, dst and
 * the workspace must not overlap: PFH_ERR_ALIASED_BUFFERS):
 * C2cPlan::{forward, inverse, backward} (distributed.rs:1448-1500),
 * R2cPlan::{forward, inverse, backward} (distributed/r2c.rs:856-1000),
 * R2rPlan/DhtPlan::{forward, inverse, backward} (distributed/r2r.rs:1271+),
 * LocalTransposePlan::execute_views (local_transpose.rs:87-132),
 * AllToAllvTransposePlan::execute_views (alltoallv_transpose.rs:107-160),
 * plus the adjoint operations. src_elems / dst_elems are the element counts
 * of the caller's buffers (checked like the Rust slice lengths). */
pfh_status pfh_execute(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst, size_t dst_elems,
                       void *workspace, size_t workspace_bytes, pfh_stream_t stream, int flags);

/* In-place execution in one buffer of capacity_bytes (C2cPlan::*_in_place,
 * distributed.rs:3672-3774; R2cPlan in-place arrays; local *_in_place
 * methods; LocalTransposePlan::execute_in_place, local_transpose.rs:159-243).
 *
 * Failure state (run_in_place_transaction, distributed.rs:3776-3794): the
 * reference's in-place array state lives here in the workspace. A call that
 * fails after it started writing `data` (an inverse whose spectrum the C2R
 * boundary check rejects: the complex stages before the check have already
 * overwritten data) poisons the workspace; every later in-place call with it
 * returns PFH_ERR_POISONED and writes nothing, until pfh_workspace_init
 * resets it. With PFH_EXEC_DEFER_STATUS the failure is reported by
 * pfh_workspace_take_status (or by the next synchronous in-place call), and
 * in-place calls issued before that write nothing. Out-of-place calls are
 * not affected by the poisoned state. */
pfh_status pfh_execute_in_place(pfh_plan plan, int op, void *data, size_t capacity_bytes, void *workspace,
                                size_t workspace_bytes, pfh_stream_t stream, int flags);

/* Local line plans: LocalC2cPlan::{forward, backward, inverse}
 * (lib.rs:498-619), LocalR2cPlan::{forward, inverse, backward}
 * (r2c.rs:428-521), LocalR2rPlan / LocalDhtPlan::{forward, backward,
 * inverse} (r2r.rs:530-621, dht.rs:133-200). The batch is
 * src_elems / (input line length); BatchCountMismatch / NonIntegralBatch /
 * BufferLengthMismatch are checked as in the reference. Local plans need no
 * workspace except C2R (its endpoint preflight status area) and plans whose
 * lines run as pieces (lines too long for one workgroup, or whose FFT length
 * has a prime factor above 128: a four-step split or Bluestein's chirp
 * convolution through workspace regions; such plans run their batch in
 * chunks of lines, so pfh_plan_workspace_bytes does not depend on the
 * batch). Any length the reference's LocalC2cPlan::new accepts
 * (lib.rs:422-438) plans. */
pfh_status pfh_local_execute(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst,
                             size_t dst_elems, void *workspace, size_t workspace_bytes, pfh_stream_t stream,
                             int flags);

/* In-place local execution: LocalC2cPlan::*_in_place (lib.rs:556-619),
 * LocalR2cPlan::{forward,inverse,backward}_in_place on the packed array
 * (r2c.rs:532-623), LocalR2rPlan / LocalDhtPlan::*_in_place. elems is the
 * number of input elements (reals for an R2C forward, complex values for a
 * C2R); capacity_bytes must hold the larger of input and output. R2C / C2R
 * stage their input in the workspace (status area + input bytes). */
pfh_status pfh_local_execute_in_place(pfh_plan plan, int op, void *data, size_t elems, size_t capacity_bytes,
                                      void *workspace, size_t workspace_bytes, pfh_stream_t stream, int flags);

/* *_with_timing (distributed.rs:290-325, 3796-3867): same as pfh_execute,
 * synchronous, with per-step device times from HIP events. */
pfh_status pfh_execute_with_timing(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst,
                                   size_t dst_elems, void *workspace, size_t workspace_bytes, pfh_stream_t stream,
                                   pfh_transform_timing *timing);

/* Reads and clears the sticky device status of a workspace after
 * PFH_EXEC_DEFER_STATUS executions (synchronizes the stream). */
pfh_status pfh_workspace_take_status(pfh_plan plan, void *workspace, pfh_stream_t stream);

/* Allocation-free pfh_execute. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The operation runs
 * on a workspace inside the library, a static device array of 128 MiB per
 * GPU (held until the process ends), instead of a caller workspace; it calls
 * no hipMalloc / hipFree. Distributed plans only; an operation whose
 * workspace (pfh_plan_workspace_bytes) exceeds 128 MiB gives
 * PFH_ERR_WORKSPACE_TOO_SMALL {required, actual = 128 MiB}. flags must be 0:
 * a rejected C2R spectrum is reported synchronously (PFH_ERR_INVALID_SPECTRUM,
 * dst untouched). All calls on one GPU share its array: calls that overlap
 * on one GPU (concurrent host threads, or unordered streams) race. */
pfh_status pfh_execute_noalloc(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst,
                               size_t dst_elems, pfh_stream_t stream, int flags);

/* ---------------------------------------------------------------------- */
/* Batched uneven transforms (no reference counterpart; the reference runs */
/* collections sequentially, distributed/collections.rs:229-311).          */
/* ---------------------------------------------------------------------- */

typedef struct pfh_batch_s *pfh_batch;

/* Builds one persistent launch per stage over `count` independent C2C
 * plans (PFH_C2C descriptors selecting every axis), all of one precision,
 * uneven shapes allowed; ops PFH_FORWARD, PFH_INVERSE, PFH_BACKWARD.
 * Tiles, radix split and lines per block come from the autotuner when
 * autotune != 0, else from the planner's defaults.
 *
 * The autotuner times every candidate tuning on the current device and
 * keeps the fastest under a key of the descriptors, the precision and the
 * device (architecture name, CU count, LDS per workgroup): in the process,
 * and appended to a cache file that later processes read, so a batch of the
 * same problems on the same kind of device is measured once. The file is
 * the environment's PFH_TUNE_CACHE (empty, "0" or "off": the process only),
 * else $XDG_CACHE_HOME/pencil-fft-hip/batch_tuning.txt, else
 * $HOME/.cache/pencil-fft-hip/batch_tuning.txt; a file that cannot be read
 * or written means measuring again. */
pfh_status pfh_batch_create(const pfh_desc *descs, int count, int autotune, pfh_batch *out);

/* Where a batch's tuning came from (pfh_batch_tuning_source). */
enum {
    PFH_TUNED_DEFAULT = 0,       /* autotune == 0: the planner's defaults */
    PFH_TUNED_MEASURED = 1,      /* timed on the device by this creation */
    PFH_TUNED_PROCESS_CACHE = 2, /* measured earlier in this process */
    PFH_TUNED_FILE_CACHE = 3     /* read from the cache file */
};
pfh_status pfh_batch_tuning_source(pfh_batch batch, int *source);
/* Forgets the tunings measured or read in this process (the cache file is
 * kept): the next autotuned creation reads the file or measures again. */
void pfh_batch_tuning_cache_clear(void);
void pfh_batch_destroy(pfh_batch batch);
pfh_status pfh_batch_workspace_bytes(pfh_batch batch, size_t *bytes);
/* srcs[i] / dsts[i]: device buffers of problem i (input / output layout of
 * its plan). use_graph != 0 captures the launches once into a hipGraph for
 * these pointers and stream and replays it (a graph asked for on the legacy
 * null stream runs on a stream of the batch, ordered after the caller's
 * earlier work and before its later work by events). Calls on one batch must
 * not overlap. */
pfh_status pfh_batch_execute(pfh_batch batch, int op, const void *const *srcs, void *const *dsts, void *workspace,
                             size_t workspace_bytes, pfh_stream_t stream, int use_graph);
/* The tuning the batch runs with (threads per block, largest radix, lines
 * per tile on column-access stages, whether the items of all compile-time
 * radix plans of a stage share one launch, whether a launch keeps one
 * resident set of workgroups looping over the tiles or takes one workgroup
 * per tile, and whether fused plane stages run: two route stages of a 3D f32
 * problem in one pass over whole planes in LDS, the planes of the inner or
 * of the outer axis pair, whichever are smaller), and on how many streams
 * the launches run: 2 runs the problems whose first pass is a plane stage on
 * a second stream, concurrently with the others), as chosen by the
 * autotuner. */
pfh_status pfh_batch_tuning(pfh_batch batch, int *block, int *max_radix, int *col_lines, int *one_launch,
                            int *persistent, int *plane, int *streams);

/* ---------------------------------------------------------------------- */
/* Pointwise operations between transforms (pseudo-spectral steps).        */
/* ---------------------------------------------------------------------- */

/* data *= prod_a factors_a[i_a] over the plan's input (side 0) or output
 * (side 1) array of the forward direction, every extra block alike: the
 * reference's LocalGrid per-axis coordinate arrays seen in the pencil's
 * memory order (crates/pencil-array/src/grid.rs:20-120) combined with
 * pointwise_many_in_place (pointwise.rs:421-490). factors: device array of
 * the plan's real type, the per-axis arrays concatenated in logical axis
 * order (length sum of the side's logical shape). Complex data has both
 * parts scaled. elems must equal that array's element count. */
pfh_status pfh_pointwise_scale_axes(pfh_plan plan, int side, void *data, size_t elems, const void *factors,
                                    pfh_stream_t stream);

/* The same with complex per-axis factors ((re, im) pairs of the plan's real
 * type, concatenated in logical axis order) on the plan's complex side: the
 * spectral derivative d/dx_a is the factor i k_a on axis a and 1 elsewhere
 * (LocalGrid wavenumbers, grid.rs:20-120, in a pointwise_many closure,
 * pointwise.rs:421-490). Real data gives PFH_ERR_INVALID_ARGUMENT. */
pfh_status pfh_pointwise_scale_axes_complex(pfh_plan plan, int side, void *data, size_t elems, const void *factors,
                                            pfh_stream_t stream);

/* u = w + a w^2 elementwise over n reals (pointwise2, pointwise.rs:76-88). */
pfh_status pfh_pointwise_quadratic(int precision, const void *w, void *u, size_t n, double a, pfh_stream_t stream);

/* out = g (1 + 2 a w): the transpose of the derivative of the step above. */
pfh_status pfh_pointwise_quadratic_adjoint(int precision, const void *w, const void *g, void *out, size_t n,
                                           double a, pfh_stream_t stream);

/* *result = sum of x_i^2 over n reals, accumulated in f64 in a fixed order
 * (the local part of l2_norm, collectives.rs:519-560). workspace: device
 * memory of at least 8192 bytes; result: device pointer to one double. */
pfh_status pfh_sum_squares(int precision, const void *x, size_t n, void *workspace, size_t workspace_bytes,
                           double *result, pfh_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Reductions over a device array (the local fold of the reference's       */
/* checked global reductions at one rank, collectives.rs:519-620).         */
/* ---------------------------------------------------------------------- */

/* Scalar type codes: the reference's on-file IoElement codes
 * (crates/pencil-io/src/format.rs:50-60, :95-126). */
typedef enum pfh_scalar_type {
    PFH_SCALAR_I8 = 1,
    PFH_SCALAR_U8 = 2,
    PFH_SCALAR_I16 = 3,
    PFH_SCALAR_U16 = 4,
    PFH_SCALAR_I32 = 5,
    PFH_SCALAR_U32 = 6,
    PFH_SCALAR_I64 = 7,
    PFH_SCALAR_U64 = 8,
    PFH_SCALAR_F32 = 9,
    PFH_SCALAR_F64 = 10,
    PFH_SCALAR_C64 = 11,
    PFH_SCALAR_C128 = 12
} pfh_scalar_type;

typedef enum pfh_reduce_op {
    PFH_REDUCE_SUM = 1,     /* global_sum  (collectives.rs:519-528) */
    PFH_REDUCE_MIN = 2,     /* global_min  (:530-539); real types only */
    PFH_REDUCE_MAX = 3,     /* global_max  (:541-550); real types only */
    PFH_REDUCE_L2_NORM = 4, /* l2_norm     (:552-565) */
    PFH_REDUCE_ANY = 5,     /* any         (:567-575) */
    PFH_REDUCE_ALL = 6      /* all         (:577-588) */
} pfh_reduce_op;

/* outcome of a reduction */
enum {
    PFH_REDUCE_VALUE = 0,           /* value holds the result */
    PFH_REDUCE_NONE = 1,            /* min / max of an empty array (Ok(None)) */
    PFH_REDUCE_INTEGER_OVERFLOW = 2 /* CollectiveError::IntegerOverflow */
};

/* Device-resident result. value: SUM / MIN / MAX in the element type
 * (complex: re, im); L2_NORM in the norm type (f32 for f32 and complex f32,
 * f64 otherwise); ANY / ALL in value[0] (0 or 1). */
typedef struct pfh_reduce_result {
    uint8_t value[16];
    int32_t outcome;
    int32_t reserved[3];
} pfh_reduce_result;

/* Workspace a reduction needs (a fixed grid of partial states). */
pfh_status pfh_reduce_workspace_bytes(size_t *bytes);

/* Reduces the n elements of x (any memory order: the reductions are
 * permutation invariant, and integer sums follow the storage order, as the
 * reference's local fold does). Rules, as the reference's:
 *  - integer SUM: every step of the storage-order fold is checked; an
 *    intermediate overflow gives PFH_REDUCE_INTEGER_OVERFLOW even when the
 *    total would fit (collectives.rs:14-17, :206-214);
 *  - floating SUM: NaN wins; +inf and -inf together give NaN; a lone infinity
 *    sign is kept; complex parts independently (:300-330, :380-420). Finite
 *    values are accumulated in double-double, then rounded once;
 *  - MIN / MAX: NaN anywhere gives NaN; an empty array gives PFH_REDUCE_NONE
 *    (mapped_extreme_layout, :2066-2094);
 *  - L2_NORM: scaled sum of squares (NormState, :2150-2208); NaN gives NaN,
 *    an infinity +inf, an empty or all-zero array 0;
 *  - ANY / ALL: x != 0 (complex: either part); any of nothing is false, all
 *    of nothing true.
 * The result is written to `result` (device memory) in stream order. */
pfh_status pfh_reduce(int op, int scalar_type, const void *x, size_t n, void *workspace, size_t workspace_bytes,
                      pfh_reduce_result *result, pfh_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Broadcast pointwise combinations (pointwise2 / pointwise3 /             */
/* pointwise_many, pointwise.rs:40-598) with fixed operations.             */
/* ---------------------------------------------------------------------- */

/* Spatial layout of a pencil at one rank (Pencil::same_layout compares
 * these, pencil.rs:238-240). */
typedef struct pfh_pencil_layout {
    int ndim;                        /* spatial rank N (1..4) */
    size_t shape[PFH_MAX_DIMS];      /* global logical shape */
    int perm[PFH_MAX_DIMS];          /* memory order (AxisPermutation) */
    int topology_dims;               /* M */
    int decomposition[PFH_MAX_DIMS]; /* ordered decomposed axes (first M used) */
} pfh_pencil_layout;

/* A device pencil array: [extra...][memory-ordered spatial] elements. */
typedef struct pfh_array {
    void *data;
    int scalar_type; /* PFH_SCALAR_F32 / F64 / C64 / C128 */
    int extra_ndim;  /* 0..4 */
    size_t extra_shape[PFH_MAX_DIMS];
    pfh_pencil_layout pencil;
} pfh_array;

typedef enum pfh_pointwise_op {
    PFH_POINTWISE_LINEAR = 1,  /* out = sum_i c_i x_i        (pointwise_many, n >= 1, at most 8) */
    PFH_POINTWISE_MUL = 2,     /* out = c_0 x_0 x_1          (pointwise2) */
    PFH_POINTWISE_MUL_ADD = 3  /* out = c_0 x_0 x_1 + c_1 x_2 (pointwise3) */
} pfh_pointwise_op;

/* Every input has the output's spatial layout and extra rank; an input's
 * extra extent is the output's or 1 (broadcast). Validation follows the
 * reference's order (pointwise2: layouts of both inputs first, then extents;
 * pointwise3 / many: input by input) and completes before any write.
 * coeffs: 2 doubles (re, im) per coefficient. The output is complex when any
 * input or coefficient is; all arrays share one precision. An input may be
 * the output itself (same pointer and extra shape: the in-place forms);
 * other overlaps are not allowed. */
pfh_status pfh_pointwise(int op, const pfh_array *inputs, int n_inputs, const pfh_array *out, const double *coeffs,
                         pfh_stream_t stream);

/* ---------------------------------------------------------------------- */
/* On-file formats for device pencil arrays at one rank.                  */
/* ---------------------------------------------------------------------- */

/* The single-dataset file of write_mpi / read_mpi (crates/pencil-io/src/
 * mpi_io.rs:198-840, header layout :1492-1930): a 96-byte little-endian
 * prefix ("PENCILIO", version 1, header length, commit marker, N, type code,
 * width, extra rank, payload offset, payload bytes, writer grid rank,
 * reserved 0), the u64 arrays global[N], extra[], writer grid[], writer
 * permutation[N], then the payload in canonical row-major
 * [extra..., logical spatial...] order, little endian. A file is created
 * exclusively (an existing file is never replaced), written with an
 * INCOMPLETE marker, flushed, then committed. The writer grid of a one-rank
 * pencil is all ones (rank = topology_dims, at least 1). Reads stage and
 * validate the whole payload before the destination is written, in the
 * reference's order of checks; a rejected read leaves it unchanged.
 * workspace: device memory of pfh_io_workspace_bytes (the payload size; a
 * write of an identity-permutation pencil needs none). Both calls are
 * synchronous with respect to the file; the final device copy of a read is
 * stream ordered. */
pfh_status pfh_io_workspace_bytes(const pfh_array *array, size_t *bytes);
pfh_status pfh_io_write(const char *path, const pfh_array *array, void *workspace, size_t workspace_bytes,
                        pfh_stream_t stream);
pfh_status pfh_io_read(const char *path, const pfh_array *array, void *workspace, size_t workspace_bytes,
                       pfh_stream_t stream);

/* RawByteOrder (crates/pencil-io/src/options.rs:37-57) */
enum { PFH_BYTE_ORDER_NATIVE = 0, PFH_BYTE_ORDER_LITTLE = 1, PFH_BYTE_ORDER_BIG = 2 };

/* read_mpi_raw (crates/pencil-io/src/raw.rs:24-238): headerless canonical
 * [extra..., logical spatial...] values starting at byte `offset`, each real
 * component (a complex value has two) in `byte_order`; the whole range must
 * lie inside the file ("truncated raw payload"). */
pfh_status pfh_io_read_raw(const char *path, const pfh_array *array, uint64_t offset, int byte_order,
                           void *workspace, size_t workspace_bytes, pfh_stream_t stream);

/* The named append-only container (v2) of write_mpi_named /
 * append_mpi_named / read_mpi_named (crates/pencil-io/src/named_mpi.rs):
 * a 32-byte header ("PIONAM02", 2, 32, 0) then records, each a 72-byte
 * header ("PIOREC02", 2, name length, metadata length, payload length, total
 * length, type code, width, 0), the UTF-8 name, metadata (spatial rank,
 * extra rank, spatial dims, extra dims), the canonical payload and a trailing
 * commit word written after the flushed record. Names are non-empty UTF-8 of
 * at most 1024 bytes (slashes are literal). write creates the file
 * exclusively; append refuses a file whose committed prefix does not reach
 * its end (InvalidTail) or that holds the name (DuplicateName); read finds
 * the name in the committed prefix (NotFound), checks type and shapes
 * (MetadataMismatch "named record") and stages the payload before writing the
 * destination. Workspace as for pfh_io_read / pfh_io_write. */
pfh_status pfh_io_write_named(const char *path, const char *name, const pfh_array *array, void *workspace,
                              size_t workspace_bytes, pfh_stream_t stream);
pfh_status pfh_io_append_named(const char *path, const char *name, const pfh_array *array, void *workspace,
                               size_t workspace_bytes, pfh_stream_t stream);
pfh_status pfh_io_read_named(const char *path, const char *name, const pfh_array *array, void *workspace,
                             size_t workspace_bytes, pfh_stream_t stream);

/* DatasetInfo of read_mpi_catalog (crates/pencil-io/src/catalog.rs:113-300):
 * the validated metadata of a committed file, without its payload. */
typedef struct pfh_dataset_info {
    int scalar_type;
    int ndim;
    int extra_ndim;
    int grid_ndim;
    uint64_t shape[PFH_MAX_DIMS];
    uint64_t extra_shape[PFH_MAX_DIMS];
    uint64_t writer_grid[PFH_MAX_DIMS];
    int writer_perm[PFH_MAX_DIMS];
    uint64_t payload_bytes;
} pfh_dataset_info;

pfh_status pfh_io_catalog(const char *path, pfh_dataset_info *info);

#ifdef __cplusplus
}
#endif

#endif /* PENCIL_FFT_HIP_H */
