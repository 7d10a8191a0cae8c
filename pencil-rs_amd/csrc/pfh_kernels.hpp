// Host-side declarations of the gfx950 kernels' launch wrappers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pfh_stage.hpp"

namespace pfh {

// Sticky device status bits (kept in the caller's workspace).
constexpr int PFH_DEVICE_INVALID_SPECTRUM = 1;
constexpr int PFH_DEVICE_INVALID_ENDPOINT = 2;
constexpr int PFH_DEVICE_POISONED = 4; // an in-place call found the workspace poisoned and wrote nothing

// Workspace status area (the first 256 bytes of every plan workspace).
constexpr int kStatusOffStatus = 0;  // int: sticky device status bits
constexpr int kStatusOffGate = 4;    // int: nonzero = later stages of this call write nothing
constexpr int kStatusOffFirst = 8;   // u64: first failing local endpoint code
constexpr int kStatusOffTag = 16;    // u64: tag of the plan the workspace is bound to (0 = unbound)
constexpr int kStatusOffPoison = 24; // int: an in-place call failed after it started writing

// Plane elements per thread of the endpoint partial kernel (256 threads):
// a chunk holds at most 256 x kEndpointPerThread elements.
constexpr int kEndpointPerThread = 8;

// One chunk's endpoint statistics of an R2C boundary plane.
struct EndpointPartial {
    double maxa; // max(|re|, |im|)
    double maxi; // max |im|
    double sre;  // sum (re / maxa)^2
    double sim;  // sum (im / maxa)^2
    unsigned bad;
    unsigned pad_[3];
};

struct EndpointArgs {
    int64_t extra;   // batches
    int64_t len;     // complex elements per batch
    int64_t outer;   // blocks before the reduction axis
    int64_t stride;  // memory stride of the reduction axis
    int64_t clen;    // complex length of the reduction axis
    int planes;      // 2 (even real length) or 1
    int nblk;        // chunks per plane
    double relative, absolute;
    int keep_gate;   // 1: the gate was set before the check (in-place calls) and is not cleared
    int *poison;     // non-null (in-place calls): a rejected spectrum sets *poison = 1
};

struct PermuteArgs {
    int nd;
    int64_t len, extra;
    int64_t dext[8];    // destination memory extents (slowest first)
    int64_t sstride[8]; // source stride of the axis at each destination position
};

template <typename T>
hipError_t launch_line_stage(const StageArgs<T> &a, bool wide, int threads, size_t lds_bytes, int64_t grid,
                             hipStream_t s);

hipError_t set_line_stage_lds_limit(size_t bytes);

// Persistent batched stage over a device work list (see k_line_batch):
// the generic kernel (narrow or wide radix set) or one compile-time radix
// plan for every item (k_line_fast_batch).
template <typename T>
hipError_t launch_line_batch(const StageArgs<T> *items, const int64_t *prefix, int nitems, int64_t total, bool wide,
                             int threads, size_t lds_bytes, int64_t grid, hipStream_t s);
template <typename T>
int line_batch_blocks_per_cu(bool wide, int threads, size_t lds_bytes);
template <typename T>
hipError_t launch_line_fast_batch(const StageArgs<T> *items, const int64_t *prefix, int nitems, int64_t total, int r0,
                                  int r1, int r2, int em, int threads, size_t lds_bytes, int64_t grid, hipStream_t s);
template <typename T>
int line_fast_batch_blocks_per_cu(int r0, int r1, int r2, int em, int threads, size_t lds_bytes);
// items with any compile-time plan of 16 (f32) / 8 (f64) values per thread
template <typename T>
hipError_t launch_line_multi_batch(const StageArgs<T> *items, const int64_t *prefix, int nitems, int64_t total,
                                   int threads, size_t lds_bytes, int64_t grid, hipStream_t s);
template <typename T>
int line_multi_batch_blocks_per_cu(int threads, size_t lds_bytes);

// Stages with a compile-time radix sequence (see k_line_fast).
template <typename T>
hipError_t launch_line_fast(const StageArgs<T> &a, int r0, int r1, int r2, int em, int threads, size_t lds_bytes,
                            int64_t grid, hipStream_t s);
bool fast_plan_available(int tbytes, int r0, int r1, int r2, int em);
bool multi_batch_plan_available(int tbytes, int r0, int r1, int r2, int em);
hipError_t set_fast_lds_limit(size_t bytes);

// Fused two-axis plane stage over a work list of items (pfh_plane.hip);
// plane_plan_available: the (radices, values per thread) plan is in its list.
// LDS: tw_cap + work_cap + raw_cap complex values (tables, working plane,
// prefetch region; tw_cap + work_cap even).
template <typename T>
hipError_t launch_plane_batch(const PlaneArgs<T> *items, const int64_t *prefix, int nitems, int64_t total,
                              int threads, int tw_cap, int work_cap, int raw_cap, int64_t grid, hipStream_t s);
template <typename T>
int plane_batch_blocks_per_cu(int threads, size_t lds_bytes);
bool plane_plan_available(int tbytes, int r0, int r1, int r2, int em);
hipError_t set_plane_lds_limit(size_t bytes);

// Pre- / post-processing of long real lines (pfh_long.hip).
template <typename T>
hipError_t launch_long_prep(const LongArgs<T> &a, hipStream_t s);
template <typename T>
hipError_t launch_long_post(const LongArgs<T> &a, hipStream_t s);

template <typename T>
hipError_t launch_endpoint_check(const EndpointArgs &e, const void *data, EndpointPartial *part, int *gate,
                                 int *status, hipStream_t s);

template <typename T>
hipError_t launch_local_endpoint_check(const void *data, int64_t batch, int64_t clen, int even,
                                       unsigned long long *first_code, int *gate, int *status, hipStream_t s);

hipError_t launch_fill_i32(int *p, int64_t n, int v, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long *p, int64_t n, unsigned long long v, hipStream_t s);

// First kernel of an in-place call: gate = poison word; a poisoned
// workspace also sets PFH_DEVICE_POISONED in the status word.
hipError_t launch_inplace_begin(int *gate, const int *poison, int *status, hipStream_t s);

hipError_t launch_permute_words(const PermuteArgs &p, const void *src, void *dst, int word, hipStream_t s);

// Pointwise operations (pfh_pointwise.hip).
struct ScaleAxesArgs {
    int nd;              // spatial positions
    int complex_data;    // interleaved complex values (both parts scaled)
    int complex_factors; // factors are complex (re, im) pairs; data must be complex
    int64_t len, extra;  // elements per extra block, extra blocks
    int64_t mext[8];     // extent of each memory position (slowest first)
    int64_t foff[8];     // offset of that position's axis in the factor array
};
constexpr int kSumSquaresParts = 1024;

template <typename T>
hipError_t launch_scale_axes(const ScaleAxesArgs &s, void *data, const void *factors, hipStream_t st);
template <typename T>
hipError_t launch_quadratic(const void *w, void *u, int64_t n, double a, hipStream_t st);
template <typename T>
hipError_t launch_quadratic_adjoint(const void *w, const void *g, void *out, int64_t n, double a, hipStream_t st);
template <typename T>
hipError_t launch_sum_squares(const void *x, int64_t n, double *partial, int nparts, double *result, hipStream_t st);

} // namespace pfh
