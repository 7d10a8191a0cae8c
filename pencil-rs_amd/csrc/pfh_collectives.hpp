// Shared declarations of the device reductions and broadcast pointwise
// combinations (pfh_collectives.hip) and their C ABI (pfh_api.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>

namespace pfh {

// Scalar type codes: the reference's on-file IoElement codes
// (crates/pencil-io/src/format.rs:50-60, :95-126).
enum : int {
    kScalarI8 = 1,
    kScalarU8 = 2,
    kScalarI16 = 3,
    kScalarU16 = 4,
    kScalarI32 = 5,
    kScalarU32 = 6,
    kScalarI64 = 7,
    kScalarU64 = 8,
    kScalarF32 = 9,
    kScalarF64 = 10,
    kScalarC64 = 11,
    kScalarC128 = 12
};

enum : int { kReduceSum = 1, kReduceMin = 2, kReduceMax = 3, kReduceL2Norm = 4, kReduceAny = 5, kReduceAll = 6 };
enum : int { kReduceValue = 0, kReduceNone = 1, kReduceIntegerOverflow = 2 };

constexpr int kReduceParts = 1024;     // partial states of the first kernel, at most
constexpr int kReduceStateBytes = 128; // workspace slot per partial state
constexpr int64_t kReduceWorkspaceBytes = (int64_t)kReduceParts * kReduceStateBytes;

// Device-side result of one reduction (32 bytes, the layout of
// pfh_reduce_result in the C header).
struct ReduceResult {
    uint8_t value[16];
    int32_t outcome;
    int32_t reserved[3];
};
static_assert(sizeof(ReduceResult) == 32, "pfh_reduce_result layout");

template <typename I>
struct IntLimits {
    static constexpr I lo = std::numeric_limits<I>::lowest();
    static constexpr I hi = std::numeric_limits<I>::max();
    __host__ __device__ static constexpr I lowest() { return lo; }
    __host__ __device__ static constexpr I highest() { return hi; }
};

enum : int { kPointwiseLinear = 1, kPointwiseMul = 2, kPointwiseMulAdd = 3 };
constexpr int kPointwiseMaxInputs = 8;

struct PointwiseArgs {
    int op = 0;
    int nin = 0;
    int out_complex = 0;
    int extra_ndim = 0;
    const void *in[kPointwiseMaxInputs] = {};
    int in_complex[kPointwiseMaxInputs] = {};
    // input i's stride (in extra blocks) along output extra axis d, 0 where
    // the input has extent 1 there
    int64_t extra_stride[kPointwiseMaxInputs][4] = {};
    int64_t extra_dims[4] = {1, 1, 1, 1};
    int64_t extra_count = 0;
    int64_t spatial = 0;
    double coef[2 * kPointwiseMaxInputs] = {};
};

hipError_t launch_reduce(int op, int scalar_type, const void *x, int64_t n, void *workspace, ReduceResult *result,
                         hipStream_t st);
hipError_t launch_pointwise(const PointwiseArgs &a, void *out, bool f64, hipStream_t st);

} // namespace pfh
