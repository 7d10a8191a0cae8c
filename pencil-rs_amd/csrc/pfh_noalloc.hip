// Allocation-free execution of a distributed plan operation.
//
// pfh_execute needs a caller-provided workspace (status area + the route's
// intermediate arrays). Callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc) call pfh_execute_noalloc instead: its workspace is the
// file-scope array `scratch` below, of kScratchBytes, which every GPU the
// library runs on holds until the process ends (each GPU has its own copy at
// its own address). Operations whose workspace exceeds it are refused.
//
// Every call on one GPU shares that GPU's array and finds the last call's
// contents there: the status area is cleared at the start of every call
// (a fill kernel), the intermediates are written before they are read.
// Calls that overlap on one GPU (several host threads, or several streams
// without ordering between them) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/pencil_fft_hip.h"

namespace pfh {
pfh_status set_error(pfh_status st, const char *msg, size_t required, size_t actual, size_t batch, size_t index,
                     int hip_error);
pfh_status execute_on_scratch(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst, size_t dst_elems,
                              void *ws, size_t ws_bytes, hipStream_t stream);
} // namespace pfh

namespace {

constexpr size_t kScratchBytes = size_t(128) << 20; // 128 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// The device address of the current GPU's `scratch`, looked up once per
// device. In host code `scratch` names a host-side stand-in, not device
// memory: only this address may reach kernels or copies.
pfh_status scratch_address(void **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return pfh::set_error(PFH_ERR_DEVICE, hipGetErrorString(e), 0, 0, 0, 0, (int)e);
    if (dev < 0 || dev >= kMaxDevices) return pfh::set_error(PFH_ERR_DEVICE, "device index", 0, 0, 0, 0, 0);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[dev] == nullptr) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return pfh::set_error(PFH_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed", 0, 0, 0, 0, (int)e);
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return PFH_OK;
}

} // namespace

extern "C" pfh_status pfh_execute_noalloc(pfh_plan plan, int op, const void *src, size_t src_elems, void *dst,
                                          size_t dst_elems, pfh_stream_t stream, int flags)
{
    if (flags != 0)
        return pfh::set_error(PFH_ERR_INVALID_ARGUMENT,
                              "pfh_execute_noalloc takes no flags (its status is always read synchronously)", 0, 0, 0,
                              0, 0);
    void *ws = nullptr;
    pfh_status st = scratch_address(&ws);
    if (st) return st;
    return pfh::execute_on_scratch(plan, op, src, src_elems, dst, dst_elems, ws, kScratchBytes,
                                   reinterpret_cast<hipStream_t>(stream));
}
