// Batched uneven-size C2C transforms (placeholder until the persistent
// batched launch lands).
#include <hip/hip_runtime.h>

#include "pfh_plan.hpp"

extern "C" {

pfh_status pfh_batch_create(const pfh_desc *, int, int, pfh_batch *out)
{
    if (out) *out = nullptr;
    return PFH_ERR_UNSUPPORTED;
}

void pfh_batch_destroy(pfh_batch) {}

pfh_status pfh_batch_workspace_bytes(pfh_batch, size_t *bytes)
{
    if (bytes) *bytes = 0;
    return PFH_ERR_UNSUPPORTED;
}

pfh_status pfh_batch_execute(pfh_batch, int, const void *const *, void *const *, void *, size_t, pfh_stream_t, int)
{
    return PFH_ERR_UNSUPPORTED;
}

} // extern "C"
