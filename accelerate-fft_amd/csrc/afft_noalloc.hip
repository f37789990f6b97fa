// afft_noalloc.hip -- afft_exec_noalloc: afft_exec without any device allocation.
//
// An execution's intermediate arrays (the work array of a split axis, the auxiliary arrays of a
// multi-kernel Bluestein transform) normally come from the plan's per-stream pool, which
// allocates a set the first time a stream needs one.  Callers that may not allocate at all (code
// under graph capture, tight serving loops, harnesses that forbid hipMalloc) use this entry point
// instead: the arrays are carved from `scratch`, one fixed device array per GPU that lives as long
// as the process.  Every call on a GPU shares that GPU's scratch, so calls that overlap on one
// GPU (two streams, or two threads) race on it; a caller serialises them.  Nothing in scratch has
// to start at zero: every pass writes each intermediate element before any pass reads it.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "../../include/afft.h"
#include "afft_plan.hpp"

namespace afft {

// 64 MiB: the work array of a 2^22-point Complex Float row (32 MiB) or a 2048 x 2048 fft2D, or
// the auxiliary arrays of Bluestein batches of a few million points.
constexpr int64_t kScratchBytes = (int64_t)AFFT_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of scratch on each GPU, looked up once per device with that device current.
// In host code `scratch` names a host-side stand-in, never the device memory.
static std::mutex g_scratch_mu;
static std::map<int, void *> g_scratch_ptr;

static int scratch_address(int device, void *&out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        out = it->second;
        return AFFT_SUCCESS;
    }
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return AFFT_ERROR_HIP;
    g_scratch_ptr[device] = ptr;
    out = ptr;
    return AFFT_SUCCESS;
}

} // namespace afft

using namespace afft;

extern "C" int afft_exec_noalloc(const afft_plan *plan, int direction, const void *d_in, void *d_out,
                                 double scale, void *stream)
{
    if (!plan || (direction != -1 && direction != 1)) return AFFT_ERROR_INVALID_VALUE;
    const Plan *p = reinterpret_cast<const Plan *>(plan);
    if (p->total == 0) return AFFT_SUCCESS;
    if (!d_in || !d_out) return AFFT_ERROR_INVALID_VALUE;
    const int64_t need = workspace_bytes(*p);
    if (need > kScratchBytes) return AFFT_ERROR_SCRATCH_TOO_SMALL;
    int old = -1;
    if (hipGetDevice(&old) != hipSuccess) return AFFT_ERROR_NO_DEVICE;
    if (old != p->device && hipSetDevice(p->device) != hipSuccess) return AFFT_ERROR_HIP;
    int st = AFFT_SUCCESS;
    Workspace ws;
    if (need > 0) {
        void *base = nullptr;
        st = scratch_address(p->device, base);
        if (!st) ws = carve_workspace(*p, base);
    }
    if (!st) st = exec_plan(*p, direction, d_in, d_out, scale, (hipStream_t)stream, nullptr, nullptr, nullptr,
                            nullptr, &ws);
    if (old != p->device) (void)hipSetDevice(old);
    return st;
}
