// Device scratch of dtfft_hip_execute_no_alloc: the aux buffer of a plan's execute sequence
// held in a fixed device array, so that the entry point allocates nothing (for callers that
// cannot call hipMalloc: code under stream capture, tight loops, allocation-free harnesses).
//
// `scratch` is a file-scope device array.  Its name is a host-side stand-in in host code, not
// a device address: the address of each device's copy comes from hipGetSymbolAddress, looked
// up once per device (with that device current) on the first call there and kept per device.
// Every call on one device shares that device's copy, which keeps the last call's contents;
// calls that overlap on one device race.  No part of it has to start at zero: every execute
// sequence writes the aux region before it reads it.
#include <mutex>

#include "internal.h"

static __device__ __align__(256) unsigned char scratch[dtfft_hip::kNoAllocAuxBytes];

namespace dtfft_hip {

static std::mutex g_scratch_mu;
static void* g_scratch[kNoAllocMaxDevices] = {};

int32_t no_alloc_scratch(void** out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        set_last_error("device index " + std::to_string(dev) + " beyond the no-alloc scratch table");
        return DTFFT_HIP_ERROR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_last_error("hipGetSymbolAddress returned no address for the no-alloc scratch");
            return DTFFT_HIP_ERROR_HIP_RUNTIME;
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return DTFFT_HIP_SUCCESS;
}

}  // namespace dtfft_hip
