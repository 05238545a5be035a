// hegj_scratch.hip -- the fixed device scratch of hegj_selfjoin_noalloc.
//
// Callers that cannot allocate device memory (code under graph capture,
// serving loops, harnesses that forbid hipMalloc) run the self-join through
// hegj_selfjoin_noalloc, whose temporaries -- work units, their counts and
// offsets, the deferral list, the counters, scan partials -- come from this
// one array instead of hipMalloc.  Each device that runs the library holds
// its own copy of it until the process ends.
#include <mutex>

#include "hegj_internal.h"

namespace hegj {

namespace {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];  // device address of `scratch`, per device

}  // namespace

hegj_status noalloc_scratch(int device, void **ptr) {
    if (device < 0 || device >= kMaxDevices) return fail(HEGJ_ERR_INVALID_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        // the caller made `device` current: the address is that device's copy
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(HEGJ_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_ptr[device] = p;
    }
    *ptr = g_ptr[device];
    return HEGJ_OK;
}

}  // namespace hegj
