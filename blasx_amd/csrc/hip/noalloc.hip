// noalloc.hip -- the device-side table store of the allocation-free grouped
// GEMM entry point (blasx_gemm_grouped_device_noalloc, grouped.c).
//
// The problem and segment tables of a grouped launch normally live in a
// device buffer the library allocates and grows (grouped.c).  Callers that
// cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc) get tables in `scratch`
// instead: a file-scope device array of fixed size, one copy per GPU, held
// until the process ends.  Host code never uses the array's name: its
// address on each device comes from hipGetSymbolAddress at set-up and is
// passed to the copies and kernels from then on.
#include <hip/hip_runtime.h>
#include <stdio.h>

#include "../hiplayer.h"
#include "blasx_ext.h"

namespace {

static __device__ __align__(256) unsigned char scratch[BXH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

struct dev_scratch {
    void *ptr;      // address of `scratch` on this device
    void *done;     // event recorded after the last call's kernels
};
dev_scratch g_dev[kMaxDevices];
thread_local char g_err[160];

} // namespace

extern "C" {

const char *bxh_noalloc_error(void) { return g_err; }

// Set-up for the current device: the address of its copy of `scratch` and
// the event that orders one call's table writes after the previous call's
// kernels.  Idempotent; called from the library's init and, for a device
// first used later, from the entry point's first call on it.
int bxh_noalloc_setup(void)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        snprintf(g_err, sizeof g_err, "hipGetDevice: %s", hipGetErrorString(e));
        (void)hipGetLastError();
        return BLASX_EHIP;
    }
    if (g_dev[dev].ptr) return 0;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        snprintf(g_err, sizeof g_err, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        (void)hipGetLastError();
        return BLASX_EHIP;
    }
    hipEvent_t ev = nullptr;
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        snprintf(g_err, sizeof g_err, "hipEventCreateWithFlags: %s", hipGetErrorString(e));
        (void)hipGetLastError();
        return BLASX_EHIP;
    }
    g_dev[dev].done = ev;
    g_dev[dev].ptr = p;
    return 0;
}

// The current device's scratch address and ordering event (set up first).
int bxh_noalloc_scratch(void **ptr, void **done_event)
{
    int rc = bxh_noalloc_setup();
    if (rc) return rc;
    int dev = 0;
    (void)hipGetDevice(&dev);
    *ptr = g_dev[dev].ptr;
    *done_event = g_dev[dev].done;
    return 0;
}

} // extern "C"
