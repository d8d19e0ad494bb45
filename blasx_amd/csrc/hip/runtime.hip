// runtime.hip -- thin C-ABI wrappers over the HIP runtime for the C host
// code (device selection, HBM and pinned allocations, streams, events,
// copies).  Replaces the reference's direct CUDA runtime calls
// (cudaMalloc/cudaFree per tile, blas/LRU.c:125,182; cudaStreamCreate /
// cudaEventCreateWithFlags, blas/blasx_internal.c:59-65;
// cublasSetMatrixAsync / cublasGetMatrixAsync, blas/blasx_mem_control.c:325,
// blas/blasx_dgemm.c:66,76).
#include <hip/hip_runtime.h>
#include <stdio.h>

#include "../hiplayer.h"
#include "blasx_ext.h"

extern "C" const char *bxh_kernel_error(void);

namespace {
thread_local char g_rt_err[256];

int rt_fail(const char *what, hipError_t e)
{
    snprintf(g_rt_err, sizeof g_rt_err, "%s: %s", what, hipGetErrorString(e));
    (void)hipGetLastError(); /* the failure is reported here; do not let it resurface later */
    return e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? BLASX_ENOMEM : BLASX_EHIP;
}

#define RT(call, what)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return rt_fail(what, e_);                                        \
    } while (0)
} // namespace

extern "C" {

const char *bxh_last_error(void) { return g_rt_err[0] ? g_rt_err : bxh_kernel_error(); }

int bxh_device_count(int *n)
{
    *n = 0;
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) {
        *n = 0;
        (void)hipGetLastError();
        return rt_fail("hipGetDeviceCount", e);
    }
    return 0;
}

int bxh_set_device(int dev)
{
    RT(hipSetDevice(dev), "hipSetDevice");
    return 0;
}

int bxh_device_info(int dev, char *name, size_t name_len, int *cus, size_t *total_mem)
{
    hipDeviceProp_t p;
    RT(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
    if (name && name_len) snprintf(name, name_len, "%s", p.gcnArchName);
    if (cus) *cus = p.multiProcessorCount;
    if (total_mem) *total_mem = p.totalGlobalMem;
    return 0;
}

int bxh_mem_get_info(size_t *free_b, size_t *total_b)
{
    RT(hipMemGetInfo(free_b, total_b), "hipMemGetInfo");
    return 0;
}

int bxh_malloc(void **p, size_t bytes)
{
    *p = nullptr;
    RT(hipMalloc(p, bytes), "hipMalloc");
    return 0;
}

int bxh_free(void *p)
{
    if (p) RT(hipFree(p), "hipFree");
    return 0;
}

int bxh_host_malloc(void **p, size_t bytes)
{
    *p = nullptr;
    RT(hipHostMalloc(p, bytes, hipHostMallocPortable), "hipHostMalloc");
    return 0;
}

int bxh_host_free(void *p)
{
    if (p) RT(hipHostFree(p), "hipHostFree");
    return 0;
}

int bxh_pointer_is_pinned(const void *p)
{
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    if (a.type == hipMemoryTypeDevice) return 2;
    if (a.type == hipMemoryTypeHost) return 1;
    return 0;
}

int bxh_stream_create(void **s)
{
    hipStream_t st;
    RT(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    *s = (void *)st;
    return 0;
}

int bxh_stream_destroy(void *s)
{
    if (s) RT(hipStreamDestroy((hipStream_t)s), "hipStreamDestroy");
    return 0;
}

int bxh_stream_sync(void *s)
{
    RT(hipStreamSynchronize((hipStream_t)s), "hipStreamSynchronize");
    return 0;
}

int bxh_device_sync(void)
{
    RT(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return 0;
}

int bxh_event_create(void **e, int timing)
{
    hipEvent_t ev;
    RT(hipEventCreateWithFlags(&ev, timing ? hipEventDefault : hipEventDisableTiming),
       "hipEventCreate");
    *e = (void *)ev;
    return 0;
}

int bxh_event_destroy(void *e)
{
    if (e) RT(hipEventDestroy((hipEvent_t)e), "hipEventDestroy");
    return 0;
}

int bxh_event_record(void *e, void *s)
{
    RT(hipEventRecord((hipEvent_t)e, (hipStream_t)s), "hipEventRecord");
    return 0;
}

int bxh_event_query(void *e)
{
    hipError_t r = hipEventQuery((hipEvent_t)e);
    if (r == hipSuccess) return 1;
    if (r == hipErrorNotReady) return 0;
    return rt_fail("hipEventQuery", r);
}

int bxh_event_sync(void *e)
{
    RT(hipEventSynchronize((hipEvent_t)e), "hipEventSynchronize");
    return 0;
}

int bxh_event_elapsed_ms(float *ms, void *start, void *stop)
{
    RT(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
    return 0;
}

int bxh_stream_wait_event(void *s, void *e)
{
    RT(hipStreamWaitEvent((hipStream_t)s, (hipEvent_t)e, 0), "hipStreamWaitEvent");
    return 0;
}

int bxh_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *s)
{
    if (!bytes) return 0;
    RT(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s), "H2D copy");
    return 0;
}

int bxh_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *s)
{
    if (!bytes) return 0;
    RT(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s), "D2H copy");
    return 0;
}

int bxh_memcpy_d2d_async(void *dst, const void *src, size_t bytes, void *s)
{
    if (!bytes) return 0;
    RT(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)s), "D2D copy");
    return 0;
}

int bxh_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    RT(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "H2D copy");
    return 0;
}

int bxh_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    RT(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "D2H copy");
    return 0;
}

int bxh_memcpy2d_h2d_async(void *dst, size_t dpitch, const void *src, size_t spitch,
                           size_t width, size_t height, void *s)
{
    if (!width || !height) return 0;
    RT(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyHostToDevice,
                        (hipStream_t)s),
       "H2D 2-D copy");
    return 0;
}

int bxh_memcpy2d_d2h_async(void *dst, size_t dpitch, const void *src, size_t spitch,
                           size_t width, size_t height, void *s)
{
    if (!width || !height) return 0;
    RT(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToHost,
                        (hipStream_t)s),
       "D2H 2-D copy");
    return 0;
}

// ---- HIP graphs: capture the work enqueued on a stream, replay it ----
int bxh_graph_begin(void *s)
{
    RT(hipStreamBeginCapture((hipStream_t)s, hipStreamCaptureModeThreadLocal),
       "hipStreamBeginCapture");
    return 0;
}

int bxh_graph_end(void *s, void **exec)
{
    hipGraph_t g = nullptr;
    RT(hipStreamEndCapture((hipStream_t)s, &g), "hipStreamEndCapture");
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return rt_fail("hipGraphInstantiate", e);
    *exec = (void *)ex;
    return 0;
}

int bxh_graph_launch(void *exec, void *s)
{
    RT(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)s), "hipGraphLaunch");
    return 0;
}

int bxh_graph_destroy(void *exec)
{
    if (exec) RT(hipGraphExecDestroy((hipGraphExec_t)exec), "hipGraphExecDestroy");
    return 0;
}

int bxh_memset_async(void *dst, int v, size_t bytes, void *s)
{
    if (!bytes) return 0;
    RT(hipMemsetAsync(dst, v, bytes, (hipStream_t)s), "hipMemsetAsync");
    return 0;
}

} // extern "C"
