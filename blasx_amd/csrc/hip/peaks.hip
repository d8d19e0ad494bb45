// peaks.hip -- microbenchmarks of the MI355X ceilings the roofline fractions
// are quoted against: dense MFMA throughput (FP64 16x16x4, FP32 16x16x4,
// INT8 16x16x64) and HBM copy bandwidth.  Each MFMA kernel keeps 8 or 16
// independent accumulators per wave in registers (no memory traffic inside
// the timed loop) and writes one value per lane at the end so nothing is
// optimised away; the copy kernel moves a buffer of known size.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../hiplayer.h"
#include "blasx_ext.h"

namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NACC = 16;

// INT8: issued in its VGPR form for the same reason as the FP64 kernel below
// (the builtin's accumulators get copied between register files in the loop).

__global__ __launch_bounds__(256) void k_peak_i8(int iters, int *__restrict__ out)
{
    i32x4 acc[NACC];
#pragma unroll
    for (int i = 0; i < NACC; ++i) acc[i] = i32x4{0, 0, 0, 0};
    const i32x4 a = i32x4{(int)threadIdx.x, 1, 2, 3}, b = i32x4{3, 2, 1, (int)blockIdx.x};
    asm volatile("s_nop 7" ::: "memory");
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < NACC; ++i)
            asm volatile("v_mfma_i32_16x16x64_i8 %0, %1, %2, %0" : "+v"(acc[i]) : "v"(a), "v"(b));
    }
    asm volatile("s_nop 15\n\ts_nop 15" ::: "memory");
    int s = 0;
#pragma unroll
    for (int i = 0; i < NACC; ++i) s += acc[i][0] ^ acc[i][1] ^ acc[i][2] ^ acc[i][3];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// FP64: the compiler keeps loop-carried FP64 MFMA accumulators in VGPRs and
// copies them to and from AGPRs around every MFMA (two moves per register per
// iteration, which halves the measured rate), so this kernel issues the
// VGPR form of the instruction itself.  8 accumulators of 8 registers; each is
// reused every 8 MFMAs, well past the instruction's latency.  The loop runs
// 2 x iters rounds so the operation count matches the other kernels.  The
// s_nop padding covers the wait states between a VALU write and an MFMA read
// of the same registers, and between the last MFMA and the VALU reads.
constexpr int NACC64 = 8;
__global__ __launch_bounds__(256) void k_peak_f64(int iters, double *__restrict__ out)
{
    d4 acc[NACC64];
#pragma unroll
    for (int i = 0; i < NACC64; ++i) acc[i] = d4{0, 0, 0, 0};
    const double a = 1.0 + threadIdx.x * 0x1p-20, b = 1.0 - blockIdx.x * 0x1p-30;
    asm volatile("s_nop 7" ::: "memory");
    for (int it = 0; it < 2 * iters; ++it) {
#pragma unroll
        for (int i = 0; i < NACC64; ++i)
            asm volatile("v_mfma_f64_16x16x4_f64 %0, %1, %2, %0" : "+v"(acc[i]) : "v"(a), "v"(b));
    }
    asm volatile("s_nop 15\n\ts_nop 15\n\ts_nop 15" ::: "memory");
    double s = 0;
#pragma unroll
    for (int i = 0; i < NACC64; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_peak_f32(int iters, float *__restrict__ out)
{
    f4 acc[NACC];
#pragma unroll
    for (int i = 0; i < NACC; ++i) acc[i] = f4{0, 0, 0, 0};
    const float a = 1.0f + threadIdx.x * 0x1p-20f, b = 1.0f - blockIdx.x * 0x1p-24f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < NACC; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
    }
    float s = 0;
#pragma unroll
    for (int i = 0; i < NACC; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_peak_copy(const i32x4 *__restrict__ src, i32x4 *__restrict__ dst, int64_t n16)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// time one launch configuration: best of `reps` launches, in seconds
template <typename F> double best_time(F launch, hipStream_t st, int reps)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -1;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return -1;
    }
    double best = -1;
    launch(); // warm up
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, st);
        launch();
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) break;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (best < 0 || ms * 1e-3 < best) best = ms * 1e-3;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

} // namespace

extern "C" {

// out[0..3] = FP64 MFMA TFLOP/s, FP32 MFMA TFLOP/s, INT8 MFMA TOP/s, HBM GB/s
// (read + write bytes of a copy of `copy_bytes`), each the best over 1, 2 and
// 4 blocks of 4 waves per CU; `cus` = compute units to fill
int bxh_measure_peaks(int cus, size_t copy_bytes, double *out, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int max_blocks = 4 * cus;
    void *buf = nullptr;
    if (hipMalloc(&buf, (size_t)max_blocks * 256 * 8) != hipSuccess) return BLASX_ENOMEM;
    const int iters = 4000;
    for (int k = 0; k < 3; ++k) out[k] = 0;
    for (int per_cu = 1; per_cu <= 4; per_cu *= 2) {
        const int nb = per_cu * cus;
        const double ops64 = (double)nb * 4 * iters * NACC * (16.0 * 16 * 4 * 2);
        const double ops8 = (double)nb * 4 * iters * NACC * (16.0 * 16 * 64 * 2);
        double t = best_time([&] { hipLaunchKernelGGL(k_peak_f64, dim3(nb), dim3(256), 0, st, iters, (double *)buf); }, st, 3);
        if (t > 0 && ops64 / t / 1e12 > out[0]) out[0] = ops64 / t / 1e12;
        t = best_time([&] { hipLaunchKernelGGL(k_peak_f32, dim3(nb), dim3(256), 0, st, iters, (float *)buf); }, st, 3);
        if (t > 0 && ops64 / t / 1e12 > out[1]) out[1] = ops64 / t / 1e12;
        t = best_time([&] { hipLaunchKernelGGL(k_peak_i8, dim3(nb), dim3(256), 0, st, iters, (int *)buf); }, st, 3);
        if (t > 0 && ops8 / t / 1e12 > out[2]) out[2] = ops8 / t / 1e12;
    }
    (void)hipFree(buf);
    out[3] = 0;
    void *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, copy_bytes) == hipSuccess && hipMalloc(&dst, copy_bytes) == hipSuccess) {
        const int64_t n16 = (int64_t)(copy_bytes / 16);
        const double t = best_time([&] { hipLaunchKernelGGL(k_peak_copy, dim3(8 * cus * 4), dim3(256), 0, st, (const i32x4 *)src, (i32x4 *)dst, n16); }, st, 5);
        if (t > 0) out[3] = 2.0 * (double)(n16 * 16) / t / 1e9;
    }
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : BLASX_EHIP;
}

} // extern "C"
