// level3.hip -- operand preparation for the level-3 routines that run as a
// GEMM on a materialised operand (SYMM, HEMM, TRMM): the full symmetric or
// Hermitian matrix from its stored triangle, or the triangular matrix with
// the other triangle zero (and a unit diagonal when asked).  HBM-bound
// element work: one pass over an n x n matrix, the mirrored half staged
// through LDS in 32 x 32 tiles so both its reads and its writes coalesce.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../hiplayer.h"
#include "blasx_ext.h"

namespace {

struct z16 {
    double re, im;
};

template <typename E> __device__ __forceinline__ E conj_of(E v) { return v; }
template <> __device__ __forceinline__ z16 conj_of(z16 v) { return z16{v.re, -v.im}; }
template <typename E> __device__ __forceinline__ E real_of(E v) { return v; }
template <> __device__ __forceinline__ z16 real_of(z16 v) { return z16{v.re, 0.0}; }
template <typename E> __device__ __forceinline__ E zero_of() { return E(0); }
template <> __device__ __forceinline__ z16 zero_of() { return z16{0.0, 0.0}; }
template <typename E> __device__ __forceinline__ E one_of() { return E(1); }
template <> __device__ __forceinline__ z16 one_of() { return z16{1.0, 0.0}; }

constexpr int TB = 32;

// Block (bx, by) owns destination tile rows [by*32, +32), columns [bx*32, +32)
// of the n x n matrix A; 256 threads as 32 (row) x 8 (column step).  Tiles
// wholly inside the stored triangle exit at once (uniformly per block).
template <typename E>
__global__ __launch_bounds__(256) void k_tri_fill(int n, int lower, int mode, E *A, int64_t lda)
{
    __shared__ E tile[TB][TB + 1];
    const int bi = blockIdx.y, bj = blockIdx.x;
    if (lower ? bi > bj : bi < bj) return; // stored side, off the diagonal
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const bool mirror = mode == BX_FILL_SYM || mode == BX_FILL_HERM;
    if (mirror) {
        // the source of tile (bi, bj) is tile (bj, bi): tile[a][b] = A(bj*32 + a, bi*32 + b)
        for (int k = ty; k < TB; k += 8) {
            const int r = bj * TB + tx, c = bi * TB + k;
            if (r < n && c < n) tile[tx][k] = A[r + (int64_t)c * lda];
        }
        __syncthreads();
    }
    for (int k = ty; k < TB; k += 8) {
        const int r = bi * TB + tx, c = bj * TB + k;
        if (r >= n || c >= n) continue;
        E *p = A + r + (int64_t)c * lda;
        if (r == c) {
            if (mode == BX_FILL_HERM) *p = real_of(*p);
            else if (mode == BX_FILL_UNIT) *p = one_of<E>();
            continue;
        }
        if (lower ? r > c : r < c) continue; // stored element
        if (!mirror) {
            *p = zero_of<E>();
        } else {
            // A(r, c) = A(c, r) = tile[c - bj*32][r - bi*32]
            const E v = tile[k][tx];
            *p = mode == BX_FILL_HERM ? conj_of(v) : v;
        }
    }
}

template <typename E> int launch_fill(int n, int lower, int mode, void *A, int64_t lda, hipStream_t st)
{
    const int nt = (n + TB - 1) / TB;
    hipLaunchKernelGGL(k_tri_fill<E>, dim3(nt, nt), dim3(256), 0, st, n, lower, mode, (E *)A, lda);
    return hipGetLastError() == hipSuccess ? 0 : BLASX_EHIP;
}

} // namespace

extern "C" {

int bxh_tri_fill(int dtype, int n, int lower, int mode, void *A, int64_t lda, void *stream)
{
    if (n <= 0) return 0;
    if (lda < n || mode < BX_FILL_SYM || mode > BX_FILL_UNIT) return BLASX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
    case BLASX_S: return launch_fill<float>(n, lower, mode, A, lda, st);
    case BLASX_D: return launch_fill<double>(n, lower, mode, A, lda, st);
    case BLASX_Z: return launch_fill<z16>(n, lower, mode, A, lda, st);
    default: return BLASX_EINVAL;
    }
}

int bxh_memcpy2d_d2d_async(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                           size_t height, void *s)
{
    if (!width || !height) return 0;
    const hipError_t e = hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height,
                                          hipMemcpyDeviceToDevice, (hipStream_t)s);
    return e == hipSuccess ? 0 : BLASX_EHIP;
}

} // extern "C"
