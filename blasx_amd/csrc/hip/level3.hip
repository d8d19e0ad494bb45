// level3.hip -- the non-GEMM kernels of the level-3 routines built on GEMMs:
//  * operand materialisation (SYMM, HEMM, TRMM, TRSM): the full symmetric or
//    Hermitian matrix from its stored triangle, or the triangular matrix with
//    the other triangle zero (and a unit diagonal when asked).  HBM-bound
//    element work: one pass over an n x n matrix, the mirrored half staged
//    through LDS in 32 x 32 tiles so both its reads and its writes coalesce;
//  * a tiled (conjugate) transpose (TRSM's op(A) and its Right side);
//  * TRSM's diagonal-block solve: substitution on 64-row blocks (32 for
//    complex), one column per thread.
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

// acc - a * b, and a / b, for the substitution
template <typename E> __device__ __forceinline__ E sub_mul(E acc, E a, E b) { return acc - a * b; }
template <> __device__ __forceinline__ z16 sub_mul(z16 acc, z16 a, z16 b)
{
    return z16{acc.re - (a.re * b.re - a.im * b.im), acc.im - (a.re * b.im + a.im * b.re)};
}
template <typename E> __device__ __forceinline__ E divide(E a, E b) { return a / b; }
// complex division by Smith's method (Smith 1962), which avoids overflow in |b|^2
template <> __device__ __forceinline__ z16 divide(z16 a, z16 b)
{
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return z16{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.im + b.re * r;
    return z16{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

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

// out(c, r) = op(in(r, c)) for an r x c block: 32 x 32 tiles through LDS so
// both sides coalesce; conj for the conjugate transpose
template <typename E>
__global__ __launch_bounds__(256) void k_transpose(int rows, int cols, const E *__restrict__ in,
                                                   int64_t ldi, E *__restrict__ out, int64_t ldo,
                                                   int conj)
{
    __shared__ E tile[TB][TB + 1];
    const int r0 = blockIdx.y * TB, c0 = blockIdx.x * TB;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < TB; k += 8) {
        const int r = r0 + tx, c = c0 + k;
        if (r < rows && c < cols) tile[tx][k] = in[r + (int64_t)c * ldi];
    }
    __syncthreads();
    for (int k = ty; k < TB; k += 8) {
        const int c = c0 + tx, r = r0 + k; // out row c, out column r
        if (r < rows && c < cols) {
            const E v = tile[k][tx];
            out[c + (int64_t)r * ldo] = conj ? conj_of(v) : v;
        }
    }
}

// Triangular solve of one diagonal block: T (m x m, m <= NB, lower or upper,
// its own diagonal) times X = B, in place on the m x n block X.  A block of
// NB threads owns NB columns: T and the columns are staged in LDS with
// coalesced loads, and each thread substitutes down (lower) or up (upper) its
// column held in registers.  Rows m..NB-1 are padded as identity rows.
template <typename E, int NB>
__global__ __launch_bounds__(NB) void k_trsm_leaf(int lower, int m, int n, const E *__restrict__ T,
                                                  int64_t ldt, E *__restrict__ X, int64_t ldx)
{
    __shared__ E t[NB][NB + 1];
    __shared__ E xs[NB][NB + 1];
    const int tid = threadIdx.x, j0 = blockIdx.x * NB;
    for (int c = 0; c < NB; ++c) {
        // t[r][c] = T(r, c); identity beyond m
        t[tid][c] = (tid < m && c < m) ? T[tid + (int64_t)c * ldt]
                                       : (tid == c ? one_of<E>() : zero_of<E>());
        const int j = j0 + c;
        xs[tid][c] = (tid < m && j < n) ? X[tid + (int64_t)j * ldx] : zero_of<E>();
    }
    __syncthreads();
    E x[NB];
#pragma unroll
    for (int r = 0; r < NB; ++r) x[r] = xs[r][tid];
    if (lower) {
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            E acc = x[r];
#pragma unroll
            for (int c = 0; c < r; ++c) acc = sub_mul(acc, t[r][c], x[c]);
            x[r] = divide(acc, t[r][r]);
        }
    } else {
#pragma unroll
        for (int r = NB - 1; r >= 0; --r) {
            E acc = x[r];
#pragma unroll
            for (int c = r + 1; c < NB; ++c) acc = sub_mul(acc, t[r][c], x[c]);
            x[r] = divide(acc, t[r][r]);
        }
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) xs[r][tid] = x[r];
    __syncthreads();
    for (int c = 0; c < NB; ++c) {
        const int j = j0 + c;
        if (tid < m && j < n) X[tid + (int64_t)j * ldx] = xs[tid][c];
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

int bxh_transpose(int dtype, int rows, int cols, const void *in, int64_t ldi, void *out, int64_t ldo,
                  int conj, void *stream)
{
    if (rows <= 0 || cols <= 0) return 0;
    const dim3 grid((cols + TB - 1) / TB, (rows + TB - 1) / TB);
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
    case BLASX_S:
        hipLaunchKernelGGL(k_transpose<float>, grid, dim3(256), 0, st, rows, cols,
                           (const float *)in, ldi, (float *)out, ldo, 0);
        break;
    case BLASX_D:
        hipLaunchKernelGGL(k_transpose<double>, grid, dim3(256), 0, st, rows, cols,
                           (const double *)in, ldi, (double *)out, ldo, 0);
        break;
    case BLASX_Z:
        hipLaunchKernelGGL(k_transpose<z16>, grid, dim3(256), 0, st, rows, cols, (const z16 *)in,
                           ldi, (z16 *)out, ldo, conj);
        break;
    default: return BLASX_EINVAL;
    }
    return hipGetLastError() == hipSuccess ? 0 : BLASX_EHIP;
}

int bxh_trsm_leaf_nb(int dtype) { return dtype == BLASX_Z ? 32 : 64; }

int bxh_trsm_leaf(int dtype, int lower, int m, int n, const void *T, int64_t ldt, void *X,
                  int64_t ldx, void *stream)
{
    const int nb = bxh_trsm_leaf_nb(dtype);
    if (m <= 0 || n <= 0) return 0;
    if (m > nb || ldt < m || ldx < m) return BLASX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((n + nb - 1) / nb);
    switch (dtype) {
    case BLASX_S:
        hipLaunchKernelGGL((k_trsm_leaf<float, 64>), grid, dim3(64), 0, st, lower, m, n,
                           (const float *)T, ldt, (float *)X, ldx);
        break;
    case BLASX_D:
        hipLaunchKernelGGL((k_trsm_leaf<double, 64>), grid, dim3(64), 0, st, lower, m, n,
                           (const double *)T, ldt, (double *)X, ldx);
        break;
    case BLASX_Z:
        hipLaunchKernelGGL((k_trsm_leaf<z16, 32>), grid, dim3(32), 0, st, lower, m, n,
                           (const z16 *)T, ldt, (z16 *)X, ldx);
        break;
    default: return BLASX_EINVAL;
    }
    return hipGetLastError() == hipSuccess ? 0 : BLASX_EHIP;
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
