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

// complex elements: T is double (complex128) or float (complex64)
template <typename T> struct cx {
    T re, im;
};
using z16 = cx<double>;
using c8 = cx<float>;

template <typename E> struct is_cx { static constexpr bool value = false; };
template <typename T> struct is_cx<cx<T>> { static constexpr bool value = true; };

template <typename E> __device__ __forceinline__ E conj_of(E v)
{
    if constexpr (is_cx<E>::value) return E{v.re, -v.im};
    else return v;
}
template <typename E> __device__ __forceinline__ E real_of(E v)
{
    if constexpr (is_cx<E>::value) return E{v.re, 0};
    else return v;
}
template <typename E> __device__ __forceinline__ E zero_of()
{
    if constexpr (is_cx<E>::value) return E{0, 0};
    else return E(0);
}
template <typename E> __device__ __forceinline__ E one_of()
{
    if constexpr (is_cx<E>::value) return E{1, 0};
    else return E(1);
}

// acc - a * b, and a / b, for the substitution
template <typename E> __device__ __forceinline__ E sub_mul(E acc, E a, E b)
{
    if constexpr (is_cx<E>::value)
        return E{acc.re - (a.re * b.re - a.im * b.im), acc.im - (a.re * b.im + a.im * b.re)};
    else return acc - a * b;
}
// complex division by Smith's method (Smith 1962), which avoids overflow in |b|^2
template <typename E> __device__ __forceinline__ E divide(E a, E b)
{
    if constexpr (is_cx<E>::value) {
        if (fabs(b.re) >= fabs(b.im)) {
            const auto r = b.im / b.re, d = b.re + b.im * r;
            return E{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
        }
        const auto r = b.re / b.im, d = b.im + b.re * r;
        return E{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
    } else {
        return a / b;
    }
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

// dst = src on the triangle tri names (BX_TRI_* bits; row >= col for lower,
// row <= col for upper), the other triangle of dst untouched; BX_TRI_HERM
// also zeroes the diagonal's imaginary part
template <typename E>
__global__ void k_tri_copy(int n, int tri, const E *__restrict__ src, int64_t lds,
                           E *__restrict__ dst, int64_t ldd)
{
    const int64_t total = (int64_t)n * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(t % n), c = (int)(t / n);
        if ((tri & BX_TRI_LOWER) ? r < c : r > c) continue;
        E v = src[r + (int64_t)c * lds];
        if ((tri & BX_TRI_HERM) && r == c) v = real_of(v);
        dst[r + (int64_t)c * ldd] = v;
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
    case BLASX_C: return launch_fill<c8>(n, lower, mode, A, lda, st);
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
    case BLASX_C:
        hipLaunchKernelGGL(k_transpose<c8>, grid, dim3(256), 0, st, rows, cols, (const c8 *)in,
                           ldi, (c8 *)out, ldo, conj);
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
    case BLASX_C:
        hipLaunchKernelGGL((k_trsm_leaf<c8, 64>), grid, dim3(64), 0, st, lower, m, n,
                           (const c8 *)T, ldt, (c8 *)X, ldx);
        break;
    default: return BLASX_EINVAL;
    }
    return hipGetLastError() == hipSuccess ? 0 : BLASX_EHIP;
}

int bxh_tri_copy(int dtype, int n, int tri, const void *src, int64_t lds, void *dst, int64_t ldd,
                 void *stream)
{
    if (n <= 0) return 0;
    const int64_t total = (int64_t)n * n;
    int grid = (int)((total + 255) / 256);
    if (grid > 65536) grid = 65536;
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
    case BLASX_S:
        hipLaunchKernelGGL(k_tri_copy<float>, dim3(grid), dim3(256), 0, st, n, tri,
                           (const float *)src, lds, (float *)dst, ldd);
        break;
    case BLASX_D:
        hipLaunchKernelGGL(k_tri_copy<double>, dim3(grid), dim3(256), 0, st, n, tri,
                           (const double *)src, lds, (double *)dst, ldd);
        break;
    case BLASX_Z:
        hipLaunchKernelGGL(k_tri_copy<z16>, dim3(grid), dim3(256), 0, st, n, tri,
                           (const z16 *)src, lds, (z16 *)dst, ldd);
        break;
    case BLASX_C:
        hipLaunchKernelGGL(k_tri_copy<c8>, dim3(grid), dim3(256), 0, st, n, tri, (const c8 *)src,
                           lds, (c8 *)dst, ldd);
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
