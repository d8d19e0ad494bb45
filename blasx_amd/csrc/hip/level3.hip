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
// c ? a : b by value, component by component (a conditional on two struct
// lvalues may select their addresses and pin a register array in memory)
template <typename E> __device__ __forceinline__ E pick(bool c, E a, E b)
{
    if constexpr (is_cx<E>::value) return E{c ? a.re : b.re, c ? a.im : b.im};
    else return c ? a : b;
}
// complex division by Smith's method (Smith 1962), which avoids overflow in
// |b|^2: with p the larger-magnitude part of b and o the other, r = o / p and
// d = p + o r; |b.re| >= |b.im| gives ((a.re + a.im r) + i (a.im - a.re r)) / d,
// else ((a.im + a.re r) + i (a.im r - a.re)) / d.  Written with selects, not
// branches, so unrolled callers stay one basic block.
template <typename E> __device__ __forceinline__ E divide(E a, E b)
{
    if constexpr (is_cx<E>::value) {
        const bool big = fabs(b.re) >= fabs(b.im);
        const auto p = big ? b.re : b.im, o = big ? b.im : b.re;
        const auto u = big ? a.re : a.im, v = big ? a.im : a.re;
        const auto r = o / p, d = p + o * r;
        const auto t = v - u * r;
        return E{(u + v * r) / d, (big ? t : -t) / d};
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

// v from lane Q of each group of four lanes, to all four (DPP quad_perm
// [Q,Q,Q,Q], one move per 32-bit word)
template <int Q> __device__ __forceinline__ int quad_bcast_word(int w)
{
    return __builtin_amdgcn_update_dpp(0, w, Q * 0x55, 0xF, 0xF, false);
}
template <int Q> __device__ __forceinline__ float quad_bcast_real(float v)
{
    return __builtin_bit_cast(float, quad_bcast_word<Q>(__builtin_bit_cast(int, v)));
}
template <int Q> __device__ __forceinline__ double quad_bcast_real(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)quad_bcast_word<Q>((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)quad_bcast_word<Q>((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <int Q, typename E> __device__ __forceinline__ E quad_bcast(E v)
{
    if constexpr (is_cx<E>::value) return E{quad_bcast_real<Q>(v.re), quad_bcast_real<Q>(v.im)};
    else return quad_bcast_real<Q>(v);
}
// f(step<0>{}), ..., f(step<N-1>{}): the loop index as a compile-time
// constant, so register arrays indexed by it stay in registers
template <int V> struct step {
    static constexpr int value = V;
};
template <int K, int N, typename F> __device__ __forceinline__ void static_for(F &f)
{
    if constexpr (K < N) {
        f(step<K>{});
        static_for<K + 1, N>(f);
    }
}

// Triangular solve of one diagonal block: T (m x m, m <= NB, lower or upper,
// its own diagonal) times X = B, in place on the m x n block X.  A block of
// LEAF_THREADS threads owns LEAF_COLS columns; T and the NB x LEAF_COLS block
// of X are staged in LDS, each thread issuing all of its global loads before
// the first is waited on.  Four lanes share a column, lane q holding rows
// 4s + q (s < NB/4) in registers.  The solve runs in the loop order of
// netlib's DTRSM (Left, NoTrans: lower k = 0..m-1, upper k = m-1..0;
// x[k] /= T(k,k), then x[i] -= T(i,k) x[k] for the rows below / above): at
// step k the lane owning row k hands x[k] to its three neighbours by a quad
// broadcast, all four divide it alike, and each updates its own rows, so one
// lane does a quarter of the column's FMAs.  T is kept by step (tq[k] holds
// column k, lane q's rows contiguous) for wide LDS reads.  Rows m..NB-1 are
// padded as identity rows and never stored.
constexpr int LEAF_THREADS = 256;
constexpr int LEAF_COLS = LEAF_THREADS / 4;

template <typename E, int NB>
__global__ __launch_bounds__(LEAF_THREADS) void k_trsm_leaf(int lower, int m, int n,
                                                            const E *__restrict__ T, int64_t ldt,
                                                            E *__restrict__ X, int64_t ldx)
{
    constexpr int S = NB / 4;                      // rows per lane
    constexpr int G = LEAF_THREADS / NB;           // staging: column groups
    constexpr int PT = NB / G, PX = LEAF_COLS / G; // staged columns per thread
    static_assert(S * 4 == NB && G * NB == LEAF_THREADS && PT * G == NB &&
                      PX * G == LEAF_COLS, "leaf tiling");
    __shared__ E tq[NB][NB + 1];        // tq[k][q*S + s] = T(4s + q, k)
    __shared__ E xs[NB][LEAF_COLS + 1]; // xs[i][c] = X(i, j0 + c)
    const int tid = threadIdx.x, j0 = blockIdx.x * LEAF_COLS;
    {
        const int r = tid % NB, g = tid / NB;
        E tv[PT], xv[PX];
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int c = g + p * G;
            tv[p] = (r < m && c < m) ? T[r + (int64_t)c * ldt]
                                     : (r == c ? one_of<E>() : zero_of<E>());
        }
#pragma unroll
        for (int p = 0; p < PX; ++p) {
            const int j = j0 + g + p * G;
            xv[p] = (r < m && j < n) ? X[r + (int64_t)j * ldx] : zero_of<E>();
        }
#pragma unroll
        for (int p = 0; p < PT; ++p) tq[g + p * G][(r & 3) * S + (r >> 2)] = tv[p];
#pragma unroll
        for (int p = 0; p < PX; ++p) xs[r][g + p * G] = xv[p];
    }
    __syncthreads();
    const int col = tid >> 2, q = tid & 3;
    E x[S];
#pragma unroll
    for (int s = 0; s < S; ++s) x[s] = xs[4 * s + q][col];
    // step k of the solve; lower runs k = 0..NB-1, upper k = NB-1..0
    auto solve_step = [&](auto kc) {
        constexpr int k = decltype(kc)::value, sk = k >> 2, qk = k & 3;
        const E xk = divide(quad_bcast<qk>(x[sk]), tq[k][qk * S + sk]);
        // slot sk: row k itself on lane qk, the rows after it (below for
        // lower, above for upper) on the lanes past qk in that direction
        const E upd = sub_mul(x[sk], tq[k][q * S + sk], xk);
        x[sk] = pick(q == qk, xk, pick(lower ? q > qk : q < qk, upd, x[sk]));
#pragma unroll
        for (int s = 0; s < S; ++s)
            if (lower ? s > sk : s < sk) x[s] = sub_mul(x[s], tq[k][q * S + s], xk);
    };
    auto lower_step = [&](auto kc) { solve_step(kc); };
    auto upper_step = [&](auto kc) { solve_step(step<NB - 1 - decltype(kc)::value>{}); };
    if (lower) static_for<0, NB>(lower_step);
    else static_for<0, NB>(upper_step);
#pragma unroll
    for (int s = 0; s < S; ++s) xs[4 * s + q][col] = x[s];
    __syncthreads();
    {
        const int r = tid % NB, g = tid / NB;
#pragma unroll
        for (int p = 0; p < PX; ++p) {
            const int c = g + p * G, j = j0 + c;
            if (r < m && j < n) X[r + (int64_t)j * ldx] = xs[r][c];
        }
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
    const dim3 grid((n + LEAF_COLS - 1) / LEAF_COLS);
    switch (dtype) {
    case BLASX_S:
        hipLaunchKernelGGL((k_trsm_leaf<float, 64>), grid, dim3(LEAF_THREADS), 0, st, lower, m, n,
                           (const float *)T, ldt, (float *)X, ldx);
        break;
    case BLASX_D:
        hipLaunchKernelGGL((k_trsm_leaf<double, 64>), grid, dim3(LEAF_THREADS), 0, st, lower, m, n,
                           (const double *)T, ldt, (double *)X, ldx);
        break;
    case BLASX_Z:
        hipLaunchKernelGGL((k_trsm_leaf<z16, 32>), grid, dim3(LEAF_THREADS), 0, st, lower, m, n,
                           (const z16 *)T, ldt, (z16 *)X, ldx);
        break;
    case BLASX_C:
        hipLaunchKernelGGL((k_trsm_leaf<c8, 64>), grid, dim3(LEAF_THREADS), 0, st, lower, m, n,
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
