// cgemm.hip -- complex64 GEMM on the FP32 matrix cores.  CDNA has no
// complex or complex-FP32 MFMA; with op(A) = Ar + i Ai (M x K) and
// op(B) = Br + i Bi (K x N) the real block product
//     [Cr  Ci] = [Ar  Ai] [ Br  Bi]
//                         [-Bi  Br]
// is one real GEMM of M x 2N x 2K -- exactly the 8MNK real FLOPs of the
// complex product, each output a sum of 2K FP32 products -- and runs on the
// SGEMM kernels.  The kernels here pack the operands into that form
// (transposing and conjugating as op() asks, through 32 x 32 LDS tiles so
// reads and writes coalesce) and combine the planar result into the
// interleaved C with alpha and beta.  HBM-bound element work.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../hiplayer.h"
#include "blasx_ext.h"

namespace {

struct c8 {
    float re, im;
};

constexpr int TB = 32;

// Ap (M x 2K, ld M): Ap[i, k] = Re op(A)[i, k], Ap[i, K + k] = Im op(A)[i, k];
// op 0: A is M x K; 1 (T), 2 (C): A is K x M.  Block (x, y) owns op(A) rows
// [x*32, +32), columns [y*32, +32).
__global__ __launch_bounds__(256) void k_cpack_a(int M, int K, int op, const c8 *__restrict__ A,
                                                 int64_t lda, float *__restrict__ Ap)
{
    __shared__ c8 tile[TB][TB + 1];
    const int i0 = blockIdx.x * TB, k0 = blockIdx.y * TB;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    if (op != 0) {
        // tile[k - k0][i - i0] = A(k, i), read along k
        for (int ii = ty; ii < TB; ii += 8) {
            const int k = k0 + tx, i = i0 + ii;
            if (k < K && i < M) tile[tx][ii] = A[k + (int64_t)i * lda];
        }
        __syncthreads();
    }
    for (int kk = ty; kk < TB; kk += 8) {
        const int i = i0 + tx, k = k0 + kk;
        if (i >= M || k >= K) continue;
        c8 v = op == 0 ? A[i + (int64_t)k * lda] : tile[kk][tx];
        if (op == 2) v.im = -v.im;
        Ap[i + (int64_t)k * M] = v.re;
        Ap[i + (int64_t)(K + k) * M] = v.im;
    }
}

// Bp (2K x 2N, ld 2K) = [[Br, Bi], [-Bi, Br]] from op(B) (K x N); op 0: B is
// K x N; 1 (T), 2 (C): B is N x K.  Block (x, y) owns op(B) rows [x*32, +32),
// columns [y*32, +32).
__global__ __launch_bounds__(256) void k_cpack_b(int K, int N, int op, const c8 *__restrict__ B,
                                                 int64_t ldb, float *__restrict__ Bp)
{
    __shared__ c8 tile[TB][TB + 1];
    const int k0 = blockIdx.x * TB, j0 = blockIdx.y * TB;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    if (op != 0) {
        // tile[j - j0][k - k0] = B(j, k), read along j
        for (int kk = ty; kk < TB; kk += 8) {
            const int j = j0 + tx, k = k0 + kk;
            if (j < N && k < K) tile[tx][kk] = B[j + (int64_t)k * ldb];
        }
        __syncthreads();
    }
    const int64_t ld = 2 * (int64_t)K;
    for (int jj = ty; jj < TB; jj += 8) {
        const int k = k0 + tx, j = j0 + jj;
        if (k >= K || j >= N) continue;
        c8 v = op == 0 ? B[k + (int64_t)j * ldb] : tile[jj][tx];
        if (op == 2) v.im = -v.im;
        Bp[k + j * ld] = v.re;
        Bp[k + (N + j) * ld] = v.im;
        Bp[K + k + j * ld] = -v.im;
        Bp[K + k + (N + j) * ld] = v.re;
    }
}

// C = alpha (P[:, :N] + i P[:, N:]) + beta C; without P (no product) C = beta C;
// beta == 0 does not read C
__global__ void k_ccombine(int M, int N, const float *__restrict__ P, int64_t ldp, float ar,
                           float ai, float br, float bi, int beta_zero, c8 *__restrict__ C,
                           int64_t ldc)
{
    const int64_t total = (int64_t)M * N;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % M, j = t / M;
        float re = 0.0f, im = 0.0f;
        if (P) {
            const float pr = P[i + j * ldp], pi = P[i + (N + j) * ldp];
            re = ar * pr - ai * pi;
            im = ar * pi + ai * pr;
        }
        c8 *c = C + i + j * ldc;
        if (!beta_zero) {
            const c8 v = *c;
            re += br * v.re - bi * v.im;
            im += br * v.im + bi * v.re;
        }
        *c = c8{re, im};
    }
}

int launched() { return hipGetLastError() == hipSuccess ? 0 : BLASX_EHIP; }

} // namespace

extern "C" {

int bxh_cpack_a(int M, int K, int op, const void *A, int64_t lda, float *Ap, void *stream)
{
    if (M <= 0 || K <= 0) return 0;
    if (op < 0 || op > 2) return BLASX_EINVAL;
    const dim3 grid((M + TB - 1) / TB, (K + TB - 1) / TB);
    hipLaunchKernelGGL(k_cpack_a, grid, dim3(256), 0, (hipStream_t)stream, M, K, op,
                       (const c8 *)A, lda, Ap);
    return launched();
}

int bxh_cpack_b(int K, int N, int op, const void *B, int64_t ldb, float *Bp, void *stream)
{
    if (K <= 0 || N <= 0) return 0;
    if (op < 0 || op > 2) return BLASX_EINVAL;
    const dim3 grid((K + TB - 1) / TB, (N + TB - 1) / TB);
    hipLaunchKernelGGL(k_cpack_b, grid, dim3(256), 0, (hipStream_t)stream, K, N, op,
                       (const c8 *)B, ldb, Bp);
    return launched();
}

int bxh_ccombine(int M, int N, const float *P, int64_t ldp, const double *alpha,
                 const double *beta, void *C, int64_t ldc, void *stream)
{
    if (M <= 0 || N <= 0) return 0;
    const int64_t total = (int64_t)M * N;
    int grid = (int)((total + 255) / 256);
    if (grid > 65536) grid = 65536;
    const int bz = beta[0] == 0.0 && beta[1] == 0.0;
    hipLaunchKernelGGL(k_ccombine, dim3(grid), dim3(256), 0, (hipStream_t)stream, M, N, P, ldp,
                       (float)alpha[0], (float)alpha[1], (float)beta[0], (float)beta[1], bz,
                       (c8 *)C, ldc);
    return launched();
}

} // extern "C"
