// ozaki_kernels.hip -- split-precision DGEMM on the INT8 matrix cores of
// gfx950 (Ozaki scheme: Ozaki, Ogita, Oishi & Rump 2012, error-free
// splitting of matrix products; INT8 realisation after Ootomo, Ozaki &
// Yokota 2024).  No counterpart in the reference.
//
// For C = alpha*op(A)*op(B) + beta*C in FP64:
//  1. scale: row i of op(A) by 2^-e_i, column j of op(B) by 2^-f_j, where
//     e_i, f_j make every scaled entry lie in (-1, 1)  (k_rowcol_exponent);
//  2. split: r <- 128 r; d_p = trunc(r); r <- r - d_p, p = 1..s.  Every
//     step is exact in FP64, |d_p| <= 127, and
//        a = 2^e (sum_p d_p 2^-7p + r_s 2^-7s),  |r_s| < 1   (k_split)
//     digits go to int8 slice matrices, K contiguous, zero padded;
//  3. products: for each digit group g = p + q <= s + 1, the exact integer
//        S_g = sum_{p+q=g} D^A_p (D^B_q)^T
//     is one INT8-MFMA launch whose K range runs over the g-1 slice pairs
//     (int32 accumulation is exact while (g-1)*K*127^2 < 2^31; the host
//     splits longer ranges), folded into an FP64 workspace as
//     W += 2^-7g * S_g, smallest groups first  (k_ozaki_group);
//  4. C = alpha * 2^(e_i + f_j) * W + beta * C  (k_ozaki_finish).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../hiplayer.h"
#include "blasx_ext.h"

namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int OZ_BM = 128, OZ_BN = 128, OZ_BK = 64;
constexpr int OZ_STR = OZ_BK + 16; // LDS row stride in bytes: conflict-free 16-byte reads
constexpr int OZ_NT = 256;

// ---------------------------------------------------------------------------
// 1. exponents: e[i] = smallest integer with max_k |x(i,k)| * 2^-e < 1
//    mode 0: x(i,k) = X[i + k*ld]  (reduce along the strided dimension:
//            one thread per i, coalesced across threads)
//    mode 1: x(i,k) = X[k + i*ld]  (reduce along the contiguous dimension:
//            one block per i)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int exp_of(double amax)
{
    if (amax == 0.0 || !isfinite(amax)) return 0;
    int e;
    frexp(amax, &e); // amax = m * 2^e, m in [0.5, 1)
    return e;
}

__global__ void k_exp_strided(const double *__restrict__ X, int64_t ld, int n, int len,
                              int *__restrict__ e)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double m = 0.0;
    for (int k = 0; k < len; ++k) m = fmax(m, fabs(X[i + (int64_t)k * ld]));
    e[i] = exp_of(m);
}

__global__ void k_exp_contig(const double *__restrict__ X, int64_t ld, int n, int len,
                             int *__restrict__ e)
{
    __shared__ double red[256];
    const int i = blockIdx.x;
    double m = 0.0;
    const double *row = X + (int64_t)i * ld;
    for (int k = threadIdx.x; k < len; k += blockDim.x) m = fmax(m, fabs(row[k]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) e[i] = exp_of(red[0]);
}

// ---------------------------------------------------------------------------
// 2. split into s int8 slices, layout D[p][i][k] (k contiguous, row stride
//    kp, slice stride np*kp), zero for i >= n or k >= len.
//    strided (mode 0): x(i,k) = X[i + k*ld]; staged through LDS so that both
//    the FP64 reads (along i) and the int8 writes (along k) are coalesced.
// ---------------------------------------------------------------------------
constexpr int SP_T = 64; // 64 x 64 element tile per block

__global__ void k_split_strided(const double *__restrict__ X, int64_t ld, int n, int len,
                                const int *__restrict__ e, int s, int8_t *__restrict__ D,
                                int np, int kp)
{
    __shared__ int8_t tile[8][SP_T][SP_T + 4];
    const int i0 = blockIdx.x * SP_T, k0 = blockIdx.y * SP_T;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6; // 64 x 4
    for (int kk = ty; kk < SP_T; kk += 4) {
        const int i = i0 + tx, k = k0 + kk;
        // |r| < 1; r*128 is exact, trunc is exact, the fractional remainder
        // is exactly representable: every digit step is error free
        double r = (i < n && k < len) ? ldexp(X[i + (int64_t)k * ld], -e[i]) : 0.0;
        for (int p = 0; p < s; ++p) {
            r *= 128.0;
            const double t = trunc(r);
            tile[p][tx][kk] = (int8_t)(int)t;
            r -= t;
        }
    }
    __syncthreads();
    // write: each thread writes 16 consecutive k bytes of one row
    const int row = threadIdx.x >> 2, kc = (threadIdx.x & 3) * 16;
    const int i = i0 + row;
    if (i >= np) return;
    for (int p = 0; p < s; ++p) {
        int8_t *dst = D + (int64_t)p * np * kp + (int64_t)i * kp + k0 + kc;
        uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < 16; ++b)
            w[b >> 2] |= ((uint32_t)(uint8_t)tile[p][row][kc + b]) << (8 * (b & 3));
        *(i32x4 *)dst = i32x4{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
    }
}

// contiguous (mode 1): x(i,k) = X[k + i*ld]; one thread per 16 consecutive k
__global__ void k_split_contig(const double *__restrict__ X, int64_t ld, int n, int len,
                               const int *__restrict__ e, int s, int8_t *__restrict__ D, int np,
                               int kp)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int chunks = kp / 16;
    const int64_t total = (int64_t)np * chunks;
    if (t >= total) return;
    const int i = (int)(t / chunks), k0 = (int)(t % chunks) * 16;
    double r[16];
    const int ei = i < n ? e[i] : 0;
#pragma unroll
    for (int b = 0; b < 16; ++b)
        r[b] = (i < n && k0 + b < len) ? ldexp(X[k0 + b + (int64_t)i * ld], -ei) : 0.0;
    for (int p = 0; p < s; ++p) {
        uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            r[b] *= 128.0;
            const double tq = trunc(r[b]);
            r[b] -= tq;
            w[b >> 2] |= ((uint32_t)(uint8_t)(int8_t)(int)tq) << (8 * (b & 3));
        }
        *(i32x4 *)(D + (int64_t)p * np * kp + (int64_t)i * kp + k0) =
            i32x4{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
    }
}

// ---------------------------------------------------------------------------
// 3. one digit group: S = sum over segments of DA_seg * DB_seg^T (int32),
//    W = (first ? 0 : W) + scale * S.  DA/DB rows are K-contiguous and padded
//    to whole block tiles, so loads need no bounds checks.
//    256 threads = 2x2 waves, wave tile 64x64 = 4x4 MFMA_I32_16X16X64_I8.
// ---------------------------------------------------------------------------
struct OzSeg {
    const int8_t *A;
    const int8_t *B;
};

__global__ __launch_bounds__(OZ_NT, 2) void k_ozaki_group(const OzSeg *__restrict__ segs,
                                                          int nseg, int kp, int m, int n,
                                                          double *__restrict__ W, int64_t ldw,
                                                          double scale, int first)
{
    __shared__ __attribute__((aligned(16))) int8_t sA[2][OZ_BM * OZ_STR];
    __shared__ __attribute__((aligned(16))) int8_t sB[2][OZ_BN * OZ_STR];
    const int tiles_n = (n + OZ_BN - 1) / OZ_BN;
    int b = blockIdx.x;
    const int nb = gridDim.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3); // contiguous tiles per XCD
    const int tm = b / tiles_n, tn = b % tiles_n;
    const int m0 = tm * OZ_BM, n0 = tn * OZ_BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int l16 = lane & 15, lg = lane >> 4;

    i32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = i32x4{0, 0, 0, 0};

    // loader: 128 rows x 64 bytes per operand = 512 x 16 B, two per thread
    const int lrow = tid >> 2, lk = (tid & 3) * 16;
    i32x4 ra[2], rb[2];
    auto gload = [&](int seg, int k0) {
        const int8_t *pa = segs[seg].A + (int64_t)(m0 + lrow) * kp + k0 + lk;
        const int8_t *pb = segs[seg].B + (int64_t)(n0 + lrow) * kp + k0 + lk;
        ra[0] = *(const i32x4 *)pa;
        ra[1] = *(const i32x4 *)(pa + (int64_t)64 * kp);
        rb[0] = *(const i32x4 *)pb;
        rb[1] = *(const i32x4 *)(pb + (int64_t)64 * kp);
    };
    auto lstore = [&](int buf) {
        *(i32x4 *)&sA[buf][lrow * OZ_STR + lk] = ra[0];
        *(i32x4 *)&sA[buf][(lrow + 64) * OZ_STR + lk] = ra[1];
        *(i32x4 *)&sB[buf][lrow * OZ_STR + lk] = rb[0];
        *(i32x4 *)&sB[buf][(lrow + 64) * OZ_STR + lk] = rb[1];
    };

    const int ktiles = kp / OZ_BK;
    const int total = nseg * ktiles;
    gload(0, 0);
    lstore(0);
    __syncthreads();
    for (int it = 0; it < total; ++it) {
        const int buf = it & 1;
        const bool more = it + 1 < total;
        if (more) gload((it + 1) / ktiles, ((it + 1) % ktiles) * OZ_BK);
        i32x4 fa[4], fb[4];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
            fa[mb] = *(const i32x4 *)&sA[buf][(wm * 64 + mb * 16 + l16) * OZ_STR + lg * 16];
#pragma unroll
        for (int nb2 = 0; nb2 < 4; ++nb2)
            fb[nb2] = *(const i32x4 *)&sB[buf][(wn * 64 + nb2 * 16 + l16) * OZ_STR + lg * 16];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int nb2 = 0; nb2 < 4; ++nb2)
                acc[mb][nb2] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[mb], fb[nb2], acc[mb][nb2], 0, 0, 0);
        if (more) lstore(buf ^ 1);
        __syncthreads();
    }

    // epilogue: output row of accumulator element v is 4*lg + v (as FP32)
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb2 = 0; nb2 < 4; ++nb2) {
            const int col = n0 + wn * 64 + nb2 * 16 + l16;
            if (col >= n) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm * 64 + mb * 16 + 4 * lg + v;
                if (row >= m) continue;
                double *w = W + row + (int64_t)col * ldw;
                const double add = scale * (double)acc[mb][nb2][v];
                *w = first ? add : *w + add;
            }
        }
}

// Same computation with a 256x256 block tile: 4 waves of 128x128 (8x8
// MFMA_I32_16X16X64_I8 accumulators, 256 registers, held in AGPRs), one
// block per CU.  Twice the INT8 operations per byte of L2 traffic of the
// 128x128 kernel (256 instead of 128 ops/byte) and half the LDS reads per
// MFMA; used when the padded problem has enough 256x256 tiles to fill the GPU.
constexpr int OZ_BIG = 256;

__global__ __launch_bounds__(OZ_NT, 1) void k_ozaki_group_big(const OzSeg *__restrict__ segs,
                                                              int nseg, int kp, int m, int n,
                                                              double *__restrict__ W, int64_t ldw,
                                                              double scale, int first)
{
    __shared__ __attribute__((aligned(16))) int8_t sA[2][OZ_BIG * OZ_STR];
    __shared__ __attribute__((aligned(16))) int8_t sB[2][OZ_BIG * OZ_STR];
    const int tiles_n = (n + OZ_BIG - 1) / OZ_BIG;
    int b = blockIdx.x;
    const int nb = gridDim.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3); // contiguous tiles per XCD
    const int tm = b / tiles_n, tn = b % tiles_n;
    const int m0 = tm * OZ_BIG, n0 = tn * OZ_BIG;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int l16 = lane & 15, lg = lane >> 4;

    i32x4 acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = i32x4{0, 0, 0, 0};

    // loader: 256 rows x 64 bytes per operand = 1024 x 16 B, four per thread
    const int lrow = tid >> 2, lk = (tid & 3) * 16;
    i32x4 ra[4], rb[4];
    auto gload = [&](int seg, int k0) {
        const int8_t *pa = segs[seg].A + (int64_t)(m0 + lrow) * kp + k0 + lk;
        const int8_t *pb = segs[seg].B + (int64_t)(n0 + lrow) * kp + k0 + lk;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ra[i] = *(const i32x4 *)(pa + (int64_t)(64 * i) * kp);
            rb[i] = *(const i32x4 *)(pb + (int64_t)(64 * i) * kp);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            *(i32x4 *)&sA[buf][(lrow + 64 * i) * OZ_STR + lk] = ra[i];
            *(i32x4 *)&sB[buf][(lrow + 64 * i) * OZ_STR + lk] = rb[i];
        }
    };

    const int ktiles = kp / OZ_BK;
    const int total = nseg * ktiles;
    gload(0, 0);
    lstore(0);
    __syncthreads();
    for (int it = 0; it < total; ++it) {
        const int buf = it & 1;
        const bool more = it + 1 < total;
        if (more) gload((it + 1) / ktiles, ((it + 1) % ktiles) * OZ_BK);
        i32x4 fb[8];
#pragma unroll
        for (int nb2 = 0; nb2 < 8; ++nb2)
            fb[nb2] = *(const i32x4 *)&sB[buf][(wn * 128 + nb2 * 16 + l16) * OZ_STR + lg * 16];
#pragma unroll
        for (int mb = 0; mb < 8; ++mb) {
            const i32x4 fa = *(const i32x4 *)&sA[buf][(wm * 128 + mb * 16 + l16) * OZ_STR + lg * 16];
#pragma unroll
            for (int nb2 = 0; nb2 < 8; ++nb2)
                acc[mb][nb2] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa, fb[nb2], acc[mb][nb2], 0, 0, 0);
        }
        if (more) lstore(buf ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int mb = 0; mb < 8; ++mb)
#pragma unroll
        for (int nb2 = 0; nb2 < 8; ++nb2) {
            const int col = n0 + wn * 128 + nb2 * 16 + l16;
            if (col >= n) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm * 128 + mb * 16 + 4 * lg + v;
                if (row >= m) continue;
                double *w = W + row + (int64_t)col * ldw;
                const double add = scale * (double)acc[mb][nb2][v];
                *w = first ? add : *w + add;
            }
        }
}

// 4. C = alpha * 2^(e_i + f_j) * W + beta * C
__global__ void k_ozaki_finish(int m, int n, const double *__restrict__ W, int64_t ldw,
                               const int *__restrict__ ea, const int *__restrict__ eb,
                               double alpha, double beta, int beta_zero, double *__restrict__ C,
                               int64_t ldc)
{
    const int64_t total = (int64_t)m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % m), j = (int)(t / m);
        const double v = alpha * ldexp(W[i + (int64_t)j * ldw], ea[i] + eb[j]);
        double *c = C + i + (int64_t)j * ldc;
        *c = beta_zero ? v : fma(beta, *c, v);
    }
}

// complex support: split a complex matrix (rows x cols, ld) into dense real
// and imaginary planes (ld = rows); conj negates the imaginary plane
__global__ void k_deinterleave(const double *__restrict__ Z, int64_t ld, int rows, int cols,
                               int conj, double *__restrict__ re, double *__restrict__ im)
{
    const int64_t total = (int64_t)rows * cols;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % rows), j = (int)(t / rows);
        const double *z = Z + 2 * (i + (int64_t)j * ld);
        re[t] = z[0];
        im[t] = conj ? -z[1] : z[1];
    }
}

// C = alpha * (Pr + i Pi) + beta * C with a fixed FMA order
__global__ void k_zcombine(int m, int n, const double *__restrict__ pr, const double *__restrict__ pi,
                           double ar, double ai, double br, double bi, int beta_zero,
                           double *__restrict__ C, int64_t ldc)
{
    const int64_t total = (int64_t)m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % m), j = (int)(t / m);
        double *c = C + 2 * (i + (int64_t)j * ldc);
        const double tr = fma(ar, pr[t], -(ai * pi[t]));
        const double ti = fma(ar, pi[t], ai * pr[t]);
        if (beta_zero) {
            c[0] = tr;
            c[1] = ti;
        } else {
            const double cr = c[0], cim = c[1];
            c[0] = fma(br, cr, fma(-bi, cim, tr));
            c[1] = fma(br, cim, fma(bi, cr, ti));
        }
    }
}

thread_local char g_oz_err[256];
int oz_fail(const char *what, hipError_t e)
{
    snprintf(g_oz_err, sizeof g_oz_err, "%s: %s", what, hipGetErrorString(e));
    return BLASX_EHIP;
}

} // namespace

extern "C" {

const char *bxh_ozaki_error(void) { return g_oz_err; }

int bxh_oz_block_m(void) { return OZ_BM; }
int bxh_oz_block_k(void) { return OZ_BK; }

// exponents of the rows of op(X) (n rows, len = reduced extent):
// contiguous != 0: x(i,k) = X[k + i*ld], else x(i,k) = X[i + k*ld]
int bxh_oz_exponents(const double *X, int64_t ld, int n, int len, int contiguous, int *e,
                     void *stream)
{
    if (n <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (contiguous)
        hipLaunchKernelGGL(k_exp_contig, dim3(n), dim3(256), 0, st, X, ld, n, len, e);
    else
        hipLaunchKernelGGL(k_exp_strided, dim3((n + 255) / 256), dim3(256), 0, st, X, ld, n, len,
                           e);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("exponent launch", r);
}

int bxh_oz_split(const double *X, int64_t ld, int n, int len, int contiguous, const int *e,
                 int s, int8_t *D, int np, int kp, void *stream)
{
    if (s < 1 || s > 8 || np % 64 || kp % 64) return BLASX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (contiguous) {
        const int64_t total = (int64_t)np * (kp / 16);
        hipLaunchKernelGGL(k_split_contig, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                           X, ld, n, len, e, s, D, np, kp);
    } else {
        hipLaunchKernelGGL(k_split_strided, dim3(np / SP_T, kp / SP_T), dim3(256), 0, st, X, ld,
                           n, len, e, s, D, np, kp);
    }
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("split launch", r);
}

// segs_dev: nseg (A slice, B slice) pointers, each slice np x kp / mp x kp
int bxh_oz_group(const void *segs_dev, int nseg, int kp, int m, int n, int mp, int np, double *W,
                 int64_t ldw, double scale, int first, void *stream)
{
    if (nseg <= 0) return 0;
    if (mp % OZ_BM || np % OZ_BN || kp % OZ_BK) return BLASX_EINVAL;
    const int big_tiles = (mp % OZ_BIG == 0 && np % OZ_BIG == 0) ? (mp / OZ_BIG) * (np / OZ_BIG) : 0;
    if (big_tiles >= 256) { // enough 256x256 tiles to give every CU one
        hipLaunchKernelGGL(k_ozaki_group_big, dim3(big_tiles), dim3(OZ_NT), 0,
                           (hipStream_t)stream, (const OzSeg *)segs_dev, nseg, kp, m, n, W, ldw,
                           scale, first);
    } else {
        const int grid = (mp / OZ_BM) * (np / OZ_BN);
        hipLaunchKernelGGL(k_ozaki_group, dim3(grid), dim3(OZ_NT), 0, (hipStream_t)stream,
                           (const OzSeg *)segs_dev, nseg, kp, m, n, W, ldw, scale, first);
    }
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("ozaki group launch", r);
}

int bxh_oz_deinterleave(const double *Z, int64_t ld, int rows, int cols, int conj, double *re,
                        double *im, void *stream)
{
    const int64_t total = (int64_t)rows * cols;
    if (total <= 0) return 0;
    int grid = (int)((total + 255) / 256);
    if (grid > 262144) grid = 262144;
    hipLaunchKernelGGL(k_deinterleave, dim3(grid), dim3(256), 0, (hipStream_t)stream, Z, ld,
                       rows, cols, conj, re, im);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("deinterleave launch", r);
}

int bxh_oz_zcombine(int m, int n, const double *pr, const double *pi, const double *alpha,
                    const double *beta, double *C, int64_t ldc, void *stream)
{
    const int64_t total = (int64_t)m * n;
    if (total <= 0) return 0;
    int grid = (int)((total + 255) / 256);
    if (grid > 262144) grid = 262144;
    const int bz = beta[0] == 0.0 && beta[1] == 0.0;
    hipLaunchKernelGGL(k_zcombine, dim3(grid), dim3(256), 0, (hipStream_t)stream, m, n, pr, pi,
                       alpha[0], alpha[1], beta[0], beta[1], bz, C, ldc);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("zcombine launch", r);
}

int bxh_oz_finish(int m, int n, const double *W, int64_t ldw, const int *ea, const int *eb,
                  double alpha, double beta, double *C, int64_t ldc, void *stream)
{
    if (m <= 0 || n <= 0) return 0;
    const int64_t total = (int64_t)m * n;
    int grid = (int)((total + 255) / 256);
    if (grid > 262144) grid = 262144;
    hipLaunchKernelGGL(k_ozaki_finish, dim3(grid), dim3(256), 0, (hipStream_t)stream, m, n, W,
                       ldw, ea, eb, alpha, beta, beta == 0.0 ? 1 : 0, C, ldc);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("ozaki finish launch", r);
}

} // extern "C"
