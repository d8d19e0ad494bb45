/*
 * abi.c -- the drop-in boundary: cblas_{s,d,c,z}gemm, {s,d,c,z}gemm_ and
 * xerbla_.
 *
 * Replaces blas/dgemm.c:18-154, blas/sgemm.c:16-153, blas/zgemm.c:16-156,
 * blas/cgemm.c:16-142 and blas/error.c:5-8.  Kept identical to the
 * reference: validation order and Fortran info numbers (blas/dgemm.c:46-53),
 * the xerbla_ message, the f77 trans-character mapping
 * (blas/blasx_internal.c:161-174), the size dispatch "any of M,N,K above
 * 1000 (S, D) / 500 (Z, C) goes to the GPU, else the CPU BLAS"
 * (blas/dgemm.c:61-64, blas/zgemm.c:59-62), blocking semantics.  CGEMM,
 * which the reference keeps on the CPU (blas/cgemm.c:59-60), runs on the GPU
 * too, as one real GEMM on the FP32 matrix cores (bx_gpu_gemm_incore,
 * level3.c).  Operands already in HBM are used in place, whatever the size
 * of the call, and are never handed to the CPU BLAS.
 *
 * Failure policy (the reference asserts on device errors): a GPU-sized call
 * whose GPU path fails is logged and rerun on the CPU BLAS, unless an
 * operand is in HBM or BLASX_CPU_FALLBACK=0 (bx_gpu_failed, core.c).
 *
 * Deliberate deviations (netlib/OpenBLAS CBLAS semantics, where the
 * reference's GPU path is wrong): RowMajor is computed as
 * C^T = op(B)^T op(A)^T (the reference ignores Order on its GPU path);
 * alpha == 0 or K == 0 gives C = beta*C without reading A and B; beta == 0
 * does not read C; offsets are 64-bit.  Kept as the reference has it: two bad
 * f77 trans characters report info 2.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "blasx.h"
#include "internal.h"

BX_EXPORT __attribute__((weak)) void xerbla_(const char *srname, const int *info)
{
    printf(" ** On entry to %6s parameter number %2d had an illegal value\n", srname, *info);
    fflush(stdout);
}

static int valid_trans(int t) { return t == CblasNoTrans || t == CblasTrans || t == CblasConjTrans; }

static int op_code(int dt, int t)
{
    if (t == CblasNoTrans) return BX_OP_N;
    if (t == CblasTrans) return BX_OP_T;
    return (dt == BLASX_Z || dt == BLASX_C) ? BX_OP_C : BX_OP_T;
}

static void to_double2(int dt, const void *v, double out[2])
{
    switch (dt) {
    case BLASX_S: out[0] = *(const float *)v; out[1] = 0; break;
    case BLASX_D: out[0] = *(const double *)v; out[1] = 0; break;
    case BLASX_C: out[0] = ((const float *)v)[0]; out[1] = ((const float *)v)[1]; break;
    default: out[0] = ((const double *)v)[0]; out[1] = ((const double *)v)[1]; break;
    }
}

/* C = beta * C on the host (alpha == 0 or K == 0); beta == 0 writes zeros */
static void host_scale(int dt, int m, int n, const double *beta, void *C, int ldc)
{
    int bz = beta[0] == 0.0 && beta[1] == 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            int64_t o = i + (int64_t)j * ldc;
            if (dt == BLASX_D) {
                double *p = (double *)C + o;
                *p = bz ? 0.0 : beta[0] * *p;
            } else if (dt == BLASX_S) {
                float *p = (float *)C + o;
                *p = bz ? 0.0f : (float)beta[0] * *p;
            } else if (dt == BLASX_Z) {
                double *p = (double *)C + 2 * o;
                double r = p[0], im = p[1];
                p[0] = bz ? 0.0 : beta[0] * r - beta[1] * im;
                p[1] = bz ? 0.0 : beta[0] * im + beta[1] * r;
            } else {
                float *p = (float *)C + 2 * o;
                float r = p[0], im = p[1], br = (float)beta[0], bi = (float)beta[1];
                p[0] = bz ? 0.0f : br * r - bi * im;
                p[1] = bz ? 0.0f : br * im + bi * r;
            }
        }
}

static const char *dt_name(int dt)
{
    return dt == BLASX_S ? "SGEMM " : dt == BLASX_D ? "DGEMM " : dt == BLASX_Z ? "ZGEMM " : "CGEMM ";
}

static void gemm_entry(int dt, int order, int TransA, int TransB, int M, int N, int K,
                       const void *alpha, const void *A, int lda, const void *B, int ldb,
                       const void *beta, void *C, int ldc)
{
    const char *name = dt_name(dt);
    int ta = TransA, tb = TransB, m = M, n = N, la = lda, lb = ldb;
    const void *a = A, *b = B;
    int info = 0;
    if (order == CblasRowMajor) {
        /* C^T = op(B)^T op(A)^T: a column-major problem with the roles of
         * A and B exchanged; info numbers refer to that problem, as in
         * OpenBLAS's cblas interface */
        ta = TransB;
        tb = TransA;
        m = N;
        n = M;
        a = B;
        b = A;
        la = ldb;
        lb = lda;
    } else if (order != CblasColMajor) {
        xerbla_(name, &info); /* info 0: bad Order, as OpenBLAS reports it */
        return;
    }
    int nrowa = (ta == CblasNoTrans) ? m : K;
    int nrowb = (tb == CblasNoTrans) ? K : n;
    if (!valid_trans(ta)) info = 1;
    else if (!valid_trans(tb)) info = 2;
    else if (m < 0) info = 3;
    else if (n < 0) info = 4;
    else if (K < 0) info = 5;
    else if (la < (nrowa > 1 ? nrowa : 1)) info = 8;
    else if (lb < (nrowb > 1 ? nrowb : 1)) info = 10;
    else if (ldc < (m > 1 ? m : 1)) info = 13;
    if (info != 0) {
        xerbla_(name, &info);
        return;
    }
    if (m == 0 || n == 0) return;

    double al[2], be[2];
    to_double2(dt, alpha, al);
    to_double2(dt, beta, be);
    const int no_product = K == 0 || (al[0] == 0.0 && al[1] == 0.0);
    if (no_product && be[0] == 1.0 && be[1] == 0.0) return; /* C unchanged */

    /* Operands already in HBM (device pointers) never reach host code: the
     * CPU BLAS and host_scale would dereference them.  A and B count only
     * when the product reads them. */
    const int c_hbm = bx_in_hbm(C);
    const int ab_hbm = !no_product && (bx_in_hbm(a) || bx_in_hbm(b));
    const int hbm = c_hbm || ab_hbm;

    bx_config();
    const int thr = (dt == BLASX_Z || dt == BLASX_C) ? g_bx.min_dim_cplx : g_bx.min_dim_real;
    const int gpu_sized = m > thr || n > thr || K > thr;
    if (!hbm && no_product) {
        host_scale(dt, m, n, be, C, ldc);
        return;
    }
    if (!hbm && !gpu_sized) {
        if (bx_cpu_gemm(dt, order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C,
                        ldc) != 0)
            bx_fatal("%s: no CPU BLAS could be loaded for a CPU-sized call (set BLASX_CPUBLAS)",
                     name);
        return;
    }

    int rc = bx_ensure_init();
    if (rc == 0) {
        const int opa = op_code(dt, ta), opb = op_code(dt, tb);
        if (dt != BLASX_C && c_hbm && (no_product || (bx_in_hbm(a) && bx_in_hbm(b)))) {
            /* every operand the call reads is in HBM: run in place */
            rc = bx_device_gemm(dt, ta, tb, m, n, no_product ? 0 : K, al, a, la, b, lb, be, C,
                                ldc, NULL);
            if (rc == 0) rc = bxh_stream_sync(g_bx.s_comp);
        } else if (dt == BLASX_C || hbm) {
            /* complex64, or operands split between host and HBM: in core */
            rc = bx_gpu_gemm_incore(dt, ta, tb, m, n, no_product ? 0 : K, al, a, la, b, lb, be, C,
                                    ldc);
            if (rc == BLASX_ENOMEM && !hbm) {
                /* host operands beyond HBM: the CPU BLAS, as for any routine run in core */
                bx_log(1, "%s: operands do not fit in HBM; running on the CPU BLAS", name);
                if (bx_cpu_gemm(dt, order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta,
                                C, ldc) != 0)
                    bx_fatal("%s: no CPU BLAS could be loaded (set BLASX_CPUBLAS)", name);
                return;
            }
        } else {
            rc = bx_tile_gemm(dt, opa, opb, m, n, K, al, a, la, b, lb, be, C, ldc);
        }
        if (rc == 0) {
            BX_STAT_ADD(calls_gpu, 1);
            return;
        }
    }
    if (bx_gpu_failed(name, rc, !hbm) &&
        bx_cpu_gemm(dt, order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc) != 0)
        bx_fatal("%s: no CPU BLAS could be loaded for the fallback (set BLASX_CPUBLAS)", name);
}

/* ------------------------------------------------------------ CBLAS */
BX_EXPORT void cblas_sgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_TRANSPOSE TransB, const int M, const int N,
                           const int K, const float alpha, const float *A, const int lda,
                           const float *B, const int ldb, const float beta, float *C,
                           const int ldc)
{
    gemm_entry(BLASX_S, Order, TransA, TransB, M, N, K, &alpha, A, lda, B, ldb, &beta, C, ldc);
}

BX_EXPORT void cblas_dgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_TRANSPOSE TransB, const int M, const int N,
                           const int K, const double alpha, const double *A, const int lda,
                           const double *B, const int ldb, const double beta, double *C,
                           const int ldc)
{
    gemm_entry(BLASX_D, Order, TransA, TransB, M, N, K, &alpha, A, lda, B, ldb, &beta, C, ldc);
}

BX_EXPORT void cblas_cgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_TRANSPOSE TransB, const int M, const int N,
                           const int K, const void *alpha, const void *A, const int lda,
                           const void *B, const int ldb, const void *beta, void *C,
                           const int ldc)
{
    gemm_entry(BLASX_C, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

BX_EXPORT void cblas_zgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_TRANSPOSE TransB, const int M, const int N,
                           const int K, const void *alpha, const void *A, const int lda,
                           const void *B, const int ldb, const void *beta, void *C,
                           const int ldc)
{
    gemm_entry(BLASX_Z, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* ------------------------------------------------------------ Fortran 77 */
static int f77_trans(const char *c, int *out)
{
    switch (*c) {
    case 'N': case 'n': *out = CblasNoTrans; return 0;
    case 'T': case 't': *out = CblasTrans; return 0;
    case 'C': case 'c': *out = CblasConjTrans; return 0;
    default: return -1;
    }
}

/* info 1 / 2 for a bad trans character; both checks run and the later one
 * wins, so two bad characters report 2 (blas/dgemm.c:141-147,
 * blas/sgemm.c:140-146, blas/zgemm.c:143-149, blas/cgemm.c:130-135);
 * otherwise forwards to the ColMajor CBLAS entry */
static int f77_prologue(int dt, const char *transa, const char *transb, int *ta, int *tb)
{
    int info = -1;
    if (f77_trans(transa, ta) < 0) info = 1;
    if (f77_trans(transb, tb) < 0) info = 2;
    if (info >= 0) {
        xerbla_(dt_name(dt), &info);
        return -1;
    }
    return 0;
}

BX_EXPORT void sgemm_(const char *transa, const char *transb, const int *m, const int *n,
                      const int *k, const float *alpha, const float *A, const int *lda,
                      const float *B, const int *ldb, const float *beta, float *C,
                      const int *ldc)
{
    int ta, tb;
    if (f77_prologue(BLASX_S, transa, transb, &ta, &tb)) return;
    gemm_entry(BLASX_S, CblasColMajor, ta, tb, *m, *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);
}

BX_EXPORT void dgemm_(const char *transa, const char *transb, const int *m, const int *n,
                      const int *k, const double *alpha, const double *A, const int *lda,
                      const double *B, const int *ldb, const double *beta, double *C,
                      const int *ldc)
{
    int ta, tb;
    if (f77_prologue(BLASX_D, transa, transb, &ta, &tb)) return;
    gemm_entry(BLASX_D, CblasColMajor, ta, tb, *m, *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);
}

BX_EXPORT void cgemm_(const char *transa, const char *transb, const int *m, const int *n,
                      const int *k, const void *alpha, const void *A, const int *lda,
                      const void *B, const int *ldb, const void *beta, void *C, const int *ldc)
{
    int ta, tb;
    if (f77_prologue(BLASX_C, transa, transb, &ta, &tb)) return;
    gemm_entry(BLASX_C, CblasColMajor, ta, tb, *m, *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);
}

BX_EXPORT void zgemm_(const char *transa, const char *transb, const int *m, const int *n,
                      const int *k, const void *alpha, const void *A, const int *lda,
                      const void *B, const int *ldb, const void *beta, void *C, const int *ldc)
{
    int ta, tb;
    if (f77_prologue(BLASX_Z, transa, transb, &ta, &tb)) return;
    gemm_entry(BLASX_Z, CblasColMajor, ta, tb, *m, *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);
}
