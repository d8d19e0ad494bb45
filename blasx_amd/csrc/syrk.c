/*
 * syrk.c -- symmetric and Hermitian rank-k updates on the GPU:
 * cblas_{s,d,c,z}syrk, cblas_{c,z}herk and their f77 forms,
 *   C = alpha op(A) op(A)^T + beta C   (op(A)^H and real alpha, beta for HERK)
 * on the triangle Uplo names.
 *
 * Replaces blas/ssyrk.c:15-95, blas/dsyrk.c:15-95, blas/csyrk.c:15-95,
 * blas/zsyrk.c:15-95, blas/cherk.c:15-108 and blas/zherk.c:15-108, which validate and then always forward to the CPU
 * BLAS (blas/dsyrk.c:46-49: both dispatch branches choose the CPU).  Kept
 * from them: the validation order and info numbers (blas/dsyrk.c:33-38,
 * lowest number wins), the xerbla_ names, the "N or K above 1000" size test
 * the reference computes (blas/dsyrk.c:48; 500 for complex types).  The
 * GPU runs what the reference's unused triangular tile helpers describe
 * (blas/blasx_mem_control.c:3-22, 189-219): only the C tiles on the stored
 * triangle, the diagonal tiles masked.
 *
 * CSYRK and CHERK (complex64) run on the FP32 matrix cores as a full product
 * whose stored triangle is copied back (bx_gpu_crank, level3.c).  For the
 * other types, host operands go through the tile runtime (bx_tile_syrk) and
 * operands in HBM run as one grouped launch here: for each 256-row panel of C, the
 * off-diagonal part is a plain GEMM problem (eligible for the DMA-staged
 * block) and the diagonal 256 x 256 block a masked one.
 *
 * Deviations (netlib semantics): RowMajor is computed (the other triangle
 * with the transposition flipped); ZSYRK rejects ConjTrans (info 2), as
 * netlib does; f77 argument errors report the lowest parameter number.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "blasx.h"
#include "cpu_forward.h"
#include "internal.h"

/* ---------------------------------------------------------------- device */
enum { PANEL = 256 };

static pthread_mutex_t g_slock = PTHREAD_MUTEX_INITIALIZER;
static struct {
    void *host, *dev, *done;
    size_t cap;
} g_sbuf;

/* one problem (C block m x n) with its K segments seg[0..nseg) */
static void add_problem(bxh_problem_t *P, bxh_seg_t *S, int idx, int nseg, const bxh_seg_t *seg,
                        int dt, void *C, int64_t ldc, int m, int n, int opa, int opb,
                        const double *alpha, const double *beta, int tri, int *total)
{
    const int bm = bxh_tile_m(dt), bn = bxh_tile_n(dt);
    memset(P, 0, sizeof *P);
    P->C = C;
    P->ldc = ldc;
    P->m = m;
    P->n = n;
    P->nseg = nseg;
    P->seg_off = idx * nseg;
    P->opa = opa;
    P->opb = opb;
    P->tiles_m = (m + bm - 1) / bm;
    P->tiles_n = (n + bn - 1) / bn;
    P->tile_start = *total;
    *total += P->tiles_m * P->tiles_n;
    P->alpha_re = alpha[0];
    P->alpha_im = alpha[1];
    P->beta_re = beta[0];
    P->beta_im = beta[1];
    P->beta_zero = beta[0] == 0.0 && beta[1] == 0.0;
    P->tri = tri;
    P->tri_off = 0;
    for (int i = 0; i < nseg; ++i) S[i] = seg[i];
}

/* operands, C and the launch on `stream` (NULL: the library's compute
 * stream); asynchronous, like the other device entry points.  With B (SYR2K)
 * C = alpha op(A) op(B)^T + alpha op(B) op(A)^T + beta C on the triangle. */
int bx_device_syrk(int dt, int tri, int opa, int N, int K, const double *alpha, const void *A,
                   int64_t lda, const void *B, int64_t ldb, const double *beta, void *C,
                   int64_t ldc, void *stream)
{
    return bx_device_syr2k(dt, tri, opa, N, K, alpha, A, lda, B, ldb, B, ldb, beta, C, ldc,
                           stream);
}

/* the same with the second term's op(B) rows taken from B2 (HER2K passes a
 * copy of B scaled so that one alpha serves both terms) */
int bx_device_syr2k(int dt, int tri, int opa, int N, int K, const double *alpha, const void *A,
                    int64_t lda, const void *B, int64_t ldb, const void *B2, int64_t ldb2,
                    const double *beta, void *C, int64_t ldc, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    void *st = stream ? stream : g_bx.s_comp;
    if (N <= 0) return 0;
    if (K == 0 || (alpha[0] == 0.0 && alpha[1] == 0.0))
        return bxh_scale_tri(dt, N, N, tri, beta, C, ldc, st);

    const int es = bx_esize(dt);
    const int opb = opa != BX_OP_N ? BX_OP_N : (tri & BX_TRI_HERM) ? BX_OP_C : BX_OP_T;
    const int npan = (N + PANEL - 1) / PANEL, maxp = 2 * npan, nseg = B ? 2 : 1;
    const size_t prob_bytes = ((size_t)maxp * sizeof(bxh_problem_t) + 255) & ~(size_t)255;
    const size_t need = prob_bytes + (size_t)maxp * nseg * sizeof(bxh_seg_t);
    int rc = 0;
    pthread_mutex_lock(&g_slock);
    if (!g_sbuf.done) rc = bxh_event_create(&g_sbuf.done, 0);
    if (!rc) rc = bxh_event_sync(g_sbuf.done); /* the previous call's tables are consumed */
    if (!rc && need > g_sbuf.cap) {
        bxh_host_free(g_sbuf.host);
        bxh_free(g_sbuf.dev);
        g_sbuf.host = g_sbuf.dev = NULL;
        g_sbuf.cap = 0;
        rc = bxh_host_malloc(&g_sbuf.host, need);
        if (!rc) rc = bxh_malloc(&g_sbuf.dev, need);
        if (!rc) g_sbuf.cap = need;
    }
    if (!rc) {
        bxh_problem_t *probs = (bxh_problem_t *)g_sbuf.host;
        bxh_seg_t *segs = (bxh_seg_t *)((char *)g_sbuf.host + prob_bytes);
        const char *a = (const char *)A, *b = B ? (const char *)B : (const char *)A;
        const char *b2 = B2 ? (const char *)B2 : b;
        char *c = (char *)C;
        /* row x of op(A) starts at element x of A (op N) or at column x (T, C) */
        const int64_t step = (opa == BX_OP_N ? 1 : lda) * es;
        const int64_t stepb = (opa == BX_OP_N ? 1 : ldb) * es;
        const int64_t stepb2 = (opa == BX_OP_N ? 1 : ldb2) * es;
        int np = 0, total = 0;
        for (int r0 = 0; r0 < N; r0 += PANEL) {
            const int rows = N - r0 < PANEL ? N - r0 : PANEL;
            /* rows r0.. of op(C) against columns c0.. (count n) */
            int c0 = r0, n = rows, with_tri = tri;
            for (int part = 0; part < 2; ++part) {
                if (part == 1) {
                    with_tri = 0;
                    if (tri & BX_TRI_LOWER)
                        c0 = 0, n = r0;
                    else
                        c0 = r0 + rows, n = N - c0;
                    if (n <= 0) break;
                }
                /* term 1: op(A) rows r0.. x op(B) rows c0..; term 2 (SYR2K): op(B)
                 * rows r0.. x op(A) rows c0..; SYRK has B = A and one term */
                bxh_seg_t sg[2];
                sg[0].A = a + r0 * step;
                sg[0].lda = lda;
                sg[0].B = B ? b + c0 * stepb : a + c0 * step;
                sg[0].ldb = B ? ldb : lda;
                sg[0].k = K;
                sg[1].A = b2 + r0 * stepb2;
                sg[1].lda = B2 ? ldb2 : ldb;
                sg[1].B = a + c0 * step;
                sg[1].ldb = lda;
                sg[1].k = K;
                add_problem(&probs[np], &segs[np * nseg], np, nseg, sg, dt,
                            c + (r0 + (int64_t)c0 * ldc) * es, ldc, rows, n, opa, opb, alpha,
                            beta, with_tri, &total);
                ++np;
            }
        }
        char *dv = (char *)g_sbuf.dev;
        rc = bxh_memcpy_h2d_async(dv, g_sbuf.host, need, st);
        if (!rc)
            rc = bxh_gemm_grouped(dt, (const bxh_problem_t *)dv, np,
                                  (const bxh_seg_t *)(dv + prob_bytes), total, 0, st);
        if (!rc) rc = bxh_event_record(g_sbuf.done, st);
        if (!rc) {
            BX_STAT_ADD(kernel_launches, 1);
            bx_stat_flops(nseg * (dt == BLASX_Z ? 4.0 : 1.0) * (double)N * (N + 1) * K);
        }
    }
    pthread_mutex_unlock(&g_slock);
    return rc;
}

/* ------------------------------------------------------------------ ABI */
static const char *syrk_name(int dt, int herk)
{
    if (herk) return dt == BLASX_C ? "CHERK" : "ZHERK"; /* blas/zherk.c:13, blas/cherk.c:13 */
    return dt == BLASX_S ? "SSYRK " : dt == BLASX_D ? "DSYRK " : dt == BLASX_C ? "CSYRK " : "ZSYRK ";
}

static int trans_ok(int dt, int herk, int t)
{
    if (t == CblasNoTrans) return 1;
    if (herk) return t == CblasConjTrans;
    if (dt == BLASX_Z || dt == BLASX_C) return t == CblasTrans;
    return t == CblasTrans || t == CblasConjTrans; /* C means T for real data */
}

static void scalar2(int dt, int herk, const void *v, double out[2])
{
    if (dt == BLASX_S || (dt == BLASX_C && herk)) {
        out[0] = *(const float *)v, out[1] = 0;
    } else if (dt == BLASX_C) {
        out[0] = ((const float *)v)[0], out[1] = ((const float *)v)[1];
    } else if (dt == BLASX_D || herk) {
        out[0] = *(const double *)v, out[1] = 0;
    } else {
        out[0] = ((const double *)v)[0], out[1] = ((const double *)v)[1];
    }
}

/* the CPU BLAS's own routine, Order passed through, as the reference does
 * (blas/dsyrk.c:58) */
static int cpu_syrk(int dt, int herk, int order, int uplo, int trans, int N, int K,
                    const void *alpha, const void *A, int lda, const void *beta, void *C, int ldc)
{
    if (bx_cpublas_load() != 0) return BLASX_ECPU;
    if (herk && dt == BLASX_C)
        bx_cpu_cherk(order, uplo, trans, N, K, *(const float *)alpha, A, lda,
                     *(const float *)beta, C, ldc);
    else if (herk)
        bx_cpu_zherk(order, uplo, trans, N, K, *(const double *)alpha, A, lda,
                     *(const double *)beta, C, ldc);
    else if (dt == BLASX_C)
        bx_cpu_csyrk(order, uplo, trans, N, K, alpha, A, lda, beta, C, ldc);
    else if (dt == BLASX_S)
        bx_cpu_ssyrk(order, uplo, trans, N, K, *(const float *)alpha, A, lda,
                     *(const float *)beta, C, ldc);
    else if (dt == BLASX_D)
        bx_cpu_dsyrk(order, uplo, trans, N, K, *(const double *)alpha, A, lda,
                     *(const double *)beta, C, ldc);
    else
        bx_cpu_zsyrk(order, uplo, trans, N, K, alpha, A, lda, beta, C, ldc);
    BX_STAT_ADD(calls_cpu, 1);
    return 0;
}

static void syrk_entry(int dt, int herk, int order, int Uplo, int Trans, int N, int K,
                       const void *alpha, const void *A, int lda, const void *beta, void *C,
                       int ldc)
{
    const char *name = syrk_name(dt, herk);
    int uplo = Uplo, trans = Trans, info = 0;
    if (order == CblasRowMajor) {
        /* a row-major C is the column-major C^T: the other triangle, with
         * op(A) and op(A)^T exchanged */
        uplo = Uplo == CblasUpper ? CblasLower : Uplo == CblasLower ? CblasUpper : Uplo;
        if (trans_ok(dt, herk, Trans))
            trans = Trans != CblasNoTrans ? CblasNoTrans : herk ? CblasConjTrans : CblasTrans;
    } else if (order != CblasColMajor) {
        xerbla_(name, &info); /* info 0: bad Order, as OpenBLAS reports it */
        return;
    }
    const int nrowa = trans == CblasNoTrans ? N : K;
    if (ldc < (N > 1 ? N : 1)) info = 10;
    if (lda < (nrowa > 1 ? nrowa : 1)) info = 7;
    if (K < 0) info = 4;
    if (N < 0) info = 3;
    if (!trans_ok(dt, herk, trans)) info = 2;
    if (uplo != CblasUpper && uplo != CblasLower) info = 1;
    if (info != 0) {
        xerbla_(name, &info);
        return;
    }
    if (N == 0) return;
    double al[2], be[2];
    scalar2(dt, herk, alpha, al);
    scalar2(dt, herk, beta, be);
    const int no_product = K == 0 || (al[0] == 0.0 && al[1] == 0.0);
    if (no_product && be[0] == 1.0 && be[1] == 0.0) return;

    bx_config();
    const int thr = (dt == BLASX_Z || dt == BLASX_C) ? g_bx.min_dim_cplx : g_bx.min_dim_real;
    if (!(N > thr || K > thr)) {
        if (cpu_syrk(dt, herk, order, Uplo, Trans, N, K, alpha, A, lda, beta, C, ldc) != 0)
            bx_fatal("%s: no CPU BLAS could be loaded for a CPU-sized call (set BLASX_CPUBLAS)",
                     name);
        return;
    }
    const int tri = (uplo == CblasLower ? BX_TRI_LOWER : BX_TRI_UPPER) | (herk ? BX_TRI_HERM : 0);
    const int opa = trans == CblasNoTrans ? BX_OP_N : herk ? BX_OP_C : BX_OP_T;
    int rc = bx_ensure_init();
    if (rc == 0 && dt == BLASX_C) {
        /* complex64 on the FP32 matrix cores (level3.c) */
        rc = bx_gpu_crank(herk, uplo, trans, N, K, al, A, lda, NULL, 0, be, C, ldc);
        if (rc == BLASX_ENOMEM) {
            bx_log(1, "%s: operands do not fit in HBM; running on the CPU BLAS", name);
            if (cpu_syrk(dt, herk, order, Uplo, Trans, N, K, alpha, A, lda, beta, C, ldc) != 0)
                bx_fatal("%s: no CPU BLAS could be loaded (set BLASX_CPUBLAS)", name);
            return;
        }
        if (rc == 0) {
            BX_STAT_ADD(calls_gpu, 1);
            return;
        }
    } else if (rc == 0) {
        if (bxh_pointer_is_pinned(C) == 2 && (no_product || bxh_pointer_is_pinned(A) == 2)) {
            /* operands already in HBM: run in place */
            rc = bx_device_syrk(dt, tri, opa, N, K, al, A, lda, NULL, 0, be, C, ldc, NULL);
            if (rc == 0) rc = bxh_stream_sync(g_bx.s_comp);
        } else if (no_product) {
            /* C = beta C on the triangle: host work, no product to run */
            if (cpu_syrk(dt, herk, order, Uplo, Trans, N, K, alpha, A, lda, beta, C, ldc) != 0)
                bx_fatal("%s: no CPU BLAS could be loaded (set BLASX_CPUBLAS)", name);
            return;
        } else {
            rc = bx_tile_syrk(dt, tri, opa, N, K, al, A, lda, be, C, ldc);
        }
        if (rc == 0) {
            BX_STAT_ADD(calls_gpu, 1);
            return;
        }
    }
    if (bx_gpu_failed(name, rc, !bx_in_hbm(A) && !bx_in_hbm(C)) &&
        cpu_syrk(dt, herk, order, Uplo, Trans, N, K, alpha, A, lda, beta, C, ldc) != 0)
        bx_fatal("%s: no CPU BLAS could be loaded for the fallback (set BLASX_CPUBLAS)", name);
}

BX_EXPORT void cblas_ssyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                           const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                           const float alpha, const float *A, const int lda, const float beta,
                           float *C, const int ldc)
{
    syrk_entry(BLASX_S, 0, Order, Uplo, Trans, N, K, &alpha, A, lda, &beta, C, ldc);
}

BX_EXPORT void cblas_dsyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                           const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                           const double alpha, const double *A, const int lda, const double beta,
                           double *C, const int ldc)
{
    syrk_entry(BLASX_D, 0, Order, Uplo, Trans, N, K, &alpha, A, lda, &beta, C, ldc);
}

BX_EXPORT void cblas_zsyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                           const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                           const void *alpha, const void *A, const int lda, const void *beta,
                           void *C, const int ldc)
{
    syrk_entry(BLASX_Z, 0, Order, Uplo, Trans, N, K, alpha, A, lda, beta, C, ldc);
}

BX_EXPORT void cblas_zherk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                           const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                           const double alpha, const void *A, const int lda, const double beta,
                           void *C, const int ldc)
{
    syrk_entry(BLASX_Z, 1, Order, Uplo, Trans, N, K, &alpha, A, lda, &beta, C, ldc);
}

BX_EXPORT void cblas_csyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                           const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                           const void *alpha, const void *A, const int lda, const void *beta,
                           void *C, const int ldc)
{
    syrk_entry(BLASX_C, 0, Order, Uplo, Trans, N, K, alpha, A, lda, beta, C, ldc);
}

BX_EXPORT void cblas_cherk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                           const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                           const float alpha, const void *A, const int lda, const float beta,
                           void *C, const int ldc)
{
    syrk_entry(BLASX_C, 1, Order, Uplo, Trans, N, K, &alpha, A, lda, &beta, C, ldc);
}

/* f77: info 1 for a bad UPLO, 2 for a bad TRANS character; both checks
 * run and the later one wins, so two bad characters report 2
 * (blas/dsyrk.c:83-85, blas/zherk.c:89-91), then the ColMajor CBLAS path */
static int f77_syrk_args(const char *name, const char *uplo, const char *trans, int *u, int *t)
{
    int info = 0;
    switch (*uplo) {
    case 'U': case 'u': *u = CblasUpper; break;
    case 'L': case 'l': *u = CblasLower; break;
    default: info = 1;
    }
    switch (*trans) {
    case 'N': case 'n': *t = CblasNoTrans; break;
    case 'T': case 't': *t = CblasTrans; break;
    case 'C': case 'c': *t = CblasConjTrans; break;
    default: info = 2;
    }
    if (info) xerbla_(name, &info);
    return info;
}

BX_EXPORT void ssyrk_(const char *uplo, const char *trans, const int *n, const int *k,
                      const float *alpha, const float *A, const int *lda, const float *beta,
                      float *C, const int *ldc)
{
    int u, t;
    if (f77_syrk_args("SSYRK ", uplo, trans, &u, &t)) return;
    syrk_entry(BLASX_S, 0, CblasColMajor, u, t, *n, *k, alpha, A, *lda, beta, C, *ldc);
}

BX_EXPORT void dsyrk_(const char *uplo, const char *trans, const int *n, const int *k,
                      const double *alpha, const double *A, const int *lda, const double *beta,
                      double *C, const int *ldc)
{
    int u, t;
    if (f77_syrk_args("DSYRK ", uplo, trans, &u, &t)) return;
    syrk_entry(BLASX_D, 0, CblasColMajor, u, t, *n, *k, alpha, A, *lda, beta, C, *ldc);
}

BX_EXPORT void zsyrk_(const char *uplo, const char *trans, const int *n, const int *k,
                      const void *alpha, const void *A, const int *lda, const void *beta, void *C,
                      const int *ldc)
{
    int u, t;
    if (f77_syrk_args("ZSYRK ", uplo, trans, &u, &t)) return;
    syrk_entry(BLASX_Z, 0, CblasColMajor, u, t, *n, *k, alpha, A, *lda, beta, C, *ldc);
}

BX_EXPORT void zherk_(const char *uplo, const char *trans, const int *n, const int *k,
                      const double *alpha, const void *A, const int *lda, const double *beta,
                      void *C, const int *ldc)
{
    int u, t;
    if (f77_syrk_args("ZHERK", uplo, trans, &u, &t)) return;
    syrk_entry(BLASX_Z, 1, CblasColMajor, u, t, *n, *k, alpha, A, *lda, beta, C, *ldc);
}

BX_EXPORT void csyrk_(const char *uplo, const char *trans, const int *n, const int *k,
                      const void *alpha, const void *A, const int *lda, const void *beta, void *C,
                      const int *ldc)
{
    int u, t;
    if (f77_syrk_args("CSYRK ", uplo, trans, &u, &t)) return;
    syrk_entry(BLASX_C, 0, CblasColMajor, u, t, *n, *k, alpha, A, *lda, beta, C, *ldc);
}

BX_EXPORT void cherk_(const char *uplo, const char *trans, const int *n, const int *k,
                      const float *alpha, const void *A, const int *lda, const float *beta,
                      void *C, const int *ldc)
{
    int u, t;
    if (f77_syrk_args("CHERK", uplo, trans, &u, &t)) return;
    syrk_entry(BLASX_C, 1, CblasColMajor, u, t, *n, *k, alpha, A, *lda, beta, C, *ldc);
}
