/*
 * level3.c -- SYMM, HEMM, TRMM, TRSM and SYR2K on the GPU, and their f77 forms:
 *   cblas_{s,d,z}symm, cblas_zhemm  C = alpha A B + beta C (Side Left) or
 *                                   alpha B A + beta C (Right), A symmetric
 *                                   (Hermitian) given by one triangle
 *   cblas_{s,d,z}trmm               B = alpha op(A) B (Left) or alpha B op(A),
 *                                   A triangular, optionally unit diagonal
 *   cblas_{s,d,z}trsm               B = alpha op(A)^-1 B (Left) or alpha B op(A)^-1
 *   cblas_{s,d,z}syr2k              C = alpha op(A) op(B)^T + alpha op(B) op(A)^T
 *                                   + beta C on one triangle
 *   cblas_zher2k                    the Hermitian form: op(B)^H, conj(alpha) on the
 *                                   second term, beta real
 *
 * Replaces blas/{s,d,z}symm.c:16-95, blas/zhemm.c:16-101,
 * blas/{s,d,z}trmm.c:17-90, blas/{s,d,z}trsm.c:16-100 and
 * blas/{s,d,z}syr2k.c:16-110 and blas/zher2k.c:15-118.  The reference
 * validates (blas/dsymm.c:29-36, blas/dtrmm.c:32-40, blas/dsyr2k.c:43-49) and
 * then always forwards to the CPU BLAS: both branches of its dispatch choose
 * the CPU (blas/dsymm.c:44-47, blas/dtrmm.c:48-51, blas/dsyr2k.c:57-60).
 * Kept: the validation order and info numbers (lowest wins), the xerbla_
 * names, the "an extent above 1000 (500 complex)" size test.
 *
 * On the GPU:
 *   SYMM/HEMM are one GEMM on an operand materialised in HBM (bxh_tri_fill
 *   mirrors the stored triangle: O(n^2) bytes against the GEMM's O(n^2 m)
 *   FLOPs).  TRMM materialises the triangular matrix the same way and runs
 *   one grouped launch of panel GEMMs, each over the K range where op(A) is
 *   nonzero (half a full GEMM's work), into an output then copied over B.
 *   TRSM solves op(A) X = alpha B recursively: op(A) materialised (transposed
 *   when asked), the triangle split in halves, one half solved, the other
 *   updated by one GEMM over the solved half, down to 64-row (32 complex)
 *   diagonal blocks solved by substitution (bxh_trsm_leaf).  Side Right runs
 *   as the Left solve of the transposed problem.
 *   SYR2K is the rank-k grouped launch with two K segments per problem
 *   (bx_device_syrk); HER2K the same with a rescaled copy of B in the
 *   second term.
 * Host operands are copied whole to HBM for the call: these routines run in
 * core, and a call whose operands do not fit goes to the CPU BLAS.  Operands
 * already in HBM are used in place.
 *
 * CGEMM (complex64, the ABI in abi.c) also runs here: one real GEMM on the
 * FP32 matrix cores over packed operands (bx_gpu_gemm_incore), and so does
 * a GEMM of any type with an operand already in HBM.
 *
 * Deviations (netlib semantics): RowMajor is computed; ZSYR2K rejects
 * ConjTrans (info 2); ZHER2K reports itself as "ZHER2K".
 */
#include <pthread.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "blasx.h"
#include "cpu_forward.h"
#include "internal.h"

/* ------------------------------------------------------------ device work */
/* the HBM buffers of one call; the most any routine here takes is 8 (complex64
 * TRSM, Right side: A, op(A), B, three packing buffers, T^T, B^T) */
enum { SCRATCH_MAX = 12 };
typedef struct {
    void *buf[SCRATCH_MAX];
    int n;
} scratch_t;

static int scratch_alloc(scratch_t *s, void **p, size_t bytes)
{
    if (s->n >= SCRATCH_MAX) return BLASX_EINVAL;
    if (bytes < 16) bytes = 16;
    if (bxh_malloc(p, bytes) != 0) {
        /* the tile cache keeps its HBM slab between calls: release it, retry
         * (tiles kept resident across calls are dropped with it) */
        bx_cache_t *c = bx_global_cache();
        if (!c) return BLASX_ENOMEM;
        bx_log(1, "releasing the tile cache's HBM for a %zu-byte operand", bytes);
        bx_cache_invalidate_all(c);
        bx_cache_free_memory(c);
        if (bxh_malloc(p, bytes) != 0) return BLASX_ENOMEM;
    }
    s->buf[s->n++] = *p;
    return 0;
}

static void scratch_free(scratch_t *s)
{
    for (int i = 0; i < s->n; ++i) bxh_free(s->buf[i]);
    s->n = 0;
}

/* the caller's rows x cols matrix as a device matrix: itself when it is in
 * HBM, else a dense copy (ld = rows), uploaded when `upload` */
static int dev_view(scratch_t *s, const void *p, int rows, int cols, int ld, int es, int upload,
                    void **dev, int64_t *dld)
{
    if (bxh_pointer_is_pinned(p) == 2) {
        *dev = (void *)p;
        *dld = ld;
        return 0;
    }
    const int rc = scratch_alloc(s, dev, (size_t)rows * cols * es);
    *dld = rows > 1 ? rows : 1;
    if (rc || !upload) return rc;
    return bxh_memcpy2d_h2d_async(*dev, (size_t)rows * es, p, (size_t)ld * es, (size_t)rows * es,
                                  (size_t)cols, g_bx.s_comp);
}

/* a private dense copy of the n x n operand (host or HBM), ld = n */
static int dev_copy(scratch_t *s, const void *p, int n, int ld, int es, void **dev)
{
    const int rc = scratch_alloc(s, dev, (size_t)n * n * es);
    if (rc) return rc;
    const size_t w = (size_t)n * es;
    if (bxh_pointer_is_pinned(p) == 2)
        return bxh_memcpy2d_d2d_async(*dev, w, p, (size_t)ld * es, w, (size_t)n, g_bx.s_comp);
    return bxh_memcpy2d_h2d_async(*dev, w, p, (size_t)ld * es, w, (size_t)n, g_bx.s_comp);
}

/* the device result (rows x cols, leading dimension dld) into the caller's matrix */
static int dev_store(void *p, int ld, const void *dev, int64_t dld, int rows, int cols, int es)
{
    if (p == dev) return 0;
    const size_t w = (size_t)rows * es;
    if (bxh_pointer_is_pinned(p) == 2)
        return bxh_memcpy2d_d2d_async(p, (size_t)ld * es, dev, (size_t)dld * es, w, (size_t)cols,
                                      g_bx.s_comp);
    return bxh_memcpy2d_d2h_async(p, (size_t)ld * es, dev, (size_t)dld * es, w, (size_t)cols,
                                  g_bx.s_comp);
}

static int finish(scratch_t *s, int rc)
{
    const int rs = bxh_stream_sync(g_bx.s_comp);
    scratch_free(s);
    pthread_mutex_unlock(&g_bx.gpu_lock);
    return rc ? rc : rs;
}

/* ---- complex64 on the FP32 matrix cores ----
 * C = alpha op(A) op(B) + beta C as one real GEMM of M x 2N x 2K over the
 * packed operands [Re op(A), Im op(A)] and [[Re, Im], [-Im, Re]] of op(B)
 * (bxh_cpack_a/b), then C = alpha P + beta C (bxh_ccombine).  The packing
 * buffers come from the caller, sized for the largest product it runs. */
typedef struct {
    float *Ap, *Bp, *P;
    size_t nAp, nBp, nP; /* floats */
} cwork_t;

static int cwork_alloc(scratch_t *s, cwork_t *w, size_t nAp, size_t nBp, size_t nP)
{
    int rc = scratch_alloc(s, (void **)&w->Ap, nAp * sizeof(float));
    if (!rc) rc = scratch_alloc(s, (void **)&w->Bp, nBp * sizeof(float));
    if (!rc) rc = scratch_alloc(s, (void **)&w->P, nP * sizeof(float));
    w->nAp = nAp, w->nBp = nBp, w->nP = nP;
    return rc;
}

static int cgemm_dev(cwork_t *w, int transa, int transb, int M, int N, int K, const double *al,
                     const void *A, int64_t lda, const void *B, int64_t ldb, const double *be,
                     void *C, int64_t ldc)
{
    const int opa = transa == CblasNoTrans ? 0 : transa == CblasTrans ? 1 : 2;
    const int opb = transb == CblasNoTrans ? 0 : transb == CblasTrans ? 1 : 2;
    const int product = K > 0 && (al[0] != 0.0 || al[1] != 0.0);
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    int rc = 0;
    if (product) {
        if ((size_t)M * 2 * K > w->nAp || (size_t)4 * K * N > w->nBp || (size_t)M * 2 * N > w->nP)
            return BLASX_EINVAL;
        rc = bxh_cpack_a(M, K, opa, A, lda, w->Ap, g_bx.s_comp);
        if (!rc) rc = bxh_cpack_b(K, N, opb, B, ldb, w->Bp, g_bx.s_comp);
        if (!rc)
            rc = bx_device_gemm(BLASX_S, CblasNoTrans, CblasNoTrans, M, 2 * N, 2 * K, one, w->Ap,
                                M, w->Bp, 2 * (int64_t)K, zero, w->P, M, g_bx.s_comp);
    }
    if (!rc) rc = bxh_ccombine(M, N, product ? w->P : NULL, M, al, be, C, ldc, g_bx.s_comp);
    return rc;
}

/* one GEMM on HBM operands of any element type */
static int gemm_any(int dt, cwork_t *w, int ta, int tb, int M, int N, int K, const double *al,
                    const void *A, int64_t lda, const void *B, int64_t ldb, const double *be,
                    void *C, int64_t ldc)
{
    if (dt == BLASX_C) return cgemm_dev(w, ta, tb, M, N, K, al, A, lda, B, ldb, be, C, ldc);
    return bx_device_gemm(dt, ta, tb, M, N, K, al, A, lda, B, ldb, be, C, ldc, g_bx.s_comp);
}

/* A GEMM through the ABI (abi.c) run in core: complex64 (on the FP32 matrix
 * cores), or any type when an operand is already in HBM.  Operands and C as
 * for the routines here: used in place when in HBM, else copied whole for
 * the call.  alpha == 0 or K == 0 reads neither A nor B, and beta == 0 does
 * not read C (C in HBM is then scaled or zeroed on the device). */
int bx_gpu_gemm_incore(int dt, int transa, int transb, int M, int N, int K, const double *al,
                       const void *A, int lda, const void *B, int ldb, const double *be, void *C,
                       int ldc)
{
    const int es = bx_esize(dt);
    const int ra = transa == CblasNoTrans ? M : K, ca = transa == CblasNoTrans ? K : M;
    const int rb = transb == CblasNoTrans ? K : N, cb = transb == CblasNoTrans ? N : K;
    const int bz = be[0] == 0.0 && be[1] == 0.0;
    const int product = K > 0 && (al[0] != 0.0 || al[1] != 0.0);
    scratch_t s = {{0}, 0};
    cwork_t w = {0};
    void *dA = NULL, *dB = NULL, *dC;
    int64_t lA = 1, lB = 1, lC;
    pthread_mutex_lock(&g_bx.gpu_lock);
    int rc = 0;
    if (product) {
        rc = dev_view(&s, A, ra, ca, lda, es, 1, &dA, &lA);
        if (!rc) rc = dev_view(&s, B, rb, cb, ldb, es, 1, &dB, &lB);
        if (!rc && dt == BLASX_C)
            rc = cwork_alloc(&s, &w, (size_t)M * 2 * K, (size_t)4 * K * N, (size_t)M * 2 * N);
    }
    if (!rc) rc = dev_view(&s, C, M, N, ldc, es, !bz, &dC, &lC);
    if (!rc && !product) /* C = beta C: shapes that pass the device API's checks with K = 0 */
        rc = gemm_any(dt, &w, CblasNoTrans, CblasNoTrans, M, N, 0, al, NULL, M > 1 ? M : 1, NULL,
                      1, be, dC, lC);
    else if (!rc)
        rc = gemm_any(dt, &w, transa, transb, M, N, K, al, dA, lA, dB, lB, be, dC, lC);
    if (!rc) rc = dev_store(C, ldc, dC, lC, M, N, es);
    return finish(&s, rc);
}

/* CSYRK, CHERK (B NULL), CSYR2K and CHER2K: the full N x N product into a copy
 * of C, then its stored triangle copied back (HERK/HER2K: real diagonal).
 * Twice the FLOPs of a triangle-only update, on the FP32 matrix cores. */
int bx_gpu_crank(int herm, int uplo, int trans, int N, int K, const double *al, const void *A,
                 int lda, const void *B, int ldb, const double *be, void *C, int ldc)
{
    const int es = 8;
    const int rows = trans == CblasNoTrans ? N : K, cols = trans == CblasNoTrans ? K : N;
    const int conjt = herm ? CblasConjTrans : CblasTrans;
    const int ta = trans == CblasNoTrans ? CblasNoTrans : conjt;
    const int tb = trans == CblasNoTrans ? conjt : CblasNoTrans;
    const double one[2] = {1.0, 0.0};
    /* the second term of HER2K carries conj(alpha) */
    const double al2[2] = {al[0], herm ? -al[1] : al[1]};
    const int tri = (uplo == CblasLower ? BX_TRI_LOWER : BX_TRI_UPPER) | (herm ? BX_TRI_HERM : 0);
    scratch_t s = {{0}, 0};
    cwork_t w = {0};
    void *dA, *dB = NULL, *dC, *dF;
    int64_t lA, lB = 1, lC;
    pthread_mutex_lock(&g_bx.gpu_lock);
    int rc = dev_view(&s, A, rows, cols, lda, es, 1, &dA, &lA);
    if (!rc && B) rc = dev_view(&s, B, rows, cols, ldb, es, 1, &dB, &lB);
    if (!rc) rc = dev_view(&s, C, N, N, ldc, es, 1, &dC, &lC);
    if (!rc) rc = scratch_alloc(&s, &dF, (size_t)N * N * es);
    if (!rc)
        rc = bxh_memcpy2d_d2d_async(dF, (size_t)N * es, dC, (size_t)lC * es, (size_t)N * es,
                                    (size_t)N, g_bx.s_comp);
    if (!rc) rc = cwork_alloc(&s, &w, (size_t)N * 2 * K, (size_t)4 * K * N, (size_t)N * 2 * N);
    if (!rc) rc = cgemm_dev(&w, ta, tb, N, N, K, al, dA, lA, B ? dB : dA, B ? lB : lA, be, dF, N);
    if (!rc && B) rc = cgemm_dev(&w, ta, tb, N, N, K, al2, dB, lB, dA, lA, one, dF, N);
    if (!rc) rc = bxh_tri_copy(BLASX_C, N, tri, dF, N, dC, lC, g_bx.s_comp);
    if (!rc) rc = dev_store(C, ldc, dC, lC, N, N, es);
    return finish(&s, rc);
}

static int gpu_symm(int dt, int herm, int side, int uplo, int M, int N, const double *al,
                    const void *A, int lda, const void *B, int ldb, const double *be, void *C,
                    int ldc)
{
    const int es = bx_esize(dt), na = side == CblasLeft ? M : N;
    const int bz = be[0] == 0.0 && be[1] == 0.0;
    scratch_t s = {{0}, 0};
    void *dA, *dB, *dC;
    int64_t lB, lC;
    pthread_mutex_lock(&g_bx.gpu_lock);
    int rc = dev_copy(&s, A, na, lda, es, &dA);
    if (!rc)
        rc = bxh_tri_fill(dt, na, uplo == CblasLower, herm ? BX_FILL_HERM : BX_FILL_SYM, dA, na,
                          g_bx.s_comp);
    if (!rc) rc = dev_view(&s, B, M, N, ldb, es, 1, &dB, &lB);
    if (!rc) rc = dev_view(&s, C, M, N, ldc, es, !bz, &dC, &lC);
    cwork_t w = {0};
    if (!rc && dt == BLASX_C)
        rc = cwork_alloc(&s, &w, (size_t)M * 2 * na, (size_t)4 * na * N, (size_t)M * 2 * N);
    if (!rc)
        rc = side == CblasLeft ? gemm_any(dt, &w, CblasNoTrans, CblasNoTrans, M, N, M, al, dA, na,
                                          dB, lB, be, dC, lC)
                               : gemm_any(dt, &w, CblasNoTrans, CblasNoTrans, M, N, N, al, dB, lB,
                                          dA, na, be, dC, lC);
    if (!rc) rc = dev_store(C, ldc, dC, lC, M, N, es);
    return finish(&s, rc);
}

enum { TRMM_PANEL = 256 };

static int gpu_trmm(int dt, int side, int uplo, int trans, int diag, int M, int N,
                    const double *al, const void *A, int lda, void *B, int ldb)
{
    const int es = bx_esize(dt), na = side == CblasLeft ? M : N;
    scratch_t s = {{0}, 0};
    void *dT, *dB, *dOut;
    int64_t lB;
    pthread_mutex_lock(&g_bx.gpu_lock);
    int rc = dev_copy(&s, A, na, lda, es, &dT);
    if (!rc)
        rc = bxh_tri_fill(dt, na, uplo == CblasLower, diag == CblasUnit ? BX_FILL_UNIT : BX_FILL_TRI,
                          dT, na, g_bx.s_comp);
    if (!rc) rc = dev_view(&s, B, M, N, ldb, es, 1, &dB, &lB);
    if (!rc) rc = scratch_alloc(&s, &dOut, (size_t)M * N * es);
    if (dt == BLASX_C) {
        /* complex64: one full product on the FP32 matrix cores */
        const double zero[2] = {0.0, 0.0};
        cwork_t w = {0};
        if (!rc) rc = cwork_alloc(&s, &w, (size_t)M * 2 * na, (size_t)4 * na * N, (size_t)M * 2 * N);
        if (!rc)
            rc = side == CblasLeft ? gemm_any(dt, &w, trans, CblasNoTrans, M, N, M, al, dT, na, dB,
                                              lB, zero, dOut, M)
                                   : gemm_any(dt, &w, CblasNoTrans, trans, M, N, N, al, dB, lB, dT,
                                              na, zero, dOut, M);
        if (!rc) rc = dev_store(B, ldb, dOut, M, M, N, es);
        return finish(&s, rc);
    }
    /* op(T) is triangular, so each 256-wide panel of the output (rows for
     * Left, columns for Right) needs only the K range where op(T) is nonzero:
     * one grouped launch of panel GEMMs does half a full GEMM's work */
    const int npan = (na + TRMM_PANEL - 1) / TRMM_PANEL;
    blasx_gemm_desc_t *d = rc ? NULL : (blasx_gemm_desc_t *)calloc(npan, sizeof *d);
    if (!rc && !d) rc = BLASX_ENOMEM;
    if (!rc) {
        const int op_lower = (uplo == CblasLower) == (trans == CblasNoTrans);
        const char *t = (const char *)dT, *b = (const char *)dB;
        char *o = (char *)dOut;
        for (int i = 0; i < npan; ++i) {
            const int x0 = i * TRMM_PANEL, w = na - x0 < TRMM_PANEL ? na - x0 : TRMM_PANEL;
            blasx_gemm_desc_t *q = &d[i];
            q->dtype = dt;
            q->alpha[0] = al[0];
            q->alpha[1] = al[1];
            int k0, k1;
            if (side == CblasLeft) {
                /* out[x0.., :] = alpha op(T)[x0.., k0:k1] B[k0:k1, :] */
                k0 = op_lower ? 0 : x0;
                k1 = op_lower ? x0 + w : M;
                q->transa = trans;
                q->transb = CblasNoTrans;
                q->M = w, q->N = N, q->K = k1 - k0;
                /* op(T)[x0, k0] is T[x0, k0] (N) or T[k0, x0] (T, C) */
                q->A = t + (trans == CblasNoTrans ? x0 + (int64_t)k0 * na : k0 + (int64_t)x0 * na) * es;
                q->lda = na;
                q->B = b + (int64_t)k0 * es;
                q->ldb = lB;
                q->C = o + (int64_t)x0 * es;
            } else {
                /* out[:, x0..] = alpha B[:, k0:k1] op(T)[k0:k1, x0..] */
                k0 = op_lower ? x0 : 0;
                k1 = op_lower ? N : x0 + w;
                q->transa = CblasNoTrans;
                q->transb = trans;
                q->M = M, q->N = w, q->K = k1 - k0;
                q->A = b + (int64_t)k0 * lB * es;
                q->lda = lB;
                q->B = t + (trans == CblasNoTrans ? k0 + (int64_t)x0 * na : x0 + (int64_t)k0 * na) * es;
                q->ldb = na;
                q->C = o + (int64_t)x0 * M * es;
            }
            q->ldc = M;
        }
        rc = blasx_gemm_grouped_device(d, npan, g_bx.s_comp);
    }
    if (!rc) rc = dev_store(B, ldb, dOut, M, M, N, es);
    rc = finish(&s, rc); /* the launch has consumed the descriptors */
    free(d);
    return rc;
}

/* X = T^-1 X in place (m x n), T triangular m x m: split T in halves at a
 * multiple of the leaf size, solve one half, update the other with one GEMM
 * (K = the solved half), recurse; leaves are substitution kernels */
static int trsm_rec(int dt, cwork_t *w, int lower, const char *T, int64_t ldt, char *X,
                    int64_t ldx, int m, int n, int es)
{
    const int nb = bxh_trsm_leaf_nb(dt);
    if (m <= nb) return bxh_trsm_leaf(dt, lower, m, n, T, ldt, X, ldx, g_bx.s_comp);
    const int m1 = (m / 2 + nb - 1) / nb * nb, m2 = m - m1;
    const double minus_one[2] = {-1.0, 0.0}, one[2] = {1.0, 0.0};
    const char *T22 = T + (m1 + (int64_t)m1 * ldt) * es;
    char *X2 = X + (int64_t)m1 * es;
    int rc;
    if (lower) {
        /* [T11 0; T21 T22] [X1; X2]: X1 first, then X2 -= T21 X1 */
        if ((rc = trsm_rec(dt, w, 1, T, ldt, X, ldx, m1, n, es))) return rc;
        if ((rc = gemm_any(dt, w, CblasNoTrans, CblasNoTrans, m2, n, m1, minus_one,
                           T + (int64_t)m1 * es, ldt, X, ldx, one, X2, ldx)))
            return rc;
        return trsm_rec(dt, w, 1, T22, ldt, X2, ldx, m2, n, es);
    }
    /* [T11 T12; 0 T22]: X2 first, then X1 -= T12 X2 */
    if ((rc = trsm_rec(dt, w, 0, T22, ldt, X2, ldx, m2, n, es))) return rc;
    if ((rc = gemm_any(dt, w, CblasNoTrans, CblasNoTrans, m1, n, m2, minus_one,
                       T + (int64_t)m1 * ldt * es, ldt, X2, ldx, one, X, ldx)))
        return rc;
    return trsm_rec(dt, w, 0, T, ldt, X, ldx, m1, n, es);
}

/* C = alpha C for any element type */
static int scale_any(int dt, int m, int n, const double *al, void *C, int64_t ldc)
{
    if (dt == BLASX_C) {
        const double zero[2] = {0.0, 0.0};
        return bxh_ccombine(m, n, NULL, 0, zero, al, C, ldc, g_bx.s_comp);
    }
    return bxh_scale(dt, m, n, al, C, ldc, g_bx.s_comp);
}

static int gpu_trsm(int dt, int side, int uplo, int trans, int diag, int M, int N,
                    const double *al, const void *A, int lda, void *B, int ldb)
{
    const int es = bx_esize(dt), na = side == CblasLeft ? M : N;
    const int alpha_one = al[0] == 1.0 && al[1] == 0.0;
    scratch_t s = {{0}, 0};
    void *dA, *dT, *dB, *dX;
    int64_t lB;
    pthread_mutex_lock(&g_bx.gpu_lock);
    /* T = op(A), dense and triangular (lower when lowerT) */
    int rc = dev_copy(&s, A, na, lda, es, &dA);
    int lowerT = uplo == CblasLower;
    dT = dA;
    if (!rc && trans != CblasNoTrans) {
        rc = scratch_alloc(&s, &dT, (size_t)na * na * es);
        if (!rc) rc = bxh_transpose(dt, na, na, dA, na, dT, na, trans == CblasConjTrans, g_bx.s_comp);
        lowerT = !lowerT;
    }
    if (!rc && diag == CblasUnit) rc = bxh_tri_fill(dt, na, lowerT, BX_FILL_UNIT, dT, na, g_bx.s_comp);
    if (!rc) rc = dev_view(&s, B, M, N, ldb, es, 1, &dB, &lB);
    /* complex64 updates: packing buffers for the largest one (2 m1 m2 <= na^2) */
    cwork_t w = {0};
    const int nn = side == CblasLeft ? N : M;
    if (!rc && dt == BLASX_C)
        rc = cwork_alloc(&s, &w, (size_t)na * na, (size_t)4 * na * nn, (size_t)2 * na * nn);
    if (!rc && side == CblasLeft) {
        /* op(A) X = alpha B, solved in place */
        if (!alpha_one) rc = scale_any(dt, M, N, al, dB, lB);
        if (!rc) rc = trsm_rec(dt, &w, lowerT, dT, na, dB, lB, M, N, es);
    } else if (!rc) {
        /* X op(A) = alpha B  <=>  op(A)^T X^T = alpha B^T: solve on B^T with T^T */
        void *dTt;
        rc = scratch_alloc(&s, &dTt, (size_t)na * na * es);
        if (!rc) rc = bxh_transpose(dt, na, na, dT, na, dTt, na, 0, g_bx.s_comp);
        if (!rc) rc = scratch_alloc(&s, &dX, (size_t)N * M * es);
        if (!rc) rc = bxh_transpose(dt, M, N, dB, lB, dX, N, 0, g_bx.s_comp);
        if (!rc && !alpha_one) rc = scale_any(dt, N, M, al, dX, N);
        if (!rc) rc = trsm_rec(dt, &w, !lowerT, dTt, na, dX, N, N, M, es);
        if (!rc) rc = bxh_transpose(dt, N, M, dX, N, dB, lB, 0, g_bx.s_comp);
    }
    if (!rc) rc = dev_store(B, ldb, dB, lB, M, N, es);
    return finish(&s, rc);
}

static int gpu_syr2k(int dt, int uplo, int trans, int N, int K, const double *al, const void *A,
                     int lda, const void *B, int ldb, const double *be, void *C, int ldc)
{
    const int es = bx_esize(dt);
    const int rows = trans == CblasNoTrans ? N : K, cols = trans == CblasNoTrans ? K : N;
    scratch_t s = {{0}, 0};
    void *dA, *dB, *dC;
    int64_t lA, lB, lC;
    pthread_mutex_lock(&g_bx.gpu_lock);
    int rc = dev_view(&s, A, rows, cols, lda, es, 1, &dA, &lA);
    if (!rc) rc = dev_view(&s, B, rows, cols, ldb, es, 1, &dB, &lB);
    /* C goes back whole, so its other triangle is uploaded even when beta == 0 */
    if (!rc) rc = dev_view(&s, C, N, N, ldc, es, 1, &dC, &lC);
    if (!rc)
        rc = bx_device_syrk(dt, uplo == CblasLower ? BX_TRI_LOWER : BX_TRI_UPPER,
                            trans == CblasNoTrans ? BX_OP_N : BX_OP_T, N, K, al, dA, lA, dB, lB,
                            be, dC, lC, g_bx.s_comp);
    if (!rc) rc = dev_store(C, ldc, dC, lC, N, N, es);
    return finish(&s, rc);
}

/* The unimodular factor of HER2K's second term: conj(a)/a when conj_ratio,
 * else a/conj(a), for a finite a != 0.  With u = a/|a| (hypot: no overflow
 * or underflow at any magnitude of a) they are conj(u)^2 and u^2. */
void bx_her2k_factor(const double *a, int conj_ratio, double *out)
{
    const double r = hypot(a[0], a[1]), ur = a[0] / r, ui = a[1] / r;
    const double im = 2.0 * ur * ui;
    out[0] = ur * ur - ui * ui;
    out[1] = conj_ratio ? -im : im;
}

/* HER2K: C = alpha op(A) op(B)^H + conj(alpha) op(B) op(A)^H + beta C, beta
 * real, the diagonal real.  The grouped problem has one alpha, so the second
 * term reads a copy of B scaled by s, |s| = 1: conj(alpha)/alpha for NoTrans
 * (alpha (sB) A^H = conj(alpha) B A^H) and alpha/conj(alpha) for ConjTrans
 * (alpha (sB)^H A = conj(alpha) B^H A). */
static int gpu_her2k(int uplo, int trans, int N, int K, const double *al, const void *A, int lda,
                     const void *B, int ldb, double beta, void *C, int ldc)
{
    const int es = 16;
    const int rows = trans == CblasNoTrans ? N : K, cols = trans == CblasNoTrans ? K : N;
    const double be[2] = {beta, 0.0};
    scratch_t s = {{0}, 0};
    void *dA, *dB, *dB2 = NULL, *dC;
    int64_t lA, lB, lC;
    pthread_mutex_lock(&g_bx.gpu_lock);
    int rc = dev_view(&s, A, rows, cols, lda, es, 1, &dA, &lA);
    if (!rc) rc = dev_view(&s, B, rows, cols, ldb, es, 1, &dB, &lB);
    const int product = K > 0 && (al[0] != 0.0 || al[1] != 0.0);
    if (!rc && product) {
        double sc[2];
        bx_her2k_factor(al, trans == CblasNoTrans, sc);
        rc = scratch_alloc(&s, &dB2, (size_t)rows * cols * es);
        if (!rc)
            rc = bxh_memcpy2d_d2d_async(dB2, (size_t)rows * es, dB, (size_t)lB * es,
                                        (size_t)rows * es, (size_t)cols, g_bx.s_comp);
        if (!rc) rc = bxh_scale(BLASX_Z, rows, cols, sc, dB2, rows, g_bx.s_comp);
    }
    if (!rc) rc = dev_view(&s, C, N, N, ldc, es, 1, &dC, &lC);
    if (!rc)
        rc = bx_device_syr2k(BLASX_Z, (uplo == CblasLower ? BX_TRI_LOWER : BX_TRI_UPPER) | BX_TRI_HERM,
                             trans == CblasNoTrans ? BX_OP_N : BX_OP_C, N, K, al, dA, lA, dB, lB,
                             dB2 ? dB2 : dB, dB2 ? rows : lB, be, dC, lC, g_bx.s_comp);
    if (!rc) rc = dev_store(C, ldc, dC, lC, N, N, es);
    return finish(&s, rc);
}

/* ------------------------------------------------------------------ ABI */
static void scalar2(int dt, const void *v, double out[2])
{
    if (dt == BLASX_S)
        out[0] = *(const float *)v, out[1] = 0;
    else if (dt == BLASX_D)
        out[0] = *(const double *)v, out[1] = 0;
    else if (dt == BLASX_C)
        out[0] = ((const float *)v)[0], out[1] = ((const float *)v)[1];
    else
        out[0] = ((const double *)v)[0], out[1] = ((const double *)v)[1];
}

static int is_cplx(int dt) { return dt == BLASX_Z || dt == BLASX_C; }

static int flip_side(int s) { return s == CblasLeft ? CblasRight : s == CblasRight ? CblasLeft : s; }
static int flip_uplo(int u) { return u == CblasUpper ? CblasLower : u == CblasLower ? CblasUpper : u; }
static int side_ok(int s) { return s == CblasLeft || s == CblasRight; }
static int uplo_ok(int u) { return u == CblasUpper || u == CblasLower; }
static int max1(int v) { return v > 1 ? v : 1; }

static int gpu_sized(int dt, int a, int b)
{
    bx_config();
    const int thr = is_cplx(dt) ? g_bx.min_dim_cplx : g_bx.min_dim_real;
    return a > thr || b > thr;
}

/* after a GPU attempt: done, or the CPU BLAS (host operands beyond HBM, or
 * a failure under the default policy), or a loud stop (bx_gpu_failed) */
static int gpu_outcome(const char *name, int rc, const void *A, const void *B, const void *C)
{
    if (rc == 0) {
        BX_STAT_ADD(calls_gpu, 1);
        return 0;
    }
    const int host = !bx_in_hbm(A) && !bx_in_hbm(B) && !bx_in_hbm(C);
    if (rc == BLASX_ENOMEM && host) {
        bx_log(1, "%s: operands do not fit in HBM; running on the CPU BLAS", name);
        return 1;
    }
    return bx_gpu_failed(name, rc, host);
}

static void need_cpublas(const char *name)
{
    if (bx_cpublas_load() != 0)
        bx_fatal("%s: no CPU BLAS could be loaded (set BLASX_CPUBLAS)", name);
    BX_STAT_ADD(calls_cpu, 1);
}

/* ---- SYMM / HEMM ---- */
static void symm_entry(int dt, int herm, int order, int Side, int Uplo, int M, int N,
                       const void *alpha, const void *A, int lda, const void *B, int ldb,
                       const void *beta, void *C, int ldc)
{
    /* blas/zhemm.c:14 and blas/chemm.c:14 name themselves without the trailing blank */
    const char *name = herm ? (dt == BLASX_C ? "CHEMM" : "ZHEMM")
                       : dt == BLASX_S ? "SSYMM " : dt == BLASX_D ? "DSYMM "
                       : dt == BLASX_C ? "CSYMM " : "ZSYMM ";
    int side = Side, uplo = Uplo, m = M, n = N, info = 0;
    if (order == CblasRowMajor) {
        /* a row-major C is the column-major C^T = B^T A^T: the other side and
         * triangle, M and N exchanged (A^T is A, or conj(A) for HEMM, which is
         * what the column-major view of the stored triangle holds) */
        side = flip_side(Side);
        uplo = flip_uplo(Uplo);
        m = N;
        n = M;
    } else if (order != CblasColMajor) {
        xerbla_(name, &info);
        return;
    }
    const int nrowa = side == CblasLeft ? m : n;
    if (!side_ok(side)) info = 1;
    else if (!uplo_ok(uplo)) info = 2;
    else if (m < 0) info = 3;
    else if (n < 0) info = 4;
    else if (lda < max1(nrowa)) info = 7;
    else if (ldb < max1(m)) info = 9;
    else if (ldc < max1(m)) info = 12;
    if (info) {
        xerbla_(name, &info);
        return;
    }
    if (m == 0 || n == 0) return;
    double al[2], be[2];
    scalar2(dt, alpha, al);
    scalar2(dt, beta, be);
    if (al[0] == 0.0 && al[1] == 0.0 && be[0] == 1.0 && be[1] == 0.0) return;
    if (gpu_sized(dt, m, n)) {
        int rc = bx_ensure_init();
        if (!rc) rc = gpu_symm(dt, herm, side, uplo, m, n, al, A, lda, B, ldb, be, C, ldc);
        if (!gpu_outcome(name, rc, A, B, C)) return;
    }
    need_cpublas(name);
    if (herm && dt == BLASX_C)
        bx_cpu_chemm(order, Side, Uplo, M, N, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (herm)
        bx_cpu_zhemm(order, Side, Uplo, M, N, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (dt == BLASX_C)
        bx_cpu_csymm(order, Side, Uplo, M, N, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (dt == BLASX_S)
        bx_cpu_ssymm(order, Side, Uplo, M, N, *(const float *)alpha, A, lda, B, ldb,
                     *(const float *)beta, C, ldc);
    else if (dt == BLASX_D)
        bx_cpu_dsymm(order, Side, Uplo, M, N, *(const double *)alpha, A, lda, B, ldb,
                     *(const double *)beta, C, ldc);
    else
        bx_cpu_zsymm(order, Side, Uplo, M, N, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* ---- TRMM ---- */
static void trmm_entry(int dt, int order, int Side, int Uplo, int TransA, int Diag, int M, int N,
                       const void *alpha, const void *A, int lda, void *B, int ldb)
{
    const char *name = dt == BLASX_S   ? "STRMM "
                       : dt == BLASX_D ? "DTRMM "
                       : dt == BLASX_C ? "CTRMM "
                                       : "ZTRMM ";
    int side = Side, uplo = Uplo, m = M, n = N, info = 0;
    if (order == CblasRowMajor) {
        /* B^T = op(A)^T B^T (or B^T op(A)^T): the other side, and A's stored
         * triangle read column-major is the other triangle of A^T */
        side = flip_side(Side);
        uplo = flip_uplo(Uplo);
        m = N;
        n = M;
    } else if (order != CblasColMajor) {
        xerbla_(name, &info);
        return;
    }
    const int nrowa = side == CblasLeft ? m : n;
    if (!side_ok(side)) info = 1;
    else if (!uplo_ok(uplo)) info = 2;
    else if (TransA != CblasNoTrans && TransA != CblasTrans && TransA != CblasConjTrans) info = 3;
    else if (Diag != CblasNonUnit && Diag != CblasUnit) info = 4;
    else if (m < 0) info = 5;
    else if (n < 0) info = 6;
    else if (lda < max1(nrowa)) info = 9;
    else if (ldb < max1(m)) info = 11;
    if (info) {
        xerbla_(name, &info);
        return;
    }
    if (m == 0 || n == 0) return;
    double al[2];
    scalar2(dt, alpha, al);
    if (gpu_sized(dt, m, n)) {
        int rc = bx_ensure_init();
        if (!rc) rc = gpu_trmm(dt, side, uplo, TransA, Diag, m, n, al, A, lda, B, ldb);
        if (!gpu_outcome(name, rc, A, B, NULL)) return;
    }
    need_cpublas(name);
    if (dt == BLASX_S)
        bx_cpu_strmm(order, Side, Uplo, TransA, Diag, M, N, *(const float *)alpha, A, lda, B, ldb);
    else if (dt == BLASX_D)
        bx_cpu_dtrmm(order, Side, Uplo, TransA, Diag, M, N, *(const double *)alpha, A, lda, B,
                     ldb);
    else if (dt == BLASX_C)
        bx_cpu_ctrmm(order, Side, Uplo, TransA, Diag, M, N, alpha, A, lda, B, ldb);
    else
        bx_cpu_ztrmm(order, Side, Uplo, TransA, Diag, M, N, alpha, A, lda, B, ldb);
}

/* ---- TRSM ---- */
static void trsm_entry(int dt, int order, int Side, int Uplo, int TransA, int Diag, int M, int N,
                       const void *alpha, const void *A, int lda, void *B, int ldb)
{
    const char *name = dt == BLASX_S   ? "STRSM "
                       : dt == BLASX_D ? "DTRSM "
                       : dt == BLASX_C ? "CTRSM "
                                       : "ZTRSM ";
    int side = Side, uplo = Uplo, m = M, n = N, info = 0;
    if (order == CblasRowMajor) {
        /* the same mapping as TRMM: B^T op(A)^T or op(A)^T B^T on the
         * column-major view, the other side and triangle */
        side = flip_side(Side);
        uplo = flip_uplo(Uplo);
        m = N;
        n = M;
    } else if (order != CblasColMajor) {
        xerbla_(name, &info);
        return;
    }
    const int nrowa = side == CblasLeft ? m : n;
    if (!side_ok(side)) info = 1;
    else if (!uplo_ok(uplo)) info = 2;
    else if (TransA != CblasNoTrans && TransA != CblasTrans && TransA != CblasConjTrans) info = 3;
    else if (Diag != CblasNonUnit && Diag != CblasUnit) info = 4;
    else if (m < 0) info = 5;
    else if (n < 0) info = 6;
    else if (lda < max1(nrowa)) info = 9;
    else if (ldb < max1(m)) info = 11;
    if (info) {
        xerbla_(name, &info);
        return;
    }
    if (m == 0 || n == 0) return;
    double al[2];
    scalar2(dt, alpha, al);
    if (gpu_sized(dt, m, n)) {
        int rc = bx_ensure_init();
        if (!rc) rc = gpu_trsm(dt, side, uplo, TransA, Diag, m, n, al, A, lda, B, ldb);
        if (!gpu_outcome(name, rc, A, B, NULL)) return;
    }
    need_cpublas(name);
    if (dt == BLASX_S)
        bx_cpu_strsm(order, Side, Uplo, TransA, Diag, M, N, *(const float *)alpha, A, lda, B, ldb);
    else if (dt == BLASX_D)
        bx_cpu_dtrsm(order, Side, Uplo, TransA, Diag, M, N, *(const double *)alpha, A, lda, B,
                     ldb);
    else if (dt == BLASX_C)
        bx_cpu_ctrsm(order, Side, Uplo, TransA, Diag, M, N, alpha, A, lda, B, ldb);
    else
        bx_cpu_ztrsm(order, Side, Uplo, TransA, Diag, M, N, alpha, A, lda, B, ldb);
}

/* ---- SYR2K ---- */
static int syr2k_trans_ok(int dt, int t)
{
    if (t == CblasNoTrans || t == CblasTrans) return 1;
    return !is_cplx(dt) && t == CblasConjTrans; /* C means T for real data */
}

static void syr2k_entry(int dt, int order, int Uplo, int Trans, int N, int K, const void *alpha,
                        const void *A, int lda, const void *B, int ldb, const void *beta, void *C,
                        int ldc)
{
    const char *name = dt == BLASX_S   ? "SSYR2K"
                       : dt == BLASX_D ? "DSYR2K"
                       : dt == BLASX_C ? "CSYR2K"
                                       : "ZSYR2K";
    int uplo = Uplo, trans = Trans, info = 0;
    if (order == CblasRowMajor) {
        uplo = flip_uplo(Uplo);
        if (syr2k_trans_ok(dt, Trans)) trans = Trans == CblasNoTrans ? CblasTrans : CblasNoTrans;
    } else if (order != CblasColMajor) {
        xerbla_(name, &info);
        return;
    }
    const int nrowa = trans == CblasNoTrans ? N : K;
    if (!uplo_ok(uplo)) info = 1;
    else if (!syr2k_trans_ok(dt, trans)) info = 2;
    else if (N < 0) info = 3;
    else if (K < 0) info = 4;
    else if (lda < max1(nrowa)) info = 7;
    else if (ldb < max1(nrowa)) info = 9;
    else if (ldc < max1(N)) info = 12;
    if (info) {
        xerbla_(name, &info);
        return;
    }
    if (N == 0) return;
    double al[2], be[2];
    scalar2(dt, alpha, al);
    scalar2(dt, beta, be);
    const int no_product = K == 0 || (al[0] == 0.0 && al[1] == 0.0);
    if (no_product && be[0] == 1.0 && be[1] == 0.0) return;
    const int in_hbm = bxh_pointer_is_pinned(C) == 2;
    if (gpu_sized(dt, N, K) && (!no_product || in_hbm)) {
        int rc = bx_ensure_init();
        if (!rc)
            rc = dt == BLASX_C
                     ? bx_gpu_crank(0, uplo, trans, N, K, al, A, lda, B, ldb, be, C, ldc)
                     : gpu_syr2k(dt, uplo, trans, N, K, al, A, lda, B, ldb, be, C, ldc);
        if (!gpu_outcome(name, rc, A, B, C)) return;
    }
    need_cpublas(name);
    if (dt == BLASX_S)
        bx_cpu_ssyr2k(order, Uplo, Trans, N, K, *(const float *)alpha, A, lda, B, ldb,
                      *(const float *)beta, C, ldc);
    else if (dt == BLASX_D)
        bx_cpu_dsyr2k(order, Uplo, Trans, N, K, *(const double *)alpha, A, lda, B, ldb,
                      *(const double *)beta, C, ldc);
    else if (dt == BLASX_C)
        bx_cpu_csyr2k(order, Uplo, Trans, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    else
        bx_cpu_zsyr2k(order, Uplo, Trans, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* ---- HER2K ---- */
static void her2k_entry(int dt, int order, int Uplo, int Trans, int N, int K, const void *alpha,
                        const void *A, int lda, const void *B, int ldb, double beta, void *C,
                        int ldc)
{
    /* netlib's names; the reference's zher2k wrapper reports "ZHERK" (blas/zher2k.c:13) */
    const char *name = dt == BLASX_C ? "CHER2K" : "ZHER2K";
    int uplo = Uplo, trans = Trans, info = 0;
    double al[2];
    scalar2(dt, alpha, al);
    if (order == CblasRowMajor) {
        /* C^T = conj(alpha) B'^H A' + alpha A'^H B' on the column-major views
         * A' = A^T, B' = B^T: the other triangle, the other transposition and
         * the conjugate alpha */
        uplo = flip_uplo(Uplo);
        if (Trans == CblasNoTrans || Trans == CblasConjTrans)
            trans = Trans == CblasNoTrans ? CblasConjTrans : CblasNoTrans;
        al[1] = -al[1];
    } else if (order != CblasColMajor) {
        xerbla_(name, &info);
        return;
    }
    const int nrowa = trans == CblasNoTrans ? N : K;
    if (!uplo_ok(uplo)) info = 1;
    else if (trans != CblasNoTrans && trans != CblasConjTrans) info = 2;
    else if (N < 0) info = 3;
    else if (K < 0) info = 4;
    else if (lda < max1(nrowa)) info = 7;
    else if (ldb < max1(nrowa)) info = 9;
    else if (ldc < max1(N)) info = 12;
    if (info) {
        xerbla_(name, &info);
        return;
    }
    if (N == 0) return;
    const int no_product = K == 0 || (al[0] == 0.0 && al[1] == 0.0);
    if (no_product && beta == 1.0) return;
    const int in_hbm = bxh_pointer_is_pinned(C) == 2;
    if (gpu_sized(dt, N, K) && (!no_product || in_hbm)) {
        int rc = bx_ensure_init();
        if (!rc) {
            const double be[2] = {beta, 0.0};
            rc = dt == BLASX_C ? bx_gpu_crank(1, uplo, trans, N, K, al, A, lda, B, ldb, be, C, ldc)
                               : gpu_her2k(uplo, trans, N, K, al, A, lda, B, ldb, beta, C, ldc);
        }
        if (!gpu_outcome(name, rc, A, B, C)) return;
    }
    need_cpublas(name);
    if (dt == BLASX_C)
        bx_cpu_cher2k(order, Uplo, Trans, N, K, alpha, A, lda, B, ldb, (float)beta, C, ldc);
    else
        bx_cpu_zher2k(order, Uplo, Trans, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* ------------------------------------------------------------ CBLAS */
BX_EXPORT void cblas_ssymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const int M, const int N, const float alpha,
                           const float *A, const int lda, const float *B, const int ldb,
                           const float beta, float *C, const int ldc)
{
    symm_entry(BLASX_S, 0, Order, Side, Uplo, M, N, &alpha, A, lda, B, ldb, &beta, C, ldc);
}

BX_EXPORT void cblas_dsymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const int M, const int N,
                           const double alpha, const double *A, const int lda, const double *B,
                           const int ldb, const double beta, double *C, const int ldc)
{
    symm_entry(BLASX_D, 0, Order, Side, Uplo, M, N, &alpha, A, lda, B, ldb, &beta, C, ldc);
}

BX_EXPORT void cblas_zsymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                           const void *A, const int lda, const void *B, const int ldb,
                           const void *beta, void *C, const int ldc)
{
    symm_entry(BLASX_Z, 0, Order, Side, Uplo, M, N, alpha, A, lda, B, ldb, beta, C, ldc);
}

BX_EXPORT void cblas_zhemm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                           const void *A, const int lda, const void *B, const int ldb,
                           const void *beta, void *C, const int ldc)
{
    symm_entry(BLASX_Z, 1, Order, Side, Uplo, M, N, alpha, A, lda, B, ldb, beta, C, ldc);
}

BX_EXPORT void cblas_strmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_DIAG Diag, const int M, const int N, const float alpha,
                           const float *A, const int lda, float *B, const int ldb)
{
    trmm_entry(BLASX_S, Order, Side, Uplo, TransA, Diag, M, N, &alpha, A, lda, B, ldb);
}

BX_EXPORT void cblas_dtrmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_DIAG Diag, const int M, const int N,
                           const double alpha, const double *A, const int lda, double *B,
                           const int ldb)
{
    trmm_entry(BLASX_D, Order, Side, Uplo, TransA, Diag, M, N, &alpha, A, lda, B, ldb);
}

BX_EXPORT void cblas_ztrmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                           const void *A, const int lda, void *B, const int ldb)
{
    trmm_entry(BLASX_Z, Order, Side, Uplo, TransA, Diag, M, N, alpha, A, lda, B, ldb);
}

BX_EXPORT void cblas_strsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_DIAG Diag, const int M, const int N, const float alpha,
                           const float *A, const int lda, float *B, const int ldb)
{
    trsm_entry(BLASX_S, Order, Side, Uplo, TransA, Diag, M, N, &alpha, A, lda, B, ldb);
}

BX_EXPORT void cblas_dtrsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_DIAG Diag, const int M, const int N,
                           const double alpha, const double *A, const int lda, double *B,
                           const int ldb)
{
    trsm_entry(BLASX_D, Order, Side, Uplo, TransA, Diag, M, N, &alpha, A, lda, B, ldb);
}

BX_EXPORT void cblas_ztrsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                           const void *A, const int lda, void *B, const int ldb)
{
    trsm_entry(BLASX_Z, Order, Side, Uplo, TransA, Diag, M, N, alpha, A, lda, B, ldb);
}

BX_EXPORT void cblas_zher2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                            const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                            const void *alpha, const void *A, const int lda, const void *B,
                            const int ldb, const double beta, void *C, const int ldc)
{
    her2k_entry(BLASX_Z, Order, Uplo, Trans, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

BX_EXPORT void cblas_cher2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                            const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                            const void *alpha, const void *A, const int lda, const void *B,
                            const int ldb, const float beta, void *C, const int ldc)
{
    her2k_entry(BLASX_C, Order, Uplo, Trans, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

BX_EXPORT void cblas_csyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                            const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                            const void *alpha, const void *A, const int lda, const void *B,
                            const int ldb, const void *beta, void *C, const int ldc)
{
    syr2k_entry(BLASX_C, Order, Uplo, Trans, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

BX_EXPORT void cblas_csymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                           const void *A, const int lda, const void *B, const int ldb,
                           const void *beta, void *C, const int ldc)
{
    symm_entry(BLASX_C, 0, Order, Side, Uplo, M, N, alpha, A, lda, B, ldb, beta, C, ldc);
}

BX_EXPORT void cblas_chemm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                           const void *A, const int lda, const void *B, const int ldb,
                           const void *beta, void *C, const int ldc)
{
    symm_entry(BLASX_C, 1, Order, Side, Uplo, M, N, alpha, A, lda, B, ldb, beta, C, ldc);
}

BX_EXPORT void cblas_ctrmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                           const void *A, const int lda, void *B, const int ldb)
{
    trmm_entry(BLASX_C, Order, Side, Uplo, TransA, Diag, M, N, alpha, A, lda, B, ldb);
}

BX_EXPORT void cblas_ctrsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                           const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                           const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                           const void *A, const int lda, void *B, const int ldb)
{
    trsm_entry(BLASX_C, Order, Side, Uplo, TransA, Diag, M, N, alpha, A, lda, B, ldb);
}

BX_EXPORT void cblas_ssyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                            const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                            const float alpha, const float *A, const int lda, const float *B,
                            const int ldb, const float beta, float *C, const int ldc)
{
    syr2k_entry(BLASX_S, Order, Uplo, Trans, N, K, &alpha, A, lda, B, ldb, &beta, C, ldc);
}

BX_EXPORT void cblas_dsyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                            const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                            const double alpha, const double *A, const int lda, const double *B,
                            const int ldb, const double beta, double *C, const int ldc)
{
    syr2k_entry(BLASX_D, Order, Uplo, Trans, N, K, &alpha, A, lda, B, ldb, &beta, C, ldc);
}

BX_EXPORT void cblas_zsyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                            const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                            const void *alpha, const void *A, const int lda, const void *B,
                            const int ldb, const void *beta, void *C, const int ldc)
{
    syr2k_entry(BLASX_Z, Order, Uplo, Trans, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* ------------------------------------------------------------ f77 */
static int f77_side(const char *c, int *out)
{
    switch (*c) {
    case 'L': case 'l': *out = CblasLeft; return 0;
    case 'R': case 'r': *out = CblasRight; return 0;
    default: return -1;
    }
}

static int f77_uplo(const char *c, int *out)
{
    switch (*c) {
    case 'U': case 'u': *out = CblasUpper; return 0;
    case 'L': case 'l': *out = CblasLower; return 0;
    default: return -1;
    }
}

static int f77_trans(const char *c, int *out)
{
    switch (*c) {
    case 'N': case 'n': *out = CblasNoTrans; return 0;
    case 'T': case 't': *out = CblasTrans; return 0;
    case 'C': case 'c': *out = CblasConjTrans; return 0;
    default: return -1;
    }
}

static int f77_diag(const char *c, int *out)
{
    switch (*c) {
    case 'N': case 'n': *out = CblasNonUnit; return 0;
    case 'U': case 'u': *out = CblasUnit; return 0;
    default: return -1;
    }
}

/* a bad character argument: xerbla_ with its position.  SYMM, TRMM and TRSM
 * check their characters last to first, so the lowest position wins
 * (blas/dsymm.c:83-84, blas/dtrmm.c:76-80, blas/dtrsm.c:86-90) */
static int f77_bad(const char *name, int info)
{
    xerbla_(name, &info);
    return 1;
}

/* SYR2K and HER2K check UPLO then TRANS, and the later check wins: two bad
 * characters report 2 (blas/dsyr2k.c:96-98, blas/zher2k.c:97-99) */
static int f77_uplo_trans(const char *name, const char *uplo, const char *trans, int *u, int *t)
{
    int info = 0;
    if (f77_uplo(uplo, u)) info = 1;
    if (f77_trans(trans, t)) info = 2;
    if (info) xerbla_(name, &info);
    return info;
}

#define F77_SYMM(NAME, DT, HERM, TNAME, T)                                                     \
    BX_EXPORT void NAME(const char *side, const char *uplo, const int *m, const int *n,         \
                        const T *alpha, const T *A, const int *lda, const T *B, const int *ldb, \
                        const T *beta, T *C, const int *ldc)                                    \
    {                                                                                          \
        int s, u;                                                                              \
        if (f77_side(side, &s) && f77_bad(TNAME, 1)) return;                                   \
        if (f77_uplo(uplo, &u) && f77_bad(TNAME, 2)) return;                                   \
        symm_entry(DT, HERM, CblasColMajor, s, u, *m, *n, alpha, A, *lda, B, *ldb, beta, C,    \
                   *ldc);                                                                      \
    }
F77_SYMM(ssymm_, BLASX_S, 0, "SSYMM ", float)
F77_SYMM(dsymm_, BLASX_D, 0, "DSYMM ", double)
F77_SYMM(zsymm_, BLASX_Z, 0, "ZSYMM ", void)
F77_SYMM(zhemm_, BLASX_Z, 1, "ZHEMM", void)

#define F77_TRMM(NAME, DT, TNAME, T)                                                           \
    BX_EXPORT void NAME(const char *side, const char *uplo, const char *transa,                \
                        const char *diag, const int *m, const int *n, const T *alpha,           \
                        const T *A, const int *lda, T *B, const int *ldb)                       \
    {                                                                                          \
        int s, u, t, d;                                                                        \
        if (f77_side(side, &s) && f77_bad(TNAME, 1)) return;                                   \
        if (f77_uplo(uplo, &u) && f77_bad(TNAME, 2)) return;                                   \
        if (f77_trans(transa, &t) && f77_bad(TNAME, 3)) return;                                \
        if (f77_diag(diag, &d) && f77_bad(TNAME, 4)) return;                                   \
        trmm_entry(DT, CblasColMajor, s, u, t, d, *m, *n, alpha, A, *lda, B, *ldb);            \
    }
F77_TRMM(strmm_, BLASX_S, "STRMM ", float)
F77_TRMM(dtrmm_, BLASX_D, "DTRMM ", double)
F77_TRMM(ztrmm_, BLASX_Z, "ZTRMM ", void)

#define F77_TRSM(NAME, DT, TNAME, T)                                                           \
    BX_EXPORT void NAME(const char *side, const char *uplo, const char *transa,                \
                        const char *diag, const int *m, const int *n, const T *alpha,           \
                        const T *A, const int *lda, T *B, const int *ldb)                       \
    {                                                                                          \
        int s, u, t, d;                                                                        \
        if (f77_side(side, &s) && f77_bad(TNAME, 1)) return;                                   \
        if (f77_uplo(uplo, &u) && f77_bad(TNAME, 2)) return;                                   \
        if (f77_trans(transa, &t) && f77_bad(TNAME, 3)) return;                                \
        if (f77_diag(diag, &d) && f77_bad(TNAME, 4)) return;                                   \
        trsm_entry(DT, CblasColMajor, s, u, t, d, *m, *n, alpha, A, *lda, B, *ldb);            \
    }
F77_TRSM(strsm_, BLASX_S, "STRSM ", float)
F77_TRSM(dtrsm_, BLASX_D, "DTRSM ", double)
F77_TRSM(ztrsm_, BLASX_Z, "ZTRSM ", void)

#define F77_SYR2K(NAME, DT, TNAME, T)                                                          \
    BX_EXPORT void NAME(const char *uplo, const char *trans, const int *n, const int *k,        \
                        const T *alpha, const T *A, const int *lda, const T *B, const int *ldb, \
                        const T *beta, T *C, const int *ldc)                                    \
    {                                                                                          \
        int u, t;                                                                              \
        if (f77_uplo_trans(TNAME, uplo, trans, &u, &t)) return;                                \
        syr2k_entry(DT, CblasColMajor, u, t, *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);  \
    }
BX_EXPORT void zher2k_(const char *uplo, const char *trans, const int *n, const int *k,
                       const void *alpha, const void *A, const int *lda, const void *B,
                       const int *ldb, const double *beta, void *C, const int *ldc)
{
    int u, t;
    if (f77_uplo_trans("ZHER2K", uplo, trans, &u, &t)) return;
    her2k_entry(BLASX_Z, CblasColMajor, u, t, *n, *k, alpha, A, *lda, B, *ldb, *beta, C, *ldc);
}

BX_EXPORT void cher2k_(const char *uplo, const char *trans, const int *n, const int *k,
                       const void *alpha, const void *A, const int *lda, const void *B,
                       const int *ldb, const float *beta, void *C, const int *ldc)
{
    int u, t;
    if (f77_uplo_trans("CHER2K", uplo, trans, &u, &t)) return;
    her2k_entry(BLASX_C, CblasColMajor, u, t, *n, *k, alpha, A, *lda, B, *ldb, *beta, C, *ldc);
}

F77_SYMM(csymm_, BLASX_C, 0, "CSYMM ", void)
F77_SYMM(chemm_, BLASX_C, 1, "CHEMM", void)
F77_TRMM(ctrmm_, BLASX_C, "CTRMM ", void)
F77_TRSM(ctrsm_, BLASX_C, "CTRSM ", void)
F77_SYR2K(csyr2k_, BLASX_C, "CSYR2K", void)
F77_SYR2K(ssyr2k_, BLASX_S, "SSYR2K", float)
F77_SYR2K(dsyr2k_, BLASX_D, "DSYR2K", double)
F77_SYR2K(zsyr2k_, BLASX_Z, "ZSYR2K", void)
