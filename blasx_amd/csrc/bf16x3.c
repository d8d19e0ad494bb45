/*
 * bf16x3.c -- the BF16x3 split DGEMM, host side: the BF16 variant of the
 * split-precision GEMMs beside the INT8 digit scheme (split.c) and the CRT
 * scheme (crt.c).  No counterpart in the reference.
 *
 * Rows of op(A) and columns of op(B) are scaled by powers of two into
 * (-1, 1) (the digit scheme's exponents), each scaled entry is cut into three
 * bf16 pieces (s1 + s2 + s3 carries about 24 significant bits; the split is
 * stated in hip/ozaki_kernels.hip), and the six piece products with
 * p + q <= 4 run on the BF16 matrix cores in three launches, one per group
 * g = p + q (4, 3, 2: the smallest first), each summing its products over
 * the whole K range in FP32 and adding the result into an FP64 workspace W.
 * C = alpha 2^(e_i + f_j) W + beta C in FP64.
 *
 * Accuracy is set by the FP32 sums, not by the pieces: about sqrt(K) 2^-24
 * relative to sum_k |a_ik b_kj| -- FP32-like, far from FP64, for a third of
 * the matrix-core work of the 6-slice INT8 digit scheme.
 *
 * Workspace (HBM, grown on demand and kept): 3 bf16 planes of op(A) (Mp x Kp)
 * and of op(B)^T (Np x Kp), W (M x N), exponents; one call at a time.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

int bxh_oz_block_m(void);
int bxh_oz_block_k(void);
int bxh_oz_exponents(const double *X, int64_t ld, int n, int len, int contiguous, int norm,
                     int *e, double *part, void *stream);
int bxh_oz_exp_chunks(void);
int bxh_oz_finish(int m, int n, const double *W, int64_t ldw, const int *ea, const int *eb,
                  double alpha, double beta, double *C, int64_t ldc, void *stream);
int bxh_bf16_split(const double *X, int64_t ld, int n, int len, int contiguous, const int *e,
                   int8_t *D, int np, int kp2, void *stream);
int bxh_bf16_group(const void *segs_dev, int nseg, int kp2, int m, int n, int mp, int np,
                   double *W, int64_t ldw, int first, void *stream);

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static struct {
    void *dev, *segs_host, *segs_dev;
    size_t cap;
    void *done;
} g_w;

static size_t rup(size_t v, size_t a) { return (v + a - 1) / a * a; }

static int bx_dgemm_bf16x3(int opa, int opb, int M, int N, int K, double alpha, const double *A,
                           int64_t lda, const double *B, int64_t ldb, double beta, double *C,
                           int64_t ldc, void *stream)
{
    if (M == 0 || N == 0) return 0;
    if (K == 0 || alpha == 0.0) {
        const double b2[2] = {beta, 0.0};
        return bxh_scale(BLASX_D, M, N, b2, C, ldc, stream);
    }
    const int pad = (M >= 4096 && N >= 4096) ? 2 * bxh_oz_block_m() : bxh_oz_block_m();
    const int mp = (int)rup(M, pad), np = (int)rup(N, pad);
    /* K in bytes of bf16 (2 per element), an even number of 64-byte K tiles */
    const int kp2 = (int)rup(2 * (size_t)K, 2 * bxh_oz_block_k());
    const size_t pa = (size_t)3 * mp * kp2, pb = (size_t)3 * np * kp2;
    const size_t off_b = rup(pa, 256), off_w = off_b + rup(pb, 256);
    const size_t off_ea = off_w + rup((size_t)M * N * sizeof(double), 256);
    const size_t off_eb = off_ea + rup((size_t)M * 4, 256);
    const size_t off_part = off_eb + rup((size_t)N * 4, 256);
    const size_t total = off_part + (size_t)bxh_oz_exp_chunks() * (M > N ? M : N) * sizeof(double);
    const size_t segs_bytes = 6 * 2 * sizeof(void *);

    pthread_mutex_lock(&g_lock);
    int rc = 0;
    if (!g_w.done) rc = bxh_event_create(&g_w.done, 0);
    if (!rc) rc = bxh_event_sync(g_w.done); /* the last call is done with the buffers */
    if (!rc && total > g_w.cap) {
        bxh_free(g_w.dev);
        g_w.dev = NULL;
        g_w.cap = 0;
        rc = bxh_malloc(&g_w.dev, total);
        if (!rc) g_w.cap = total;
    }
    if (!rc && !g_w.segs_host) rc = bxh_host_malloc(&g_w.segs_host, segs_bytes);
    if (!rc && !g_w.segs_dev) rc = bxh_malloc(&g_w.segs_dev, segs_bytes);
    if (rc) {
        pthread_mutex_unlock(&g_lock);
        return rc;
    }
    char *base = (char *)g_w.dev;
    int8_t *PA = (int8_t *)base, *PB = (int8_t *)(base + off_b);
    double *W = (double *)(base + off_w);
    int *ea = (int *)(base + off_ea), *eb = (int *)(base + off_eb);
    double *part = (double *)(base + off_part);

    const int a_contig = opa != BX_OP_N, b_contig = opb == BX_OP_N;
    rc = bxh_oz_exponents(A, lda, M, K, a_contig, 0, ea, part, stream);
    if (!rc) rc = bxh_oz_exponents(B, ldb, N, K, b_contig, 0, eb, part, stream);
    if (!rc) rc = bxh_bf16_split(A, lda, M, K, a_contig, ea, PA, mp, kp2, stream);
    if (!rc) rc = bxh_bf16_split(B, ldb, N, K, b_contig, eb, PB, np, kp2, stream);

    /* segment table: groups g = 4, 3, 2 (pairs (p, q), p + q = g) */
    const int8_t **tab = (const int8_t **)g_w.segs_host;
    const int gs[3] = {4, 3, 2};
    int start[3], count[3], nent = 0;
    for (int gi = 0; gi < 3; ++gi) {
        start[gi] = nent;
        for (int p = 1; p <= 3; ++p) {
            const int q = gs[gi] - p;
            if (q < 1 || q > 3) continue;
            tab[2 * nent] = PA + (size_t)(p - 1) * mp * kp2;
            tab[2 * nent + 1] = PB + (size_t)(q - 1) * np * kp2;
            ++nent;
        }
        count[gi] = nent - start[gi];
    }
    if (!rc) rc = bxh_memcpy_h2d_async(g_w.segs_dev, g_w.segs_host, segs_bytes, stream);
    for (int gi = 0; gi < 3 && !rc; ++gi)
        rc = bxh_bf16_group((const char *)g_w.segs_dev + (size_t)start[gi] * 2 * sizeof(void *),
                            count[gi], kp2, M, N, mp, np, W, M, gi == 0, stream);
    if (!rc) rc = bxh_oz_finish(M, N, W, M, ea, eb, alpha, beta, C, ldc, stream);
    if (!rc) rc = bxh_event_record(g_w.done, stream);
    pthread_mutex_unlock(&g_lock);
    if (!rc) {
        BX_STAT_ADD(kernel_launches, 8);
        bx_stat_flops(2.0 * M * (double)N * K);
    }
    return rc;
}

BX_EXPORT int blasx_dgemm_bf16x3_device(int transa, int transb, int M, int N, int K,
                                        double alpha, const double *dA, int64_t lda,
                                        const double *dB, int64_t ldb, double beta, double *dC,
                                        int64_t ldc, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    const int opa = transa == 111 ? BX_OP_N : (transa == 112 || transa == 113) ? BX_OP_T : -1;
    const int opb = transb == 111 ? BX_OP_N : (transb == 112 || transb == 113) ? BX_OP_T : -1;
    if (opa < 0 || opb < 0 || M < 0 || N < 0 || K < 0) return BLASX_EINVAL;
    const int64_t nrowa = opa == BX_OP_N ? M : K, nrowb = opb == BX_OP_N ? K : N;
    if (lda < (nrowa > 1 ? nrowa : 1) || ldb < (nrowb > 1 ? nrowb : 1) || ldc < (M > 1 ? M : 1))
        return BLASX_EINVAL;
    return bx_dgemm_bf16x3(opa, opb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc,
                           stream ? stream : g_bx.s_comp);
}
