/*
 * split.c -- split-precision DGEMM (Ozaki scheme on INT8 MFMA), host side.
 * No counterpart in the reference; the algorithm and its exactness argument
 * are stated in hip/ozaki_kernels.hip.
 *
 * Workspace (HBM, grown on demand and kept): s int8 slices of op(A)
 * (Mp x Kp) and of op(B)^T (Np x Kp), the FP64 accumulator W (M x N), the
 * row/column exponents and the segment table.  One call at a time per
 * process (a mutex guards the workspace).
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
int bxh_oz_split(const double *X, int64_t ld, int n, int len, int contiguous, const int *e,
                 int s, int8_t *D, int np, int kp, void *stream);
int bxh_oz_group(const void *segs_dev, int nseg, int kp, int m, int n, int mp, int np, double *W,
                 int64_t ldw, double scale, int first, void *stream);
int bxh_oz_finish(int m, int n, const double *W, int64_t ldw, const int *ea, const int *eb,
                  double alpha, double beta, double *C, int64_t ldc, void *stream);

static pthread_mutex_t g_oz_lock = PTHREAD_MUTEX_INITIALIZER;
static struct {
    void *dev;
    size_t dev_cap;
    void *segs_host, *segs_dev;
    size_t segs_cap;
    void *done; /* event: the last call using the workspace finished */
} g_oz;

static size_t rup(size_t v, size_t a) { return (v + a - 1) / a * a; }

static int ensure(void **p, size_t *cap, size_t bytes, int host)
{
    if (bytes <= *cap) return 0;
    if (*p) {
        if (g_oz.done) bxh_event_sync(g_oz.done); /* the old buffer may still be read */
        if (host) bxh_host_free(*p); else bxh_free(*p);
        *p = NULL;
        *cap = 0;
    }
    int rc = host ? bxh_host_malloc(p, bytes) : bxh_malloc(p, bytes);
    if (rc == 0) *cap = bytes;
    return rc;
}

/* exact int32 accumulation bound: nseg * Kp * 127^2 < 2^31 */
static int max_segments(int kp) { return (int)(2147483647LL / ((int64_t)kp * 127 * 127)); }

int bx_dgemm_split(int opa, int opb, int M, int N, int K, double alpha, const double *A,
                   int64_t lda, const double *B, int64_t ldb, double beta, double *C,
                   int64_t ldc, int s, void *stream)
{
    if (M == 0 || N == 0) return 0;
    if (K == 0 || alpha == 0.0) {
        double b2[2] = {beta, 0.0};
        return bxh_scale(BLASX_D, M, N, b2, C, ldc, stream);
    }
    /* large problems pad to the 256x256 tile of the big INT8 kernel */
    const int pad = (M >= 4096 && N >= 4096) ? 2 * bxh_oz_block_m() : bxh_oz_block_m();
    const int mp = (int)rup(M, pad), np = (int)rup(N, pad);
    const int kp = (int)rup(K, 2 * bxh_oz_block_k()); /* an even number of K tiles */
    const int segmax = max_segments(kp);
    if (segmax < 1) return BLASX_ERANGE; /* K above ~133k: int32 products could overflow */

    const size_t a8 = (size_t)s * mp * kp, b8 = (size_t)s * np * kp;
    const size_t wbytes = (size_t)M * N * sizeof(double);
    const size_t off_b8 = rup(a8, 256), off_w = off_b8 + rup(b8, 256);
    const size_t off_ea = off_w + rup(wbytes, 256), off_eb = off_ea + rup((size_t)M * 4, 256);
    const size_t off_part = off_eb + rup((size_t)N * 4, 256);
    const size_t total = off_part + (size_t)bxh_oz_exp_chunks() * (M > N ? M : N) * sizeof(double);

    pthread_mutex_lock(&g_oz_lock);
    int rc = 0;
    if (!g_oz.done) rc = bxh_event_create(&g_oz.done, 0);
    if (!rc) rc = ensure(&g_oz.dev, &g_oz.dev_cap, total, 0);
    /* segment table: every (p,q) pair once */
    const size_t segs_bytes = (size_t)s * s * 2 * sizeof(void *);
    if (!rc && segs_bytes > g_oz.segs_cap) {
        size_t cap = g_oz.segs_cap;
        rc = ensure(&g_oz.segs_host, &cap, segs_bytes, 1);
        size_t cap2 = g_oz.segs_cap;
        if (!rc) rc = ensure(&g_oz.segs_dev, &cap2, segs_bytes, 0);
        if (!rc) g_oz.segs_cap = segs_bytes;
    }
    if (rc) {
        pthread_mutex_unlock(&g_oz_lock);
        return rc;
    }
    /* the segment table is rewritten below: the previous call must be done with it */
    bxh_event_sync(g_oz.done);

    char *base = (char *)g_oz.dev;
    int8_t *DA = (int8_t *)base, *DB = (int8_t *)(base + off_b8);
    double *W = (double *)(base + off_w);
    int *ea = (int *)(base + off_ea), *eb = (int *)(base + off_eb);
    double *part = (double *)(base + off_part);

    /* exponents and slices: rows of op(A), columns of op(B) */
    const int a_contig = opa != BX_OP_N; /* A^T: row i of op(A) is column i of A */
    const int b_contig = opb == BX_OP_N; /* B: column j of op(B) is column j of B */
    rc = bxh_oz_exponents(A, lda, M, K, a_contig, 0, ea, part, stream);
    if (!rc) rc = bxh_oz_exponents(B, ldb, N, K, b_contig, 0, eb, part, stream);
    if (!rc) rc = bxh_oz_split(A, lda, M, K, a_contig, ea, s, DA, mp, kp, stream);
    if (!rc) rc = bxh_oz_split(B, ldb, N, K, b_contig, eb, s, DB, np, kp, stream);

    /* segment table, grouped by g = p + q, smallest contributions first */
    const int8_t **tab = (const int8_t **)g_oz.segs_host;
    int nent = 0;
    int gstart[20], gcount[20], ng = 0;
    for (int g = s + 1; g >= 2 && !rc; --g) {
        gstart[ng] = nent;
        for (int p = 1; p <= s; ++p) {
            const int q = g - p;
            if (q < 1 || q > s) continue;
            tab[2 * nent] = DA + (size_t)(p - 1) * mp * kp;
            tab[2 * nent + 1] = DB + (size_t)(q - 1) * np * kp;
            ++nent;
        }
        gcount[ng] = nent - gstart[ng];
        ++ng;
    }
    if (!rc) rc = bxh_memcpy_h2d_async(g_oz.segs_dev, g_oz.segs_host,
                                       (size_t)nent * 2 * sizeof(void *), stream);
    int first = 1, launches = 0;
    for (int gi = 0; gi < ng && !rc; ++gi) {
        const int g = s + 1 - gi;
        double scale = 1.0;
        for (int k = 0; k < g; ++k) scale *= 0x1.0p-7;
        for (int off = 0; off < gcount[gi] && !rc; off += segmax) {
            const int cnt = gcount[gi] - off < segmax ? gcount[gi] - off : segmax;
            const char *segp = (const char *)g_oz.segs_dev + (size_t)(gstart[gi] + off) * 2 * sizeof(void *);
            rc = bxh_oz_group(segp, cnt, kp, M, N, mp, np, W, M, scale, first, stream);
            first = 0;
            ++launches;
        }
    }
    if (!rc) rc = bxh_oz_finish(M, N, W, M, ea, eb, alpha, beta, C, ldc, stream);
    if (!rc) rc = bxh_event_record(g_oz.done, stream);
    pthread_mutex_unlock(&g_oz_lock);
    if (!rc) {
        BX_STAT_ADD(kernel_launches, launches + 5);
        bx_stat_flops(2.0 * M * (double)N * K);
    }
    return rc;
}

int bxh_oz_deinterleave(const double *Z, int64_t ld, int rows, int cols, int conj, double *re,
                        double *im, void *stream);
int bxh_oz_zcombine(int m, int n, const double *pr, const double *pi, const double *alpha,
                    const double *beta, double *C, int64_t ldc, void *stream);

/* complex split GEMM: four real split products on de-interleaved planes
 *   Pr = Ar Br - Ai Bi,  Pi = Ar Bi + Ai Br,  C = alpha (Pr + i Pi) + beta C
 * (a conjugated operand has its imaginary plane negated) */
static pthread_mutex_t g_zs_lock = PTHREAD_MUTEX_INITIALIZER;
static struct {
    void *dev;
    size_t cap;
    void *done;
} g_zs;

int bx_zgemm_split(int opa, int opb, int M, int N, int K, const double *alpha, const double *A,
                   int64_t lda, const double *B, int64_t ldb, const double *beta, double *C,
                   int64_t ldc, int s, void *stream)
{
    if (M == 0 || N == 0) return 0;
    if (K == 0 || (alpha[0] == 0.0 && alpha[1] == 0.0))
        return bxh_scale(BLASX_Z, M, N, beta, C, ldc, stream);
    const int ra = opa == BX_OP_N ? M : K, ca = opa == BX_OP_N ? K : M;
    const int rb = opb == BX_OP_N ? K : N, cb = opb == BX_OP_N ? N : K;
    const size_t na = (size_t)ra * ca, nb = (size_t)rb * cb, nc = (size_t)M * N;
    const size_t bytes = 8 * (2 * na + 2 * nb + 2 * nc);
    pthread_mutex_lock(&g_zs_lock);
    int rc = 0;
    if (!g_zs.done) rc = bxh_event_create(&g_zs.done, 0);
    if (!rc && bytes > g_zs.cap) {
        bxh_event_sync(g_zs.done);
        bxh_free(g_zs.dev);
        g_zs.dev = NULL;
        g_zs.cap = 0;
        rc = bxh_malloc(&g_zs.dev, bytes);
        if (!rc) g_zs.cap = bytes;
    }
    if (!rc) rc = bxh_event_sync(g_zs.done);
    double *Ar = (double *)g_zs.dev, *Ai = Ar + na, *Br = Ai + na, *Bi = Br + nb;
    double *Pr = Bi + nb, *Pi = Pr + nc;
    const int ropa = opa == BX_OP_N ? BX_OP_N : BX_OP_T, ropb = opb == BX_OP_N ? BX_OP_N : BX_OP_T;
    if (!rc) rc = bxh_oz_deinterleave(A, lda, ra, ca, opa == BX_OP_C, Ar, Ai, stream);
    if (!rc) rc = bxh_oz_deinterleave(B, ldb, rb, cb, opb == BX_OP_C, Br, Bi, stream);
    if (!rc) rc = bx_dgemm_split(ropa, ropb, M, N, K, 1.0, Ar, ra, Br, rb, 0.0, Pr, M, s, stream);
    if (!rc) rc = bx_dgemm_split(ropa, ropb, M, N, K, -1.0, Ai, ra, Bi, rb, 1.0, Pr, M, s, stream);
    if (!rc) rc = bx_dgemm_split(ropa, ropb, M, N, K, 1.0, Ar, ra, Bi, rb, 0.0, Pi, M, s, stream);
    if (!rc) rc = bx_dgemm_split(ropa, ropb, M, N, K, 1.0, Ai, ra, Br, rb, 1.0, Pi, M, s, stream);
    if (!rc) rc = bxh_oz_zcombine(M, N, Pr, Pi, alpha, beta, C, ldc, stream);
    if (!rc) rc = bxh_event_record(g_zs.done, stream);
    pthread_mutex_unlock(&g_zs_lock);
    return rc;
}

BX_EXPORT int blasx_zgemm_split_device(int transa, int transb, int M, int N, int K,
                                       const double *alpha, const double *dA, int64_t lda,
                                       const double *dB, int64_t ldb, const double *beta,
                                       double *dC, int64_t ldc, int slices, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    const int opa = transa == 111 ? BX_OP_N : transa == 112 ? BX_OP_T : transa == 113 ? BX_OP_C : -1;
    const int opb = transb == 111 ? BX_OP_N : transb == 112 ? BX_OP_T : transb == 113 ? BX_OP_C : -1;
    if (opa < 0 || opb < 0 || M < 0 || N < 0 || K < 0) return BLASX_EINVAL;
    int64_t nrowa = opa == BX_OP_N ? M : K, nrowb = opb == BX_OP_N ? K : N;
    if (lda < (nrowa > 1 ? nrowa : 1) || ldb < (nrowb > 1 ? nrowb : 1) || ldc < (M > 1 ? M : 1))
        return BLASX_EINVAL;
    if (slices <= 0) slices = 7;
    if (slices < 2 || slices > 9) return BLASX_EINVAL;
    return bx_zgemm_split(opa, opb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, slices,
                          stream ? stream : g_bx.s_comp);
}

BX_EXPORT int blasx_dgemm_split_device(int transa, int transb, int M, int N, int K, double alpha,
                                       const double *dA, int64_t lda, const double *dB,
                                       int64_t ldb, double beta, double *dC, int64_t ldc,
                                       int slices, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    const int opa = transa == 111 ? BX_OP_N : (transa == 112 || transa == 113) ? BX_OP_T : -1;
    const int opb = transb == 111 ? BX_OP_N : (transb == 112 || transb == 113) ? BX_OP_T : -1;
    if (opa < 0 || opb < 0 || M < 0 || N < 0 || K < 0) return BLASX_EINVAL;
    int64_t nrowa = opa == BX_OP_N ? M : K, nrowb = opb == BX_OP_N ? K : N;
    if (lda < (nrowa > 1 ? nrowa : 1) || ldb < (nrowb > 1 ? nrowb : 1) || ldc < (M > 1 ? M : 1))
        return BLASX_EINVAL;
    if (slices <= 0) slices = 7;
    if (slices < 2 || slices > 9) return BLASX_EINVAL;
    return bx_dgemm_split(opa, opb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, slices,
                          stream ? stream : g_bx.s_comp);
}
