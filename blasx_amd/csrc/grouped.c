/*
 * grouped.c -- grouped GEMM: many uneven GEMMs (mixed shapes, transposes
 * and element types) in one launch per element type.  No counterpart in the
 * reference, which issues one cublas call per tile product
 * (blas/blasx_dgemm.c:87-94).
 *
 * Descriptors are turned into a device-resident table of bxh_problem_t
 * (one block-tile range each) and bxh_seg_t (one K segment each) consumed
 * by the grouped kernel; blocks walk the table with a grid stride, so a
 * 64^3 request and an 8192^3 request share one launch without a tail.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static int op_of(int dt, int t)
{
    if (t == 111) return BX_OP_N;
    if (t == 112) return BX_OP_T;
    if (t == 113) return dt == BLASX_Z ? BX_OP_C : BX_OP_T;
    return -1;
}

int bx_desc_check(const blasx_gemm_desc_t *d)
{
    if (d->dtype != BLASX_S && d->dtype != BLASX_D && d->dtype != BLASX_Z) return BLASX_EINVAL;
    const int opa = op_of(d->dtype, d->transa), opb = op_of(d->dtype, d->transb);
    if (opa < 0 || opb < 0 || d->M < 0 || d->N < 0 || d->K < 0) return BLASX_EINVAL;
    const int64_t nrowa = opa == BX_OP_N ? d->M : d->K, nrowb = opb == BX_OP_N ? d->K : d->N;
    if (d->lda < (nrowa > 1 ? nrowa : 1) || d->ldb < (nrowb > 1 ? nrowb : 1) ||
        d->ldc < (d->M > 1 ? d->M : 1))
        return BLASX_EINVAL;
    return 0;
}

static int desc_runs(const blasx_gemm_desc_t *x, int dtype)
{
    if (x->dtype != dtype || x->M == 0 || x->N == 0) return 0;
    const int az = x->alpha[0] == 0.0 && (dtype != BLASX_Z || x->alpha[1] == 0.0);
    return x->K != 0 && !az;
}

/* Fill problem/segment tables for the descriptors of one dtype (device
 * pointers in the descriptors).  Descriptors with alpha == 0 or K == 0 are
 * skipped (the caller scales C separately).  Each problem's block tile
 * shape comes from bx_grouped_shape, given the launch's load: its tile
 * count at the large tile.  Returns the tile total. */
int bx_grouped_fill(const blasx_gemm_desc_t *d, int n, int dtype, bxh_problem_t *probs,
                    bxh_seg_t *segs, int *nprob)
{
    const int bm0 = bxh_tile_m(dtype), bn0 = bxh_tile_n(dtype);
    int64_t load = 0;
    for (int i = 0; i < n; ++i)
        if (desc_runs(&d[i], dtype))
            load += (int64_t)((d[i].M + bm0 - 1) / bm0) * ((d[i].N + bn0 - 1) / bn0);
    int np = 0, total = 0;
    /* the previous problem's choice, reused while M, N and K repeat: the
     * table lookup takes a lock and interpolates, which on launches of many
     * small equal problems cost more host time than the kernel took */
    int lm = -1, ln = -1, lk = -1, shape = 0;
    for (int i = 0; i < n; ++i) {
        const blasx_gemm_desc_t *x = &d[i];
        if (!desc_runs(x, dtype)) continue;
        bxh_problem_t *P = &probs[np];
        int bm, bn;
        if (x->M != lm || x->N != ln || x->K != lk) {
            shape = bx_grouped_shape(dtype, x->M, x->N, x->K, load);
            lm = x->M;
            ln = x->N;
            lk = x->K;
        }
        bxh_tile_dims(dtype, shape, &bm, &bn);
        memset(P, 0, sizeof *P);
        P->C = x->C;
        P->ldc = x->ldc;
        P->m = x->M;
        P->n = x->N;
        P->nseg = 1;
        P->seg_off = np;
        P->opa = op_of(dtype, x->transa);
        P->opb = op_of(dtype, x->transb);
        P->shape = shape;
        P->tiles_m = (x->M + bm - 1) / bm;
        P->tiles_n = (x->N + bn - 1) / bn;
        P->tile_start = total;
        total += P->tiles_m * P->tiles_n;
        P->alpha_re = x->alpha[0];
        P->alpha_im = dtype == BLASX_Z ? x->alpha[1] : 0.0;
        P->beta_re = x->beta[0];
        P->beta_im = dtype == BLASX_Z ? x->beta[1] : 0.0;
        P->beta_zero = P->beta_re == 0.0 && P->beta_im == 0.0;
        segs[np].A = x->A;
        segs[np].B = x->B;
        segs[np].lda = x->lda;
        segs[np].ldb = x->ldb;
        segs[np].k = x->K;
        ++np;
    }
    *nprob = np;
    return total;
}

/* ---- blasx_gemm_grouped_device ---- */
static pthread_mutex_t g_glock = PTHREAD_MUTEX_INITIALIZER;
static struct {
    void *host, *dev, *done;
    size_t cap;
} g_gbuf;

BX_EXPORT int blasx_gemm_grouped_device(const blasx_gemm_desc_t *d, int n, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    if (n < 0 || (n > 0 && !d)) return BLASX_EINVAL;
    for (int i = 0; i < n; ++i)
        if (bx_desc_check(&d[i])) return BLASX_EINVAL;
    void *st = stream ? stream : g_bx.s_comp;
    int rc = 0;
    /* degenerate descriptors: C = beta C */
    for (int i = 0; i < n && !rc; ++i) {
        const int az = d[i].alpha[0] == 0.0 && (d[i].dtype != BLASX_Z || d[i].alpha[1] == 0.0);
        if (d[i].M > 0 && d[i].N > 0 && (d[i].K == 0 || az))
            rc = bxh_scale(d[i].dtype, d[i].M, d[i].N, d[i].beta, d[i].C, d[i].ldc, st);
    }
    if (rc || n == 0) return rc;

    const size_t per = (size_t)n * (sizeof(bxh_problem_t) + sizeof(bxh_seg_t));
    const size_t need = 3 * ((per + 255) & ~(size_t)255);
    pthread_mutex_lock(&g_glock);
    if (!g_gbuf.done) rc = bxh_event_create(&g_gbuf.done, 0);
    if (!rc) rc = bxh_event_sync(g_gbuf.done); /* the staging buffers are free again */
    if (!rc && need > g_gbuf.cap) {
        bxh_host_free(g_gbuf.host);
        bxh_free(g_gbuf.dev);
        g_gbuf.host = g_gbuf.dev = NULL;
        g_gbuf.cap = 0;
        rc = bxh_host_malloc(&g_gbuf.host, need);
        if (!rc) rc = bxh_malloc(&g_gbuf.dev, need);
        if (!rc) g_gbuf.cap = need;
    }
    const int dts[3] = {BLASX_S, BLASX_D, BLASX_Z};
    const size_t slot = (per + 255) & ~(size_t)255;
    int launches = 0;
    double flops = 0;
    for (int k = 0; k < 3 && !rc; ++k) {
        char *h = (char *)g_gbuf.host + k * slot, *dv = (char *)g_gbuf.dev + k * slot;
        bxh_problem_t *probs = (bxh_problem_t *)h;
        bxh_seg_t *segs = (bxh_seg_t *)(h + (size_t)n * sizeof(bxh_problem_t));
        int np = 0;
        const int total = bx_grouped_fill(d, n, dts[k], probs, segs, &np);
        if (np == 0) continue;
        rc = bxh_memcpy_h2d_async(dv, h, per, st);
        if (!rc)
            rc = bxh_gemm_grouped(dts[k], (const bxh_problem_t *)dv, np,
                                  (const bxh_seg_t *)(dv + (size_t)n * sizeof(bxh_problem_t)),
                                  total, 0, st);
        ++launches;
        for (int i = 0; i < n; ++i)
            if (d[i].dtype == dts[k])
                flops += (dts[k] == BLASX_Z ? 8.0 : 2.0) * d[i].M * (double)d[i].N * d[i].K;
    }
    if (!rc) rc = bxh_event_record(g_gbuf.done, st);
    pthread_mutex_unlock(&g_glock);
    if (!rc) {
        BX_STAT_ADD(kernel_launches, launches);
        bx_stat_flops(flops);
    }
    return rc;
}

/* ---- blasx_gemm_grouped_device_noalloc ----
 * The same launches as blasx_gemm_grouped_device with the problem and
 * segment tables in the static per-GPU table store of hip/noalloc.hip
 * instead of a device buffer the library allocates: no hipMalloc, hipFree or
 * other allocating call on this path.  The host side builds the tables in a
 * static (pageable) host buffer and copies them on `stream`.  Calls are
 * ordered on the device: a call's table copy waits for the event the
 * previous call recorded after its kernels, on whatever stream that ran, so
 * the store is never overwritten under a running launch.  The store keeps the
 * last call's tables; nothing in it needs to start at zero. */
static pthread_mutex_t g_nalock = PTHREAD_MUTEX_INITIALIZER;
static unsigned char g_na_host[BXH_NOALLOC_SCRATCH_BYTES] __attribute__((aligned(256)));
static void *g_na_copied; /* recorded after the host tables were last copied */

BX_EXPORT int blasx_gemm_grouped_device_noalloc(const blasx_gemm_desc_t *d, int n, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    if (n < 0 || (n > 0 && !d)) return BLASX_EINVAL;
    if (n > BXH_NOALLOC_MAX_PROBLEMS) {
        bx_log(0, "blasx_gemm_grouped_device_noalloc: %d problems exceed the static table of "
                  "%d (use blasx_gemm_grouped_device)",
               n, BXH_NOALLOC_MAX_PROBLEMS);
        return BLASX_ERANGE;
    }
    for (int i = 0; i < n; ++i)
        if (bx_desc_check(&d[i])) return BLASX_EINVAL;
    void *st = stream ? stream : g_bx.s_comp;
    pthread_mutex_lock(&g_nalock);
    void *store = NULL, *done = NULL;
    int rc = bxh_noalloc_scratch(&store, &done);
    if (rc) bx_log(0, "noalloc table store: %s", bxh_noalloc_error());
    if (!rc && !g_na_copied) rc = bxh_event_create(&g_na_copied, 0);
    /* the host tables are rewritten only after their last copy has read them */
    if (!rc) rc = bxh_event_sync(g_na_copied);
    /* and the device tables only after the previous call's kernels */
    if (!rc) rc = bxh_stream_wait_event(st, done);
    for (int i = 0; i < n && !rc; ++i) {
        const int az = d[i].alpha[0] == 0.0 && (d[i].dtype != BLASX_Z || d[i].alpha[1] == 0.0);
        if (d[i].M > 0 && d[i].N > 0 && (d[i].K == 0 || az))
            rc = bxh_scale(d[i].dtype, d[i].M, d[i].N, d[i].beta, d[i].C, d[i].ldc, st);
    }
    const int dts[3] = {BLASX_S, BLASX_D, BLASX_Z};
    int launches = 0;
    double flops = 0;
    for (int k = 0; k < 3 && !rc && n > 0; ++k) {
        unsigned char *h = g_na_host + (size_t)k * BXH_NOALLOC_SLOT_BYTES;
        char *dv = (char *)store + (size_t)k * BXH_NOALLOC_SLOT_BYTES;
        bxh_problem_t *probs = (bxh_problem_t *)h;
        bxh_seg_t *segs = (bxh_seg_t *)(h + BXH_NOALLOC_SEG_OFFSET);
        int np = 0;
        const int total = bx_grouped_fill(d, n, dts[k], probs, segs, &np);
        if (np == 0) continue;
        rc = bxh_memcpy_h2d_async(dv, h, (size_t)np * sizeof(bxh_problem_t), st);
        if (!rc)
            rc = bxh_memcpy_h2d_async(dv + BXH_NOALLOC_SEG_OFFSET, h + BXH_NOALLOC_SEG_OFFSET,
                                      (size_t)np * sizeof(bxh_seg_t), st);
        if (!rc)
            rc = bxh_gemm_grouped(dts[k], (const bxh_problem_t *)dv, np,
                                  (const bxh_seg_t *)(dv + BXH_NOALLOC_SEG_OFFSET), total, 0, st);
        ++launches;
        for (int i = 0; i < n; ++i)
            if (d[i].dtype == dts[k])
                flops += (dts[k] == BLASX_Z ? 8.0 : 2.0) * d[i].M * (double)d[i].N * d[i].K;
    }
    if (!rc) rc = bxh_event_record(g_na_copied, st);
    if (!rc) rc = bxh_event_record(done, st);
    pthread_mutex_unlock(&g_nalock);
    if (!rc) {
        BX_STAT_ADD(kernel_launches, launches);
        bx_stat_flops(flops);
    }
    return rc;
}
