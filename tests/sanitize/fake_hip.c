/*
 * fake_hip.c -- a host stand-in for the HIP layer (hiplayer.h) and the
 * library context (core.c), so that the host runtime's threaded code
 * (tile_cache.c, queue.c, hostcopy.c, grouped.c, server.c) can be built
 * and run under AddressSanitizer and ThreadSanitizer on the CPU.
 *
 * "Device" memory is host memory; every copy and kernel runs synchronously
 * when issued, so streams and events are trivially complete.  The grouped
 * kernels are a naive GEMM over the problem and segment tables; a captured
 * graph records the grouped launches between begin and end and replays them.
 * Test infrastructure only: nothing here is part of libblasx.
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

bx_ctx_t g_bx;
bx_stats_atomic_t g_stats = {.flops_lock = PTHREAD_MUTEX_INITIALIZER};

void bx_config(void)
{
    if (g_bx.stage_threads == 0) g_bx.stage_threads = 6;
}
int bx_ensure_init(void)
{
    static pthread_once_t once = PTHREAD_ONCE_INIT;
    pthread_once(&once, bx_config);
    return 0;
}
void bx_log(int level, const char *fmt, ...)
{
    if (level > 0) return;
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "[fake blasx] ");
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
}
void bx_fatal(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    abort();
}
void bx_stat_flops(double f)
{
    pthread_mutex_lock(&g_stats.flops_lock);
    g_stats.flops += f;
    pthread_mutex_unlock(&g_stats.flops_lock);
}

/* ---- runtime ---- */
const char *bxh_last_error(void) { return "fake device"; }
int bxh_set_device(int dev) { (void)dev; return 0; }
int bxh_device_sync(void) { return 0; }
int bxh_malloc(void **p, size_t bytes)
{
    *p = malloc(bytes ? bytes : 1);
    return *p ? 0 : BLASX_ENOMEM;
}
int bxh_free(void *p) { free(p); return 0; }
int bxh_host_malloc(void **p, size_t bytes) { return bxh_malloc(p, bytes); }
int bxh_host_free(void *p) { free(p); return 0; }
int bxh_stream_create(void **s) { *s = malloc(1); return 0; }
int bxh_stream_destroy(void *s) { free(s); return 0; }
int bxh_stream_sync(void *s) { (void)s; return 0; }
int bxh_event_create(void **e, int timing) { (void)timing; *e = malloc(1); return 0; }
int bxh_event_destroy(void *e) { free(e); return 0; }
int bxh_event_record(void *e, void *s) { (void)e; (void)s; return 0; }
int bxh_event_query(void *e) { (void)e; return 1; }
int bxh_event_sync(void *e) { (void)e; return 0; }
int bxh_stream_wait_event(void *s, void *e) { (void)s; (void)e; return 0; }
int bxh_memcpy_h2d_async(void *d, const void *s, size_t n, void *st)
{
    (void)st;
    memcpy(d, s, n);
    return 0;
}
int bxh_memcpy_h2d(void *d, const void *s, size_t n) { memcpy(d, s, n); return 0; }
static void copy2d(void *dst, size_t dp, const void *src, size_t sp, size_t w, size_t h)
{
    for (size_t c = 0; c < h; ++c) memcpy((char *)dst + c * dp, (const char *)src + c * sp, w);
}
int bxh_memcpy2d_h2d_async(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h,
                           void *st)
{
    (void)st;
    copy2d(d, dp, s, sp, w, h);
    return 0;
}
int bxh_memcpy2d_d2h_async(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h,
                           void *st)
{
    (void)st;
    copy2d(d, dp, s, sp, w, h);
    return 0;
}

/* ---- kernels ---- */
int bxh_tile_m(int dtype) { (void)dtype; return 128; }
int bxh_tile_n(int dtype) { return dtype == BLASX_Z ? 64 : 128; }
void bxh_tile_dims(int dtype, int shape, int *bm, int *bn)
{
    *bm = dtype != BLASX_Z && shape == BXH_SHAPE_64x64 ? 64 : 128;
    *bn = dtype == BLASX_Z || shape != BXH_SHAPE_128x128 ? 64 : 128;
}
/* stand-in for the measured shape table (autotune.c): every shape occurs */
int bx_grouped_shape(int dt, int M, int N, int K, int64_t load)
{
    return dt == BLASX_Z ? 0 : (int)(((int64_t)M + N + K + load) % BXH_NSHAPES);
}

typedef struct { double re, im; } zc;

static zc ld_elem(int dt, const void *p, int64_t i)
{
    zc v = {0, 0};
    if (dt == BLASX_S) v.re = ((const float *)p)[i];
    else if (dt == BLASX_D) v.re = ((const double *)p)[i];
    else v = ((const zc *)p)[i];
    return v;
}

static void st_elem(int dt, void *p, int64_t i, zc v)
{
    if (dt == BLASX_S) ((float *)p)[i] = (float)v.re;
    else if (dt == BLASX_D) ((double *)p)[i] = v.re;
    else ((zc *)p)[i] = v;
}

/* op(X)(r, c) of a stored matrix with leading dimension ld */
static zc op_at(int dt, int op, const void *X, int64_t ld, int r, int c)
{
    zc v = op == 0 ? ld_elem(dt, X, r + (int64_t)c * ld) : ld_elem(dt, X, c + (int64_t)r * ld);
    if (op == 2) v.im = -v.im;
    return v;
}

static void naive_problems(int dt, const bxh_problem_t *probs, int nprob, const bxh_seg_t *segs)
{
    for (int p = 0; p < nprob; ++p) {
        const bxh_problem_t *P = &probs[p];
        for (int j = 0; j < P->n; ++j)
            for (int i = 0; i < P->m; ++i) {
                zc acc = {0, 0};
                for (int s = 0; s < P->nseg; ++s) {
                    const bxh_seg_t *S = &segs[P->seg_off + s];
                    for (int k = 0; k < S->k; ++k) {
                        const zc a = op_at(dt, P->opa, S->A, S->lda, i, k);
                        const zc b = op_at(dt, P->opb, S->B, S->ldb, k, j);
                        acc.re += a.re * b.re - a.im * b.im;
                        acc.im += a.re * b.im + a.im * b.re;
                    }
                }
                zc out = {P->alpha_re * acc.re - P->alpha_im * acc.im,
                          P->alpha_re * acc.im + P->alpha_im * acc.re};
                const int64_t ci = i + (int64_t)j * P->ldc;
                if (!P->beta_zero) {
                    const zc c = ld_elem(dt, P->C, ci);
                    out.re += P->beta_re * c.re - P->beta_im * c.im;
                    out.im += P->beta_re * c.im + P->beta_im * c.re;
                }
                st_elem(dt, P->C, ci, out);
            }
    }
}

int bxh_gemm_grouped(int dtype, const bxh_problem_t *probs, int nprob, const bxh_seg_t *segs,
                     int total, int grid, void *stream)
{
    (void)total, (void)grid, (void)stream;
    naive_problems(dtype, probs, nprob, segs);
    return 0;
}

/* graph capture: the grouped launches between begin and end, per stream */
typedef struct rec {
    int dtype;
    const bxh_problem_t *probs;
    const bxh_seg_t *segs;
    const int *counts;
} rec_t;
typedef struct graph {
    rec_t r[8];
    int n;
} graph_t;
static __thread graph_t *g_capturing;

int bxh_graph_begin(void *s)
{
    (void)s;
    g_capturing = (graph_t *)calloc(1, sizeof(graph_t));
    return g_capturing ? 0 : BLASX_ENOMEM;
}
int bxh_graph_end(void *s, void **exec)
{
    (void)s;
    *exec = g_capturing;
    g_capturing = NULL;
    return 0;
}
int bxh_graph_launch(void *exec, void *s)
{
    (void)s;
    graph_t *g = (graph_t *)exec;
    for (int i = 0; i < g->n; ++i)
        naive_problems(g->r[i].dtype, g->r[i].probs, g->r[i].counts[0], g->r[i].segs);
    return 0;
}
int bxh_graph_destroy(void *exec) { free(exec); return 0; }

int bxh_gemm_grouped_dyn(int dtype, const bxh_problem_t *probs, const bxh_seg_t *segs,
                         const int *counts, int grid, void *stream)
{
    (void)grid, (void)stream;
    if (g_capturing) {
        if (g_capturing->n >= 8) return BLASX_ERANGE;
        g_capturing->r[g_capturing->n++] = (rec_t){dtype, probs, segs, counts};
        return 0;
    }
    naive_problems(dtype, probs, counts[0], segs);
    return 0;
}

int bxh_scale(int dtype, int m, int n, const double *beta, void *C, int64_t ldc, void *stream)
{
    (void)stream;
    const int bz = beta[0] == 0.0 && (dtype != BLASX_Z || beta[1] == 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            const int64_t ci = i + (int64_t)j * ldc;
            zc v = {0, 0};
            if (!bz) {
                const zc c = ld_elem(dtype, C, ci);
                v.re = beta[0] * c.re - (dtype == BLASX_Z ? beta[1] * c.im : 0);
                v.im = dtype == BLASX_Z ? beta[0] * c.im + beta[1] * c.re : 0;
            }
            st_elem(dtype, C, ci, v);
        }
    return 0;
}

/* the allocation-free grouped entry point's table store */
static unsigned char g_store[BXH_NOALLOC_SCRATCH_BYTES];
int bxh_noalloc_setup(void) { return 0; }
int bxh_noalloc_scratch(void **ptr, void **done)
{
    static char ev;
    *ptr = g_store;
    *done = &ev;
    return 0;
}
const char *bxh_noalloc_error(void) { return ""; }
