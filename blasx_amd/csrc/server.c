/*
 * server.c -- GEMM request server: continuous batching of independent GEMM
 * requests (mixed shapes, transposes and element types) with operands in
 * host memory.  No counterpart in the reference, whose ABI is one blocking
 * call at a time and not reentrant (shared global streams and buffers,
 * blas/blasx_tile_resource.c:18-21).
 *
 * A scheduler thread owns three streams (copy-in, compute, copy-out) and a
 * device arena of a fixed byte budget:
 *   * admission: queued requests are admitted in arrival order while their
 *     operands fit in the arena; a request larger than the whole arena is
 *     rejected at submit, so admission never deadlocks;
 *   * batching: what was admitted since the last launch forms the next
 *     batch, closed at BATCH_BYTES of copies; up to NBATCH batches are in
 *     flight, so requests that arrive while others compute are copied in
 *     during that compute and launched as soon as a batch slot frees up;
 *   * launch: each batch slot owns a captured hipGraph of three persistent
 *     grouped kernels (S, D, Z) that read their problem tables from that
 *     slot's device buffer; a launch uploads the tables and replays the
 *     graph;
 *   * resident operands (blasx_server_register_operand) are uploaded once
 *     and shared by every request that points into them;
 *   * completion: the scheduler polls each batch's final event, copies
 *     have already written C back, latencies are stamped and waiters woken.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "internal.h"

#define NBATCH 3
#define MAXRES 64
/* a batch closes once its copies reach this many bytes, so that under load
 * several batches are in flight and one batch's copies overlap another's
 * compute (one huge batch would serialise copy-in, compute and copy-out) */
#define BATCH_BYTES ((size_t)384 << 20)

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + t.tv_nsec * 1e-9;
}

/* ---------------- arena: first fit over a sorted free list ---------------- */
typedef struct fblk {
    size_t off, len;
    struct fblk *next;
} fblk_t;

typedef struct arena {
    char *base;
    size_t bytes, used;
    fblk_t *free;
} arena_t;

static int arena_init(arena_t *a, size_t bytes)
{
    memset(a, 0, sizeof *a);
    if (bxh_malloc((void **)&a->base, bytes) != 0) return BLASX_ENOMEM;
    a->bytes = bytes;
    a->free = (fblk_t *)malloc(sizeof(fblk_t));
    a->free->off = 0;
    a->free->len = bytes;
    a->free->next = NULL;
    return 0;
}

static void *arena_alloc(arena_t *a, size_t len)
{
    len = (len + 255) & ~(size_t)255;
    if (len == 0) len = 256;
    for (fblk_t **pp = &a->free; *pp; pp = &(*pp)->next) {
        fblk_t *b = *pp;
        if (b->len < len) continue;
        void *p = a->base + b->off;
        b->off += len;
        b->len -= len;
        if (b->len == 0) {
            *pp = b->next;
            free(b);
        }
        a->used += len;
        return p;
    }
    return NULL;
}

static void arena_free(arena_t *a, void *p, size_t len)
{
    if (!p) return;
    len = (len + 255) & ~(size_t)255;
    if (len == 0) len = 256;
    const size_t off = (size_t)((char *)p - a->base);
    a->used -= len;
    fblk_t **pp = &a->free;
    while (*pp && (*pp)->off < off) pp = &(*pp)->next;
    fblk_t *nb = (fblk_t *)malloc(sizeof(fblk_t));
    nb->off = off;
    nb->len = len;
    nb->next = *pp;
    *pp = nb;
    /* coalesce with the next block, then with the previous one */
    if (nb->next && nb->off + nb->len == nb->next->off) {
        fblk_t *x = nb->next;
        nb->len += x->len;
        nb->next = x->next;
        free(x);
    }
    for (fblk_t *q = a->free; q && q != nb; q = q->next)
        if (q->next == nb && q->off + q->len == nb->off) {
            q->len += nb->len;
            q->next = nb->next;
            free(nb);
            break;
        }
}

static void arena_destroy(arena_t *a)
{
    while (a->free) {
        fblk_t *x = a->free;
        a->free = x->next;
        free(x);
    }
    bxh_free(a->base);
}

/* ---------------- server ---------------- */
typedef struct sreq {
    blasx_gemm_desc_t d;   /* host pointers from the caller */
    double t_submit, t_done;
    int state;             /* 0 queued, 1 in flight, 2 done, <0 error */
    void *dA, *dB, *dC;
    size_t bA, bB, bC;     /* arena bytes; 0 for resident operands */
} sreq_t;

typedef struct batch {
    int in_use;
    int64_t *ids;
    sreq_t *snap;          /* copies of the admitted requests, read while unlocked */
    int n;
    void *ev_h2d, *ev_comp, *ev_done;
    void *desc_host, *desc_dev;
    void *graph;
} batch_t;

struct blasx_server {
    pthread_t thr;
    pthread_mutex_t mu;
    pthread_cond_t cv_work, cv_done;
    int stop, started;
    sreq_t *req;
    int64_t cap, nreq, next_admit;
    arena_t arena;
    struct {
        const char *host;
        size_t bytes;
        char *dev;
    } res[MAXRES];
    int nres;
    void *s_h2d, *s_comp, *s_d2h;
    batch_t b[NBATCH];
    int max_batch, grid;
    size_t slot_bytes; /* per dtype: counts + problems + segments */
    uint64_t batches_launched, requests_done;
    double busy_flops;
};

static size_t elem(int dt) { return dt == BLASX_S ? 4 : dt == BLASX_D ? 8 : 16; }

/* stored extents of the operands of a request */
static void extents(const blasx_gemm_desc_t *d, int *ra, int *ca, int *rb, int *cb)
{
    *ra = d->transa == 111 ? d->M : d->K;
    *ca = d->transa == 111 ? d->K : d->M;
    *rb = d->transb == 111 ? d->K : d->N;
    *cb = d->transb == 111 ? d->N : d->K;
}

static char *resident_dev(blasx_server_t *s, const void *host, size_t span)
{
    const char *h = (const char *)host;
    for (int i = 0; i < s->nres; ++i)
        if (h >= s->res[i].host && h + span <= s->res[i].host + s->res[i].bytes)
            return s->res[i].dev + (h - s->res[i].host);
    return NULL;
}

static size_t span_bytes(int rows, int cols, int64_t ld, size_t es)
{
    return cols > 0 ? ((size_t)(cols - 1) * ld + rows) * es : 0;
}

/* device layout of one batch slot: per dtype k: [counts(2 ints) | probs | segs] */
static char *slot_part(blasx_server_t *s, char *base, int k) { return base + (size_t)k * s->slot_bytes; }

static int capture_graph(blasx_server_t *s, batch_t *b)
{
    const int dts[3] = {BLASX_S, BLASX_D, BLASX_Z};
    int rc = bxh_graph_begin(s->s_comp);
    for (int k = 0; k < 3 && !rc; ++k) {
        char *p = slot_part(s, (char *)b->desc_dev, k);
        rc = bxh_gemm_grouped_dyn(dts[k], (const bxh_problem_t *)(p + 256),
                                  (const bxh_seg_t *)(p + 256 + (size_t)s->max_batch * sizeof(bxh_problem_t)),
                                  (const int *)p, s->grid, s->s_comp);
    }
    void *exec = NULL;
    int rc2 = bxh_graph_end(s->s_comp, &exec);
    if (!rc) rc = rc2;
    b->graph = exec;
    return rc;
}

/* copy in, launch and copy out one batch.  Runs WITHOUT the lock: it reads
 * only the batch's snapshot of its requests and the batch's own buffers, so
 * submitters are never blocked behind copies. */
static int launch_batch(blasx_server_t *s, batch_t *b)
{
    int rc = 0;
    blasx_gemm_desc_t *dev = (blasx_gemm_desc_t *)malloc(sizeof(blasx_gemm_desc_t) * b->n);
    for (int i = 0; i < b->n && !rc; ++i) {
        const sreq_t *r = &b->snap[i];
        const blasx_gemm_desc_t *d = &r->d;
        const size_t es = elem(d->dtype);
        int ra, ca, rb, cb;
        extents(d, &ra, &ca, &rb, &cb);
        dev[i] = *d;
        if (r->bA)
            rc = bxh_memcpy2d_h2d_async(r->dA, (size_t)ra * es, d->A, (size_t)d->lda * es,
                                        (size_t)ra * es, ca, s->s_h2d);
        if (!rc && r->bB)
            rc = bxh_memcpy2d_h2d_async(r->dB, (size_t)rb * es, d->B, (size_t)d->ldb * es,
                                        (size_t)rb * es, cb, s->s_h2d);
        const int bz = d->beta[0] == 0.0 && d->beta[1] == 0.0;
        if (!rc && !bz)
            rc = bxh_memcpy2d_h2d_async(r->dC, (size_t)d->M * es, d->C, (size_t)d->ldc * es,
                                        (size_t)d->M * es, d->N, s->s_h2d);
        dev[i].A = r->dA;
        dev[i].lda = r->bA ? ra : d->lda;
        dev[i].B = r->dB;
        dev[i].ldb = r->bB ? rb : d->ldb;
        dev[i].C = r->dC;
        dev[i].ldc = d->M;
    }
    if (!rc) rc = bxh_event_record(b->ev_h2d, s->s_h2d);
    if (!rc) rc = bxh_stream_wait_event(s->s_comp, b->ev_h2d);
    /* problem tables; degenerate requests (alpha = 0 or K = 0) scale C */
    const int dts[3] = {BLASX_S, BLASX_D, BLASX_Z};
    for (int k = 0; k < 3 && !rc; ++k) {
        char *p = slot_part(s, (char *)b->desc_host, k);
        bxh_problem_t *probs = (bxh_problem_t *)(p + 256);
        bxh_seg_t *segs = (bxh_seg_t *)(p + 256 + (size_t)s->max_batch * sizeof(bxh_problem_t));
        int np = 0;
        const int total = bx_grouped_fill(dev, b->n, dts[k], probs, segs, &np);
        ((int *)p)[0] = np;
        ((int *)p)[1] = total;
    }
    if (!rc) rc = bxh_memcpy_h2d_async(b->desc_dev, b->desc_host, 3 * s->slot_bytes, s->s_comp);
    for (int i = 0; i < b->n && !rc; ++i) {
        const blasx_gemm_desc_t *d = &dev[i];
        const int az = d->alpha[0] == 0.0 && (d->dtype != BLASX_Z || d->alpha[1] == 0.0);
        if (d->M > 0 && d->N > 0 && (d->K == 0 || az))
            rc = bxh_scale(d->dtype, d->M, d->N, d->beta, d->C, d->ldc, s->s_comp);
    }
    if (!rc) rc = bxh_graph_launch(b->graph, s->s_comp);
    if (!rc) rc = bxh_event_record(b->ev_comp, s->s_comp);
    if (!rc) rc = bxh_stream_wait_event(s->s_d2h, b->ev_comp);
    for (int i = 0; i < b->n && !rc; ++i) {
        const sreq_t *r = &b->snap[i];
        const size_t es = elem(r->d.dtype);
        rc = bxh_memcpy2d_d2h_async(r->d.C, (size_t)r->d.ldc * es, r->dC, (size_t)r->d.M * es,
                                    (size_t)r->d.M * es, r->d.N, s->s_d2h);
    }
    if (!rc) rc = bxh_event_record(b->ev_done, s->s_d2h);
    free(dev);
    if (!rc) {
        s->batches_launched++;
        BX_STAT_ADD(kernel_launches, 3);
    }
    return rc;
}

/* host<->device bytes of one request if none of its operands is resident */
static size_t copy_bytes(const blasx_gemm_desc_t *d)
{
    const size_t es = elem(d->dtype);
    int ra, ca, rb, cb;
    extents(d, &ra, &ca, &rb, &cb);
    const int bz = d->beta[0] == 0.0 && d->beta[1] == 0.0;
    return es * ((size_t)ra * ca + (size_t)rb * cb + (size_t)(bz ? 1 : 2) * d->M * d->N);
}

static void release(blasx_server_t *s, sreq_t *r)
{
    arena_free(&s->arena, r->bA ? r->dA : NULL, r->bA);
    arena_free(&s->arena, r->bB ? r->dB : NULL, r->bB);
    arena_free(&s->arena, r->dC, r->bC);
    r->dA = r->dB = r->dC = NULL;
}

/* reserve arena space for request r; 0 on success */
static int admit(blasx_server_t *s, sreq_t *r)
{
    const blasx_gemm_desc_t *d = &r->d;
    const size_t es = elem(d->dtype);
    int ra, ca, rb, cb;
    extents(d, &ra, &ca, &rb, &cb);
    r->bA = r->bB = 0;
    r->dA = resident_dev(s, d->A, span_bytes(ra, ca, d->lda, es));
    r->dB = resident_dev(s, d->B, span_bytes(rb, cb, d->ldb, es));
    if (!r->dA) {
        /* owned allocations record a nonzero size (an empty operand still
         * holds one granule); 0 means "resident, not ours to free" */
        r->bA = (size_t)ra * ca * es;
        if (r->bA == 0) r->bA = 1;
        if (!(r->dA = arena_alloc(&s->arena, r->bA))) return -1;
    }
    if (!r->dB) {
        r->bB = (size_t)rb * cb * es;
        if (r->bB == 0) r->bB = 1;
        if (!(r->dB = arena_alloc(&s->arena, r->bB))) {
            arena_free(&s->arena, r->bA ? r->dA : NULL, r->bA);
            return -1;
        }
    }
    r->bC = (size_t)d->M * d->N * es;
    if (!(r->dC = arena_alloc(&s->arena, r->bC))) {
        arena_free(&s->arena, r->bA ? r->dA : NULL, r->bA);
        arena_free(&s->arena, r->bB ? r->dB : NULL, r->bB);
        return -1;
    }
    /* when the stored operand is not dense the copy-in packs it (ld = rows) */
    return 0;
}

static void *scheduler(void *arg)
{
    blasx_server_t *s = (blasx_server_t *)arg;
    bxh_set_device(g_bx.device);
    pthread_mutex_lock(&s->mu);
    while (!s->stop || s->next_admit < s->nreq) {
        int progress = 0;
        /* retire finished batches */
        for (int k = 0; k < NBATCH; ++k) {
            batch_t *b = &s->b[k];
            if (!b->in_use || bxh_event_query(b->ev_done) != 1) continue;
            const double t = now_s();
            for (int i = 0; i < b->n; ++i) {
                sreq_t *r = &s->req[b->ids[i]];
                release(s, r);
                r->t_done = t;
                r->state = 2;
                s->requests_done++;
            }
            b->in_use = 0;
            progress = 1;
            pthread_cond_broadcast(&s->cv_done);
        }
        /* form and launch a new batch */
        batch_t *fb = NULL;
        for (int k = 0; k < NBATCH && !fb; ++k)
            if (!s->b[k].in_use) fb = &s->b[k];
        if (fb && s->next_admit < s->nreq) {
            fb->n = 0;
            size_t bytes = 0;
            while (s->next_admit < s->nreq && fb->n < s->max_batch) {
                sreq_t *r = &s->req[s->next_admit];
                const size_t rb = copy_bytes(&r->d);
                if (fb->n > 0 && bytes + rb > BATCH_BYTES) break;
                if (admit(s, r) != 0) break; /* wait for memory: keep arrival order */
                bytes += rb;
                r->state = 1;
                fb->snap[fb->n] = *r;
                fb->ids[fb->n++] = s->next_admit++;
            }
            if (fb->n > 0) {
                fb->in_use = 1;
                /* only this thread launches or retires batches, so the slot
                 * stays ours while the lock is dropped for the copies */
                pthread_mutex_unlock(&s->mu);
                const int lrc = launch_batch(s, fb);
                pthread_mutex_lock(&s->mu);
                if (lrc != 0) {
                    bx_log(0, "server: batch launch failed: %s", bxh_last_error());
                    bxh_device_sync();
                    for (int i = 0; i < fb->n; ++i) {
                        sreq_t *r = &s->req[fb->ids[i]];
                        release(s, r);
                        r->state = BLASX_EHIP;
                    }
                    fb->in_use = 0;
                    pthread_cond_broadcast(&s->cv_done);
                }
                progress = 1;
            }
        }
        if (!progress) {
            int inflight = 0;
            for (int k = 0; k < NBATCH; ++k) inflight |= s->b[k].in_use;
            struct timespec ts;
            clock_gettime(CLOCK_REALTIME, &ts);
            ts.tv_nsec += inflight ? 50000 : 2000000; /* poll while busy, sleep when idle */
            if (ts.tv_nsec >= 1000000000) {
                ts.tv_sec++;
                ts.tv_nsec -= 1000000000;
            }
            pthread_cond_timedwait(&s->cv_work, &s->mu, &ts);
        }
        if (s->stop) {
            int inflight = 0;
            for (int k = 0; k < NBATCH; ++k) inflight |= s->b[k].in_use;
            if (!inflight && s->next_admit >= s->nreq) break;
        }
    }
    pthread_mutex_unlock(&s->mu);
    return NULL;
}

BX_EXPORT blasx_server_t *blasx_server_create(size_t hbm_bytes, int max_batch)
{
    if (bx_ensure_init() != 0) return NULL;
    if (max_batch <= 0) max_batch = 256;
    blasx_server_t *s = (blasx_server_t *)calloc(1, sizeof *s);
    if (!s) return NULL;
    pthread_mutex_init(&s->mu, NULL);
    pthread_cond_init(&s->cv_work, NULL);
    pthread_cond_init(&s->cv_done, NULL);
    s->max_batch = max_batch;
    s->grid = 2 * (g_bx.cus > 0 ? g_bx.cus : 256);
    s->slot_bytes = (256 + (size_t)max_batch * (sizeof(bxh_problem_t) + sizeof(bxh_seg_t)) + 255) &
                    ~(size_t)255;
    int rc = arena_init(&s->arena, hbm_bytes);
    if (!rc) rc = bxh_stream_create(&s->s_h2d);
    if (!rc) rc = bxh_stream_create(&s->s_comp);
    if (!rc) rc = bxh_stream_create(&s->s_d2h);
    for (int k = 0; k < NBATCH && !rc; ++k) {
        batch_t *b = &s->b[k];
        b->ids = (int64_t *)malloc(sizeof(int64_t) * max_batch);
        b->snap = (sreq_t *)malloc(sizeof(sreq_t) * max_batch);
        rc = bxh_event_create(&b->ev_h2d, 0);
        if (!rc) rc = bxh_event_create(&b->ev_comp, 0);
        if (!rc) rc = bxh_event_create(&b->ev_done, 0);
        if (!rc) rc = bxh_host_malloc(&b->desc_host, 3 * s->slot_bytes);
        if (!rc) rc = bxh_malloc(&b->desc_dev, 3 * s->slot_bytes);
        if (!rc) {
            /* empty tables until the first launch */
            memset(b->desc_host, 0, 3 * s->slot_bytes);
            rc = bxh_memcpy_h2d(b->desc_dev, b->desc_host, 3 * s->slot_bytes);
        }
        if (!rc) rc = capture_graph(s, b);
    }
    s->cap = 1024;
    s->req = (sreq_t *)calloc(s->cap, sizeof(sreq_t));
    if (!rc && pthread_create(&s->thr, NULL, scheduler, s) == 0) s->started = 1;
    if (rc || !s->started) {
        bx_log(0, "server: creation failed: %s", bxh_last_error());
        blasx_server_destroy(s);
        return NULL;
    }
    return s;
}

BX_EXPORT int blasx_server_register_operand(blasx_server_t *s, const void *host, size_t bytes)
{
    if (!s || !host || !bytes) return BLASX_EINVAL;
    pthread_mutex_lock(&s->mu);
    int rc = 0;
    if (s->nres >= MAXRES) rc = BLASX_ERANGE;
    char *dev = NULL;
    if (!rc && !(dev = (char *)arena_alloc(&s->arena, bytes))) rc = BLASX_ENOMEM;
    if (!rc) rc = bxh_memcpy_h2d(dev, host, bytes);
    if (!rc) {
        s->res[s->nres].host = (const char *)host;
        s->res[s->nres].bytes = bytes;
        s->res[s->nres].dev = dev;
        s->nres++;
    }
    pthread_mutex_unlock(&s->mu);
    return rc;
}

BX_EXPORT int64_t blasx_server_submit(blasx_server_t *s, const blasx_gemm_desc_t *d)
{
    if (!s || !d) return BLASX_EINVAL;
    if (bx_desc_check(d)) return BLASX_EINVAL;
    const size_t es = elem(d->dtype);
    int ra, ca, rb, cb;
    extents(d, &ra, &ca, &rb, &cb);
    const size_t need = (size_t)ra * ca * es + (size_t)rb * cb * es + (size_t)d->M * d->N * es +
                        3 * 256;
    pthread_mutex_lock(&s->mu);
    if (need > s->arena.bytes) {
        pthread_mutex_unlock(&s->mu);
        return BLASX_ERANGE; /* could never be admitted */
    }
    if (s->nreq == s->cap) {
        sreq_t *nr = (sreq_t *)realloc(s->req, sizeof(sreq_t) * s->cap * 2);
        if (!nr) {
            pthread_mutex_unlock(&s->mu);
            return BLASX_ENOMEM;
        }
        memset(nr + s->cap, 0, sizeof(sreq_t) * s->cap);
        s->req = nr;
        s->cap *= 2;
    }
    const int64_t id = s->nreq++;
    s->req[id].d = *d;
    s->req[id].t_submit = now_s();
    s->req[id].state = 0;
    pthread_cond_signal(&s->cv_work);
    pthread_mutex_unlock(&s->mu);
    return id;
}

BX_EXPORT int blasx_server_poll(blasx_server_t *s, int64_t id)
{
    if (!s || id < 0) return BLASX_EINVAL;
    pthread_mutex_lock(&s->mu);
    int st = id < s->nreq ? s->req[id].state : BLASX_EINVAL;
    pthread_mutex_unlock(&s->mu);
    return st == 2 ? 1 : st < 0 ? st : 0;
}

BX_EXPORT int blasx_server_wait(blasx_server_t *s, int64_t id, double *latency_ms)
{
    if (!s || id < 0) return BLASX_EINVAL;
    pthread_mutex_lock(&s->mu);
    if (id >= s->nreq) {
        pthread_mutex_unlock(&s->mu);
        return BLASX_EINVAL;
    }
    while (s->req[id].state == 0 || s->req[id].state == 1) pthread_cond_wait(&s->cv_done, &s->mu);
    const int st = s->req[id].state;
    if (latency_ms) *latency_ms = (s->req[id].t_done - s->req[id].t_submit) * 1e3;
    pthread_mutex_unlock(&s->mu);
    return st == 2 ? 0 : st;
}

BX_EXPORT void blasx_server_destroy(blasx_server_t *s)
{
    if (!s) return;
    if (s->started) {
        pthread_mutex_lock(&s->mu);
        s->stop = 1;
        pthread_cond_signal(&s->cv_work);
        pthread_mutex_unlock(&s->mu);
        pthread_join(s->thr, NULL);
    }
    bxh_device_sync();
    for (int k = 0; k < NBATCH; ++k) {
        batch_t *b = &s->b[k];
        bxh_graph_destroy(b->graph);
        bxh_event_destroy(b->ev_h2d);
        bxh_event_destroy(b->ev_comp);
        bxh_event_destroy(b->ev_done);
        bxh_host_free(b->desc_host);
        bxh_free(b->desc_dev);
        free(b->ids);
        free(b->snap);
    }
    bxh_stream_destroy(s->s_h2d);
    bxh_stream_destroy(s->s_comp);
    bxh_stream_destroy(s->s_d2h);
    if (s->arena.base) arena_destroy(&s->arena);
    free(s->req);
    pthread_mutex_destroy(&s->mu);
    pthread_cond_destroy(&s->cv_work);
    pthread_cond_destroy(&s->cv_done);
    free(s);
}

BX_EXPORT int blasx_server_counters(blasx_server_t *s, uint64_t *batches, uint64_t *done,
                                    size_t *arena_used)
{
    if (!s) return BLASX_EINVAL;
    pthread_mutex_lock(&s->mu);
    if (batches) *batches = s->batches_launched;
    if (done) *done = s->requests_done;
    if (arena_used) *arena_used = s->arena.used;
    pthread_mutex_unlock(&s->mu);
    return 0;
}
