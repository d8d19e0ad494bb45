/*
 * stress.c -- the host runtime's threaded parts under AddressSanitizer /
 * ThreadSanitizer (built by tests/sanitize/Makefile against fake_hip.c):
 *   1. the tile cache: random acquire / scratch / invalidate sequences
 *      against its structural invariants (bx_cache_validate);
 *   2. the task queue: 8 threads drain 200000 ids, each handed out once;
 *   3. the host copy pool: concurrent 2-D copies from several threads
 *      (one job at a time inside the pool), checked against memcpy;
 *   4. the GEMM server: 4 submitting threads, 240 requests of mixed type,
 *      shape and transpose through admission under a small arena (so
 *      requests wait for memory), continuous batching and completion;
 *      every result checked against a naive product.
 * Exit status 0 when every check passes.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define CHECK(c)                                                                              \
    do {                                                                                      \
        if (!(c)) {                                                                           \
            fprintf(stderr, "check failed: %s (%s:%d)\n", #c, __FILE__, __LINE__);            \
            exit(1);                                                                          \
        }                                                                                     \
    } while (0)

static uint64_t rng_state = 88172645463325252ull;
static uint64_t xr(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return rng_state;
}

/* ---- 1. cache ---- */
static int host_alloc(void *u, void **p, size_t b)
{
    (void)u;
    *p = malloc(b);
    return *p ? 0 : BLASX_ENOMEM;
}
static void host_free(void *u, void *p)
{
    (void)u;
    free(p);
}

static void test_cache(void)
{
    bx_cache_ops_t ops = {host_alloc, host_free, NULL};
    bx_cache_t *c = bx_cache_create(4096, 4096 * 8, 4096 * 50, &ops);
    CHECK(c);
    int scratch[8], ns = 0;
    for (int it = 0; it < 20000; ++it) {
        const int64_t round = it / 7;
        const uint64_t r = xr();
        if (r % 97 == 0) {
            bx_cache_invalidate_all(c);
        } else if (r % 13 == 0 && ns < 8) {
            const int s = bx_cache_take_scratch(c, round);
            if (s >= 0) scratch[ns++] = s;
        } else if (r % 11 == 0 && ns > 0) {
            bx_cache_return_scratch(c, scratch[--ns]);
        } else {
            bx_tile_key_t k = {(uintptr_t)(0x10000 + 64 * (r % 90)), 100, 10, 10, 8, 0};
            int hit;
            int64_t ev;
            const int s = bx_cache_acquire(c, &k, round, &hit, &ev);
            if (s >= 0) memset(bx_cache_slot_ptr(c, s), (int)(r & 255), 4096);
        }
        if (it % 64 == 0) CHECK(bx_cache_validate(c) == 0);
    }
    CHECK(bx_cache_validate(c) == 0);
    bx_cache_destroy(c);
}

/* ---- 2. queue ---- */
#define QN 200000
static bx_queue_t g_q;
static unsigned char g_seen[QN];
static void *q_worker(void *a)
{
    (void)a;
    for (int64_t t; (t = bx_queue_next(&g_q)) >= 0;) __atomic_fetch_add(&g_seen[t], 1, __ATOMIC_RELAXED);
    return NULL;
}
static void test_queue(void)
{
    bx_queue_init(&g_q, QN);
    pthread_t th[8];
    for (int i = 0; i < 8; ++i) pthread_create(&th[i], NULL, q_worker, NULL);
    for (int i = 0; i < 8; ++i) pthread_join(th[i], NULL);
    for (int i = 0; i < QN; ++i) CHECK(g_seen[i] == 1);
    CHECK(bx_queue_next(&g_q) == -1);
}

/* ---- 3. host copy pool ---- */
void bx_host_copy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                    size_t height);
static void *copy_worker(void *a)
{
    const int id = (int)(intptr_t)a;
    for (int it = 0; it < 6; ++it) {
        const size_t w = 1000 + 4096 * (size_t)((id + it) % 5), h = 300 + 97 * (size_t)it;
        const size_t sp = w + 64, dp = w + 128;
        unsigned char *src = (unsigned char *)malloc(sp * h), *dst = (unsigned char *)malloc(dp * h);
        for (size_t i = 0; i < sp * h; ++i) src[i] = (unsigned char)(i * 31 + id);
        bx_host_copy2d(dst, dp, src, sp, w, h);
        for (size_t c = 0; c < h; ++c) CHECK(memcmp(dst + c * dp, src + c * sp, w) == 0);
        free(src);
        free(dst);
    }
    return NULL;
}
static void test_hostcopy(void)
{
    pthread_t th[4];
    for (int i = 0; i < 4; ++i) pthread_create(&th[i], NULL, copy_worker, (void *)(intptr_t)i);
    for (int i = 0; i < 4; ++i) pthread_join(th[i], NULL);
}

/* ---- 4. server ---- */
#define NREQ 240
typedef struct {
    blasx_gemm_desc_t d;
    void *A, *B, *C, *C0;
    int64_t id;
} req_t;
static req_t g_req[NREQ];
static blasx_server_t *g_srv;

static double val(int i) { return (double)((int)(xr() % 17) - 8) + 0.25 * (i % 3); }

static void make_req(req_t *r, int i)
{
    const int dt = (int)(xr() % 3); /* S, D, Z */
    const size_t es = dt == BLASX_S ? 4 : dt == BLASX_D ? 8 : 16;
    const int M = 1 + (int)(xr() % 40), N = 1 + (int)(xr() % 40), K = (int)(xr() % 30);
    const int ta = 111 + (int)(xr() % (dt == BLASX_Z ? 3 : 2));
    const int tb = 111 + (int)(xr() % (dt == BLASX_Z ? 3 : 2));
    const int ra = ta == 111 ? M : K, ca = ta == 111 ? K : M;
    const int rb = tb == 111 ? K : N, cb = tb == 111 ? N : K;
    memset(&r->d, 0, sizeof r->d);
    r->d.dtype = dt;
    r->d.transa = ta;
    r->d.transb = tb;
    r->d.M = M;
    r->d.N = N;
    r->d.K = K;
    r->d.alpha[0] = 1.5;
    r->d.beta[0] = (i % 5 == 0) ? 0.0 : -0.5;
    r->A = calloc((size_t)(ra ? ra : 1) * (ca ? ca : 1), es);
    r->B = calloc((size_t)(rb ? rb : 1) * (cb ? cb : 1), es);
    r->C = calloc((size_t)M * N, es);
    r->C0 = calloc((size_t)M * N, es);
    const size_t na = (size_t)ra * ca, nb = (size_t)rb * cb, nc = (size_t)M * N;
    const int per = dt == BLASX_Z ? 2 : 1;
    for (size_t k = 0; k < na * per; ++k)
        if (dt == BLASX_S) ((float *)r->A)[k] = (float)val(i); else ((double *)r->A)[k] = val(i);
    for (size_t k = 0; k < nb * per; ++k)
        if (dt == BLASX_S) ((float *)r->B)[k] = (float)val(i); else ((double *)r->B)[k] = val(i);
    for (size_t k = 0; k < nc * per; ++k)
        if (dt == BLASX_S) ((float *)r->C0)[k] = (float)val(i); else ((double *)r->C0)[k] = val(i);
    memcpy(r->C, r->C0, nc * es);
    r->d.A = r->A;
    r->d.lda = ra > 1 ? ra : 1;
    r->d.B = r->B;
    r->d.ldb = rb > 1 ? rb : 1;
    r->d.C = r->C;
    r->d.ldc = M;
}

static void check_req(const req_t *r)
{
    const blasx_gemm_desc_t *d = &r->d;
    const int dt = d->dtype, cplx = dt == BLASX_Z;
    for (int j = 0; j < d->N; ++j)
        for (int i = 0; i < d->M; ++i) {
            double re = 0, im = 0;
            for (int k = 0; k < d->K; ++k) {
                const int64_t ia = d->transa == 111 ? i + (int64_t)k * d->lda : k + (int64_t)i * d->lda;
                const int64_t ib = d->transb == 111 ? k + (int64_t)j * d->ldb : j + (int64_t)k * d->ldb;
                double ar, ai = 0, br, bi = 0;
                if (dt == BLASX_S) ar = ((float *)r->A)[ia], br = ((float *)r->B)[ib];
                else if (dt == BLASX_D) ar = ((double *)r->A)[ia], br = ((double *)r->B)[ib];
                else {
                    ar = ((double *)r->A)[2 * ia], ai = ((double *)r->A)[2 * ia + 1];
                    br = ((double *)r->B)[2 * ib], bi = ((double *)r->B)[2 * ib + 1];
                    if (d->transa == 113) ai = -ai;
                    if (d->transb == 113) bi = -bi;
                }
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
            }
            const int64_t ci = i + (int64_t)j * d->ldc;
            double c0r, c0i = 0, gr, gi = 0;
            if (dt == BLASX_S) c0r = ((float *)r->C0)[ci], gr = ((float *)r->C)[ci];
            else if (dt == BLASX_D) c0r = ((double *)r->C0)[ci], gr = ((double *)r->C)[ci];
            else {
                c0r = ((double *)r->C0)[2 * ci], c0i = ((double *)r->C0)[2 * ci + 1];
                gr = ((double *)r->C)[2 * ci], gi = ((double *)r->C)[2 * ci + 1];
            }
            const double er = d->alpha[0] * re + (d->beta[0] != 0.0 ? d->beta[0] * c0r : 0.0);
            const double ei = d->alpha[0] * im + (d->beta[0] != 0.0 ? d->beta[0] * c0i : 0.0);
            /* small integers and quarters: exact in every type at these sizes */
            CHECK(gr == er);
            if (cplx) CHECK(gi == ei);
        }
}

static void *submit_worker(void *a)
{
    const int t = (int)(intptr_t)a;
    for (int i = t; i < NREQ; i += 4) {
        g_req[i].id = blasx_server_submit(g_srv, &g_req[i].d);
        CHECK(g_req[i].id >= 0);
        if (i % 9 == 0) {
            double ms;
            CHECK(blasx_server_wait(g_srv, g_req[i].id, &ms) == 0);
        }
    }
    return NULL;
}

static void test_server(void)
{
    for (int i = 0; i < NREQ; ++i) make_req(&g_req[i], i);
    /* an arena of 48 KiB: most requests fit alone, few together, so
     * admission has to wait for completions */
    g_srv = blasx_server_create(48 << 10, 16);
    CHECK(g_srv);
    pthread_t th[4];
    for (int i = 0; i < 4; ++i) pthread_create(&th[i], NULL, submit_worker, (void *)(intptr_t)i);
    for (int i = 0; i < 4; ++i) pthread_join(th[i], NULL);
    for (int i = 0; i < NREQ; ++i) {
        double ms;
        CHECK(blasx_server_wait(g_srv, g_req[i].id, &ms) == 0);
        check_req(&g_req[i]);
    }
    uint64_t batches, done;
    size_t used;
    CHECK(blasx_server_counters(g_srv, &batches, &done, &used) == 0);
    CHECK(done == NREQ && used == 0 && batches > 1);
    blasx_server_destroy(g_srv);
    for (int i = 0; i < NREQ; ++i) {
        free(g_req[i].A);
        free(g_req[i].B);
        free(g_req[i].C);
        free(g_req[i].C0);
    }
}

int main(void)
{
    test_cache();
    printf("cache ok\n");
    test_queue();
    printf("queue ok\n");
    test_hostcopy();
    printf("host copy pool ok\n");
    test_server();
    printf("server ok\n");
    return 0;
}
