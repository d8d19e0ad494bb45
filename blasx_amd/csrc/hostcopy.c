/*
 * hostcopy.c -- parallel strided host copies for staging pageable memory.
 *
 * DMA engines can only read or write pinned host memory; HIP stages
 * pageable copies itself, synchronously and on one thread.  The tile
 * runtime instead packs a tile from the caller's pageable matrix into a
 * pinned staging buffer with this pool (columns split over the workers and
 * the calling thread), then DMAs it asynchronously; copy-out runs the other
 * way.  The pool's workers live for the life of the process.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define MAXW 32

static struct {
    pthread_once_t once;
    pthread_mutex_t mu;
    pthread_cond_t go, done;
    int nw;          /* worker threads (the caller works too) */
    uint64_t gen;    /* job generation */
    int pending;     /* workers still busy on the current job */
    char *dst;
    const char *src;
    size_t dpitch, spitch, width, height;
} P = {PTHREAD_ONCE_INIT, PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER,
       PTHREAD_COND_INITIALIZER, 0, 0, 0, NULL, NULL, 0, 0, 0, 0};

static void copy_range(size_t c0, size_t c1)
{
    for (size_t c = c0; c < c1; ++c) memcpy(P.dst + c * P.dpitch, P.src + c * P.spitch, P.width);
}

static void *worker(void *arg)
{
    const int me = (int)(intptr_t)arg; /* 1..nw; part 0 is the caller's */
    uint64_t seen = 0;
    pthread_mutex_lock(&P.mu);
    for (;;) {
        while (P.gen == seen) pthread_cond_wait(&P.go, &P.mu);
        seen = P.gen;
        const size_t h = P.height, parts = (size_t)P.nw + 1;
        pthread_mutex_unlock(&P.mu);
        copy_range(h * me / parts, h * (me + 1) / parts);
        pthread_mutex_lock(&P.mu);
        if (--P.pending == 0) pthread_cond_signal(&P.done);
    }
    return NULL;
}

static void start(void)
{
    bx_config();
    int n = g_bx.stage_threads;
    if (n < 0) n = 0;
    if (n > MAXW) n = MAXW;
    int started = 0;
    for (int i = 0; i < n; ++i) {
        pthread_t t;
        if (pthread_create(&t, NULL, worker, (void *)(intptr_t)(i + 1)) != 0) break;
        pthread_detach(t);
        ++started;
    }
    P.nw = started;
}

/* dst/src: `height` columns of `width` bytes at the given pitches */
void bx_host_copy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                    size_t height)
{
    pthread_once(&P.once, start);
    const size_t bytes = width * height;
    if (P.nw == 0 || bytes < ((size_t)1 << 20) || height < 2) {
        for (size_t c = 0; c < height; ++c)
            memcpy((char *)dst + c * dpitch, (const char *)src + c * spitch, width);
        return;
    }
    static pthread_mutex_t one_job = PTHREAD_MUTEX_INITIALIZER; /* one job at a time */
    pthread_mutex_lock(&one_job);
    pthread_mutex_lock(&P.mu);
    P.dst = (char *)dst;
    P.src = (const char *)src;
    P.dpitch = dpitch;
    P.spitch = spitch;
    P.width = width;
    P.height = height;
    P.pending = P.nw;
    P.gen++;
    pthread_cond_broadcast(&P.go);
    pthread_mutex_unlock(&P.mu);
    copy_range(0, height / ((size_t)P.nw + 1));
    pthread_mutex_lock(&P.mu);
    while (P.pending > 0) pthread_cond_wait(&P.done, &P.mu);
    pthread_mutex_unlock(&P.mu);
    pthread_mutex_unlock(&one_job);
}
