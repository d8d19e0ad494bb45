/*
 * hostcopy.c -- parallel strided host copies for staging pageable memory.
 *
 * DMA engines can only read or write pinned host memory; HIP stages
 * pageable copies itself, synchronously and on one thread.  The tile
 * runtime instead packs a tile from the caller's pageable matrix into a
 * pinned staging buffer with this pool, then DMAs it asynchronously;
 * copy-out runs the other way.  The pool's workers live for the life of the
 * process.
 *
 * Work distribution is the reference's task queue (blas/blasx_task_queue.c:
 * 3-17, bx_queue in queue.c): a job is cut into column chunks, and the
 * workers and the calling thread take chunk ids from one shared
 * compare-and-swap counter, highest first, until it reports -1 -- the
 * reference hands C tiles to its GPU threads the same way
 * (blas/blasx_dgemm.c:174-183).  A worker that is descheduled simply takes
 * fewer chunks.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define MAXW 32
#define CHUNK_BYTES ((size_t)1 << 20) /* about 1 MiB of columns per task */

static struct {
    pthread_once_t once;
    pthread_mutex_t mu;
    pthread_cond_t go, done;
    int nw;          /* worker threads (the caller works too) */
    uint64_t gen;    /* job generation */
    int pending;     /* workers still busy on the current job */
    char *dst;
    const char *src;
    size_t dpitch, spitch, width, height, cols_per_task;
    bx_queue_t tasks; /* chunk ids of the current job */
} P = {PTHREAD_ONCE_INIT, PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER,
       PTHREAD_COND_INITIALIZER, 0, 0, 0, NULL, NULL, 0, 0, 0, 0, 0, {0}};

/* take chunks until the queue is empty */
static void drain(void)
{
    for (int64_t t; (t = bx_queue_next(&P.tasks)) >= 0;) {
        const size_t c0 = (size_t)t * P.cols_per_task;
        size_t c1 = c0 + P.cols_per_task;
        if (c1 > P.height) c1 = P.height;
        for (size_t c = c0; c < c1; ++c)
            memcpy(P.dst + c * P.dpitch, P.src + c * P.spitch, P.width);
    }
}

static void *worker(void *arg)
{
    (void)arg;
    uint64_t seen = 0;
    pthread_mutex_lock(&P.mu);
    for (;;) {
        while (P.gen == seen) pthread_cond_wait(&P.go, &P.mu);
        seen = P.gen;
        pthread_mutex_unlock(&P.mu);
        drain();
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
        if (pthread_create(&t, NULL, worker, NULL) != 0) break;
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
    P.cols_per_task = CHUNK_BYTES / width ? CHUNK_BYTES / width : 1;
    bx_queue_init(&P.tasks, (int64_t)((height + P.cols_per_task - 1) / P.cols_per_task));
    P.pending = P.nw;
    P.gen++;
    pthread_cond_broadcast(&P.go);
    pthread_mutex_unlock(&P.mu);
    drain();
    pthread_mutex_lock(&P.mu);
    while (P.pending > 0) pthread_cond_wait(&P.done, &P.mu);
    pthread_mutex_unlock(&P.mu);
    pthread_mutex_unlock(&one_job);
}
