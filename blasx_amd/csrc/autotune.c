/*
 * autotune.c -- tile-size table for the host-operand runtime.
 *
 * The reference fixes the tile edge at compile time (BLOCKSIZE_DGEMM 2048,
 * BLOCKSIZE_ZGEMM 1024, include/blasx.h:5-7).  Here the edge is chosen per
 * call from a table of measured winners (written by blasx_autotune_tiles)
 * and otherwise by the heuristic in tile_rt.c.  The table starts with the
 * entries measured on an MI355X by blasx_amd/tune.py (tune_table.inc, built
 * in), then those of $BLASX_TUNE_FILE (one entry per line, "dtype M N K T"),
 * then those added at run time; the nearest (log-size) entry of the same
 * dtype wins, the later one on a tie.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "internal.h"

#define MAX_ENTRIES 256
static struct {
    int dt, M, N, K, T;
} g_tab[MAX_ENTRIES];
static int g_n = -1;
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;

/* measured defaults: {dtype, M, N, K, T} */
static const int g_builtin[][5] = {
#include "tune_table.inc"
    {-1, 0, 0, 0, 0}};

static void load_table(void)
{
    g_n = 0;
    for (size_t i = 0; i < sizeof g_builtin / sizeof g_builtin[0] && g_n < MAX_ENTRIES; ++i) {
        if (g_builtin[i][0] < 0) continue;
        g_tab[g_n].dt = g_builtin[i][0];
        g_tab[g_n].M = g_builtin[i][1];
        g_tab[g_n].N = g_builtin[i][2];
        g_tab[g_n].K = g_builtin[i][3];
        g_tab[g_n].T = g_builtin[i][4];
        ++g_n;
    }
    const char *path = getenv("BLASX_TUNE_FILE");
    if (!path || !*path) return;
    FILE *f = fopen(path, "r");
    if (!f) return;
    while (g_n < MAX_ENTRIES && fscanf(f, "%d %d %d %d %d", &g_tab[g_n].dt, &g_tab[g_n].M,
                                       &g_tab[g_n].N, &g_tab[g_n].K, &g_tab[g_n].T) == 5)
        ++g_n;
    fclose(f);
}

int bx_tune_lookup(int dt, int M, int N, int K)
{
    pthread_mutex_lock(&g_lock);
    if (g_n < 0) load_table();
    int best = 0;
    double bd = 1e30;
    for (int i = 0; i < g_n; ++i) {
        if (g_tab[i].dt != dt) continue;
        double d = fabs(log((double)M / g_tab[i].M)) + fabs(log((double)N / g_tab[i].N)) +
                   fabs(log((double)K / g_tab[i].K));
        if (d <= bd) { /* later entries win ties: run-time over file over built-in */
            bd = d;
            best = g_tab[i].T;
        }
    }
    pthread_mutex_unlock(&g_lock);
    return bd < 1.0 ? best : 0; /* only near matches (within ~e in size) */
}

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + t.tv_nsec * 1e-9;
}

/* Measure the host-operand runtime for each candidate tile edge on
 * synthetic pinned operands of the given shape (NN, alpha = beta = 1), keep
 * the fastest in the table (and append it to $BLASX_TUNE_FILE if set).
 * times_out[i] receives the seconds of candidate 256 << i (0 if skipped). */
BX_EXPORT int blasx_autotune_tile_size(int dtype, int M, int N, int K, int reps, int *best_T,
                                       double times_out[4])
{
    if (dtype != BLASX_S && dtype != BLASX_D && dtype != BLASX_Z) return BLASX_EINVAL;
    if (M <= 0 || N <= 0 || K <= 0) return BLASX_EINVAL;
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    if (reps < 1) reps = 2;
    const size_t es = bx_esize(dtype);
    void *A = NULL, *B = NULL, *C = NULL;
    int rc = bxh_host_malloc(&A, (size_t)M * K * es);
    if (!rc) rc = bxh_host_malloc(&B, (size_t)K * N * es);
    if (!rc) rc = bxh_host_malloc(&C, (size_t)M * N * es);
    if (!rc) rc = blasx_fill_uniform_host(dtype, A, M, K, M, 1, 0);
    if (!rc) rc = blasx_fill_uniform_host(dtype, B, K, N, K, 1, 1);
    if (!rc) rc = blasx_fill_uniform_host(dtype, C, M, N, M, 1, 2);
    const int saved = g_bx.tile_override[dtype];
    double best = 1e30;
    int bestT = 0;
    const double one[2] = {1.0, 0.0};
    for (int i = 0; i < 4 && !rc; ++i) {
        const int T = 256 << i;
        if (times_out) times_out[i] = 0.0;
        if ((size_t)T * T * es > (size_t)2048 * 2048 * 8) continue; /* larger than a cache slot */
        g_bx.tile_override[dtype] = T;
        rc = bx_tile_gemm(dtype, BX_OP_N, BX_OP_N, M, N, K, one, A, M, B, K, one, C, M); /* warm-up */
        double t0 = now_s();
        for (int r = 0; r < reps && !rc; ++r)
            rc = bx_tile_gemm(dtype, BX_OP_N, BX_OP_N, M, N, K, one, A, M, B, K, one, C, M);
        const double t = (now_s() - t0) / reps;
        if (times_out) times_out[i] = t;
        if (!rc && t < best) {
            best = t;
            bestT = T;
        }
    }
    g_bx.tile_override[dtype] = saved;
    bxh_host_free(A);
    bxh_host_free(B);
    bxh_host_free(C);
    if (rc) return rc;
    if (best_T) *best_T = bestT;
    blasx_tune_table_add(dtype, M, N, K, bestT);
    const char *path = getenv("BLASX_TUNE_FILE");
    if (path && *path) {
        FILE *f = fopen(path, "a");
        if (f) {
            fprintf(f, "%d %d %d %d %d\n", dtype, M, N, K, bestT);
            fclose(f);
        }
    }
    return 0;
}

BX_EXPORT int blasx_tune_table_add(int dtype, int M, int N, int K, int T)
{
    pthread_mutex_lock(&g_lock);
    if (g_n < 0) load_table();
    int rc = BLASX_ERANGE;
    if (g_n < MAX_ENTRIES) {
        g_tab[g_n].dt = dtype;
        g_tab[g_n].M = M;
        g_tab[g_n].N = N;
        g_tab[g_n].K = K;
        g_tab[g_n].T = T;
        ++g_n;
        rc = 0;
    }
    pthread_mutex_unlock(&g_lock);
    return rc;
}
