/*
 * autotune.c -- tile-size table for the host-operand runtime.
 *
 * The reference fixes the tile edge at compile time (BLOCKSIZE_DGEMM 2048,
 * BLOCKSIZE_ZGEMM 1024, include/blasx.h:5-7).  Here the edge is chosen per
 * call from a table of measured winners (written by blasx_autotune_tiles)
 * and otherwise by the heuristic in tile_rt.c.  Table file: one entry per
 * line, "dtype M N K T", nearest (log-size) entry of the same dtype wins.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>

#include "internal.h"

#define MAX_ENTRIES 256
static struct {
    int dt, M, N, K, T;
} g_tab[MAX_ENTRIES];
static int g_n = -1;
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;

static void load_table(void)
{
    g_n = 0;
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
        if (d < bd) {
            bd = d;
            best = g_tab[i].T;
        }
    }
    pthread_mutex_unlock(&g_lock);
    return bd < 1.0 ? best : 0; /* only near matches (within ~e in size) */
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
