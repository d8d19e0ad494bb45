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

/* ---- block tile shapes of the grouped launches ------------------------------
 * A grouped launch runs each real problem on 128 x 128, 128 x 64 or 64 x 64
 * block tiles (BXH_SHAPE_*); complex problems keep 128 x 64.  Smaller tiles
 * give a launch with few tiles more blocks to spread over the CUs, at a
 * lower rate per tile.  The table holds measured launch times per shape for
 * workloads of equal problems (tune_grouped.inc, written by
 * blasx_amd/tune.py --grouped on an MI355X; entries added at run time carry
 * only a winner).  A query takes the workload nearest in log M, N, K, then
 * the shape with the least time at the query's load (the launch's tile count
 * at the large tile, same element type), each shape's log time interpolated
 * linearly in log load between the two measured loads around it (the
 * nearest one outside them).  A workload far from every measured one, or
 * no table, goes to a heuristic on the load and the problem's shape. */
#define MAX_GENTRIES 512
typedef struct {
    int dt;
    double lm, ln, lk, ll; /* logs of M, N, K, load */
    double lt[BXH_NSHAPES]; /* logs of the launch times */
} gentry_t;
static gentry_t g_gtab[MAX_GENTRIES];
static int g_gn = -1;
static int g_gforce[3] = {-1, -1, -1};

/* measured: {dtype, M, N, K, load, t0, t1, t2} (microseconds) */
static const long long g_gbuiltin[][8] = {
#include "tune_grouped.inc"
    {-1, 0, 0, 0, 0, 0, 0, 0}};

static void gadd(int dt, double M, double N, double K, double load, const double *t)
{
    if (g_gn >= MAX_GENTRIES) return;
    gentry_t *e = &g_gtab[g_gn++];
    e->dt = dt;
    e->lm = log(M);
    e->ln = log(N);
    e->lk = log(K);
    e->ll = log(load);
    for (int s = 0; s < BXH_NSHAPES; ++s) e->lt[s] = log(t[s]);
}

static void load_gtable(void)
{
    g_gn = 0;
    for (size_t i = 0; i < sizeof g_gbuiltin / sizeof g_gbuiltin[0]; ++i) {
        const long long *r = g_gbuiltin[i];
        if (r[0] < 0 || r[1] <= 0 || r[2] <= 0 || r[3] <= 0 || r[4] <= 0) continue;
        if (r[5] <= 0 || r[6] <= 0 || r[7] <= 0) continue;
        const double t[3] = {(double)r[5], (double)r[6], (double)r[7]};
        gadd((int)r[0], (double)r[1], (double)r[2], (double)r[3], (double)r[4], t);
    }
}

static int argmin3(const double *v)
{
    int b = 0;
    for (int s = 1; s < BXH_NSHAPES; ++s)
        if (v[s] < v[b]) b = s;
    return b;
}

/* the table's choice, or -1 without an entry of this dtype; g_lock held */
static int gtable_choice(int dt, int M, int N, int K, int64_t load)
{
    const double lm = log((double)M), ln = log((double)N), lk = log((double)K),
                 ll = log((double)load);
    /* nearest workload (later entries win ties) */
    int key = -1;
    double bd = 1e30;
    for (int i = 0; i < g_gn; ++i) {
        const gentry_t *e = &g_gtab[i];
        if (e->dt != dt) continue;
        const double d = fabs(lm - e->lm) + fabs(ln - e->ln) + fabs(lk - e->lk);
        if (d <= bd) {
            bd = d;
            key = i;
        }
    }
    /* a workload far from every measured one (any dimension off by more than
     * about 4x, or several by less) is left to the heuristic */
    if (key < 0 || bd > 1.5) return -1;
    const gentry_t *k = &g_gtab[key];
    /* that workload's measured loads around the query (later entries win
     * equal loads) */
    int lo = -1, hi = -1;
    for (int i = 0; i < g_gn; ++i) {
        const gentry_t *e = &g_gtab[i];
        if (e->dt != dt || e->lm != k->lm || e->ln != k->ln || e->lk != k->lk) continue;
        if (e->ll <= ll && (lo < 0 || e->ll >= g_gtab[lo].ll)) lo = i;
        if (e->ll >= ll && (hi < 0 || e->ll <= g_gtab[hi].ll)) hi = i;
    }
    if (lo < 0) return argmin3(g_gtab[hi].lt);
    if (hi < 0 || g_gtab[hi].ll == g_gtab[lo].ll) return argmin3(g_gtab[lo].lt);
    const double f = (ll - g_gtab[lo].ll) / (g_gtab[hi].ll - g_gtab[lo].ll);
    double t[BXH_NSHAPES];
    for (int s = 0; s < BXH_NSHAPES; ++s)
        t[s] = g_gtab[lo].lt[s] + f * (g_gtab[hi].lt[s] - g_gtab[lo].lt[s]);
    return argmin3(t);
}

int bx_grouped_shape(int dt, int M, int N, int K, int64_t load)
{
    if (dt != BLASX_S && dt != BLASX_D) return BXH_SHAPE_128x128;
    if (M <= 0 || N <= 0 || K <= 0 || load <= 0) return BXH_SHAPE_128x128;
    pthread_mutex_lock(&g_lock);
    if (g_gn < 0) load_gtable();
    int shape = g_gforce[dt];
    if (shape < 0) shape = gtable_choice(dt, M, N, K, load);
    pthread_mutex_unlock(&g_lock);
    if (shape >= 0) return shape;
    /* no nearby measurement: once the launch fills every CU twice, the large
     * tile, or the narrow ones for a problem that would leave half of each
     * large tile empty; below that, smaller tiles for more blocks */
    const int64_t slots = 2 * (int64_t)(g_bx.cus > 0 ? g_bx.cus : 256);
    if (load >= slots) {
        if (M <= 64) return BXH_SHAPE_64x64;
        return N <= 64 ? BXH_SHAPE_128x64 : BXH_SHAPE_128x128;
    }
    return 2 * load >= slots ? BXH_SHAPE_128x64 : BXH_SHAPE_64x64;
}

/* force one shape for a real element type (-1: back to the table) */
BX_EXPORT int blasx_set_grouped_shape(int dtype, int shape)
{
    if (dtype != BLASX_S && dtype != BLASX_D) return BLASX_EINVAL;
    if (shape < -1 || shape >= BXH_NSHAPES) return BLASX_EINVAL;
    pthread_mutex_lock(&g_lock);
    g_gforce[dtype] = shape;
    pthread_mutex_unlock(&g_lock);
    return 0;
}

BX_EXPORT int blasx_grouped_shape(int dtype, int M, int N, int K, int64_t load)
{
    return bx_grouped_shape(dtype, M, N, K, load);
}

BX_EXPORT int blasx_grouped_shape_add(int dtype, int M, int N, int K, int64_t load, int shape)
{
    if (dtype != BLASX_S && dtype != BLASX_D) return BLASX_EINVAL;
    if (M <= 0 || N <= 0 || K <= 0 || load <= 0 || shape < 0 || shape >= BXH_NSHAPES)
        return BLASX_EINVAL;
    pthread_mutex_lock(&g_lock);
    if (g_gn < 0) load_gtable();
    const int rc = g_gn < MAX_GENTRIES ? 0 : BLASX_ERANGE;
    double t[BXH_NSHAPES];
    for (int i = 0; i < BXH_NSHAPES; ++i) t[i] = i == shape ? 1.0 : 2.0; /* a winner only */
    if (!rc) gadd(dtype, M, N, K, (double)load, t);
    pthread_mutex_unlock(&g_lock);
    return rc;
}
