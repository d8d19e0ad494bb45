/*
 * core.c -- library state: lazy GPU initialisation, configuration,
 * logging, statistics, the residency registry and the extension API's
 * memory and device-GEMM entry points.
 *
 * Replaces the reference's lazy init and resource creation
 * (blasx_init / blasx_tile_init / blasx_resource_init,
 * blas/blasx_internal.c:51-100, global resources blas/blasx_tile_resource.c:10-26,
 * init mutex blas/blasx_mutex.c:3-37).  Differences by design: one
 * process-wide set of three streams (copy-in, compute, copy-out) on one GPU;
 * no GPU is not an assert (blas/blasx_internal.c:82) but a reported state.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "internal.h"

bx_ctx_t g_bx = {.gpu_lock = PTHREAD_MUTEX_INITIALIZER};
bx_stats_atomic_t g_stats = {.flops_lock = PTHREAD_MUTEX_INITIALIZER};

static pthread_once_t g_once = PTHREAD_ONCE_INIT;
static int g_init_rc = BLASX_ENOGPU;

static long env_long(const char *name, long dflt)
{
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    char *end = NULL;
    long r = strtol(v, &end, 10);
    if (end == v) return dflt;
    /* accept K/M/G suffixes for byte sizes */
    if (*end == 'k' || *end == 'K') r <<= 10;
    else if (*end == 'm' || *end == 'M') r <<= 20;
    else if (*end == 'g' || *end == 'G') r <<= 30;
    return r;
}

void bx_log(int level, const char *fmt, ...)
{
    if (level > g_bx.log_level) return;
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "[blasx] ");
    vfprintf(stderr, fmt, ap);
    fputc('\n', stderr);
    va_end(ap);
}

void bx_fatal(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "[blasx] FATAL: ");
    vfprintf(stderr, fmt, ap);
    fputc('\n', stderr);
    va_end(ap);
    fflush(stderr);
    abort();
}

/* A GPU-sized call whose GPU path failed with rc.  Returns 1 when the caller
 * is to run the CPU BLAS instead: the failure is logged (always printed) and
 * counted in cpu_fallbacks.  That is the default policy.  It cannot apply when
 * an operand is in HBM (the CPU BLAS cannot read it: host_operands == 0) or
 * when BLASX_CPU_FALLBACK=0 asks for the strict policy; both end the process
 * with a message naming the routine and the failing HIP call. */
int bx_gpu_failed(const char *name, int rc, int host_operands)
{
    if (host_operands && g_bx.cpu_fallback) {
        bx_log(0, "%s: GPU path failed (code %d: %s); running on the CPU BLAS", name, rc,
               bxh_last_error());
        BX_STAT_ADD(cpu_fallbacks, 1);
        return 1;
    }
    if (!host_operands)
        bx_fatal("%s: GPU path failed (code %d: %s) on operands in HBM, which the CPU BLAS "
                 "cannot read",
                 name, rc, bxh_last_error());
    bx_fatal("%s: GPU path failed (code %d: %s) and BLASX_CPU_FALLBACK=0 forbids the CPU BLAS",
             name, rc, bxh_last_error());
}

/* 1 if p is device memory (HBM) of the GPU this library drives; only asked
 * once that GPU is up, so a machine without one sees host memory only */
int bx_in_hbm(const void *p)
{
    return p != NULL && bx_ensure_init() == 0 && bxh_pointer_is_pinned(p) == 2;
}

static void read_config(void)
{
    g_bx.log_level = (int)env_long("BLASX_LOG", 0);
    g_bx.min_dim_real = (int)env_long("BLASX_GPU_MIN_DIM", 1000);
    g_bx.min_dim_cplx = (int)env_long("BLASX_GPU_MIN_DIM_Z", 500);
    g_bx.cpu_fallback = (int)env_long("BLASX_CPU_FALLBACK", 1);
    long t = env_long("BLASX_TILE", 0);
    for (int i = 0; i < 4; ++i) g_bx.tile_override[i] = (int)t;
    g_bx.hbm_budget = (size_t)env_long("BLASX_HBM_BUDGET", 0);
    g_bx.stage_threads = (int)env_long("BLASX_STAGE_THREADS", 8);
}

static pthread_once_t g_cfg_once = PTHREAD_ONCE_INIT;

void bx_config(void) { pthread_once(&g_cfg_once, read_config); }

static void do_init(void)
{
    bx_config();
    int n = 0;
    if (bxh_device_count(&n) != 0 || n <= 0) {
        g_init_rc = BLASX_ENOGPU;
        bx_log(1, "no GPU visible: %s", bxh_last_error());
        return;
    }
    g_bx.device = (int)env_long("BLASX_DEVICE", 0);
    if (g_bx.device >= n) g_bx.device = 0;
    if (bxh_set_device(g_bx.device) != 0) {
        g_init_rc = BLASX_ENOGPU;
        return;
    }
    bxh_device_info(g_bx.device, g_bx.arch, sizeof g_bx.arch, &g_bx.cus, &g_bx.hbm_total);
    if (bxh_stream_create(&g_bx.s_comp) || bxh_stream_create(&g_bx.s_h2d) ||
        bxh_stream_create(&g_bx.s_d2h)) {
        g_init_rc = BLASX_EHIP;
        bx_log(0, "stream creation failed: %s", bxh_last_error());
        return;
    }
    /* the allocation-free grouped entry point's table store on this GPU */
    if (bxh_noalloc_setup() != 0) bx_log(0, "noalloc table store: %s", bxh_noalloc_error());
    if (g_bx.hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        bxh_mem_get_info(&fr, &tot);
        g_bx.hbm_budget = (size_t)(fr * 0.80);
    }
    g_bx.gpu_ok = 1;
    g_init_rc = 0;
    bx_log(1, "GPU %d (%s, %d CUs, %zu MiB HBM), tile-cache budget %zu MiB", g_bx.device,
           g_bx.arch, g_bx.cus, g_bx.hbm_total >> 20, g_bx.hbm_budget >> 20);
}

int bx_ensure_init(void)
{
    pthread_once(&g_once, do_init);
    if (g_init_rc == 0) bxh_set_device(g_bx.device); /* device is per host thread */
    return g_init_rc;
}

BX_EXPORT int blasx_init(void) { return bx_ensure_init(); }

BX_EXPORT void blasx_finalize(void)
{
    if (!g_bx.gpu_ok) return;
    pthread_mutex_lock(&g_bx.gpu_lock);
    bxh_device_sync();
    bx_cache_t *c = bx_global_cache();
    if (c) bx_cache_free_memory(c);
    pthread_mutex_unlock(&g_bx.gpu_lock);
}

BX_EXPORT int blasx_gpu_available(void) { return bx_ensure_init() == 0; }

BX_EXPORT const char *blasx_version(void) { return "blasx-mi355x 0.1 (gfx950)"; }

/* ------------------------------------------------------------ statistics */
void bx_stat_flops(double f)
{
    pthread_mutex_lock(&g_stats.flops_lock);
    g_stats.flops += f;
    pthread_mutex_unlock(&g_stats.flops_lock);
}

BX_EXPORT void blasx_get_stats(blasx_stats_t *o)
{
    if (!o) return;
    o->calls_gpu = __atomic_load_n(&g_stats.calls_gpu, __ATOMIC_RELAXED);
    o->calls_cpu = __atomic_load_n(&g_stats.calls_cpu, __ATOMIC_RELAXED);
    o->cpu_fallbacks = __atomic_load_n(&g_stats.cpu_fallbacks, __ATOMIC_RELAXED);
    o->kernel_launches = __atomic_load_n(&g_stats.kernel_launches, __ATOMIC_RELAXED);
    o->bytes_h2d = __atomic_load_n(&g_stats.bytes_h2d, __ATOMIC_RELAXED);
    o->bytes_d2h = __atomic_load_n(&g_stats.bytes_d2h, __ATOMIC_RELAXED);
    o->bytes_lower_bound = __atomic_load_n(&g_stats.bytes_lower_bound, __ATOMIC_RELAXED);
    o->tile_hits = __atomic_load_n(&g_stats.tile_hits, __ATOMIC_RELAXED);
    o->tile_misses = __atomic_load_n(&g_stats.tile_misses, __ATOMIC_RELAXED);
    o->tile_evictions = __atomic_load_n(&g_stats.tile_evictions, __ATOMIC_RELAXED);
    o->tiles_generated = __atomic_load_n(&g_stats.tiles_generated, __ATOMIC_RELAXED);
    pthread_mutex_lock(&g_stats.flops_lock);
    o->flops = g_stats.flops;
    pthread_mutex_unlock(&g_stats.flops_lock);
}

BX_EXPORT void blasx_reset_stats(void)
{
    pthread_mutex_lock(&g_stats.flops_lock);
    g_stats.calls_gpu = g_stats.calls_cpu = g_stats.cpu_fallbacks = 0;
    g_stats.kernel_launches = g_stats.bytes_h2d = g_stats.bytes_d2h = 0;
    g_stats.bytes_lower_bound = g_stats.tile_hits = g_stats.tile_misses = 0;
    g_stats.tile_evictions = g_stats.tiles_generated = 0;
    g_stats.flops = 0;
    pthread_mutex_unlock(&g_stats.flops_lock);
}

/* ------------------------------------------------------------ configuration */
BX_EXPORT int blasx_set_hbm_budget(size_t bytes)
{
    if (bx_ensure_init() != 0) {
        g_bx.hbm_budget = bytes;
        return BLASX_ENOGPU;
    }
    pthread_mutex_lock(&g_bx.gpu_lock);
    g_bx.hbm_budget = bytes;
    bx_cache_t *c = bx_global_cache();
    int rc = 0;
    if (c) {
        bxh_device_sync();
        rc = bx_cache_set_budget(c, bytes);
    }
    pthread_mutex_unlock(&g_bx.gpu_lock);
    return rc;
}

BX_EXPORT size_t blasx_get_hbm_budget(void)
{
    bx_ensure_init();
    return g_bx.hbm_budget;
}

BX_EXPORT int blasx_set_tile_size(int dtype, int tile)
{
    if (dtype < 0 || dtype > 3) return BLASX_EINVAL;
    if (tile != 0 && tile != 256 && tile != 512 && tile != 1024 && tile != 2048)
        return BLASX_EINVAL;
    if (tile * (int64_t)tile * bx_esize(dtype) > (int64_t)2048 * 2048 * 8) return BLASX_ERANGE;
    bx_ensure_init();
    g_bx.tile_override[dtype] = tile;
    return 0;
}

BX_EXPORT int blasx_get_tile_size(int dtype, int M, int N, int K)
{
    if (dtype < 0 || dtype > 3) return BLASX_EINVAL;
    bx_ensure_init();
    if (g_bx.tile_override[dtype]) return g_bx.tile_override[dtype];
    return bx_autotune_tile(dtype, M, N, K);
}

/* ------------------------------------------------------------ residency */
#define BX_MAX_RESIDENT 64
static struct {
    uintptr_t begin, end;
} g_res[BX_MAX_RESIDENT];
static int g_nres;
static pthread_mutex_t g_res_lock = PTHREAD_MUTEX_INITIALIZER;

int bx_resident_contains(uintptr_t begin, uintptr_t end)
{
    int r = 0;
    pthread_mutex_lock(&g_res_lock);
    for (int i = 0; i < g_nres; ++i)
        if (begin >= g_res[i].begin && end <= g_res[i].end) {
            r = 1;
            break;
        }
    pthread_mutex_unlock(&g_res_lock);
    return r;
}

BX_EXPORT int blasx_resident_register(const void *host_ptr, size_t bytes)
{
    if (!host_ptr || !bytes) return BLASX_EINVAL;
    pthread_mutex_lock(&g_res_lock);
    if (g_nres >= BX_MAX_RESIDENT) {
        pthread_mutex_unlock(&g_res_lock);
        return BLASX_ERANGE;
    }
    g_res[g_nres].begin = (uintptr_t)host_ptr;
    g_res[g_nres].end = (uintptr_t)host_ptr + bytes;
    ++g_nres;
    pthread_mutex_unlock(&g_res_lock);
    return 0;
}

static int keep_none(const bx_tile_key_t *k, void *u)
{
    (void)k;
    (void)u;
    return 0;
}

static int keep_outside(const bx_tile_key_t *k, void *u)
{
    /* keep tiles that do NOT overlap the range being unregistered */
    const uintptr_t *r = (const uintptr_t *)u;
    uintptr_t b = k->addr;
    uintptr_t e = k->addr + (uintptr_t)((k->cols - 1) * k->ld + k->rows) * k->esize;
    return e <= r[0] || b >= r[1];
}

BX_EXPORT int blasx_resident_unregister(const void *host_ptr)
{
    uintptr_t range[2] = {0, 0};
    int found = 0;
    pthread_mutex_lock(&g_res_lock);
    for (int i = 0; i < g_nres; ++i)
        if (g_res[i].begin == (uintptr_t)host_ptr) {
            range[0] = g_res[i].begin;
            range[1] = g_res[i].end;
            g_res[i] = g_res[--g_nres];
            found = 1;
            break;
        }
    pthread_mutex_unlock(&g_res_lock);
    if (!found) return BLASX_EINVAL;
    if (bx_ensure_init() == 0) {
        pthread_mutex_lock(&g_bx.gpu_lock);
        bx_cache_t *c = bx_global_cache();
        if (c) bx_cache_invalidate(c, keep_outside, range);
        pthread_mutex_unlock(&g_bx.gpu_lock);
    }
    return 0;
}

BX_EXPORT void blasx_cache_flush(void)
{
    if (bx_ensure_init() != 0) return;
    pthread_mutex_lock(&g_bx.gpu_lock);
    bx_cache_t *c = bx_global_cache();
    if (c) bx_cache_invalidate(c, keep_none, NULL);
    pthread_mutex_unlock(&g_bx.gpu_lock);
}

/* ------------------------------------------------------------ memory */
BX_EXPORT void *blasx_host_alloc(size_t bytes)
{
    if (bx_ensure_init() != 0) return NULL;
    void *p = NULL;
    if (bxh_host_malloc(&p, bytes) != 0) return NULL;
    return p;
}

BX_EXPORT void blasx_host_free(void *p)
{
    if (p && bx_ensure_init() == 0) bxh_host_free(p);
}

BX_EXPORT void *blasx_dev_alloc(size_t bytes)
{
    if (bx_ensure_init() != 0) return NULL;
    void *p = NULL;
    if (bxh_malloc(&p, bytes) != 0) return NULL;
    return p;
}

BX_EXPORT void blasx_dev_free(void *p)
{
    if (p && bx_ensure_init() == 0) bxh_free(p);
}

BX_EXPORT int blasx_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    return bxh_memcpy_h2d(dst, src, bytes);
}

BX_EXPORT int blasx_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    return bxh_memcpy_d2h(dst, src, bytes);
}

BX_EXPORT int blasx_memset_device(void *dst, int value, size_t bytes)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    int rc = bxh_memset_async(dst, value, bytes, g_bx.s_comp);
    return rc ? rc : bxh_stream_sync(g_bx.s_comp);
}

BX_EXPORT int blasx_device_copy(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    return bxh_copy_kernel(dst, src, bytes, stream ? stream : g_bx.s_comp);
}

int bxh_measure_peaks(int cus, size_t copy_bytes, double *out, void *stream);

BX_EXPORT int blasx_measure_peaks(double *out)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    return bxh_measure_peaks(g_bx.cus > 0 ? g_bx.cus : 256, (size_t)4 << 30, out, g_bx.s_comp);
}

BX_EXPORT int blasx_device_synchronize(void)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    return bxh_device_sync();
}

BX_EXPORT void *blasx_compute_stream(void)
{
    if (bx_ensure_init() != 0) return NULL;
    return g_bx.s_comp;
}

BX_EXPORT void *blasx_event_create(void)
{
    if (bx_ensure_init() != 0) return NULL;
    void *e = NULL;
    return bxh_event_create(&e, 1) == 0 ? e : NULL;
}

BX_EXPORT void blasx_event_destroy(void *ev)
{
    if (ev) bxh_event_destroy(ev);
}

BX_EXPORT int blasx_event_record(void *ev, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    return bxh_event_record(ev, stream ? stream : g_bx.s_comp);
}

BX_EXPORT int blasx_event_synchronize(void *ev) { return bxh_event_sync(ev); }

BX_EXPORT int blasx_event_elapsed_ms(float *ms, void *start, void *stop)
{
    return bxh_event_elapsed_ms(ms, start, stop);
}

/* ------------------------------------------------------------ device GEMM */
static int trans_code(int t)
{
    /* CBLAS enum values, include/cblas.h:12 of the reference */
    return t == 111 ? BX_OP_N : t == 112 ? BX_OP_T : t == 113 ? BX_OP_C : -1;
}

int bx_device_gemm(int dt, int transa, int transb, int M, int N, int K, const double *alpha,
                   const void *A, int64_t lda, const void *B, int64_t ldb, const double *beta,
                   void *C, int64_t ldc, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    int opa = trans_code(transa), opb = trans_code(transb);
    if (opa < 0 || opb < 0 || M < 0 || N < 0 || K < 0) return BLASX_EINVAL;
    if (dt != BLASX_Z) {
        if (opa == BX_OP_C) opa = BX_OP_T;
        if (opb == BX_OP_C) opb = BX_OP_T;
    }
    int64_t nrowa = opa == BX_OP_N ? M : K, nrowb = opb == BX_OP_N ? K : N;
    if (lda < (nrowa > 1 ? nrowa : 1) || ldb < (nrowb > 1 ? nrowb : 1) || ldc < (M > 1 ? M : 1))
        return BLASX_EINVAL;
    if (M == 0 || N == 0) return 0;
    void *st = stream ? stream : g_bx.s_comp;
    int alpha_zero = alpha[0] == 0.0 && (dt != BLASX_Z || alpha[1] == 0.0);
    if (K == 0 || alpha_zero) {
        return bxh_scale(dt, M, N, beta, C, ldc, st);
    }
    int rc = bxh_gemm(dt, opa, opb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st);
    if (rc == 0) {
        BX_STAT_ADD(kernel_launches, 1);
        bx_stat_flops((dt == BLASX_Z ? 8.0 : 2.0) * M * (double)N * K);
    }
    return rc;
}

BX_EXPORT int blasx_dgemm_device(int transa, int transb, int M, int N, int K, double alpha,
                                 const double *dA, int64_t lda, const double *dB, int64_t ldb,
                                 double beta, double *dC, int64_t ldc, void *stream)
{
    double a[2] = {alpha, 0}, b[2] = {beta, 0};
    return bx_device_gemm(BLASX_D, transa, transb, M, N, K, a, dA, lda, dB, ldb, b, dC, ldc,
                          stream);
}

BX_EXPORT int blasx_sgemm_device(int transa, int transb, int M, int N, int K, float alpha,
                                 const float *dA, int64_t lda, const float *dB, int64_t ldb,
                                 float beta, float *dC, int64_t ldc, void *stream)
{
    double a[2] = {alpha, 0}, b[2] = {beta, 0};
    return bx_device_gemm(BLASX_S, transa, transb, M, N, K, a, dA, lda, dB, ldb, b, dC, ldc,
                          stream);
}

BX_EXPORT int blasx_zgemm_device(int transa, int transb, int M, int N, int K,
                                 const double *alpha, const double *dA, int64_t lda,
                                 const double *dB, int64_t ldb, const double *beta, double *dC,
                                 int64_t ldc, void *stream)
{
    return bx_device_gemm(BLASX_Z, transa, transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC,
                          ldc, stream);
}

/* ------------------------------------------------------------ seeded inputs */
static inline uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline double u11(uint64_t key)
{
    return ((double)(splitmix64(key) >> 11) * 0x1.0p-53) * 2.0 - 1.0;
}

typedef struct fill_job {
    int dt;
    void *p;
    int64_t rows, cols, ld;
    uint64_t base;
    int64_t c0, c1;
} fill_job_t;

static void *fill_worker(void *arg)
{
    fill_job_t *j = (fill_job_t *)arg;
    for (int64_t c = j->c0; c < j->c1; ++c)
        for (int64_t r = 0; r < j->rows; ++r) {
            uint64_t e = (uint64_t)(r + c * j->rows);
            int64_t o = r + c * j->ld;
            if (j->dt == BLASX_D)
                ((double *)j->p)[o] = u11(j->base ^ e);
            else if (j->dt == BLASX_S)
                ((float *)j->p)[o] = (float)u11(j->base ^ e);
            else {
                ((double *)j->p)[2 * o] = u11(j->base ^ (2 * e));
                ((double *)j->p)[2 * o + 1] = u11(j->base ^ (2 * e + 1));
            }
        }
    return NULL;
}

BX_EXPORT int blasx_fill_uniform_host(int dtype, void *hptr, int64_t rows, int64_t cols,
                                      int64_t ld, uint64_t seed, uint32_t id)
{
    if (dtype != BLASX_S && dtype != BLASX_D && dtype != BLASX_Z) return BLASX_EINVAL;
    if (rows < 0 || cols < 0 || ld < rows) return BLASX_EINVAL;
    long nthr = sysconf(_SC_NPROCESSORS_ONLN);
    if (nthr < 1) nthr = 1;
    if (nthr > 64) nthr = 64;
    if (rows * cols < (1 << 20)) nthr = 1;
    pthread_t th[64];
    fill_job_t jobs[64];
    uint64_t base = seed ^ ((uint64_t)id << 56);
    for (long t = 0; t < nthr; ++t) {
        jobs[t] = (fill_job_t){dtype, hptr, rows, cols, ld, base, cols * t / nthr,
                               cols * (t + 1) / nthr};
        if (nthr == 1)
            fill_worker(&jobs[t]);
        else
            pthread_create(&th[t], NULL, fill_worker, &jobs[t]);
    }
    if (nthr > 1)
        for (long t = 0; t < nthr; ++t) pthread_join(th[t], NULL);
    return 0;
}

BX_EXPORT int blasx_fill_uniform_device(int dtype, void *dptr, int64_t rows, int64_t cols,
                                        int64_t ld, uint64_t seed, uint32_t id, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    if (rows < 0 || cols < 0 || ld < rows) return BLASX_EINVAL;
    return bxh_fill_uniform(dtype, dptr, rows, cols, ld, seed, id, stream ? stream : g_bx.s_comp);
}
