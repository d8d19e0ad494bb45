/*
 * internal.h -- declarations shared by the C host runtime of libblasx.so.
 */
#ifndef BLASX_INTERNAL_H
#define BLASX_INTERNAL_H

#include <pthread.h>
#include <stddef.h>
#include <stdint.h>

#include "blasx_ext.h"
#include "hiplayer.h"

#define BX_EXPORT __attribute__((visibility("default")))

/* operand codes used below the ABI: 0 = N, 1 = T, 2 = C */
enum { BX_OP_N = 0, BX_OP_T = 1, BX_OP_C = 2 };

static inline int bx_esize(int dt)
{
    return dt == BLASX_S ? 4 : dt == BLASX_D ? 8 : dt == BLASX_C ? 8 : 16;
}

/* ---------------- global context (core.c) ---------------- */
typedef struct bx_ctx {
    int gpu_ok;              /* a GPU was found and streams exist */
    int device;
    int cus;
    char arch[64];
    size_t hbm_total;
    void *s_comp;            /* compute stream */
    void *s_h2d;             /* host -> device copy stream */
    void *s_d2h;             /* device -> host copy stream */
    pthread_mutex_t gpu_lock;   /* one host-operand GEMM at a time owns the streams */
    /* configuration */
    size_t hbm_budget;       /* bytes the tile cache may hold */
    int tile_override[4];    /* per dtype, 0 = autotune */
    int min_dim_real;        /* dispatch threshold, S and D */
    int min_dim_cplx;        /* dispatch threshold, Z */
    int cpu_fallback;        /* BLASX_CPU_FALLBACK (default 1: log, then the CPU BLAS) */
    int log_level;           /* BLASX_LOG */
    int stage_threads;       /* host threads for pageable staging copies */
} bx_ctx_t;

extern bx_ctx_t g_bx;

void bx_config(void);                /* read the environment once */
int bx_ensure_init(void);            /* returns 0 if the GPU path is usable */
void bx_log(int level, const char *fmt, ...);
void bx_fatal(const char *fmt, ...) __attribute__((noreturn));
/* failure policy of a GPU-sized call: 1 = run the CPU BLAS (logged, counted),
 * else the process ends with a message (operands in HBM, or strict policy) */
int bx_gpu_failed(const char *name, int rc, int host_operands);
int bx_in_hbm(const void *p);        /* device memory of the library's GPU */

/* ---------------- statistics (core.c) ---------------- */
typedef struct bx_stats_atomic {
    uint64_t calls_gpu, calls_cpu, cpu_fallbacks, kernel_launches;
    uint64_t bytes_h2d, bytes_d2h, bytes_lower_bound;
    uint64_t tile_hits, tile_misses, tile_evictions, tiles_generated;
    double flops;
    pthread_mutex_t flops_lock;
} bx_stats_atomic_t;
extern bx_stats_atomic_t g_stats;
#define BX_STAT_ADD(field, v) __atomic_fetch_add(&g_stats.field, (uint64_t)(v), __ATOMIC_RELAXED)
void bx_stat_flops(double f);

/* ---------------- CPU BLAS (cpublas.c) ---------------- */
int bx_cpublas_load(void);
void *bx_cpublas_sym(const char *lp64_name);
int bx_cpublas_ilp64(void);
int bx_cpu_gemm(int dt, int order, int ta, int tb, int M, int N, int K, const void *alpha,
                const void *A, int lda, const void *B, int ldb, const void *beta, void *C,
                int ldc);

/* ---------------- tile cache (tile_cache.c) ---------------- */
typedef struct bx_tile_key {
    uintptr_t addr;      /* host address of the tile's first element */
    int64_t ld;          /* host leading dimension (elements) */
    int32_t rows, cols;  /* stored extents */
    int32_t esize;
    int32_t pad_;
} bx_tile_key_t;

typedef struct bx_cache bx_cache_t;

typedef struct bx_cache_ops {
    /* allocate one chunk of `bytes` device memory; 0 on success */
    int (*chunk_alloc)(void *user, void **p, size_t bytes);
    void (*chunk_free)(void *user, void *p);
    void *user;
} bx_cache_ops_t;

bx_cache_t *bx_cache_create(size_t slot_bytes, size_t chunk_bytes, size_t budget_bytes,
                            const bx_cache_ops_t *ops);
void bx_cache_destroy(bx_cache_t *c);
size_t bx_cache_slot_bytes(const bx_cache_t *c);
int bx_cache_capacity_slots(const bx_cache_t *c);   /* budget / slot_bytes */
int bx_cache_live_slots(const bx_cache_t *c);       /* slots holding a valid tile */
int bx_cache_set_budget(bx_cache_t *c, size_t budget_bytes);
/* Look a tile up.  On a hit returns the slot index with *hit = 1.  On a miss
 * assigns a slot (free, newly grown, or evicted least-recently-used among
 * slots whose last use is older than `round`), installs the key and returns
 * it with *hit = 0; *evicted_round gets the last round of the evicted tile
 * (-1 if none).  Returns -1 (BLASX_EBUSY) if every slot is in use by `round`
 * and the budget is exhausted. */
int bx_cache_acquire(bx_cache_t *c, const bx_tile_key_t *key, int64_t round, int *hit,
                     int64_t *evicted_round);
void *bx_cache_slot_ptr(const bx_cache_t *c, int slot);
int bx_cache_lookup(const bx_cache_t *c, const bx_tile_key_t *key);   /* -1 if absent */
/* Scratch slots: taken out of the cache for the duration of a call */
int bx_cache_take_scratch(bx_cache_t *c, int64_t round);
void bx_cache_return_scratch(bx_cache_t *c, int slot);
/* Drop entries whose host range is not inside a resident range. */
typedef int (*bx_keep_fn)(const bx_tile_key_t *key, void *user);
void bx_cache_invalidate(bx_cache_t *c, bx_keep_fn keep, void *user);
void bx_cache_invalidate_all(bx_cache_t *c);
void bx_cache_reset_rounds(bx_cache_t *c);   /* forget round numbers of kept tiles */
int bx_cache_validate(const bx_cache_t *c);  /* 0 if the structural invariants hold */
uintptr_t bx_cache_slot_addr_key(const bx_cache_t *c, int slot);
void bx_cache_free_memory(bx_cache_t *c);   /* release every chunk (cache must be idle) */
/* statistics of the cache itself */
void bx_cache_counters(const bx_cache_t *c, uint64_t *hits, uint64_t *misses, uint64_t *evictions);

/* ---------------- residency registry (core.c) ---------------- */
int bx_resident_contains(uintptr_t begin, uintptr_t end);

/* ---------------- runtimes ---------------- */
/* host-operand tiled GEMM (tile_rt.c); alpha/beta point to (re[, im]) doubles */
int bx_tile_gemm(int dt, int opa, int opb, int M, int N, int K, const double *alpha,
                 const void *A, int64_t lda, const void *B, int64_t ldb, const double *beta,
                 void *C, int64_t ldc);
int bx_tile_syrk(int dt, int tri, int opa, int N, int K, const double *alpha, const void *A,
                 int64_t lda, const double *beta, void *C, int64_t ldc);
int bx_device_syrk(int dt, int tri, int opa, int N, int K, const double *alpha, const void *A,
                   int64_t lda, const void *B, int64_t ldb, const double *beta, void *C,
                   int64_t ldc, void *stream);
int bx_device_syr2k(int dt, int tri, int opa, int N, int K, const double *alpha, const void *A,
                    int64_t lda, const void *B, int64_t ldb, const void *B2, int64_t ldb2,
                    const double *beta, void *C, int64_t ldc, void *stream);
bx_cache_t *bx_global_cache(void);
/* a GEMM run in core on HBM: host operands copied whole for the call,
 * operands in HBM used in place (level3.c); complex64 always comes here */
int bx_gpu_gemm_incore(int dt, int transa, int transb, int M, int N, int K, const double *al,
                       const void *A, int lda, const void *B, int ldb, const double *be, void *C,
                       int ldc);
int bx_gpu_crank(int herm, int uplo, int trans, int N, int K, const double *al, const void *A,
                 int lda, const void *B, int ldb, const double *be, void *C, int ldc);
/* C = alpha op(A) op(B) + beta C on HBM operands (transa, transb: CBLAS enums) */
int bx_device_gemm(int dt, int transa, int transb, int M, int N, int K, const double *alpha,
                   const void *A, int64_t lda, const void *B, int64_t ldb, const double *beta,
                   void *C, int64_t ldc, void *stream);
int bx_autotune_tile(int dt, int M, int N, int K);

/* grouped GEMM tables (grouped.c) */
int bx_desc_check(const blasx_gemm_desc_t *d);
int bx_grouped_fill(const blasx_gemm_desc_t *d, int n, int dtype, bxh_problem_t *probs,
                    bxh_seg_t *segs, int *nprob);
/* block tile shape (BXH_SHAPE_*) of a grouped problem in a launch of `load`
 * large tiles of its element type (autotune.c) */
int bx_grouped_shape(int dt, int M, int N, int K, int64_t load);

/* task queue (queue.c): ids task_num-1 .. 0 exactly once across threads */
typedef struct bx_queue {
    volatile int64_t tail;
} bx_queue_t;
void bx_queue_init(bx_queue_t *q, int64_t task_num);
int64_t bx_queue_next(bx_queue_t *q);

#endif
