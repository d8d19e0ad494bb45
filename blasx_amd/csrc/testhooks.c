/*
 * testhooks.c -- exported hooks that let the CPU test-suite drive the
 * host-side policies of the library without a GPU: the tile cache over
 * host memory and the task queue.  Not part of the BLAS ABI.
 */
#include <stdlib.h>

#include "internal.h"

static int host_chunk_alloc(void *u, void **p, size_t bytes)
{
    (void)u;
    *p = malloc(bytes ? bytes : 1);
    return *p ? 0 : BLASX_ENOMEM;
}

static void host_chunk_free(void *u, void *p)
{
    (void)u;
    free(p);
}

BX_EXPORT void *bxtest_cache_create(size_t slot_bytes, size_t chunk_bytes, size_t budget)
{
    bx_cache_ops_t ops = {host_chunk_alloc, host_chunk_free, NULL};
    return bx_cache_create(slot_bytes, chunk_bytes, budget, &ops);
}

BX_EXPORT void bxtest_cache_destroy(void *c) { bx_cache_destroy((bx_cache_t *)c); }

BX_EXPORT int bxtest_cache_acquire(void *c, uintptr_t addr, int64_t ld, int rows, int cols,
                                   int esize, int64_t round, int *hit, int64_t *evicted_round)
{
    bx_tile_key_t k = {addr, ld, rows, cols, esize, 0};
    return bx_cache_acquire((bx_cache_t *)c, &k, round, hit, evicted_round);
}

BX_EXPORT int bxtest_cache_lookup(void *c, uintptr_t addr, int64_t ld, int rows, int cols,
                                  int esize)
{
    bx_tile_key_t k = {addr, ld, rows, cols, esize, 0};
    return bx_cache_lookup((bx_cache_t *)c, &k);
}

BX_EXPORT int bxtest_cache_capacity(void *c) { return bx_cache_capacity_slots((bx_cache_t *)c); }
BX_EXPORT int bxtest_cache_live(void *c) { return bx_cache_live_slots((bx_cache_t *)c); }
BX_EXPORT int bxtest_cache_take_scratch(void *c, int64_t round)
{
    return bx_cache_take_scratch((bx_cache_t *)c, round);
}
BX_EXPORT void bxtest_cache_return_scratch(void *c, int slot)
{
    bx_cache_return_scratch((bx_cache_t *)c, slot);
}
BX_EXPORT void bxtest_cache_invalidate_all(void *c) { bx_cache_invalidate_all((bx_cache_t *)c); }
BX_EXPORT void bxtest_cache_counters(void *c, uint64_t *h, uint64_t *m, uint64_t *e)
{
    bx_cache_counters((bx_cache_t *)c, h, m, e);
}
BX_EXPORT void *bxtest_cache_slot_ptr(void *c, int slot)
{
    return bx_cache_slot_ptr((bx_cache_t *)c, slot);
}

BX_EXPORT void *bxtest_queue_create(int64_t n)
{
    bx_queue_t *q = (bx_queue_t *)malloc(sizeof *q);
    if (q) bx_queue_init(q, n);
    return q;
}
BX_EXPORT int64_t bxtest_queue_next(void *q) { return bx_queue_next((bx_queue_t *)q); }
BX_EXPORT void bxtest_queue_destroy(void *q) { free(q); }

/* tile-size heuristic, without a GPU */
BX_EXPORT int bxtest_autotune_tile(int dt, int M, int N, int K)
{
    return bx_autotune_tile(dt, M, N, K);
}

BX_EXPORT int bxtest_cache_validate(void *c) { return bx_cache_validate((bx_cache_t *)c); }
BX_EXPORT int bxtest_cache_set_budget(void *c, size_t bytes)
{
    return bx_cache_set_budget((bx_cache_t *)c, bytes);
}
BX_EXPORT uintptr_t bxtest_cache_slot_key(void *c, int slot)
{
    return bx_cache_slot_addr_key((bx_cache_t *)c, slot);
}

void bx_host_copy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                    size_t height);
BX_EXPORT void bxtest_host_copy2d(void *dst, size_t dpitch, const void *src, size_t spitch,
                                  size_t width, size_t height)
{
    bx_host_copy2d(dst, dpitch, src, spitch, width, height);
}

/* CRT split: the host-computed reconstruction constants and the residue
 * helpers the kernels use (compiled for the host from the same source) */
typedef struct {
    int L;
    double W[4][20];
    double Mk[4];
    double invM;
} crt_recon_view_t;
void bx_crt_consts(int L, void *c);
int bxh_crt_resid_host(double v, int m);
int bxh_crt_resid2_host(double v, int m);
int bxh_crt_norm_exp_host(double ss);
int bxh_crt_smod_i32_host(int x, int m);
void bxh_crt_tile_origin_host(int b, int nb, int np, int *m0, int *n0);

BX_EXPORT void bxtest_crt_consts(int L, double *W /* [4][20] */, double *Mk, double *invM)
{
    crt_recon_view_t c;
    bx_crt_consts(L, &c);
    for (int k = 0; k < 4; ++k) {
        Mk[k] = c.Mk[k];
        for (int l = 0; l < 20; ++l) W[k * 20 + l] = c.W[k][l];
    }
    *invM = c.invM;
}

BX_EXPORT int bxtest_crt_resid(double v, int m) { return bxh_crt_resid_host(v, m); }
BX_EXPORT int bxtest_crt_resid2(double v, int m) { return bxh_crt_resid2_host(v, m); }
BX_EXPORT int bxtest_crt_norm_exp(double ss) { return bxh_crt_norm_exp_host(ss); }
BX_EXPORT int bxtest_crt_smod_i32(int x, int m) { return bxh_crt_smod_i32_host(x, m); }
BX_EXPORT void bxtest_crt_tile_origin(int b, int nb, int np, int *m0, int *n0)
{
    bxh_crt_tile_origin_host(b, nb, np, m0, n0);
}

/* the tile runtime's round order and operand traffic, replayed on the host */
int bx_plan_simulate(int dt, int opa, int opb, int M, int N, int K, int T, int tri, int same_ab,
                     int cap_slots, uint64_t *op_bytes, uint64_t *evictions, int *mode,
                     int *nbands);
BX_EXPORT int bxtest_plan_simulate(int dt, int opa, int opb, int M, int N, int K, int T, int tri,
                                   int same_ab, int cap_slots, uint64_t *op_bytes,
                                   uint64_t *evictions, int *mode, int *nbands)
{
    return bx_plan_simulate(dt, opa, opb, M, N, K, T, tri, same_ab, cap_slots, op_bytes,
                            evictions, mode, nbands);
}

/* HER2K's unimodular factor (level3.c) */
void bx_her2k_factor(const double *a, int conj_ratio, double *out);
BX_EXPORT void bxtest_her2k_factor(double re, double im, int conj_ratio, double *out)
{
    const double a[2] = {re, im};
    bx_her2k_factor(a, conj_ratio, out);
}

/* the BF16x3 piece split, host build of the kernels' code (ozaki_kernels.hip) */
void bxh_bf16x3_pieces_host(double x, uint16_t *out);
BX_EXPORT void bxtest_bf16x3_pieces(double x, uint16_t *out) { bxh_bf16x3_pieces_host(x, out); }
