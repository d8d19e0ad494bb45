/*
 * tile_rt.c -- tiled GEMM over host-resident operands, one MI355X.
 *
 * Replaces the reference's tiled runtime: blasx_dgemm / blasx_gpu_dgemm /
 * blasx_gpu_dgemm_kernel (blas/blasx_dgemm.c:5-356 and its s/z twins),
 * the fetch path of mem_control_kernel_* (blas/blasx_mem_control.c:222-340),
 * margin_adjustment (:24-42) and collect_final_result_* (:65-156).
 *
 * Same decomposition as the reference: C is cut into T x T tiles; the
 * product for C tile (i,k) sums op(A) tile (i,j) * op(B) tile (j,k) over
 * the K tiles j in order; for a transposed operand the stored tile (j,i)
 * is fetched (blas/blasx_dgemm.c:42-53); edge tiles carry the remainder
 * extents; tiles are fetched through the device cache keyed on the host
 * address.  Redesigned for one GPU:
 *   * the work unit is a ROUND: an RM x RN block of C tiles; in core the
 *     rounds go block row by block row with alternating direction, out of
 *     core in bands that keep one operand's panels resident and stream the
 *     other past them (see plan_order), so operand bytes stay near the
 *     reuse bound e(MK + KN) instead of refetching per round
 *     (blas/blasx_dgemm.c:164-239 fetches every tile of every round);
 *   * one grouped kernel launch per round computes every C tile of the
 *     round over all of K (one segment per cached K tile), so C is read
 *     and written once per call, not once per K step with
 *     beta_inner = 1 (blas/blasx_dgemm.c:82);
 *   * three streams: copy-in (operand and C tiles), compute, copy-out;
 *     they are ordered with events, the host never waits per tile
 *     (blas/blasx_dgemm.c:79 waits per tile, :186 spins per round);
 *   * up to DEPTH rounds are in flight; C staging and descriptor buffers
 *     rotate over DEPTH slots;
 *   * a tile evicted while a recent round may still read it is
 *     overwritten only after that round's compute event;
 *   * all offsets are 64-bit (the reference's int offsets overflow at
 *     N = 131072, blas/blasx_mem_control.c:137-138).
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define BX_SLOT_BYTES ((size_t)2048 * 2048 * 8) /* one FP64 2048^2 tile: 32 MiB */
#define BX_CHUNK_BYTES ((size_t)1 << 30)
#define DEPTH 3

static bx_cache_t *g_cache;

static int dev_chunk_alloc(void *u, void **p, size_t bytes)
{
    (void)u;
    return bxh_malloc(p, bytes);
}

static void dev_chunk_free(void *u, void *p)
{
    (void)u;
    bxh_free(p);
}

bx_cache_t *bx_global_cache(void) { return g_cache; }

static bx_cache_t *get_cache(void)
{
    if (!g_cache) {
        bx_cache_ops_t ops = {dev_chunk_alloc, dev_chunk_free, NULL};
        g_cache = bx_cache_create(BX_SLOT_BYTES, BX_CHUNK_BYTES, g_bx.hbm_budget, &ops);
    }
    return g_cache;
}

/* events and descriptor buffers, created once, guarded by g_bx.gpu_lock */
static struct {
    int ready;
    void *ev_h2d[DEPTH], *ev_comp[DEPTH], *ev_d2h[DEPTH];
    void *desc_host[DEPTH], *desc_dev[DEPTH];
    size_t desc_cap;
} g_rt;

static int ensure_events(void)
{
    if (g_rt.ready) return 0;
    for (int d = 0; d < DEPTH; ++d) {
        int rc = bxh_event_create(&g_rt.ev_h2d[d], 0);
        if (!rc) rc = bxh_event_create(&g_rt.ev_comp[d], 0);
        if (!rc) rc = bxh_event_create(&g_rt.ev_d2h[d], 0);
        if (rc) return rc;
    }
    g_rt.ready = 1;
    return 0;
}

static int ensure_desc(size_t bytes)
{
    if (bytes <= g_rt.desc_cap) return 0;
    size_t cap = g_rt.desc_cap ? g_rt.desc_cap : 65536;
    while (cap < bytes) cap *= 2;
    for (int d = 0; d < DEPTH; ++d) {
        bxh_host_free(g_rt.desc_host[d]);
        bxh_free(g_rt.desc_dev[d]);
        g_rt.desc_host[d] = g_rt.desc_dev[d] = NULL;
    }
    g_rt.desc_cap = 0;
    for (int d = 0; d < DEPTH; ++d) {
        int rc = bxh_host_malloc(&g_rt.desc_host[d], cap);
        if (!rc) rc = bxh_malloc(&g_rt.desc_dev[d], cap);
        if (rc) return rc;
    }
    g_rt.desc_cap = cap;
    return 0;
}

int bx_tune_lookup(int dt, int M, int N, int K); /* autotune.c */

int bx_autotune_tile(int dt, int M, int N, int K)
{
    int t = bx_tune_lookup(dt, M, N, K);
    if (t > 0) return t;
    const int es = bx_esize(dt);
    int T = es >= 16 ? 1024 : 2048;
    /* a problem inside one tile: one tile product (the per-tile costs of
     * a pipeline would outweigh its overlap; measured at 1024^3 and 2048^3) */
    const int big = M > N ? (M > K ? M : K) : (N > K ? N : K);
    if (big <= T) {
        while (T > 512 && big <= T / 2) T /= 2;
        return T;
    }
    /* enough C tiles to form pipelined rounds */
    while (T > 512 && (int64_t)((M + T - 1) / T) * ((N + T - 1) / T) < 16) T /= 2;
    return T;
}

/* entries that survive the call: tiles inside a registered resident range */
static int keep_resident(const bx_tile_key_t *k, void *u)
{
    (void)u;
    uintptr_t end = k->addr + (uintptr_t)((k->cols - 1) * k->ld + k->rows) * k->esize;
    return bx_resident_contains(k->addr, end);
}

/* ---- pageable host memory: pinned staging rings ----
 * copy-in: the caller's tile is packed into a pinned buffer by the host copy
 * pool, then DMA'd; a buffer is reused once its previous DMA has finished.
 * copy-out: DMA into a pinned buffer; the host unpacks it into the caller's
 * C when the ring wraps around to it, or when the call drains. */
#define NS_IN 8
#define NS_OUT 24
void bx_host_copy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                    size_t height);

static struct {
    int ready;
    void *in[NS_IN], *ev_in[NS_IN];
    int in_next;
    void *out[NS_OUT], *ev_out[NS_OUT];
    int out_next;
    struct {
        char *dst;
        size_t dpitch, width, cols;
        int busy;
    } pend[NS_OUT];
} g_st;

static int ensure_staging(void)
{
    if (g_st.ready) return 0;
    int rc = 0;
    for (int i = 0; i < NS_IN && !rc; ++i) {
        rc = bxh_host_malloc(&g_st.in[i], BX_SLOT_BYTES);
        if (!rc) rc = bxh_event_create(&g_st.ev_in[i], 0);
    }
    for (int i = 0; i < NS_OUT && !rc; ++i) {
        rc = bxh_host_malloc(&g_st.out[i], BX_SLOT_BYTES);
        if (!rc) rc = bxh_event_create(&g_st.ev_out[i], 0);
    }
    if (!rc) g_st.ready = 1;
    return rc;
}

static int copy_in(int pageable, void *dst, size_t dpitch, const void *src, size_t spitch,
                   size_t width, size_t cols)
{
    if (!pageable) return bxh_memcpy2d_h2d_async(dst, dpitch, src, spitch, width, cols, g_bx.s_h2d);
    const int i = g_st.in_next;
    g_st.in_next = (i + 1) % NS_IN;
    int rc = bxh_event_sync(g_st.ev_in[i]); /* its previous DMA has read it */
    if (rc) return rc;
    bx_host_copy2d(g_st.in[i], width, src, spitch, width, cols);
    rc = bxh_memcpy2d_h2d_async(dst, dpitch, g_st.in[i], width, width, cols, g_bx.s_h2d);
    if (!rc) rc = bxh_event_record(g_st.ev_in[i], g_bx.s_h2d);
    return rc;
}

static int finish_out(int i, int discard)
{
    if (!g_st.pend[i].busy) return 0;
    int rc = bxh_event_sync(g_st.ev_out[i]);
    if (!rc && !discard)
        bx_host_copy2d(g_st.pend[i].dst, g_st.pend[i].dpitch, g_st.out[i], g_st.pend[i].width,
                       g_st.pend[i].width, g_st.pend[i].cols);
    g_st.pend[i].busy = 0;
    return rc;
}

static int copy_out(int pageable, void *dst, size_t dpitch, const void *src, size_t spitch,
                    size_t width, size_t cols)
{
    if (!pageable) return bxh_memcpy2d_d2h_async(dst, dpitch, src, spitch, width, cols, g_bx.s_d2h);
    const int i = g_st.out_next;
    g_st.out_next = (i + 1) % NS_OUT;
    int rc = finish_out(i, 0);
    if (!rc) rc = bxh_memcpy2d_d2h_async(g_st.out[i], width, src, spitch, width, cols, g_bx.s_d2h);
    if (!rc) rc = bxh_event_record(g_st.ev_out[i], g_bx.s_d2h);
    if (!rc) {
        g_st.pend[i].dst = (char *)dst;
        g_st.pend[i].dpitch = dpitch;
        g_st.pend[i].width = width;
        g_st.pend[i].cols = cols;
        g_st.pend[i].busy = 1;
    }
    return rc;
}

/* unpack every staged C tile, oldest first */
static int drain_out(int discard)
{
    if (!g_st.ready) return 0;
    int rc = 0;
    for (int k = 0; k < NS_OUT; ++k) {
        const int r = finish_out((g_st.out_next + k) % NS_OUT, discard);
        if (!rc) rc = r;
    }
    return rc;
}

typedef struct plan {
    int dt, es, opa, opb, M, N, K, T;
    int pg_a, pg_b, pg_c; /* operand in pageable host memory: staged copies */
    int mt, nt, kt, RM, RN;
    const char *A, *B;
    char *C;
    int64_t lda, ldb, ldc;
    double alpha[2], beta[2];
    int beta_zero;
    /* rank-k update (SYRK/HERK, B = A): only C tiles on the stored triangle
     * run, diagonal tiles masked (BX_TRI_* bits); 0 for a GEMM */
    int tri;
} plan_t;

/* whether C tile (i, k) lies on the stored triangle of a rank-k update */
static int tile_on_tri(const plan_t *p, int i, int k)
{
    if (p->tri & BX_TRI_LOWER) return k <= i;
    if (p->tri & BX_TRI_UPPER) return k >= i;
    return 1;
}

/* cache key of the stored tile holding op(A) tile (i, j) (block row i, K
 * tile j); for a transposed operand that is stored tile (j, i)
 * (blas/blasx_dgemm.c:42-53) */
static bx_tile_key_t a_key(const plan_t *p, int i, int j)
{
    const int T = p->T;
    const int mc = (p->M - i * T < T) ? p->M - i * T : T;
    const int kext = (p->K - j * T < T) ? p->K - j * T : T;
    const int n = p->opa == BX_OP_N;
    const int64_t r0 = (int64_t)(n ? i : j) * T, c0 = (int64_t)(n ? j : i) * T;
    bx_tile_key_t key = {(uintptr_t)(p->A + (r0 + c0 * p->lda) * p->es), p->lda,
                         n ? mc : kext, n ? kext : mc, p->es, 0};
    return key;
}

/* the same for op(B) tile (j, k) */
static bx_tile_key_t b_key(const plan_t *p, int k, int j)
{
    const int T = p->T;
    const int nc = (p->N - k * T < T) ? p->N - k * T : T;
    const int kext = (p->K - j * T < T) ? p->K - j * T : T;
    const int n = p->opb == BX_OP_N;
    const int64_t r0 = (int64_t)(n ? j : k) * T, c0 = (int64_t)(n ? k : j) * T;
    bx_tile_key_t key = {(uintptr_t)(p->B + (r0 + c0 * p->ldb) * p->es), p->ldb,
                         n ? kext : nc, n ? nc : kext, p->es, 0};
    return key;
}

/* ---- round order ----
 * A round is the block (bi, bj) of RM x RN C tiles; it reads the A row panel
 * of block row bi (RM x kt tiles) and the B column panel of block column bj
 * (RN x kt tiles).  When every operand tile fits in the cache the order is
 * block row by block row, alternating direction.  Out of core the tile-reuse
 * lower bound needs each operand read once; with the cache holding only part
 * of them the order keeps one operand's panels resident and streams the
 * other: C is cut into BANDS of block columns whose B panels fit in the cache
 * beside two A row panels (the one being read and the next one being
 * fetched) and the C staging; a band is walked block row by block row, so B
 * is fetched once and A once per band.  LRU eviction then only ever drops A
 * panels of earlier block rows and B panels of finished bands: within a band
 * every B panel was used during the previous block row, after the A panel
 * two rows back was last read.  The roles swap (bands of block rows, A
 * resident, B streamed) when that moves fewer tiles.  Bands alternate the
 * direction of their outer walk, so the last panel of one band is the first
 * of the next.  With one band this is the in-core order. */
typedef struct round_xy {
    int bi, bj;
} round_xy_t;

enum { BX_BAND_COLS = 0, BX_BAND_ROWS = 1 };

/* fills ord[] (nbm * nbn entries) and returns the number of bands; *mode
 * gets BX_BAND_COLS or BX_BAND_ROWS */
static int plan_order(const plan_t *p, int cap_slots, round_xy_t *ord, int *mode)
{
    const int nbm = (p->mt + p->RM - 1) / p->RM, nbn = (p->nt + p->RN - 1) / p->RN;
    const int64_t pa = (int64_t)p->RM * p->kt, pb = (int64_t)p->RN * p->kt;
    const int64_t room = (int64_t)cap_slots - (int64_t)DEPTH * p->RM * p->RN - 2;
    int64_t wc = (room - 2 * pa) / pb, wr = (room - 2 * pb) / pa;
    if (wc < 1) wc = 1;
    if (wr < 1) wr = 1;
    const int64_t nbc = (nbn + wc - 1) / wc, nbr = (nbm + wr - 1) / wr;
    /* operand tiles fetched: the resident operand once, the streamed one per band */
    const int64_t cost_c = (int64_t)p->nt * p->kt + nbc * (int64_t)p->mt * p->kt;
    const int64_t cost_r = (int64_t)p->mt * p->kt + nbr * (int64_t)p->nt * p->kt;
    const int cols = cost_c <= cost_r;
    const int outer_n = cols ? nbn : nbm, inner_n = cols ? nbm : nbn;
    const int nb = (int)(cols ? nbc : nbr);
    const int w = (outer_n + nb - 1) / nb; /* balanced band width */
    int n = 0, flip = 0;
    for (int b = 0; b < nb; ++b) {
        const int o0 = b * w, o1 = (o0 + w < outer_n) ? o0 + w : outer_n;
        for (int q = 0; q < inner_n; ++q) {
            const int in = (b & 1) ? inner_n - 1 - q : q;
            for (int r = 0; r < o1 - o0; ++r) {
                const int out = flip ? o1 - 1 - r : o0 + r;
                ord[n].bi = cols ? in : out;
                ord[n].bj = cols ? out : in;
                ++n;
            }
            flip ^= 1;
        }
    }
    *mode = cols ? BX_BAND_COLS : BX_BAND_ROWS;
    return nb;
}

static void choose_round_shape(plan_t *p, int cap_slots)
{
    const int bm = bxh_tile_m(p->dt), bn = bxh_tile_n(p->dt);
    const int64_t blocks = (int64_t)((p->T + bm - 1) / bm) * ((p->T + bn - 1) / bn);
    int R = (int)((2048 + blocks - 1) / blocks);
    if (R < 1) R = 1;
    if (R > 16) R = 16;
    int RM = 1;
    while ((RM + 1) * (RM + 1) <= R) ++RM;
    int RN = (R + RM - 1) / RM;
    if (RM > p->mt) RM = p->mt;
    if (RN > p->nt) RN = p->nt;
    /* two rounds of operand tiles plus DEPTH rounds of C staging must fit */
    while ((RM > 1 || RN > 1) &&
           (int64_t)2 * (RM + RN) * p->kt + (int64_t)DEPTH * RM * RN + 2 > cap_slots) {
        if (RN >= RM && RN > 1)
            --RN;
        else
            --RM;
    }
    p->RM = RM;
    p->RN = RN;
}

static int run_plan(plan_t *p, bx_cache_t *c)
{
    const int es = p->es, T = p->T;
    const int R = p->RM * p->RN;
    const int bm = bxh_tile_m(p->dt), bn = bxh_tile_n(p->dt);
    const size_t prob_bytes = (size_t)R * sizeof(bxh_problem_t);
    const size_t seg_off_bytes = (prob_bytes + 255) & ~(size_t)255;
    const size_t desc_bytes = seg_off_bytes + (size_t)R * p->kt * sizeof(bxh_seg_t);
    int rc = ensure_desc(desc_bytes);
    if (rc) return rc;

    if ((int64_t)2 * (p->RM + p->RN) * p->kt + (int64_t)DEPTH * R + 2 >
        bx_cache_capacity_slots(c)) {
        bx_log(0, "tile cache budget too small: %d slots for K/T = %d", bx_cache_capacity_slots(c),
               p->kt);
        return BLASX_ERANGE;
    }

    /* C staging slots: DEPTH x R, taken out of the cache for this call */
    int *cst = (int *)malloc(sizeof(int) * DEPTH * R);
    int *sa = (int *)malloc(sizeof(int) * p->RM);
    int *sb = (int *)malloc(sizeof(int) * p->RN);
    /* this round's C tiles (row, column within the round block) and which
     * operand tiles they use */
    int *tii = (int *)malloc(sizeof(int) * R), *tkk = (int *)malloc(sizeof(int) * R);
    char *use_a = (char *)malloc(p->RM), *use_b = (char *)malloc(p->RN);
    const int64_t nord = (int64_t)((p->mt + p->RM - 1) / p->RM) * ((p->nt + p->RN - 1) / p->RN);
    round_xy_t *ord = (round_xy_t *)malloc(sizeof(round_xy_t) * nord);
    if (!cst || !sa || !sb || !tii || !tkk || !use_a || !use_b || !ord) {
        free(cst);
        free(sa);
        free(sb);
        free(tii);
        free(tkk);
        free(use_a);
        free(use_b);
        free(ord);
        return BLASX_ENOMEM;
    }
    int mode;
    const int nbands = plan_order(p, bx_cache_capacity_slots(c), ord, &mode);
    bx_log(2, "tile plan: %lld rounds in %d band(s) of block %s", (long long)nord, nbands,
           mode == BX_BAND_COLS ? "columns (B resident)" : "rows (A resident)");
    int ntaken = 0;
    for (; ntaken < DEPTH * R; ++ntaken) {
        cst[ntaken] = bx_cache_take_scratch(c, 0);
        if (cst[ntaken] < 0) {
            rc = BLASX_EBUSY;
            break;
        }
    }

    uint64_t h2d = 0, d2h = 0, hits = 0, misses = 0, evictions = 0;
    int64_t round = 0;
    for (int64_t oi = 0; oi < nord && !rc; ++oi) {
        const int bi = ord[oi].bi, bj = ord[oi].bj;
        const int i0 = bi * p->RM, i1 = (i0 + p->RM < p->mt) ? i0 + p->RM : p->mt;
        const int k0 = bj * p->RN, k1 = (k0 + p->RN < p->nt) ? k0 + p->RN : p->nt;
        const int ni = i1 - i0, nk = k1 - k0;
        /* the round's C tiles: all of the block, or those on the triangle */
        int ntile = 0;
        memset(use_a, 0, ni);
        memset(use_b, 0, nk);
        for (int ii = 0; ii < ni; ++ii)
            for (int kk = 0; kk < nk; ++kk)
                if (tile_on_tri(p, i0 + ii, k0 + kk)) {
                    tii[ntile] = ii;
                    tkk[ntile] = kk;
                    use_a[ii] = use_b[kk] = 1;
                    ++ntile;
                }
        if (ntile == 0) continue; /* a block off the triangle: no round */
        const int d = (int)(round % DEPTH);
        if (round >= DEPTH) {
            /* at most DEPTH rounds in flight: round-DEPTH has computed */
            if ((rc = bxh_event_sync(g_rt.ev_comp[d]))) break;
            /* its C staging must be copied out before reuse */
            if ((rc = bxh_stream_wait_event(g_bx.s_h2d, g_rt.ev_d2h[d]))) break;
            if ((rc = bxh_stream_wait_event(g_bx.s_comp, g_rt.ev_d2h[d]))) break;
        }
        char *desc = (char *)g_rt.desc_host[d];
        bxh_problem_t *prob = (bxh_problem_t *)desc;
        bxh_seg_t *seg = (bxh_seg_t *)(desc + seg_off_bytes);

        /* C tiles in, problem descriptors */
        int tiles_total = 0;
        for (int t = 0; t < ntile; ++t) {
            const int ii = tii[t], kk = tkk[t];
            const int i = i0 + ii, k = k0 + kk;
            const int mc = (p->M - i * T < T) ? p->M - i * T : T;
            const int nc = (p->N - k * T < T) ? p->N - k * T : T;
            void *cs = bx_cache_slot_ptr(c, cst[d * R + t]);
            /* a masked diagonal tile goes back whole: its other
             * triangle must hold the caller's values, so it is read
             * even when beta == 0 */
            const int diag = p->tri && i == k;
            if (!p->beta_zero || diag) {
                const char *hc = p->C + ((int64_t)i * T + (int64_t)k * T * p->ldc) * es;
                rc = copy_in(p->pg_c, cs, (size_t)T * es, hc, (size_t)p->ldc * es,
                             (size_t)mc * es, (size_t)nc);
                if (rc) break;
                h2d += (uint64_t)mc * nc * es;
            }
            bxh_problem_t *P = &prob[t];
            memset(P, 0, sizeof *P);
            P->C = cs;
            P->ldc = T;
            P->m = mc;
            P->n = nc;
            P->nseg = p->kt;
            P->seg_off = t * p->kt;
            P->opa = p->opa;
            P->opb = p->opb;
            P->tiles_m = (mc + bm - 1) / bm;
            P->tiles_n = (nc + bn - 1) / bn;
            P->tile_start = tiles_total;
            tiles_total += P->tiles_m * P->tiles_n;
            P->beta_zero = p->beta_zero;
            P->alpha_re = p->alpha[0];
            P->alpha_im = p->alpha[1];
            P->beta_re = p->beta[0];
            P->beta_im = p->beta[1];
            P->tri = diag ? p->tri : 0;
            P->tri_off = 0;
        }
        if (rc) break;

        /* operand tiles of every K step, through the cache */
        for (int j = 0; j < p->kt && !rc; ++j) {
            const int kext = (p->K - j * T < T) ? p->K - j * T : T;
            for (int ii = 0; ii < ni && !rc; ++ii) {
                if (!use_a[ii]) continue;
                const bx_tile_key_t key = a_key(p, i0 + ii, j);
                const int rows = key.rows, cols = key.cols;
                int hit;
                int64_t ev;
                sa[ii] = bx_cache_acquire(c, &key, round, &hit, &ev);
                if (sa[ii] < 0) {
                    rc = BLASX_EBUSY;
                    break;
                }
                if (hit) {
                    ++hits;
                    continue;
                }
                ++misses;
                if (ev >= 0) {
                    ++evictions;
                    if (round - ev < DEPTH)
                        rc = bxh_stream_wait_event(g_bx.s_h2d, g_rt.ev_comp[ev % DEPTH]);
                }
                if (!rc)
                    rc = copy_in(p->pg_a, bx_cache_slot_ptr(c, sa[ii]), (size_t)T * es,
                                 (const void *)key.addr, (size_t)p->lda * es,
                                 (size_t)rows * es, (size_t)cols);
                h2d += (uint64_t)rows * cols * es;
            }
            for (int kk = 0; kk < nk && !rc; ++kk) {
                if (!use_b[kk]) continue;
                const bx_tile_key_t key = b_key(p, k0 + kk, j);
                const int rows = key.rows, cols = key.cols;
                int hit;
                int64_t ev;
                sb[kk] = bx_cache_acquire(c, &key, round, &hit, &ev);
                if (sb[kk] < 0) {
                    rc = BLASX_EBUSY;
                    break;
                }
                if (hit) {
                    ++hits;
                    continue;
                }
                ++misses;
                if (ev >= 0) {
                    ++evictions;
                    if (round - ev < DEPTH)
                        rc = bxh_stream_wait_event(g_bx.s_h2d, g_rt.ev_comp[ev % DEPTH]);
                }
                if (!rc)
                    rc = copy_in(p->pg_b, bx_cache_slot_ptr(c, sb[kk]), (size_t)T * es,
                                 (const void *)key.addr, (size_t)p->ldb * es,
                                 (size_t)rows * es, (size_t)cols);
                h2d += (uint64_t)rows * cols * es;
            }
            if (rc) break;
            for (int t = 0; t < ntile; ++t) {
                bxh_seg_t *S = &seg[t * p->kt + j];
                S->A = bx_cache_slot_ptr(c, sa[tii[t]]);
                S->B = bx_cache_slot_ptr(c, sb[tkk[t]]);
                S->lda = T;
                S->ldb = T;
                S->k = kext;
            }
        }
        if (rc) break;

        /* compute after this round's copies; descriptors ride the compute stream */
        const int nprob = ntile;
        if ((rc = bxh_event_record(g_rt.ev_h2d[d], g_bx.s_h2d))) break;
        if ((rc = bxh_stream_wait_event(g_bx.s_comp, g_rt.ev_h2d[d]))) break;
        if ((rc = bxh_memcpy_h2d_async(g_rt.desc_dev[d], g_rt.desc_host[d],
                                       seg_off_bytes + (size_t)nprob * p->kt * sizeof(bxh_seg_t),
                                       g_bx.s_comp)))
            break;
        char *ddev = (char *)g_rt.desc_dev[d];
        if ((rc = bxh_gemm_grouped(p->dt, (const bxh_problem_t *)ddev, nprob,
                                   (const bxh_seg_t *)(ddev + seg_off_bytes), tiles_total, 0,
                                   g_bx.s_comp)))
            break;
        BX_STAT_ADD(kernel_launches, 1);
        if ((rc = bxh_event_record(g_rt.ev_comp[d], g_bx.s_comp))) break;

        /* C tiles out */
        if ((rc = bxh_stream_wait_event(g_bx.s_d2h, g_rt.ev_comp[d]))) break;
        for (int t = 0; t < ntile && !rc; ++t) {
            const int i = i0 + tii[t], k = k0 + tkk[t];
            const int mc = (p->M - i * T < T) ? p->M - i * T : T;
            const int nc = (p->N - k * T < T) ? p->N - k * T : T;
            char *hc = p->C + ((int64_t)i * T + (int64_t)k * T * p->ldc) * es;
            rc = copy_out(p->pg_c, hc, (size_t)p->ldc * es,
                          bx_cache_slot_ptr(c, cst[d * R + t]), (size_t)T * es,
                          (size_t)mc * es, (size_t)nc);
            d2h += (uint64_t)mc * nc * es;
        }
        if (!rc) rc = bxh_event_record(g_rt.ev_d2h[d], g_bx.s_d2h);
        ++round;
    }

    /* drain every stream before the buffers are reused or the caller reads C */
    int rc2 = bxh_stream_sync(g_bx.s_h2d);
    int rc3 = bxh_stream_sync(g_bx.s_comp);
    int rc4 = bxh_stream_sync(g_bx.s_d2h);
    if (!rc) rc = rc2 ? rc2 : rc3 ? rc3 : rc4;
    /* unpack staged C tiles into the caller's pageable C */
    const int rc5 = drain_out(rc != 0);
    if (!rc) rc = rc5;

    for (int t = 0; t < ntaken; ++t) bx_cache_return_scratch(c, cst[t]);
    free(cst);
    free(sa);
    free(sb);
    free(tii);
    free(tkk);
    free(use_a);
    free(use_b);
    free(ord);
    bx_cache_invalidate(c, keep_resident, NULL);

    BX_STAT_ADD(bytes_h2d, h2d);
    BX_STAT_ADD(bytes_d2h, d2h);
    BX_STAT_ADD(tile_hits, hits);
    BX_STAT_ADD(tile_misses, misses);
    BX_STAT_ADD(tile_evictions, evictions);
    return rc;
}

/* ---- host-only replay of a plan's cache traffic ----
 * Runs the round order and the operand acquisitions of run_plan against a
 * cache of cap_slots slots with no device memory behind it, and counts the
 * operand bytes a real call would copy host to device.  For tests and for
 * choosing budgets without a GPU.  Host addresses are synthetic (operands at
 * disjoint ranges; A == B when same_ab). */
static int fake_chunk_alloc(void *u, void **p, size_t bytes)
{
    uintptr_t *next = (uintptr_t *)u;
    *p = (void *)*next;
    *next += bytes;
    return 0;
}

int bx_plan_simulate(int dt, int opa, int opb, int M, int N, int K, int T, int tri, int same_ab,
                     int cap_slots, uint64_t *op_bytes, uint64_t *evictions, int *mode,
                     int *nbands)
{
    plan_t p;
    memset(&p, 0, sizeof p);
    p.dt = dt;
    p.es = bx_esize(dt);
    p.opa = opa;
    p.opb = opb;
    p.M = M;
    p.N = N;
    p.K = K;
    p.T = T;
    p.tri = tri;
    p.mt = (M + T - 1) / T;
    p.nt = (N + T - 1) / T;
    p.kt = (K + T - 1) / T;
    p.lda = opa == BX_OP_N ? M : K;
    p.ldb = opb == BX_OP_N ? K : N;
    p.A = (const char *)((uintptr_t)1 << 44);
    p.B = same_ab ? p.A : (const char *)((uintptr_t)1 << 45);
    choose_round_shape(&p, cap_slots);
    if ((int64_t)2 * (p.RM + p.RN) * p.kt + (int64_t)DEPTH * p.RM * p.RN + 2 > cap_slots)
        return BLASX_ERANGE;
    uintptr_t fake = (uintptr_t)1 << 40;
    bx_cache_ops_t ops = {fake_chunk_alloc, NULL, &fake};
    const size_t slot = (size_t)T * T * p.es;
    bx_cache_t *c = bx_cache_create(slot, slot * 32, (size_t)cap_slots * slot, &ops);
    const int64_t nord = (int64_t)((p.mt + p.RM - 1) / p.RM) * ((p.nt + p.RN - 1) / p.RN);
    round_xy_t *ord = (round_xy_t *)malloc(sizeof(round_xy_t) * nord);
    char *use_a = (char *)malloc(p.RM), *use_b = (char *)malloc(p.RN);
    int rc = (!c || !ord || !use_a || !use_b) ? BLASX_ENOMEM : 0;
    for (int t = 0; t < DEPTH * p.RM * p.RN && !rc; ++t)
        if (bx_cache_take_scratch(c, 0) < 0) rc = BLASX_EBUSY;
    uint64_t bytes = 0, ev_n = 0;
    int64_t round = 0;
    if (!rc) *nbands = plan_order(&p, cap_slots, ord, mode);
    for (int64_t oi = 0; oi < nord && !rc; ++oi) {
        const int i0 = ord[oi].bi * p.RM, k0 = ord[oi].bj * p.RN;
        const int ni = (i0 + p.RM < p.mt ? i0 + p.RM : p.mt) - i0;
        const int nk = (k0 + p.RN < p.nt ? k0 + p.RN : p.nt) - k0;
        int ntile = 0;
        memset(use_a, 0, ni);
        memset(use_b, 0, nk);
        for (int ii = 0; ii < ni; ++ii)
            for (int kk = 0; kk < nk; ++kk)
                if (tile_on_tri(&p, i0 + ii, k0 + kk)) {
                    use_a[ii] = use_b[kk] = 1;
                    ++ntile;
                }
        if (ntile == 0) continue;
        for (int j = 0; j < p.kt && !rc; ++j) {
            for (int r = 0; r < ni + nk && !rc; ++r) {
                const int isa = r < ni;
                if (isa ? !use_a[r] : !use_b[r - ni]) continue;
                const bx_tile_key_t key = isa ? a_key(&p, i0 + r, j) : b_key(&p, k0 + r - ni, j);
                int hit;
                int64_t ev;
                if (bx_cache_acquire(c, &key, round, &hit, &ev) < 0) {
                    rc = BLASX_EBUSY;
                    break;
                }
                if (!hit) bytes += (uint64_t)key.rows * key.cols * p.es;
                if (ev >= 0) ++ev_n;
            }
        }
        ++round;
    }
    *op_bytes = bytes;
    *evictions = ev_n;
    free(ord);
    free(use_a);
    free(use_b);
    bx_cache_destroy(c);
    return rc;
}

static int tile_call(int dt, int opa, int opb, int M, int N, int K, const double *alpha,
                     const void *A, int64_t lda, const void *B, int64_t ldb, const double *beta,
                     void *C, int64_t ldc, int tri)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    pthread_mutex_lock(&g_bx.gpu_lock);
    int rc = 0;
    bx_cache_t *c = get_cache();
    if (!c) rc = BLASX_ENOMEM;
    if (!rc) rc = ensure_events();
    if (!rc) {
        plan_t p;
        memset(&p, 0, sizeof p);
        p.dt = dt;
        p.es = bx_esize(dt);
        p.opa = opa;
        p.opb = opb;
        p.M = M;
        p.N = N;
        p.K = K;
        p.tri = tri;
        /* DMA needs pinned host memory; anything else goes through staging */
        const int pageable = bxh_pointer_is_pinned(A) == 0 || bxh_pointer_is_pinned(B) == 0 ||
                             bxh_pointer_is_pinned(C) == 0;
        int T = g_bx.tile_override[dt];
        if (!T) {
            /* the table's edges were measured on pinned operands; through the
             * staging pool each tile costs a pack and a hand-off, and 1024
             * or 2048 won at every size measured (pageable DGEMM 1024^3,
             * 2048^3, 4096^3: T = 512 1.6-1.9x slower than 1024) */
            T = bx_autotune_tile(dt, M, N, K);
            if (pageable && T < 1024) T = 1024;
        }
        while ((size_t)T * T * p.es > bx_cache_slot_bytes(c)) T /= 2;
        p.T = T;
        p.mt = (M + T - 1) / T;
        p.nt = (N + T - 1) / T;
        p.kt = (K + T - 1) / T;
        p.A = (const char *)A;
        p.B = (const char *)B;
        p.C = (char *)C;
        p.lda = lda;
        p.ldb = ldb;
        p.ldc = ldc;
        p.alpha[0] = alpha[0];
        p.alpha[1] = alpha[1];
        p.beta[0] = beta[0];
        p.beta[1] = beta[1];
        p.beta_zero = beta[0] == 0.0 && beta[1] == 0.0;
        /* DMA needs pinned host memory; anything else goes through staging */
        p.pg_a = bxh_pointer_is_pinned(A) == 0;
        p.pg_b = bxh_pointer_is_pinned(B) == 0;
        p.pg_c = bxh_pointer_is_pinned(C) == 0;
        if (p.pg_a || p.pg_b || p.pg_c) rc = ensure_staging();
        choose_round_shape(&p, bx_cache_capacity_slots(c));
        bx_cache_reset_rounds(c);
        bx_log(2, "tile %s dt=%d %dx%dx%d T=%d rounds of %dx%d tiles", tri ? "syrk" : "gemm", dt,
               M, N, K, T, p.RM, p.RN);
        if (!rc) rc = run_plan(&p, c);
        if (!rc) {
            const double cm = dt == BLASX_Z ? 4.0 : 1.0; /* real flops per multiply-add / 2 */
            if (tri) {
                /* A read once, the stored triangle of C read and written */
                const uint64_t tri_elems = (uint64_t)N * (N + 1) / 2;
                BX_STAT_ADD(bytes_lower_bound,
                            (uint64_t)p.es * ((uint64_t)N * K + tri_elems * (p.beta_zero ? 1 : 2)));
                bx_stat_flops(cm * (double)N * (N + 1) * K);
            } else {
                BX_STAT_ADD(bytes_lower_bound,
                            (uint64_t)p.es * ((uint64_t)M * K + (uint64_t)K * N +
                                              (uint64_t)M * N * (p.beta_zero ? 1 : 2)));
                bx_stat_flops(2.0 * cm * M * (double)N * K);
            }
        }
    }
    pthread_mutex_unlock(&g_bx.gpu_lock);
    return rc;
}

int bx_tile_gemm(int dt, int opa, int opb, int M, int N, int K, const double *alpha,
                 const void *A, int64_t lda, const void *B, int64_t ldb, const double *beta,
                 void *C, int64_t ldc)
{
    return tile_call(dt, opa, opb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 0);
}

/* C = alpha op(A) op(A)^T + beta C (op(A)^H for HERK, tri & BX_TRI_HERM) on
 * the triangle tri names, for host-resident operands: the GEMM tile plan
 * restricted to C tiles on that triangle, B = A, so each A tile is fetched
 * once per round for both of its roles.  The reference's dsyrk forwards to
 * the CPU BLAS (blas/dsyrk.c:46-58); its tile helpers for a GPU version enumerate
 * the same triangle (blas/blasx_mem_control.c:3-22, 189-219). */
int bx_tile_syrk(int dt, int tri, int opa, int N, int K, const double *alpha, const void *A,
                 int64_t lda, const double *beta, void *C, int64_t ldc)
{
    const int opb = opa != BX_OP_N ? BX_OP_N : (tri & BX_TRI_HERM) ? BX_OP_C : BX_OP_T;
    return tile_call(dt, opa, opb, N, N, K, alpha, A, lda, A, lda, beta, C, ldc, tri);
}
