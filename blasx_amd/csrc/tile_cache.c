/*
 * tile_cache.c -- device tile cache for host-resident operands.
 *
 * Replaces the reference's per-GPU LRU list + red-black tree
 * (blas/LRU.c:28-265, blas/red_black_tree.c:122-351,
 * blas/blasx_mem_control.c:222-340) and its disabled sub-allocator
 * (blas/gpu_malloc.c:5-156).  Same contract: a tile is identified by the
 * host address of its first element; a hit moves it to the MRU end; a miss
 * takes a free slot or evicts the least-recently-used tile that no pending
 * work reads.  Redesigned for one MI355X with 288 GB of HBM:
 *   * slots are fixed-size pieces of large chunks (default 1 GiB) grown on
 *     demand up to a byte budget: no hipMalloc/hipFree per tile
 *     (blas/LRU.c:125,182) and no allocator fragmentation;
 *   * the key carries the leading dimension, extents and element size as
 *     well as the address (the reference keys on the address alone,
 *     include/RBT_LRU_common.h:11-14, so two views of one buffer collide);
 *   * O(1) hash index with chaining instead of an O(log n) tree;
 *   * "readers" (blas/blasx_mem_control.c:45-63) become the last round that
 *     uses a slot: a slot used by the round being planned is never evicted,
 *     one used by an earlier round may be, and the caller orders the
 *     overwrite after that round's kernel with a device-side event wait;
 *   * every slot pinned returns BLASX_EBUSY instead of assert
 *     (blas/LRU.c:134).
 * This file holds no HIP calls: device memory comes from the ops table,
 * so the policy runs unchanged on the CPU in the tests.
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

typedef struct slot {
    bx_tile_key_t key;
    void *ptr;        /* device address, fixed when the slot's chunk is allocated */
    int64_t last_round;
    int valid;        /* holds a keyed tile */
    int scratch;      /* lent out by take_scratch */
    int prev, next;   /* LRU list (valid slots) or free list (next only) */
    int hnext;        /* hash chain */
} slot_t;

struct bx_cache {
    size_t slot_bytes, chunk_bytes, budget;
    int slots_per_chunk;
    int nslots;       /* slots backed by memory */
    int max_slots;    /* budget / slot_bytes */
    void **chunks;
    size_t *chunk_len;  /* bytes of each chunk (the last before a budget change may be partial) */
    int nchunks, cap_chunks;
    slot_t *s;
    int s_cap;
    int *bucket;
    int nbuckets;     /* power of two */
    int lru_head, lru_tail, free_head;
    int live;
    bx_cache_ops_t ops;
    uint64_t hits, misses, evictions;
};

static uint64_t key_hash(const bx_tile_key_t *k)
{
    uint64_t h = (uint64_t)k->addr * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)k->ld * 0xC2B2AE3D27D4EB4Full + ((uint64_t)k->rows << 32) + (uint64_t)k->cols;
    h ^= (uint64_t)k->esize * 0x165667B19E3779F9ull;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    return h ^ (h >> 32);
}

static int key_eq(const bx_tile_key_t *a, const bx_tile_key_t *b)
{
    return a->addr == b->addr && a->ld == b->ld && a->rows == b->rows && a->cols == b->cols &&
           a->esize == b->esize;
}

static void lru_unlink(bx_cache_t *c, int i)
{
    slot_t *s = &c->s[i];
    if (s->prev >= 0) c->s[s->prev].next = s->next; else c->lru_head = s->next;
    if (s->next >= 0) c->s[s->next].prev = s->prev; else c->lru_tail = s->prev;
    s->prev = s->next = -1;
}

static void lru_push_head(bx_cache_t *c, int i)
{
    slot_t *s = &c->s[i];
    s->prev = -1;
    s->next = c->lru_head;
    if (c->lru_head >= 0) c->s[c->lru_head].prev = i;
    c->lru_head = i;
    if (c->lru_tail < 0) c->lru_tail = i;
}

static void hash_insert(bx_cache_t *c, int i)
{
    int b = (int)(key_hash(&c->s[i].key) & (uint64_t)(c->nbuckets - 1));
    c->s[i].hnext = c->bucket[b];
    c->bucket[b] = i;
}

static void hash_remove(bx_cache_t *c, int i)
{
    int b = (int)(key_hash(&c->s[i].key) & (uint64_t)(c->nbuckets - 1));
    int *pp = &c->bucket[b];
    while (*pp >= 0) {
        if (*pp == i) {
            *pp = c->s[i].hnext;
            c->s[i].hnext = -1;
            return;
        }
        pp = &c->s[*pp].hnext;
    }
}

static int rehash(bx_cache_t *c, int want)
{
    int nb = 64;
    while (nb < 2 * want) nb <<= 1;
    if (nb <= c->nbuckets) return 0;
    int *nbk = (int *)malloc(sizeof(int) * nb);
    if (!nbk) return BLASX_ENOMEM;
    free(c->bucket);
    c->bucket = nbk;
    c->nbuckets = nb;
    for (int b = 0; b < nb; ++b) c->bucket[b] = -1;
    for (int i = 0; i < c->nslots; ++i)
        if (c->s[i].valid) hash_insert(c, i);
    return 0;
}

bx_cache_t *bx_cache_create(size_t slot_bytes, size_t chunk_bytes, size_t budget_bytes,
                            const bx_cache_ops_t *ops)
{
    if (slot_bytes == 0 || chunk_bytes < slot_bytes || !ops || !ops->chunk_alloc) return NULL;
    bx_cache_t *c = (bx_cache_t *)calloc(1, sizeof *c);
    if (!c) return NULL;
    c->slot_bytes = slot_bytes;
    c->slots_per_chunk = (int)(chunk_bytes / slot_bytes);
    c->chunk_bytes = (size_t)c->slots_per_chunk * slot_bytes;
    c->ops = *ops;
    c->lru_head = c->lru_tail = c->free_head = -1;
    bx_cache_set_budget(c, budget_bytes);
    if (rehash(c, 64) != 0) {
        free(c);
        return NULL;
    }
    return c;
}

void bx_cache_free_memory(bx_cache_t *c)
{
    for (int i = 0; i < c->nchunks; ++i)
        if (c->chunks[i] && c->ops.chunk_free) c->ops.chunk_free(c->ops.user, c->chunks[i]);
    c->nchunks = 0;
    c->nslots = 0;
    c->lru_head = c->lru_tail = c->free_head = -1;
    c->live = 0;
    for (int b = 0; b < c->nbuckets; ++b) c->bucket[b] = -1;
}

void bx_cache_destroy(bx_cache_t *c)
{
    if (!c) return;
    bx_cache_free_memory(c);
    free(c->chunks);
    free(c->chunk_len);
    free(c->s);
    free(c->bucket);
    free(c);
}

size_t bx_cache_slot_bytes(const bx_cache_t *c) { return c->slot_bytes; }
int bx_cache_capacity_slots(const bx_cache_t *c) { return c->max_slots; }
int bx_cache_live_slots(const bx_cache_t *c) { return c->live; }

int bx_cache_set_budget(bx_cache_t *c, size_t budget_bytes)
{
    size_t ms = budget_bytes / c->slot_bytes;
    if (ms > (1u << 30)) ms = 1u << 30;
    if ((int)ms < c->nslots) {
        /* shrinking below what is allocated: drop everything (caller made
         * the device idle) and regrow lazily */
        bx_cache_free_memory(c);
    }
    c->budget = budget_bytes;
    c->max_slots = (int)ms;
    return 0;
}

/* add one chunk of slots to the free list; 0 on success */
static int grow(bx_cache_t *c)
{
    if (c->nslots >= c->max_slots) return BLASX_EBUSY;
    int add = c->slots_per_chunk;
    if (c->nslots + add > c->max_slots) add = c->max_slots - c->nslots;
    if (c->nchunks == c->cap_chunks) {
        int nc = c->cap_chunks ? 2 * c->cap_chunks : 16;
        void **p = (void **)realloc(c->chunks, sizeof(void *) * nc);
        if (!p) return BLASX_ENOMEM;
        c->chunks = p;
        size_t *l = (size_t *)realloc(c->chunk_len, sizeof(size_t) * nc);
        if (!l) return BLASX_ENOMEM;
        c->chunk_len = l;
        c->cap_chunks = nc;
    }
    if (c->nslots + add > c->s_cap) {
        int ns = c->s_cap ? c->s_cap : 64;
        while (ns < c->nslots + add) ns *= 2;
        slot_t *p = (slot_t *)realloc(c->s, sizeof(slot_t) * ns);
        if (!p) return BLASX_ENOMEM;
        c->s = p;
        c->s_cap = ns;
    }
    void *mem = NULL;
    if (c->ops.chunk_alloc(c->ops.user, &mem, (size_t)add * c->slot_bytes) != 0 || !mem) {
        /* out of device memory: stop growing, the budget is what we have */
        c->max_slots = c->nslots;
        return BLASX_ENOMEM;
    }
    c->chunk_len[c->nchunks] = (size_t)add * c->slot_bytes;
    c->chunks[c->nchunks++] = mem;
    for (int k = 0; k < add; ++k) {
        int i = c->nslots + k;
        memset(&c->s[i], 0, sizeof(slot_t));
        c->s[i].ptr = (char *)mem + (size_t)k * c->slot_bytes;
        c->s[i].prev = -1;
        c->s[i].hnext = -1;
        c->s[i].last_round = -1;
        c->s[i].next = c->free_head;
        c->free_head = i;
    }
    c->nslots += add;
    return rehash(c, c->nslots);
}

void *bx_cache_slot_ptr(const bx_cache_t *c, int i)
{
    if (i < 0 || i >= c->nslots) return NULL;
    /* chunks may be partial (the last one before a budget change), so the
     * address is stored per slot rather than derived from the index */
    return c->s[i].ptr;
}

int bx_cache_lookup(const bx_cache_t *c, const bx_tile_key_t *key)
{
    int b = (int)(key_hash(key) & (uint64_t)(c->nbuckets - 1));
    for (int i = c->bucket[b]; i >= 0; i = c->s[i].hnext)
        if (key_eq(&c->s[i].key, key)) return i;
    return -1;
}

/* a slot to (re)use: free list, growth, or LRU eviction among slots not
 * used by `round`; *evicted_round is the evicted tile's last round or -1 */
static int obtain(bx_cache_t *c, int64_t round, int64_t *evicted_round)
{
    *evicted_round = -1;
    if (c->free_head < 0) grow(c);
    if (c->free_head >= 0) {
        int i = c->free_head;
        c->free_head = c->s[i].next;
        c->s[i].next = -1;
        return i;
    }
    for (int i = c->lru_tail; i >= 0; i = c->s[i].prev) {
        if (c->s[i].last_round < round) {
            *evicted_round = c->s[i].last_round;
            lru_unlink(c, i);
            hash_remove(c, i);
            c->s[i].valid = 0;
            c->live--;
            c->evictions++;
            return i;
        }
    }
    return -1;
}

int bx_cache_acquire(bx_cache_t *c, const bx_tile_key_t *key, int64_t round, int *hit,
                     int64_t *evicted_round)
{
    *evicted_round = -1;
    int i = bx_cache_lookup(c, key);
    if (i >= 0) {
        *hit = 1;
        c->hits++;
        if (c->s[i].last_round < round) c->s[i].last_round = round;
        lru_unlink(c, i);
        lru_push_head(c, i);
        return i;
    }
    *hit = 0;
    i = obtain(c, round, evicted_round);
    if (i < 0) return BLASX_EBUSY;
    c->misses++;
    c->s[i].key = *key;
    c->s[i].valid = 1;
    c->s[i].scratch = 0;
    c->s[i].last_round = round;
    hash_insert(c, i);
    lru_push_head(c, i);
    c->live++;
    return i;
}

int bx_cache_take_scratch(bx_cache_t *c, int64_t round)
{
    int64_t ev;
    int i = obtain(c, round, &ev);
    if (i < 0) return BLASX_EBUSY;
    c->s[i].scratch = 1;
    c->s[i].valid = 0;
    c->s[i].last_round = ev; /* remembered so the caller can order reuse */
    return i;
}

void bx_cache_return_scratch(bx_cache_t *c, int i)
{
    if (i < 0 || i >= c->nslots || !c->s[i].scratch) return;
    c->s[i].scratch = 0;
    c->s[i].last_round = -1;
    c->s[i].next = c->free_head;
    c->free_head = i;
}

void bx_cache_invalidate(bx_cache_t *c, bx_keep_fn keep, void *user)
{
    int i = c->lru_head;
    while (i >= 0) {
        int nx = c->s[i].next;
        if (!keep || !keep(&c->s[i].key, user)) {
            lru_unlink(c, i);
            hash_remove(c, i);
            c->s[i].valid = 0;
            c->s[i].last_round = -1;
            c->s[i].next = c->free_head;
            c->free_head = i;
            c->live--;
        }
        i = nx;
    }
}

void bx_cache_invalidate_all(bx_cache_t *c) { bx_cache_invalidate(c, NULL, NULL); }

/* rounds restart at 0 in every call: forget old round numbers of kept tiles */
void bx_cache_reset_rounds(bx_cache_t *c)
{
    for (int i = c->lru_head; i >= 0; i = c->s[i].next) c->s[i].last_round = -1;
}

void bx_cache_counters(const bx_cache_t *c, uint64_t *hits, uint64_t *misses, uint64_t *evictions)
{
    if (hits) *hits = c->hits;
    if (misses) *misses = c->misses;
    if (evictions) *evictions = c->evictions;
}

/* Structural invariants, for the tests: every slot's memory lies inside one
 * allocated chunk and no two slots overlap; LRU and hash agree with the
 * valid flags; the live count matches.  Returns 0 or the first failed check. */
int bx_cache_validate(const bx_cache_t *c)
{
    int live = 0;
    for (int i = c->lru_head, prev = -1; i >= 0; prev = i, i = c->s[i].next) {
        if (c->s[i].prev != prev) return 1;
        if (!c->s[i].valid) return 2;
        if (bx_cache_lookup(c, &c->s[i].key) != i) return 3;
        ++live;
    }
    if (live != c->live) return 4;
    for (int i = 0; i < c->nslots; ++i) {
        const char *p = (const char *)c->s[i].ptr;
        if (!p) return 5;
        int inside = 0;
        for (int k = 0; k < c->nchunks && !inside; ++k) {
            const char *b = (const char *)c->chunks[k];
            if (p >= b && p + c->slot_bytes <= b + c->chunk_len[k]) inside = 1;
        }
        if (!inside) return 6;
        for (int j = i + 1; j < c->nslots; ++j) {
            const char *q = (const char *)c->s[j].ptr;
            if (p < q + c->slot_bytes && q < p + c->slot_bytes) return 7;
        }
    }
    return 0;
}

uintptr_t bx_cache_slot_addr_key(const bx_cache_t *c, int i)
{
    return (i >= 0 && i < c->nslots && c->s[i].valid) ? c->s[i].key.addr : 0;
}
