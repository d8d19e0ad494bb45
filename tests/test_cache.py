"""Host-side policies of the tile runtime, on the CPU: the tile cache
(blas/LRU.c, blas/red_black_tree.c, blas/blasx_mem_control.c:222-340 in the
reference) over host memory, and the task queue (blas/blasx_task_queue.c).

Differential tests run against (a) a plain Python model of the intended
policy and (b) the reference's own red-black tree and queue, compiled from
its sources into oracle/_ref by oracle/Makefile.
"""
import ctypes as C
import random
import threading

import pytest

SLOT = 1024


class Cache:
    def __init__(self, bx, slots, chunk_slots=32):
        self.L = bx.test_lib()
        self.L.bxtest_cache_validate.argtypes = [C.c_void_p]
        self.L.bxtest_cache_set_budget.argtypes = [C.c_void_p, C.c_size_t]
        self.L.bxtest_cache_slot_key.argtypes = [C.c_void_p, C.c_int]
        self.L.bxtest_cache_slot_key.restype = C.c_uint64
        self.c = self.L.bxtest_cache_create(SLOT, SLOT * chunk_slots, SLOT * slots)
        assert self.c

    def acquire(self, addr, rnd, ld=100, rows=10, cols=10, es=8):
        hit, ev = C.c_int(), C.c_int64()
        s = self.L.bxtest_cache_acquire(self.c, addr, ld, rows, cols, es, rnd, C.byref(hit),
                                        C.byref(ev))
        return s, bool(hit.value), ev.value

    def lookup(self, addr, ld=100, rows=10, cols=10, es=8):
        return self.L.bxtest_cache_lookup(self.c, addr, ld, rows, cols, es)

    def validate(self):
        return self.L.bxtest_cache_validate(self.c)

    def counters(self):
        h, m, e = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.L.bxtest_cache_counters(self.c, C.byref(h), C.byref(m), C.byref(e))
        return h.value, m.value, e.value

    def close(self):
        self.L.bxtest_cache_destroy(self.c)


def test_partial_chunk_then_growth_regression(bx):
    """A budget of 48 slots leaves a partial second chunk (16 of 32 slots);
    after the budget grows, new slots must come from a new chunk, not from
    past the end of the partial one."""
    c = Cache(bx, 48, chunk_slots=32)
    for a in range(48):
        c.acquire(0x1000 + a * 8, 0)
    assert c.validate() == 0
    c.L.bxtest_cache_set_budget(c.c, SLOT * 200)
    for a in range(48, 150):
        s, hit, _ = c.acquire(0x1000 + a * 8, 1)
        assert s >= 0 and not hit
    assert c.validate() == 0
    ptrs = [c.L.bxtest_cache_slot_ptr(c.c, i) for i in range(150)]
    assert len(set(ptrs)) == 150
    c.close()


def test_hit_miss_and_key_fields(bx):
    c = Cache(bx, 16)
    s1, hit, _ = c.acquire(0x5000, 0)
    assert not hit
    s2, hit, _ = c.acquire(0x5000, 0)
    assert hit and s2 == s1
    # same address, different ld / extents / element size: different tiles
    # (the reference keys on the address alone, include/RBT_LRU_common.h:11-14)
    for kw in ({"ld": 101}, {"rows": 9}, {"cols": 11}, {"es": 16}):
        s, hit, _ = c.acquire(0x5000, 0, **kw)
        assert not hit and s != s1
    assert c.validate() == 0
    assert c.counters() == (1, 5, 0)
    c.close()


def test_lru_policy_matches_model(bx):
    """Randomised differential test against a simple model: on a miss with
    no free slot, evict the least recently used tile whose last use is older
    than the current round; a hit refreshes recency."""
    rng = random.Random(1)
    cap = 24
    c = Cache(bx, cap, chunk_slots=8)
    model = {}  # addr -> [last_round, last_use_time]
    t = 0
    for rnd in range(200):
        for _ in range(rng.randint(1, 12)):
            addr = 0x10000 + 64 * rng.randint(0, 60)
            t += 1
            s, hit, ev = c.acquire(addr, rnd)
            if addr in model:
                assert hit and s >= 0
                model[addr] = [rnd, t]
                continue
            assert not hit
            if len(model) < cap:
                assert s >= 0 and ev == -1
            else:
                cands = [(v[1], a) for a, v in model.items() if v[0] < rnd]
                if not cands:
                    assert s < 0  # every slot is read by this round: busy
                    continue
                _, victim = min(cands)
                assert s >= 0 and ev == model[victim][0]
                assert c.lookup(victim) == -1
                del model[victim]
            model[addr] = [rnd, t]
            assert c.lookup(addr) == s
        assert c.validate() == 0
    assert c.L.bxtest_cache_live(c.c) == len(model)
    c.close()


def test_busy_when_round_pins_everything(bx):
    c = Cache(bx, 4, chunk_slots=4)
    for a in range(4):
        assert c.acquire(0x100 * (a + 1), 7)[0] >= 0
    s, hit, _ = c.acquire(0x9999, 7)
    assert s < 0 and not hit  # BLASX_EBUSY instead of the reference's assert (blas/LRU.c:134)
    s, hit, ev = c.acquire(0x9999, 8)  # next round: the LRU tile of round 7 goes
    assert s >= 0 and ev == 7
    assert c.lookup(0x100) == -1
    c.close()


def test_scratch_slots_leave_the_index(bx):
    c = Cache(bx, 8, chunk_slots=8)
    sc = [c.L.bxtest_cache_take_scratch(c.c, 0) for _ in range(3)]
    assert len(set(sc)) == 3 and min(sc) >= 0
    for a in range(5):
        assert c.acquire(0x40 * (a + 1), 0)[0] not in sc
    assert c.acquire(0x7777, 0)[0] < 0  # 3 lent + 5 pinned by round 0
    for s in sc:
        c.L.bxtest_cache_return_scratch(c.c, s)
    assert c.acquire(0x7777, 0)[0] >= 0
    c.L.bxtest_cache_invalidate_all(c.c)
    assert c.L.bxtest_cache_live(c.c) == 0 and c.validate() == 0
    c.close()


def test_index_agrees_with_reference_red_black_tree(bx, orc):
    """Membership after a random insert/evict sequence matches the
    reference's red-black tree (blas/red_black_tree.c) driven with the same
    inserts and deletes."""
    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference tree here)")
    root = C.c_void_p(ref.rbt_init())
    rng = random.Random(2)
    cap = 32
    c = Cache(bx, cap, chunk_slots=8)
    present = set()
    for rnd in range(300):
        addr = 0x100000 + 16 * rng.randint(0, 200)
        s, hit, ev = c.acquire(addr, rnd)
        assert s >= 0
        if not hit:
            if ev >= 0 or len(present) >= cap:
                # the slot now holds addr; find which key left the index
                gone = [a for a in present if c.lookup(a) == -1]
                assert len(gone) == 1
                node = ref.rbt_find(gone[0], C.byref(root))
                assert node
                ref.rbt_delete(node, C.byref(root))
                present.discard(gone[0])
            assert not ref.rbt_find(addr, C.byref(root))
            ref.rbt_insert(addr, None, C.byref(root))
            present.add(addr)
        for probe in rng.sample(range(0x100000, 0x100000 + 16 * 201, 16), 8):
            in_ref = bool(ref.rbt_find(probe, C.byref(root)))
            assert in_ref == (c.lookup(probe) >= 0) == (probe in present)
    c.close()


def test_queue_exactly_once_under_threads(bx):
    L = bx.test_lib()
    n = 200000
    q = L.bxtest_queue_create(n)
    got = [[] for _ in range(8)]

    def worker(k):
        while True:
            v = L.bxtest_queue_next(q)
            if v < 0:
                return
            got[k].append(v)

    th = [threading.Thread(target=worker, args=(k,)) for k in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    allv = sorted(v for g in got for v in g)
    assert allv == list(range(n))
    for g in got:  # each consumer sees a descending sequence
        assert g == sorted(g, reverse=True)
    L.bxtest_queue_destroy(q)


def test_host_copy_pool_strided(bx):
    """The parallel staging copy moves `height` columns of `width` bytes
    between pitched buffers exactly (multi-threaded above 1 MiB)."""
    import numpy as np

    L = bx.test_lib()
    L.bxtest_host_copy2d.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t,
                                     C.c_size_t]
    rng = np.random.default_rng(3)
    for rows, cols, lds, ldd in [(5, 3, 7, 5), (1000, 700, 1003, 1000), (2048, 513, 4000, 2050)]:
        src = rng.standard_normal((lds, cols))
        src = np.asfortranarray(src)
        dst = np.asfortranarray(np.full((ldd, cols), -7.0))
        L.bxtest_host_copy2d(dst.ctypes.data, ldd * 8, src.ctypes.data, lds * 8, rows * 8, cols)
        assert np.array_equal(dst[:rows], src[:rows])
        assert np.all(dst[rows:] == -7.0)


def test_queue_order_matches_reference(bx, orc):
    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference tree here)")
    L = bx.test_lib()
    for n in (1, 2, 17, 256):
        rq = (C.c_int * 1)()
        # the reference's blasx_dgemm initialises with task_num = tiles - 1
        ref.init_queue(rq, n - 1)
        ref_seq = []
        while True:
            v = ref.dequeue(rq)
            if v < 0:
                break
            ref_seq.append(v)
        q = L.bxtest_queue_create(n)
        mine = []
        while True:
            v = L.bxtest_queue_next(q)
            if v < 0:
                break
            mine.append(v)
        L.bxtest_queue_destroy(q)
        assert mine == ref_seq == list(range(n - 1, -1, -1))


def _plan_sim(bx, M, N, K, T, cap_slots, dt=1, opa=0, opb=0, tri=0, same_ab=0):
    L = bx.test_lib()
    ob, ev = C.c_uint64(), C.c_uint64()
    mode, nb = C.c_int(), C.c_int()
    rc = L.bxtest_plan_simulate(dt, opa, opb, M, N, K, T, tri, same_ab, cap_slots, C.byref(ob),
                                C.byref(ev), C.byref(mode), C.byref(nb))
    return rc, ob.value, ev.value, mode.value, nb.value


@pytest.mark.parametrize("n", [65536, 131072])
@pytest.mark.parametrize("frac", [0.4, 0.75, 1.2])
def test_out_of_core_traffic_bound(bx, n, frac):
    es, T = 8, 2048
    slots = int(frac * 2 * n * n * es // (T * T * es))
    rc, ob, ev, mode, nb = _plan_sim(bx, n, n, n, T, slots)
    assert rc == 0
    ab = 2 * n * n * es
    if frac >= 0.75:
        assert ob == ab and nb == 1  # every operand tile once
    else:
        assert nb == 2 and ab * 1.4 <= ob <= ab * 1.5, (ob / ab, nb)
        total = ob + 2 * n * n * es
        assert total <= 1.25 * (ab + 2 * n * n * es)
    if frac > 1:
        assert ev == 0


@pytest.mark.parametrize("opa,opb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_out_of_core_order_rectangular_picks_cheaper_role(bx, opa, opb):
    """T = 512 (rounds of 4 x 4 tiles), ragged edges, every transpose.  A
    tall A (21 x 7 tiles) with a cache of 190 slots: A does not fit beside
    the staging, B (7 x 10 tiles) does with two A row panels, so B stays
    resident (bands of block columns) and every operand tile is fetched
    exactly once.  The mirrored shape keeps A resident instead."""
    T, es = 512, 8
    M, N, K = 20 * T + 37, 9 * T + 5, 6 * T + 11
    rc, ob, ev, mode, nb = _plan_sim(bx, M, N, K, T, 190, opa=opa, opb=opb)
    assert rc == 0 and (mode, nb) == (0, 1) and ev > 0
    assert ob == es * (M * K + K * N)
    rc, ob, ev, mode, nb = _plan_sim(bx, N, M, K, T, 190, opa=opa, opb=opb)
    assert rc == 0 and (mode, nb) == (1, 1) and ev > 0
    assert ob == es * (M * K + K * N)


def test_out_of_core_order_matches_brute_force_reuse(bx):
    """Against an exhaustive count on a small grid: operand tiles fetched
    by the replay never exceed what a per-round refetch of every tile
    (the reference's pattern, blas/blasx_dgemm.c:164-239) would move, and
    equal |A|+|B| when the cache holds both."""
    T, es = 256, 8
    for (M, N, K) in [(9 * T, 7 * T, 5 * T), (9 * T + 3, 7 * T + 100, 5 * T + 1)]:
        mt, nt, kt = -(-M // T), -(-N // T), -(-K // T)
        full = es * (M * K + K * N)
        rc, ob, _, _, _ = _plan_sim(bx, M, N, K, T, 4096)
        assert rc == 0 and ob == full
        rc, ob, _, _, nb = _plan_sim(bx, M, N, K, T, 3 * 16 + 2 * 8 * kt + 4 * kt + 10)
        assert rc == 0
        refetch = es * T * T * kt * (mt * nt) * 2
        assert full <= ob < refetch
