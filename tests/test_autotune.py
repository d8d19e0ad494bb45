"""Tile-edge autotuner for host-resident GEMMs (the reference fixes the edge
at compile time, include/blasx.h:5-7)."""
import ctypes as C

import numpy as np
import pytest


def test_tune_table_lookup_cpu(bx):
    """A measured entry overrides the heuristic (or the built-in MI355X table,
    csrc/tune_table.inc, whose entries are cubes) for nearby shapes only.
    Shapes here are far from any cube."""
    L = bx.test_lib()
    L.blasx_tune_table_add.argtypes = [C.c_int] * 5
    assert L.bxtest_autotune_tile(bx.D, 200000, 3000, 3000) == 2048  # heuristic
    assert L.blasx_tune_table_add(bx.D, 200000, 3000, 3000, 512) == 0
    assert L.bxtest_autotune_tile(bx.D, 202000, 2970, 3000) == 512
    assert L.bxtest_autotune_tile(bx.D, 200000, 300000, 3000) == 2048
    assert L.bxtest_autotune_tile(bx.S, 200000, 3000, 3000) == 2048


def test_builtin_tune_table_is_consulted_by_default(bx):
    """Every entry of the built-in table (measured on an MI355X by
    blasx_amd/tune.py) is what the per-call choice returns at its own shape
    with no environment table: the tuned edge is on the default path."""
    import os
    import re

    L = bx.test_lib()
    path = os.path.join(os.path.dirname(bx.__file__), "csrc", "tune_table.inc")
    rows = re.findall(r"\{(\d+), (\d+), (\d+), (\d+), (\d+)\}", open(path).read())
    for dt, M, N, K, T in (tuple(int(v) for v in r) for r in rows):
        assert T in (256, 512, 1024, 2048)
        assert L.bxtest_autotune_tile(dt, M, N, K) == T, (dt, M, N, K, T)


def _grouped_table(bx):
    import os
    import re

    path = os.path.join(os.path.dirname(bx.__file__), "csrc", "tune_grouped.inc")
    body = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)
    rows = []
    for r in re.findall(r"\{(-?\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+)\}",
                        body):
        dt, M, N, K, load, *t = (int(v) for v in r)
        rows.append((dt, M, N, K, load, min(range(3), key=lambda i: t[i]), t))
    return rows


def test_grouped_shape_choice_cpu(bx):
    """Block tile shape per grouped problem: complex problems always keep the
    large tile; a forced shape wins; a run-time entry at the query point wins
    (later entries on ties); every built-in entry's fastest shape is returned
    at its own point, and between two measured loads of one workload the
    shape whose log-interpolated time is least; far from every measured
    workload (or with no table) the load and the problem's width decide."""
    L = bx.lib()
    assert L.blasx_grouped_shape(bx.Z, 100, 100, 100, 1) == 0
    assert L.blasx_set_grouped_shape(bx.Z, 1) != 0
    assert L.blasx_set_grouped_shape(bx.D, 3) != 0
    try:
        assert L.blasx_set_grouped_shape(bx.D, 2) == 0
        assert L.blasx_grouped_shape(bx.D, 5000, 5000, 5000, 100000) == 2
    finally:
        L.blasx_set_grouped_shape(bx.D, -1)
    import math

    table = _grouped_table(bx)
    for dt, M, N, K, load, shape, _ in table:
        assert L.blasx_grouped_shape(dt, M, N, K, load) == shape, (dt, M, N, K, load, shape)
    # between consecutive measured loads of one workload: the interpolated
    # winner, computed here from the table
    by_key = {}
    for dt, M, N, K, load, shape, t in table:
        by_key.setdefault((dt, M, N, K), []).append((load, t))
    checked = 0
    for (dt, M, N, K), pts in by_key.items():
        pts.sort()
        for (l0, t0), (l1, t1) in zip(pts, pts[1:]):
            q = int(round(math.sqrt(l0 * l1)))
            if q in (l0, l1):
                continue
            f = (math.log(q) - math.log(l0)) / (math.log(l1) - math.log(l0))
            lt = [math.log(a) + f * (math.log(b) - math.log(a)) for a, b in zip(t0, t1)]
            best = sorted(lt)
            if best[1] - best[0] < 1e-6:
                continue  # a tie: either answer is right
            assert L.blasx_grouped_shape(dt, M, N, K, q) == lt.index(best[0]), (dt, M, N, K, q)
            checked += 1
    assert checked > 0 or not table
    if not [r for r in table if r[0] == bx.S]:
        cus = 256  # no GPU here: the library assumes an MI355X's count
        assert L.blasx_grouped_shape(bx.S, 300, 300, 300, 2 * cus) == 0
        assert L.blasx_grouped_shape(bx.S, 300, 300, 300, cus) == 1
        assert L.blasx_grouped_shape(bx.S, 300, 300, 300, cus // 2 - 1) == 2
    # workloads far from every measured one take the heuristic: a full
    # launch keeps the large tile (TRMM's 256-row panels) unless a problem
    # is 64 wide, a small one takes the small tiles
    assert L.blasx_grouped_shape(bx.D, 256, 32768, 16384, 65536) == 0
    assert L.blasx_grouped_shape(bx.D, 40000, 48, 1000, 65536) == 1
    assert L.blasx_grouped_shape(bx.D, 48, 40000, 1000, 65536) == 2
    assert L.blasx_grouped_shape(bx.D, 100, 20000, 300, 100) == 2
    # run-time entries (an unusual point, so no built-in entry is nearer)
    assert L.blasx_grouped_shape_add(bx.S, 77, 7777, 7, 3, 1) == 0
    assert L.blasx_grouped_shape(bx.S, 77, 7777, 7, 3) == 1
    assert L.blasx_grouped_shape_add(bx.S, 77, 7777, 7, 3, 2) == 0
    assert L.blasx_grouped_shape(bx.S, 77, 7777, 7, 3) == 2
    assert L.blasx_grouped_shape_add(bx.S, 77, 7777, 7, 3, 5) != 0
    assert L.blasx_grouped_shape_add(bx.Z, 77, 7777, 7, 3, 1) != 0


@pytest.mark.gpu
def test_autotuner_measures_and_records(gpu):
    bx = gpu
    L = bx.lib()
    L.blasx_autotune_tile_size.argtypes = [C.c_int] * 5 + [C.POINTER(C.c_int),
                                                           C.POINTER(C.c_double)]
    best = C.c_int()
    times = (C.c_double * 4)()
    print("[case] autotune D 3000x2500x2000", flush=True)
    assert L.blasx_autotune_tile_size(bx.D, 3000, 2500, 2000, 1, C.byref(best), times) == 0
    assert best.value in (256, 512, 1024, 2048)
    measured = [times[i] for i in range(4)]
    assert all(t > 0 for t in measured)
    assert measured[[256, 512, 1024, 2048].index(best.value)] == min(measured)
    L.blasx_get_tile_size.argtypes = [C.c_int] * 4
    assert L.blasx_get_tile_size(bx.D, 3000, 2500, 2000) == best.value
