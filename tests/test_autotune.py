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
