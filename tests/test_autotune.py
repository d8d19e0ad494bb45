"""Tile-edge autotuner for host-resident GEMMs (the reference fixes the edge
at compile time, include/blasx.h:5-7)."""
import ctypes as C

import numpy as np
import pytest


def test_tune_table_lookup_cpu(bx):
    L = bx.test_lib()
    L.blasx_tune_table_add.argtypes = [C.c_int] * 5
    # a measured entry overrides the heuristic for nearby shapes only
    assert L.bxtest_autotune_tile(bx.D, 10000, 10000, 10000) == 2048  # heuristic
    assert L.blasx_tune_table_add(bx.D, 10000, 10000, 10000, 512) == 0
    assert L.bxtest_autotune_tile(bx.D, 10100, 9900, 10000) == 512
    assert L.bxtest_autotune_tile(bx.D, 40000, 40000, 40000) == 2048
    assert L.bxtest_autotune_tile(bx.S, 10000, 10000, 10000) == 2048


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
