"""GPU parity of the rank-k updates (cblas_{s,d,z}syrk, cblas_zherk, the
f77 forms) against the oracle's naive GEMM with B = A on the stored
triangle.  The reference computes these on the CPU BLAS (blas/dsyrk.c:46-58);
here host operands run on the tile runtime (C tiles of the triangle only,
diagonal tiles masked) and HBM operands as one grouped launch.

Tolerance: ||C_gpu - C_ref||_F / (|alpha| ||A||_F^2 + |beta| ||C0||_F) <= c K eps
on the stored triangle, c = 1 real, 2 complex.  The other triangle must come
back bit for bit, and integer-valued inputs must match the naive loop
exactly."""
import sys

import numpy as np
import pytest

from syrk_common import COL, CTR, LO, NOTR, ROW, TR, UP, cblas_syrk, f77_syrk, mask, reference

pytestmark = pytest.mark.gpu

EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps,
       np.complex128: np.finfo(np.float64).eps, np.complex64: np.finfo(np.float32).eps}
DTS = [(np.float64, 0), (np.float32, 0), (np.complex128, 0), (np.complex128, 1),
       (np.complex64, 0), (np.complex64, 1)]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _data(orc, dt, herk, trans, N, K, seed):
    shape = (N, K) if trans == NOTR else (K, N)
    gdt = np.complex128 if dt == np.complex64 else dt
    A = np.asfortranarray(orc.fill_uniform(shape[0], shape[1], seed, 0, dtype=gdt).astype(dt))
    C0 = np.asfortranarray(orc.fill_uniform(N, N, seed, 2, dtype=gdt).astype(dt))
    return A, C0, shape[0]


def _scalars(dt, herk):
    if herk or np.dtype(dt).kind != "c":
        return 1.5, 0.5
    return 1.5 - 0.5j, 0.5 + 0.25j


def _check(out, ref, C0, A, alpha, beta, uplo, K, dt, what):
    m = mask(out.shape[0], uplo)
    num = np.linalg.norm((out[m] - ref[m]).ravel())
    den = abs(alpha) * np.linalg.norm(A.ravel()) ** 2 + abs(beta) * np.linalg.norm(C0.ravel())
    tol = (2 if np.dtype(dt).kind == "c" else 1) * K * EPS[dt]
    assert num / den <= tol, f"{what}: rel err {num / den:.3e} > {tol:.3e}"
    assert np.array_equal(out[~m].view(np.uint8), C0[~m].view(np.uint8)), \
        f"{what}: the other triangle changed"


def _trans_list(herk):
    return [NOTR, CTR if herk else TR]


def _nk(dt):
    # above the GPU thresholds (1000 real, 500 complex), ragged against 256 tiles
    return (600, 300) if np.dtype(dt).kind == "c" else (1100, 300)


@pytest.mark.parametrize("dt,herk", DTS)
def test_gpu_syrk_host_operands_tiles(gpu, orc, dt, herk):
    bx = gpu
    L = bx.lib()
    code = bx.dtype_code(dt)
    assert L.blasx_set_tile_size(code, 256) == 0
    try:
        N, K = _nk(dt)
        alpha, beta = _scalars(dt, herk)
        for uplo in (UP, LO):
            for trans in _trans_list(herk):
                say("host syrk", np.dtype(dt).name, herk, uplo, trans, N, K)
                A, C0, lda = _data(orc, dt, herk, trans, N, K, 3)
                ref = reference(orc, dt, herk, uplo, trans, N, K, alpha, A, lda, beta, C0)
                bx.reset_stats()
                out = C0.copy(order="F")
                cblas_syrk(L, dt, herk, COL, uplo, trans, N, K, alpha, A, lda, beta, out, N)
                st = bx.stats()
                assert st["calls_gpu"] == 1 and st["calls_cpu"] == 0
                _check(out, ref, C0, A, alpha, beta, uplo, K, dt, "cblas")
                out = C0.copy(order="F")
                f77_syrk(L, dt, herk, uplo, trans, N, K, alpha, A, lda, beta, out, N)
                _check(out, ref, C0, A, alpha, beta, uplo, K, dt, "f77")
    finally:
        L.blasx_set_tile_size(code, 0)


def test_gpu_syrk_row_major_host(gpu, orc):
    bx = gpu
    L = bx.lib()
    N, K = 1100, 300
    A, C0, lda = _data(orc, np.float64, 0, NOTR, N, K, 4)
    ref = reference(orc, np.float64, 0, LO, NOTR, N, K, 1.5, A, lda, 0.5, C0)
    say("row-major dsyrk", N, K)
    out = C0.copy(order="F")
    # the same bytes read row-major: C^T, the upper triangle, A^T
    cblas_syrk(L, np.float64, 0, ROW, UP, TR, N, K, 1.5, A, lda, 0.5, out, N)
    _check(out, ref, C0, A, 1.5, 0.5, LO, K, np.float64, "row-major")


@pytest.mark.parametrize("dt,herk", DTS)
def test_gpu_syrk_device_operands(gpu, orc, dt, herk):
    """operands in HBM: one grouped launch of off-diagonal panels and masked
    diagonal blocks; K = 320 makes the aligned panels DMA-eligible, N leaves
    a ragged last panel"""
    bx = gpu
    L = bx.lib()
    N, K = (600, 320) if np.dtype(dt).kind == "c" else (1100, 320)
    alpha, beta = _scalars(dt, herk)
    for uplo in (UP, LO):
        for trans in _trans_list(herk):
            say("device syrk", np.dtype(dt).name, herk, uplo, trans, N, K)
            A, C0, lda = _data(orc, dt, herk, trans, N, K, 5)
            ref = reference(orc, dt, herk, uplo, trans, N, K, alpha, A, lda, beta, C0)
            dA, dC = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(C0)
            cblas_syrk(L, dt, herk, COL, uplo, trans, N, K, alpha, dA.ptr, lda, beta, dC.ptr, N)
            _check(dC.download(), ref, C0, A, alpha, beta, uplo, K, dt, "device")
            dA.free()
            dC.free()


def test_gpu_syrk_exact_integers(gpu, orc):
    """integer-valued A: every partial sum is exact, so both GPU paths must
    equal the naive loop bit for bit (ragged and aligned shapes)"""
    bx = gpu
    L = bx.lib()
    rng = np.random.default_rng(11)
    assert L.blasx_set_tile_size(bx.D, 256) == 0
    try:
        for (N, K) in ((1100, 300), (1024, 320)):
            for uplo in (UP, LO):
                for trans in (NOTR, TR):
                    say("exact dsyrk", N, K, uplo, trans)
                    shape = (N, K) if trans == NOTR else (K, N)
                    A = np.asfortranarray(rng.integers(-8, 9, shape).astype(np.float64))
                    C0 = np.asfortranarray(rng.integers(-64, 65, (N, N)).astype(np.float64))
                    ref = reference(orc, np.float64, 0, uplo, trans, N, K, 3.0, A, shape[0],
                                    -2.0, C0)
                    out = C0.copy(order="F")
                    cblas_syrk(L, np.float64, 0, COL, uplo, trans, N, K, 3.0, A, shape[0], -2.0,
                               out, N)
                    assert np.array_equal(out, ref), "host path"
                    dA, dC = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(C0)
                    cblas_syrk(L, np.float64, 0, COL, uplo, trans, N, K, 3.0, dA.ptr, shape[0],
                               -2.0, dC.ptr, N)
                    assert np.array_equal(dC.download(), ref), "device path"
                    dA.free()
                    dC.free()
    finally:
        L.blasx_set_tile_size(bx.D, 0)


def test_gpu_syrk_beta_zero_does_not_read_c(gpu, orc):
    """beta = 0: NaN in the stored triangle of C must not reach the result;
    NaN in the other triangle must come back unchanged"""
    bx = gpu
    L = bx.lib()
    N, K = 1100, 320
    for herk, dt in ((0, np.float64), (1, np.complex128)):
        n = 600 if herk else N
        A, _, lda = _data(orc, dt, herk, NOTR, n, K, 6)
        C0 = np.full((n, n), np.nan, dtype=dt, order="F")
        ref = reference(orc, dt, herk, LO, NOTR, n, K, 1.0, A, lda, 0.0, np.zeros_like(C0))
        m = mask(n, LO)
        for where in ("host", "device"):
            say("beta=0", where, np.dtype(dt).name, herk)
            if where == "host":
                out = C0.copy(order="F")
                cblas_syrk(L, dt, herk, COL, LO, NOTR, n, K, 1.0, A, lda, 0.0, out, n)
            else:
                dA, dC = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(C0)
                cblas_syrk(L, dt, herk, COL, LO, NOTR, n, K, 1.0, dA.ptr, lda, 0.0, dC.ptr, n)
                out = dC.download()
                dA.free()
                dC.free()
            assert not np.isnan(out[m]).any()
            assert np.max(np.abs(out[m] - ref[m])) <= 1e-12 * K
            assert np.isnan(out[~m]).all()


def test_gpu_syrk_alpha_zero_device(gpu, orc):
    """alpha = 0 on HBM operands: C = beta C on the triangle only (and a real
    diagonal for HERK), A not read"""
    bx = gpu
    L = bx.lib()
    n = 1100  # above the GPU threshold of every type (CPU-sized calls go to the CPU BLAS)
    for herk, dt, beta in ((0, np.float64, 0.5), (1, np.complex128, 0.5), (0, np.complex128,
                                                                            0.5 - 1j)):
        say("alpha=0 device", np.dtype(dt).name, herk)
        C0 = orc.fill_uniform(n, n, 7, 2, dtype=dt)
        A = orc.fill_uniform(n, 8, 7, 0, dtype=dt)
        ref = reference(orc, dt, herk, UP, NOTR, n, 8, 0.0, A, n, beta, C0)
        dA, dC = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(C0)
        cblas_syrk(L, dt, herk, COL, UP, NOTR, n, 8, 0.0, dA.ptr, n, beta, dC.ptr, n)
        out = dC.download()
        m = mask(n, UP)
        assert np.max(np.abs(out[m] - ref[m])) <= 1e-15
        assert np.array_equal(out[~m].view(np.uint8), C0[~m].view(np.uint8))
        dA.free()
        dC.free()
