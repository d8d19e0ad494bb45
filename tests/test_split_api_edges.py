"""Degenerate calls of the split-precision device entry points (no reference
counterpart; BLAS semantics for the scalars): K = 0 or alpha = 0 scale C by
beta, beta = 0 overwrites C without reading it (NaN in C does not survive),
M = 0 or N = 0 leaves C alone, invalid sizes and leading dimensions are
rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest


def say(*a):
    print("[case]", *a, flush=True)


def _real_calls(bx):
    """(name, call(M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc) -> rc)"""
    L = bx.lib()
    return [
        ("crt", lambda *a: L.blasx_dgemm_crt_device(111, 111, *a, 16, None)),
        ("digit", lambda *a: L.blasx_dgemm_split_device(111, 111, *a, 7, None)),
        ("bf16x3", lambda *a: L.blasx_dgemm_bf16x3_device(111, 111, *a, None)),
    ]


def _cplx_calls(bx):
    L = bx.lib()

    def wrap(fn, last):
        def call(M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc):
            al = (C.c_double * 2)(alpha.real, alpha.imag)
            be = (C.c_double * 2)(beta.real, beta.imag)
            return fn(111, 111, M, N, K, C.cast(al, C.c_void_p), dA, lda, dB, ldb,
                      C.cast(be, C.c_void_p), dC, ldc, last, None)
        return call

    return [("zcrt", wrap(L.blasx_zgemm_crt_device, 16)),
            ("zdigit", wrap(L.blasx_zgemm_split_device, 7))]


@pytest.mark.gpu
def test_gpu_split_degenerate_real(gpu):
    bx = gpu
    rng = np.random.default_rng(41)
    M, N, K = 50, 37, 40
    A = np.asfortranarray(rng.standard_normal((M, K)))
    B = np.asfortranarray(rng.standard_normal((K, N)))
    C0 = np.asfortranarray(rng.standard_normal((M, N)))
    dA, dB = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(B)
    for name, call in _real_calls(bx):
        # K = 0: C = beta C
        dC = bx.DeviceArray.from_host(C0)
        say(name, "K=0")
        assert call(M, N, 0, 1.5, dA.ptr, M, dB.ptr, K, 0.5, dC.ptr, M) == 0
        bx.synchronize()
        assert np.array_equal(dC.download(), 0.5 * C0), name
        # alpha = 0: C = beta C, A and B not used
        say(name, "alpha=0")
        assert call(M, N, K, 0.0, dA.ptr, M, dB.ptr, K, -2.0, dC.ptr, M) == 0
        bx.synchronize()
        assert np.array_equal(dC.download(), -2.0 * (0.5 * C0)), name
        # beta = 0 over NaN: C is overwritten
        dN = bx.DeviceArray.from_host(np.full((M, N), np.nan, order="F"))
        say(name, "alpha=0 beta=0 over NaN")
        assert call(M, N, K, 0.0, dA.ptr, M, dB.ptr, K, 0.0, dN.ptr, M) == 0
        bx.synchronize()
        assert np.array_equal(dN.download(), np.zeros((M, N))), name
        say(name, "beta=0 over NaN, product")
        assert call(M, N, K, 1.0, dA.ptr, M, dB.ptr, K, 0.0, dN.ptr, M) == 0
        bx.synchronize()
        out = dN.download()
        assert np.all(np.isfinite(out)), name
        assert np.all(np.abs(out - A @ B) <= 1e-3 * (np.abs(A) @ np.abs(B))), name
        # M = 0, N = 0: nothing happens
        dC2 = bx.DeviceArray.from_host(C0)
        say(name, "M=0, N=0")
        assert call(0, N, K, 1.0, dA.ptr, 1, dB.ptr, K, 0.0, dC2.ptr, 1) == 0
        assert call(M, 0, K, 1.0, dA.ptr, M, dB.ptr, K, 0.0, dC2.ptr, M) == 0
        bx.synchronize()
        assert np.array_equal(dC2.download(), C0), name
        # invalid: negative size, short leading dimensions
        say(name, "invalid arguments")
        assert call(-1, N, K, 1.0, dA.ptr, M, dB.ptr, K, 0.0, dC2.ptr, M) != 0
        assert call(M, N, K, 1.0, dA.ptr, M - 1, dB.ptr, K, 0.0, dC2.ptr, M) != 0
        assert call(M, N, K, 1.0, dA.ptr, M, dB.ptr, K - 1, 0.0, dC2.ptr, M) != 0
        assert call(M, N, K, 1.0, dA.ptr, M, dB.ptr, K, 0.0, dC2.ptr, M - 1) != 0
        bx.synchronize()
        assert np.array_equal(dC2.download(), C0), name


@pytest.mark.gpu
def test_gpu_split_degenerate_complex(gpu):
    bx = gpu
    rng = np.random.default_rng(42)
    M, N, K = 30, 21, 17
    A = np.asfortranarray(rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K)))
    B = np.asfortranarray(rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N)))
    C0 = np.asfortranarray(rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N)))
    dA, dB = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(B)
    beta = 0.5 - 0.25j
    for name, call in _cplx_calls(bx):
        dC = bx.DeviceArray.from_host(C0)
        say(name, "K=0")
        assert call(M, N, 0, 1 + 1j, dA.ptr, M, dB.ptr, K, beta, dC.ptr, M) == 0
        bx.synchronize()
        out, want = dC.download(), beta * C0
        # beta C in complex arithmetic: within a few ulps of numpy's form
        assert np.all(np.abs(out - want) <= 2.0 ** -50 * abs(beta) * np.abs(C0)), name
        dN = bx.DeviceArray.from_host(np.full((M, N), np.nan + 0j, order="F"))
        say(name, "alpha=0 beta=0 over NaN")
        assert call(M, N, K, 0j, dA.ptr, M, dB.ptr, K, 0j, dN.ptr, M) == 0
        bx.synchronize()
        assert np.array_equal(dN.download(), np.zeros((M, N), dtype=complex)), name
        say(name, "invalid arguments")
        assert call(M, -1, K, 1 + 0j, dA.ptr, M, dB.ptr, K, 0j, dC.ptr, M) != 0
        assert call(M, N, K, 1 + 0j, dA.ptr, M, dB.ptr, K, 0j, dC.ptr, M - 1) != 0
