"""GEMMs through the cblas_?gemm ABI whose operands are already in HBM.

Device pointers must never reach host code: not the host-side C = beta*C of
alpha == 0 or K == 0, and not the CPU BLAS that serves CPU-sized calls (the
reference's blas/dgemm.c:61-74 sends M,N,K <= 1000 to the CPU).  Each case
runs on the device, is counted as a GPU call, and is checked against the
oracle (closed forms exactly; products within c*K*eps, c = 1 real, 2 complex).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [np.float64, np.float32, np.complex128, np.complex64]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _mat(orc, rows, cols, dt, seed, mid):
    if np.dtype(dt) == np.complex64:
        return np.asfortranarray(orc.fill_uniform(rows, cols, seed, mid, dtype=np.complex128)
                                 .astype(np.complex64))
    return orc.fill_uniform(rows, cols, seed, mid, dtype=dt)


def _wide(dt):
    return np.complex128 if np.dtype(dt).kind == "c" else np.float64


def _eps(dt):
    return np.finfo(np.float32 if np.dtype(dt) in (np.float32, np.complex64) else np.float64).eps


def _ref(orc, ta, tb, M, N, K, alpha, A, B, beta, C0):
    """the oracle on the exactly widened inputs (FP64 / complex128)"""
    w = _wide(C0.dtype)
    ref = C0.astype(w).copy(order="F")
    orc.gemm(ta, tb, M, N, K, alpha, A.astype(w), A.shape[0], B.astype(w), B.shape[0], beta,
             ref, M)
    return ref


def _call(bx, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, Cd, ldc):
    bx.reset_stats()
    bx.cblas_gemm(bx.COL_MAJOR, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, Cd, ldc)
    st = bx.stats()
    assert st["calls_gpu"] == 1 and st["calls_cpu"] == 0 and st["cpu_fallbacks"] == 0, st


def _assert_scaled(out, C0, beta, dt, what):
    """C = beta*C0: one rounded product per element for real types (exact
    equality); complex within two roundings of each part"""
    if np.dtype(dt).kind != "c":
        assert np.array_equal(out, (np.asarray(beta, dtype=dt) * C0).astype(dt)), what
        return
    w = _wide(dt)
    expect = beta * C0.astype(w)
    tol = 4 * _eps(dt) * abs(beta) * np.abs(C0.astype(w))
    assert np.all(np.abs(out.astype(w) - expect) <= tol), what


@pytest.mark.parametrize("dt", CASES)
def test_alpha_zero_and_k_zero_scale_c_in_hbm(gpu, orc, dt):
    bx = gpu
    M, N, K = 1100, 700, 900
    cplx = np.dtype(dt).kind == "c"
    A = _mat(orc, M, K, dt, 21, 0)
    B = _mat(orc, K, N, dt, 21, 1)
    C0 = _mat(orc, M, N, dt, 21, 2)
    beta = (0.5 - 0.25j) if cplx else 0.5
    dA, dB = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(B)
    for what, (alpha, k) in (("alpha=0", (0.0, K)), ("K=0", (1.5, 0))):
        say("hbm", np.dtype(dt).name, what)
        dC = bx.DeviceArray.from_host(C0)
        _call(bx, "N", "N", M, N, k, alpha, dA, M, dB, max(k, 1), beta, dC, M)
        out = dC.download()
        _assert_scaled(out, C0, beta, dt, what)
        # beta = 0 writes zeros without reading C (NaN in C)
        dC.upload(np.full((M, N), np.nan, dtype=dt))
        _call(bx, "N", "N", M, N, k, alpha, dA, M, dB, max(k, 1), 0.0, dC, M)
        assert np.array_equal(dC.download(), np.zeros((M, N), dt)), what + " beta=0"
        dC.free()
    # alpha = 0 with C in HBM never reads A or B: host NaN buffers are fine
    nan = np.full((M, K), np.nan, dtype=dt, order="F")
    dC = bx.DeviceArray.from_host(C0)
    say("hbm", np.dtype(dt).name, "alpha=0 host operands")
    _call(bx, "N", "N", M, N, K, 0.0, nan, M, nan, K, beta, dC, M)
    _assert_scaled(dC.download(), C0, beta, dt, "alpha=0 host NaN operands")
    for d in (dA, dB, dC):
        d.free()


@pytest.mark.parametrize("dt", CASES)
def test_cpu_sized_call_on_hbm_runs_on_the_device(gpu, orc, dt):
    """M = N = K = 300 is below both size thresholds; with operands in HBM it
    must still run on the GPU, against the oracle; beta = 0 over NaN"""
    bx = gpu
    n = 300
    cplx = np.dtype(dt).kind == "c"
    alpha, beta = ((1.5 - 0.5j), (0.25 + 1j)) if cplx else (1.5, -0.5)
    trans = [("N", "N"), ("T", "N"), ("N", "C" if cplx else "T")]
    for ta, tb in trans:
        say("hbm cpu-sized", np.dtype(dt).name, ta, tb)
        A, B, C0 = (_mat(orc, n, n, dt, 22, i) for i in range(3))
        ref = _ref(orc, ta, tb, n, n, n, alpha, A, B, beta, C0)
        dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
        _call(bx, ta, tb, n, n, n, alpha, dA, n, dB, n, beta, dC, n)
        w = _wide(dt)
        err = orc.relerr_fro(dC.download().astype(w), ref, alpha, A.astype(w), B.astype(w), beta,
                             C0.astype(w))
        assert err <= (2 if cplx else 1) * n * _eps(dt), (ta, tb, err)
        dC.upload(np.full((n, n), np.nan, dtype=dt))
        _call(bx, ta, tb, n, n, n, alpha, dA, n, dB, n, 0.0, dC, n)
        out = dC.download()
        assert np.isfinite(out).all()  # complex: both parts
        ref0 = _ref(orc, ta, tb, n, n, n, alpha, A, B, 0.0, C0)
        err = orc.relerr_fro(out.astype(w), ref0, alpha, A.astype(w), B.astype(w), 0.0,
                             C0.astype(w))
        assert err <= (2 if cplx else 1) * n * _eps(dt), (ta, tb, "beta=0", err)
        for d in (dA, dB, dC):
            d.free()


@pytest.mark.parametrize("dt", [np.float64, np.complex128])
def test_operands_split_between_host_and_hbm(gpu, orc, dt):
    """A in HBM, B and C on the host (and the reverse): in core on the GPU"""
    bx = gpu
    M, N, K = 1200, 650, 1030
    cplx = np.dtype(dt).kind == "c"
    alpha, beta = ((0.5 + 1j), (-1 + 0.5j)) if cplx else (1.25, -0.75)
    A, B, C0 = _mat(orc, M, K, dt, 23, 0), _mat(orc, K, N, dt, 23, 1), _mat(orc, M, N, dt, 23, 2)
    ref = _ref(orc, "N", "N", M, N, K, alpha, A, B, beta, C0)
    tol = (2 if cplx else 1) * K * _eps(dt)
    say("split A hbm", np.dtype(dt).name)
    dA = bx.DeviceArray.from_host(A)
    out = C0.copy(order="F")
    _call(bx, "N", "N", M, N, K, alpha, dA, M, B, K, beta, out, M)
    assert orc.relerr_fro(out, ref, alpha, A, B, beta, C0) <= tol
    say("split C hbm", np.dtype(dt).name)
    dC = bx.DeviceArray.from_host(C0)
    _call(bx, "N", "N", M, N, K, alpha, A, M, B, K, beta, dC, M)
    assert orc.relerr_fro(dC.download(), ref, alpha, A, B, beta, C0) <= tol
    dA.free()
    dC.free()
