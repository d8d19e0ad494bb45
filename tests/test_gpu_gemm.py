"""GPU parity of the GEMM hot path against the CPU oracle.

Tolerance (stated once, used everywhere): relative Frobenius error
  ||C_gpu - C_ref||_F / (|alpha| ||A||_F ||B||_F + |beta| ||C0||_F) <= c * K * eps
with c = 1 (real) or 2 (complex) -- the gamma_K bound of Higham 3.5 -- and
eps of the working precision.  Integer-valued inputs make every partial sum
exact, so there the GPU must equal the naive loop bit for bit.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps,
       np.complex128: np.finfo(np.float64).eps}


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _op(X, t):
    return X if t == "N" else (X.T if t == "T" else X.conj().T)


def _mat(orc, rows, cols, dt, seed, mid):
    return orc.fill_uniform(rows, cols, seed, mid, dtype=dt)


def _tol(dt, K):
    return (2 if np.dtype(dt).kind == "c" else 1) * max(K, 1) * EPS[dt]


def _check(orc, out, ref, alpha, A, B, beta, C0, K, dt, what):
    err = orc.relerr_fro(out, ref, alpha, A, B, beta, C0)
    tol = _tol(dt, K)
    assert err <= tol, f"{what}: rel err {err:.3e} > {tol:.3e}"
    return err


TRANS_REAL = [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")]
TRANS_CPLX = [(a, b) for a in "NTC" for b in "NTC"]


# ------------------------------------------------------------------ device API
@pytest.mark.parametrize("dt", [np.float64, np.float32, np.complex128])
def test_device_gemm_all_transposes_ragged(gpu, orc, dt):
    bx = gpu
    M, N, K = 261, 197, 133
    alpha, beta = (1.5 - 0.5j, 0.5 + 0.25j) if np.dtype(dt).kind == "c" else (1.5, 0.5)
    for (ta, tb) in (TRANS_CPLX if np.dtype(dt).kind == "c" else TRANS_REAL):
        say("device", np.dtype(dt).name, ta, tb, M, N, K)
        A = _mat(orc, M if ta == "N" else K, K if ta == "N" else M, dt, 1, 0)
        B = _mat(orc, K if tb == "N" else N, N if tb == "N" else K, dt, 1, 1)
        C0 = _mat(orc, M, N, dt, 1, 2)
        ref = C0.copy(order="F")
        orc.gemm(ta, tb, M, N, K, alpha, A, A.shape[0], B, B.shape[0], beta, ref, M,
                 acc64=(dt == np.float32))
        dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
        bx.device_gemm(ta, tb, M, N, K, alpha, dA.ptr, A.shape[0], dB.ptr, B.shape[0], beta,
                       dC.ptr, M, dt)
        bx.synchronize()
        out = dC.download()
        _check(orc, out, ref, alpha, A, B, beta, C0, K, dt, f"{ta}{tb}")


@pytest.mark.parametrize("shape", [(128, 128, 16), (256, 384, 512), (1, 1, 1), (7, 300, 5),
                                   (300, 1, 257), (129, 65, 17)])
def test_device_dgemm_shapes_integer_exact(gpu, orc, shape):
    bx = gpu
    M, N, K = shape
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    for ta, tb in TRANS_REAL:
        say("device int-exact", ta, tb, shape)
        A = np.asfortranarray(rng.integers(-256, 257, (M, K) if ta == "N" else (K, M)), dtype=np.float64)
        B = np.asfortranarray(rng.integers(-256, 257, (K, N) if tb == "N" else (N, K)), dtype=np.float64)
        C0 = np.asfortranarray(rng.integers(-1000, 1001, (M, N)), dtype=np.float64)
        ref = C0.copy(order="F")
        orc.gemm(ta, tb, M, N, K, 3.0, A, A.shape[0], B, B.shape[0], -2.0, ref, M)
        dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
        bx.device_gemm(ta, tb, M, N, K, 3.0, dA.ptr, A.shape[0], dB.ptr, B.shape[0], -2.0,
                       dC.ptr, M, np.float64)
        bx.synchronize()
        assert np.array_equal(dC.download(), ref), (ta, tb, shape)


@pytest.mark.parametrize("ta,tb", TRANS_REAL)
def test_device_dgemm_ragged_dma_path(gpu, orc, ta, tb):
    """Ragged DGEMMs on the DMA-staged kernel, one launch over every tile:
    edge tiles clamp their operand rows to the last ones and mask their
    stores, and the last K % 16 columns are staged by bounds-checked loads;
    operands with an odd leading dimension or an 8-byte-aligned base are
    first copied to an aligned workspace when N (for A) / M (for B) reaches
    4096.  Bit-exact on
    integer inputs, including K < 16 (no full K tile), exact multiples of 128
    with a K tail, and two calls in a row sharing the workspace; beta = 0
    ignores NaN in C."""
    bx = gpu
    rng = np.random.default_rng(37)
    # (M, N, K, ld padding, base offset in elements)
    cases = ((300, 270, 1007, "even", 0), (256, 384, 40, "even", 0), (131, 129, 9, "even", 0),
             (4097, 4100, 61, "none", 0), (4100, 4097, 45, "even", 1), (640, 4096, 33, "none", 0))
    for (M, N, K, padk, off) in cases:
        say("device dgemm ragged", ta, tb, M, N, K, padk, off)
        ra, ca = (M, K) if ta == "N" else (K, M)
        rb, cb = (K, N) if tb == "N" else (N, K)
        lda = ra + (ra % 2 if padk == "even" else 0)
        ldb = rb + (rb % 2 if padk == "even" else 0)
        Ai = np.asfortranarray(rng.integers(-256, 257, (lda, ca)), dtype=np.float64)
        Bi = np.asfortranarray(rng.integers(-256, 257, (ldb, cb)), dtype=np.float64)
        C0 = np.asfortranarray(rng.integers(-1000, 1001, (M, N)), dtype=np.float64)
        ref = C0.copy(order="F")
        orc.gemm(ta, tb, M, N, K, 3.0, Ai, lda, Bi, ldb, -2.0, ref, M)
        # the operands at an element offset inside a larger buffer
        dA = bx.DeviceArray.from_host(np.concatenate([np.zeros(off), Ai.ravel(order="F")])[:, None])
        dB = bx.DeviceArray.from_host(np.concatenate([np.zeros(off), Bi.ravel(order="F")])[:, None])
        dC = bx.DeviceArray.from_host(C0)
        for rep in range(2):
            if rep:
                dC.upload(C0)
            bx.device_gemm(ta, tb, M, N, K, 3.0, dA.ptr + 8 * off, lda, dB.ptr + 8 * off, ldb,
                           -2.0, dC.ptr, M, np.float64)
            bx.synchronize()
            assert np.array_equal(dC.download(), ref), (ta, tb, M, N, K, rep)
        for d in (dA, dB, dC):
            d.free()
    # uniform inputs, beta = 0 over NaN, a K tail and edge tiles in both dimensions
    M, N, K = 4100, 4099, 1003
    say("device dgemm ragged uniform", ta, tb, M, N, K)
    ra, ca = (M, K) if ta == "N" else (K, M)
    rb, cb = (K, N) if tb == "N" else (N, K)
    A = _mat(orc, ra, ca, np.float64, 5, 0)
    B = _mat(orc, rb, cb, np.float64, 5, 1)
    ref = np.zeros((M, N), order="F")
    orc.openblas_gemm(ta, tb, M, N, K, 0.75, A, ra, B, rb, 0.0, ref, M)
    dA, dB = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(B)
    dC = bx.DeviceArray.from_host(np.full((M, N), np.nan, order="F"))
    bx.device_gemm(ta, tb, M, N, K, 0.75, dA.ptr, ra, dB.ptr, rb, 0.0, dC.ptr, M, np.float64)
    bx.synchronize()
    out = dC.download()
    assert np.isfinite(out).all()
    _check(orc, out, ref, 0.75, A, B, 0.0, np.zeros((M, N)), K, np.float64, f"{ta}{tb} ragged")
    for d in (dA, dB, dC):
        d.free()
    # edge tiles store nothing outside C: rows M.. of a larger ldc keep their values
    M, N, K, pad = 259, 131, 35, 5
    say("device dgemm ragged ldc padding", ta, tb, M, N, K, pad)
    ra, ca = (M, K) if ta == "N" else (K, M)
    rb, cb = (K, N) if tb == "N" else (N, K)
    Ai = np.asfortranarray(rng.integers(-256, 257, (ra + ra % 2, ca)), dtype=np.float64)
    Bi = np.asfortranarray(rng.integers(-256, 257, (rb + rb % 2, cb)), dtype=np.float64)
    Cbig = np.asfortranarray(np.full((M + pad, N), 7.0))
    Cbig[:M] = rng.integers(-1000, 1001, (M, N))
    ref = Cbig.copy(order="F")
    orc.gemm(ta, tb, M, N, K, 3.0, Ai, Ai.shape[0], Bi, Bi.shape[0], -2.0, ref, M + pad)
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (Ai, Bi, Cbig))
    bx.device_gemm(ta, tb, M, N, K, 3.0, dA.ptr, Ai.shape[0], dB.ptr, Bi.shape[0], -2.0, dC.ptr,
                   M + pad, np.float64)
    bx.synchronize()
    assert np.array_equal(dC.download(), ref), (ta, tb, "ldc padding")
    for d in (dA, dB, dC):
        d.free()


@pytest.mark.parametrize("ta,tb", TRANS_REAL)
def test_device_dgemm_dma_path(gpu, orc, ta, tb):
    """Problems without edge tiles (M, N multiples of 128, K of 16, aligned
    operands) take the DMA-staged FP64 kernel: padded leading dimensions,
    several K tiles and a grid that is not a multiple of 8 blocks; bit-exact
    on integer inputs, within the tolerance on uniform inputs, and beta = 0
    ignores NaN in C."""
    bx = gpu
    rng = np.random.default_rng(31)
    for (M, N, K, pad) in ((384, 640, 1040, 2), (1024, 512, 64, 0)):
        say("device dgemm dma", ta, tb, M, N, K, pad)
        ra, ca = (M, K) if ta == "N" else (K, M)
        rb, cb = (K, N) if tb == "N" else (N, K)
        Ai = np.asfortranarray(rng.integers(-256, 257, (ra + pad, ca)), dtype=np.float64)
        Bi = np.asfortranarray(rng.integers(-256, 257, (rb + pad, cb)), dtype=np.float64)
        C0 = np.asfortranarray(rng.integers(-1000, 1001, (M, N)), dtype=np.float64)
        ref = C0.copy(order="F")
        orc.gemm(ta, tb, M, N, K, 3.0, Ai, ra + pad, Bi, rb + pad, -2.0, ref, M)
        dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (Ai, Bi, C0))
        bx.device_gemm(ta, tb, M, N, K, 3.0, dA.ptr, ra + pad, dB.ptr, rb + pad, -2.0, dC.ptr, M,
                       np.float64)
        bx.synchronize()
        assert np.array_equal(dC.download(), ref), (ta, tb, M, N, K)
        A = _mat(orc, ra, ca, np.float64, 3, 0)
        B = _mat(orc, rb, cb, np.float64, 3, 1)
        Cn = np.full((M, N), np.nan, order="F")
        ref = np.zeros((M, N), order="F")
        orc.gemm(ta, tb, M, N, K, 0.75, A, ra, B, rb, 0.0, ref, M)
        dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, Cn))
        bx.device_gemm(ta, tb, M, N, K, 0.75, dA.ptr, ra, dB.ptr, rb, 0.0, dC.ptr, M, np.float64)
        bx.synchronize()
        out = dC.download()
        assert not np.isnan(out).any()
        _check(orc, out, ref, 0.75, A, B, 0.0, np.zeros((M, N)), K, np.float64, f"{ta}{tb} dma")


def test_device_sgemm_integer_exact(gpu, orc):
    bx = gpu
    rng = np.random.default_rng(11)
    M, N, K = 200, 150, 1000
    for ta, tb in TRANS_REAL:
        say("device sgemm int-exact", ta, tb)
        A = np.asfortranarray(rng.integers(-16, 17, (M, K) if ta == "N" else (K, M)), dtype=np.float32)
        B = np.asfortranarray(rng.integers(-16, 17, (K, N) if tb == "N" else (N, K)), dtype=np.float32)
        C0 = np.asfortranarray(rng.integers(-100, 101, (M, N)), dtype=np.float32)
        ref = C0.copy(order="F")
        orc.gemm(ta, tb, M, N, K, 1.0, A, A.shape[0], B, B.shape[0], 1.0, ref, M)
        dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
        bx.device_gemm(ta, tb, M, N, K, 1.0, dA.ptr, A.shape[0], dB.ptr, B.shape[0], 1.0,
                       dC.ptr, M, np.float32)
        bx.synchronize()
        assert np.array_equal(dC.download(), ref)


def test_device_zgemm_integer_exact_conj(gpu, orc):
    bx = gpu
    rng = np.random.default_rng(12)
    M, N, K = 70, 90, 110
    for ta, tb in TRANS_CPLX:
        say("device zgemm int-exact", ta, tb)
        def ci(shape):
            return np.asfortranarray(rng.integers(-64, 65, shape) + 1j * rng.integers(-64, 65, shape))
        A = ci((M, K) if ta == "N" else (K, M))
        B = ci((K, N) if tb == "N" else (N, K))
        C0 = ci((M, N))
        al, be = 2 - 1j, 1 + 3j
        exact = al * (_op(A, ta) @ _op(B, tb)) + be * C0
        dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
        bx.device_gemm(ta, tb, M, N, K, al, dA.ptr, A.shape[0], dB.ptr, B.shape[0], be,
                       dC.ptr, M, np.complex128)
        bx.synchronize()
        assert np.array_equal(dC.download(), exact), (ta, tb)


@pytest.mark.parametrize("ta,tb", TRANS_REAL)
def test_device_sgemm_dma_path(gpu, orc, ta, tb):
    """SGEMMs without edge tiles (M, N multiples of 128, K of 32, aligned
    operands) take the DMA-staged FP32 kernel: padded leading dimensions,
    several K tiles; bit-exact on small-integer inputs (every partial sum is
    exact in FP32), within the tolerance on uniform inputs, and beta = 0
    ignores NaN in C."""
    bx = gpu
    rng = np.random.default_rng(43)
    M, N, K, pad = 256, 384, 96, 4
    ra, ca = (M, K) if ta == "N" else (K, M)
    rb, cb = (K, N) if tb == "N" else (N, K)
    say("device sgemm dma", ta, tb, M, N, K)
    A = np.asfortranarray(rng.integers(-16, 17, (ra + pad, ca)), dtype=np.float32)
    B = np.asfortranarray(rng.integers(-16, 17, (rb + pad, cb)), dtype=np.float32)
    C0 = np.asfortranarray(rng.integers(-64, 65, (M, N)), dtype=np.float32)
    exact = (3 * (_op(A[:ra], ta).astype(np.int64) @ _op(B[:rb], tb).astype(np.int64))
             - 2 * C0.astype(np.int64)).astype(np.float32)
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
    bx.device_gemm(ta, tb, M, N, K, 3.0, dA.ptr, ra + pad, dB.ptr, rb + pad, -2.0, dC.ptr, M,
                   np.float32)
    bx.synchronize()
    assert np.array_equal(dC.download(), exact), (ta, tb)
    A = _mat(orc, ra, ca, np.float32, 5, 0)
    B = _mat(orc, rb, cb, np.float32, 5, 1)
    Cn = np.full((M, N), np.nan, dtype=np.float32, order="F")
    ref = np.zeros((M, N), dtype=np.float32, order="F")
    orc.gemm(ta, tb, M, N, K, 0.75, A, ra, B, rb, 0.0, ref, M, acc64=True)
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, Cn))
    bx.device_gemm(ta, tb, M, N, K, 0.75, dA.ptr, ra, dB.ptr, rb, 0.0, dC.ptr, M, np.float32)
    bx.synchronize()
    out = dC.download()
    assert not np.isnan(out).any()
    _check(orc, out, ref, 0.75, A, B, 0.0, np.zeros((M, N), dtype=np.float32), K, np.float32,
           f"{ta}{tb} sdma")


@pytest.mark.parametrize("ta,tb", TRANS_CPLX)
def test_device_zgemm_dma_path(gpu, orc, ta, tb):
    """ZGEMMs without edge tiles (M a multiple of 128, N of 64, K of 8) take
    the DMA-staged kernel: every transpose and conjugate pair, padded leading
    dimensions, several K tiles; bit-exact on Gaussian-integer inputs, within
    the tolerance on uniform inputs, and beta = 0 ignores NaN in C."""
    bx = gpu
    rng = np.random.default_rng(41)
    M, N, K, pad = 256, 192, 72, 3
    ra, ca = (M, K) if ta == "N" else (K, M)
    rb, cb = (K, N) if tb == "N" else (N, K)

    def ci(shape):
        return np.asfortranarray(rng.integers(-64, 65, shape) + 1j * rng.integers(-64, 65, shape))
    say("device zgemm dma", ta, tb, M, N, K)
    A, B, C0 = ci((ra + pad, ca)), ci((rb + pad, cb)), ci((M, N))
    al, be = 2 - 1j, 1 + 3j
    exact = al * (_op(A[:ra], ta) @ _op(B[:rb], tb)) + be * C0
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
    bx.device_gemm(ta, tb, M, N, K, al, dA.ptr, ra + pad, dB.ptr, rb + pad, be, dC.ptr, M,
                   np.complex128)
    bx.synchronize()
    assert np.array_equal(dC.download(), exact), (ta, tb)
    A = _mat(orc, ra, ca, np.complex128, 4, 0)
    B = _mat(orc, rb, cb, np.complex128, 4, 1)
    Cn = np.full((M, N), np.nan + 1j * np.nan, order="F")
    ref = np.zeros((M, N), dtype=np.complex128, order="F")
    orc.gemm(ta, tb, M, N, K, 0.5 + 0.25j, A, ra, B, rb, 0.0, ref, M)
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, Cn))
    bx.device_gemm(ta, tb, M, N, K, 0.5 + 0.25j, dA.ptr, ra, dB.ptr, rb, 0.0, dC.ptr, M,
                   np.complex128)
    bx.synchronize()
    out = dC.download()
    assert not np.isnan(out).any()
    _check(orc, out, ref, 0.5 + 0.25j, A, B, 0.0, np.zeros((M, N), dtype=np.complex128), K,
           np.complex128, f"{ta}{tb} zdma")


def test_device_beta_zero_ignores_nan(gpu, orc):
    bx = gpu
    M, N, K = 300, 200, 100
    A = _mat(orc, M, K, np.float64, 3, 0)
    B = _mat(orc, K, N, np.float64, 3, 1)
    Cn = np.full((M, N), np.nan, order="F")
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, Cn))
    say("device beta=0 NaN C")
    bx.device_gemm("N", "N", M, N, K, 1.0, dA.ptr, M, dB.ptr, K, 0.0, dC.ptr, M, np.float64)
    bx.synchronize()
    out = dC.download()
    assert np.all(np.isfinite(out))
    ref = np.zeros((M, N), order="F")
    orc.gemm("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, ref, M)
    _check(orc, out, ref, 1.0, A, B, 0.0, ref * 0, K, np.float64, "beta0")


def test_device_fill_matches_host_generator(gpu, orc):
    bx = gpu
    for dt in (np.float64, np.float32, np.complex128):
        say("device fill", np.dtype(dt).name)
        d = bx.DeviceArray(37, 23, dt)
        d.fill_uniform(99, 5)
        bx.synchronize()
        assert np.array_equal(d.download(), orc.fill_uniform(37, 23, 99, 5, dtype=dt))


# ------------------------------------------------------------------ ABI, GPU-sized
def test_cfg0_dgemm_abi_gpu_path(gpu, orc):
    """The CPU-path config: M=N=K=1024, NN, alpha 1.5, beta 0.5, uniform
    seed 0 (ids 0,1,2), ld 1024.  1024 > 1000 routes to the GPU, as in the
    reference (blas/dgemm.c:62).  Checked against the naive triple loop and
    against OpenBLAS on the same inputs."""
    bx = gpu
    n = 1024
    A = orc.fill_uniform(n, n, 0, 0)
    B = orc.fill_uniform(n, n, 0, 1)
    C0 = orc.fill_uniform(n, n, 0, 2)
    ref = C0.copy(order="F")
    orc.gemm("N", "N", n, n, n, 1.5, A, n, B, n, 0.5, ref, n)
    ob = C0.copy(order="F")
    orc.openblas_gemm("N", "N", n, n, n, 1.5, A, n, B, n, 0.5, ob, n)
    out = C0.copy(order="F")
    bx.reset_stats()
    say("cfg0 cblas_dgemm 1024^3")
    bx.cblas_gemm(bx.COL_MAJOR, "N", "N", n, n, n, 1.5, A, n, B, n, 0.5, out, n)
    st = bx.stats()
    assert st["calls_gpu"] == 1 and st["calls_cpu"] == 0 and st["cpu_fallbacks"] == 0
    e1 = _check(orc, out, ref, 1.5, A, B, 0.5, C0, n, np.float64, "vs naive")
    e2 = _check(orc, out, ob, 1.5, A, B, 0.5, C0, n, np.float64, "vs OpenBLAS")
    print(f"cfg0: rel err vs naive {e1:.3e}, vs OpenBLAS {e2:.3e}")


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.complex128])
def test_tile_runtime_ragged_all_transposes(gpu, orc, dt):
    """Host-resident operands through the tiled runtime with a small tile
    edge, so there are ragged edge tiles in every dimension."""
    bx = gpu
    code = bx.dtype_code(dt)
    assert bx.lib().blasx_set_tile_size(code, 256) == 0
    try:
        cplx = np.dtype(dt).kind == "c"
        M, N, K = (700, 530, 610) if cplx else (1100, 780, 1030)
        alpha, beta = (0.5 + 1j, -1 + 0.5j) if cplx else (1.25, -0.75)
        for ta, tb in (TRANS_CPLX if cplx else TRANS_REAL):
            say("tile-rt", np.dtype(dt).name, ta, tb, M, N, K)
            A = _mat(orc, M if ta == "N" else K, K if ta == "N" else M, dt, 4, 0)
            B = _mat(orc, K if tb == "N" else N, N if tb == "N" else K, dt, 4, 1)
            C0 = _mat(orc, M, N, dt, 4, 2)
            ref = C0.copy(order="F")
            orc.gemm(ta, tb, M, N, K, alpha, A, A.shape[0], B, B.shape[0], beta, ref, M,
                     acc64=(dt == np.float32))
            out = C0.copy(order="F")
            bx.reset_stats()
            bx.cblas_gemm(bx.COL_MAJOR, ta, tb, M, N, K, alpha, A, A.shape[0], B, B.shape[0],
                          beta, out, M)
            st = bx.stats()
            assert st["calls_gpu"] == 1 and st["kernel_launches"] >= 1
            _check(orc, out, ref, alpha, A, B, beta, C0, K, dt, f"{ta}{tb}")
    finally:
        bx.lib().blasx_set_tile_size(code, 0)


def test_tile_runtime_integer_exact_and_ld(gpu, orc):
    bx = gpu
    assert bx.lib().blasx_set_tile_size(bx.D, 256) == 0
    try:
        rng = np.random.default_rng(13)
        M, N, K, ld = 1030, 777, 1290, 1311
        for ta, tb in TRANS_REAL:
            say("tile-rt int-exact ld", ta, tb)
            ra, ca = (M, K) if ta == "N" else (K, M)
            rb, cb = (K, N) if tb == "N" else (N, K)
            A = np.asfortranarray(rng.integers(-256, 257, (ld + 3, ca)), dtype=np.float64)
            B = np.asfortranarray(rng.integers(-256, 257, (ld + 5, cb)), dtype=np.float64)
            C0 = np.asfortranarray(rng.integers(-999, 1000, (ld, N)), dtype=np.float64)
            ref = C0.copy(order="F")
            orc.gemm(ta, tb, M, N, K, 2.0, A, ld + 3, B, ld + 5, 3.0, ref, ld)
            out = C0.copy(order="F")
            bx.cblas_gemm(bx.COL_MAJOR, ta, tb, M, N, K, 2.0, A, ld + 3, B, ld + 5, 3.0, out, ld)
            assert np.array_equal(out, ref), (ta, tb)
    finally:
        bx.lib().blasx_set_tile_size(bx.D, 0)


def test_abi_row_major_and_f77(gpu, orc):
    bx = gpu
    M, N, K = 1200, 900, 1001
    A = orc.fill_uniform(M, K, 5, 0)
    B = orc.fill_uniform(K, N, 5, 1)
    C0 = orc.fill_uniform(M, N, 5, 2)
    ref = 0.75 * (A @ B) + 1.5 * C0
    # RowMajor: pass the row-major views (C-ordered arrays)
    Ar, Br, Cr = np.ascontiguousarray(A), np.ascontiguousarray(B), np.ascontiguousarray(C0)
    say("RowMajor cblas_dgemm")
    bx.cblas_gemm(bx.ROW_MAJOR, "N", "N", M, N, K, 0.75, Ar, K, Br, N, 1.5, Cr, N)
    _check(orc, Cr, ref, 0.75, A, B, 1.5, C0, K, np.float64, "rowmajor")
    say("f77 dgemm_")
    out = C0.copy(order="F")
    bx.f77_gemm("N", "N", M, N, K, 0.75, A, M, B, K, 1.5, out, M)
    _check(orc, out, ref, 0.75, A, B, 1.5, C0, K, np.float64, "dgemm_")
    Az = orc.fill_uniform(600, 520, 6, 0, dtype=np.complex128)
    Bz = orc.fill_uniform(600, 510, 6, 1, dtype=np.complex128)
    Cz = orc.fill_uniform(520, 510, 6, 2, dtype=np.complex128)
    refz = Cz.copy(order="F")
    orc.gemm("C", "N", 520, 510, 600, 1 + 1j, Az, 600, Bz, 600, 0.5, refz, 520)
    outz = Cz.copy(order="F")
    say("f77 zgemm_ C N")
    bx.f77_gemm("C", "N", 520, 510, 600, 1 + 1j, Az, 600, Bz, 600, 0.5, outz, 520)
    _check(orc, outz, refz, 1 + 1j, Az, Bz, 0.5, Cz, 600, np.complex128, "zgemm_")


def test_out_of_core_small_budget(gpu, orc):
    """Working set (A+B tiles) larger than the cache budget: tiles are
    evicted and re-fetched, results stay exact on integer inputs."""
    bx = gpu
    L = bx.lib()
    old = L.blasx_get_hbm_budget()
    assert L.blasx_set_tile_size(bx.D, 512) == 0
    try:
        L.blasx_set_hbm_budget(48 * (32 << 20))  # 48 slots of 32 MiB
        rng = np.random.default_rng(14)
        M = N = K = 4096  # 8x8 A tiles + 8x8 B tiles = 128 > 48 slots
        A = np.asfortranarray(rng.integers(-64, 65, (M, K)), dtype=np.float64)
        B = np.asfortranarray(rng.integers(-64, 65, (K, N)), dtype=np.float64)
        C0 = np.asfortranarray(rng.integers(-64, 65, (M, N)), dtype=np.float64)
        out = C0.copy(order="F")
        bx.reset_stats()
        say("out-of-core 4096^3 budget 48 slots")
        bx.cblas_gemm(bx.COL_MAJOR, "N", "N", M, N, K, 1.0, A, M, B, K, 1.0, out, M)
        st = bx.stats()
        assert st["tile_evictions"] > 0, st
        assert st["bytes_h2d"] > st["bytes_lower_bound"] // 2
        # exact check on sampled rows/cols (full naive 4096^3 is too slow for the CPU here)
        rows = rng.choice(M, 64, replace=False)
        cols = rng.choice(N, 64, replace=False)
        exact = A[rows].astype(np.int64) @ B[:, cols].astype(np.int64) + C0[np.ix_(rows, cols)]
        assert np.array_equal(out[np.ix_(rows, cols)], exact.astype(np.float64))
    finally:
        L.blasx_set_hbm_budget(old)
        L.blasx_set_tile_size(bx.D, 0)


def test_resident_operand_reused_across_calls(gpu, orc):
    """Cross-call residency (blasx_resident_register): the registered A is
    not sent again, while the unregistered B -- mutated in place between the
    calls -- is re-read; unregistering A drops its tiles, so a changed A is
    seen.  Integer inputs: every result must equal the exact product, and a
    sampled block is also checked against the oracle's naive loop."""
    bx = gpu
    L = bx.lib()
    M = N = K = 2048
    rng = np.random.default_rng(15)
    A = np.asfortranarray(rng.integers(-64, 65, (M, K)), dtype=np.float64)
    B = np.asfortranarray(rng.integers(-64, 65, (K, N)), dtype=np.float64)

    def exact(A_, B_):   # integer sums below 2^53: any summation order is exact
        return A_ @ B_

    def sampled_oracle(C_, A_, B_):
        rows, cols = np.arange(0, M, 97), np.arange(5, N, 89)
        a = np.asfortranarray(A_[rows])
        b = np.asfortranarray(B_[:, cols])
        ref = np.zeros((len(rows), len(cols)), order="F")
        orc.gemm("N", "N", len(rows), len(cols), K, 1.0, a, len(rows), b, K, 0.0, ref, len(rows))
        assert np.array_equal(C_[np.ix_(rows, cols)], ref)

    assert L.blasx_set_tile_size(bx.D, 512) == 0
    assert L.blasx_resident_register(A.ctypes.data, A.nbytes) == 0
    registered = True
    try:
        L.blasx_cache_flush()
        C1 = np.zeros((M, N), order="F")
        bx.reset_stats()
        say("resident call 1")
        bx.gemm("N", "N", 1.0, A, B, 0.0, C1)
        s1 = bx.stats()
        assert np.array_equal(C1, exact(A, B))
        sampled_oracle(C1, A, B)
        assert s1["bytes_h2d"] >= A.nbytes + B.nbytes

        B[:, ::3] += 1.0          # B is not registered: the next call must re-read it
        C2 = np.zeros((M, N), order="F")
        say("resident call 2 (B mutated)")
        bx.gemm("N", "N", 1.0, A, B, 0.0, C2)
        s2 = bx.stats()
        assert np.array_equal(C2, exact(A, B))
        sampled_oracle(C2, A, B)
        h2d_second = s2["bytes_h2d"] - s1["bytes_h2d"]
        assert B.nbytes <= h2d_second < A.nbytes + B.nbytes, (s1, s2)
        assert s2["tile_hits"] - s1["tile_hits"] >= (M // 512) * (K // 512)

        assert L.blasx_resident_unregister(A.ctypes.data) == 0
        registered = False
        A[::2] -= 1.0             # A's tiles were dropped on unregister: the change is seen
        C3 = np.zeros((M, N), order="F")
        say("resident call 3 (A unregistered and mutated)")
        bx.gemm("N", "N", 1.0, A, B, 0.0, C3)
        assert np.array_equal(C3, exact(A, B))
        sampled_oracle(C3, A, B)
    finally:
        if registered:
            assert L.blasx_resident_unregister(A.ctypes.data) == 0
        L.blasx_cache_flush()
        L.blasx_set_tile_size(bx.D, 0)


# ------------------------------------------------------------------ CGEMM
def _c64(orc, rows, cols, seed, mid):
    return np.asfortranarray(orc.fill_uniform(rows, cols, seed, mid, dtype=np.complex128)
                             .astype(np.complex64))


@pytest.mark.parametrize("where", ["host", "device"])
def test_abi_cgemm_on_fp32_matrix_cores(gpu, orc, where):
    """complex64 through cblas_cgemm above the complex threshold: one real GEMM
    on the FP32 matrix cores over packed operands, every transpose/conjugate
    pair, against the oracle's naive complex GEMM on the same (exactly
    widened) inputs; tolerance 2 K eps(float32)"""
    bx = gpu
    M, N, K = 600, 520, 300
    alpha, beta = 1.5 - 0.5j, 0.5 + 0.25j
    eps = np.finfo(np.float32).eps
    for ta, tb in TRANS_CPLX:
        say("cgemm", where, ta, tb, M, N, K)
        A = _c64(orc, M if ta == "N" else K, K if ta == "N" else M, 3, 0)
        B = _c64(orc, K if tb == "N" else N, N if tb == "N" else K, 3, 1)
        C0 = _c64(orc, M, N, 3, 2)
        ref = C0.astype(np.complex128).copy(order="F")
        orc.gemm(ta, tb, M, N, K, alpha, A.astype(np.complex128), A.shape[0],
                 B.astype(np.complex128), B.shape[0], beta, ref, M)
        bx.reset_stats()
        if where == "host":
            out = C0.copy(order="F")
            bx.cblas_gemm(bx.COL_MAJOR, ta, tb, M, N, K, alpha, A, A.shape[0], B, B.shape[0],
                          beta, out, M)
        else:
            import ctypes as C
            dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
            al = (C.c_float * 2)(alpha.real, alpha.imag)
            be = (C.c_float * 2)(beta.real, beta.imag)
            fn = bx.lib().cblas_cgemm
            fn(bx.COL_MAJOR, bx.TRANS[ta], bx.TRANS[tb], M, N, K, C.cast(al, C.c_void_p),
               C.c_void_p(dA.ptr), A.shape[0], C.c_void_p(dB.ptr), B.shape[0],
               C.cast(be, C.c_void_p), C.c_void_p(dC.ptr), M)
            out = dC.download()
            for d in (dA, dB, dC):
                d.free()
        assert bx.stats()["calls_gpu"] == 1
        err = orc.relerr_fro(out.astype(np.complex128), ref, alpha, A.astype(np.complex128),
                             B.astype(np.complex128), beta, C0.astype(np.complex128))
        assert err <= 2 * K * eps, f"cgemm {ta}{tb}: rel err {err:.3e}"


def test_abi_cgemm_exact_integers_beta_zero_row_major(gpu, orc):
    """small complex integers: every FP32 partial sum is exact, so the result
    equals the naive loop bit for bit; beta = 0 ignores NaN in C; RowMajor is
    the transposed column-major problem"""
    bx = gpu
    rng = np.random.default_rng(13)
    M, N, K = 640, 576, 320

    def ints(*shape):
        return np.asfortranarray((rng.integers(-4, 5, shape) + 1j * rng.integers(-4, 5, shape))
                                 .astype(np.complex64))

    A, B = ints(M, K), ints(N, K)
    ref = np.zeros((M, N), np.complex128, order="F")
    orc.gemm("N", "C", M, N, K, 2 - 1j, A.astype(np.complex128), M, B.astype(np.complex128), N,
             0.0, ref, M)
    out = np.full((M, N), np.nan + 1j * np.nan, np.complex64, order="F")
    say("cgemm exact N C beta=0")
    bx.cblas_gemm(bx.COL_MAJOR, "N", "C", M, N, K, 2 - 1j, A, M, B, N, 0.0, out, M)
    assert np.array_equal(out.astype(np.complex128), ref)
    # row-major: C (M x N, ldc = N) is the column-major N x M matrix C^T
    out2 = np.zeros((N, M), np.complex64, order="F")
    say("cgemm row-major")
    bx.cblas_gemm(bx.ROW_MAJOR, "N", "T", M, N, K, 2 - 1j, np.ascontiguousarray(A), K,
                  np.ascontiguousarray(B.conj()), K, 0.0, out2, N)
    assert np.array_equal(out2.T.astype(np.complex128), ref)


@pytest.mark.parametrize("dt,T", [(np.float64, 1024), (np.float64, 2048), (np.float32, 2048),
                                  (np.complex128, 1024)])
def test_tile_runtime_pinned_operands_large_tiles(gpu, orc, dt, T):
    """Pinned host operands (hipHostMalloc: DMA'd directly, no staging copy)
    through the tiled runtime at the tile edges the 32768 host-resident
    runs use, with ragged edge tiles in every dimension and a K tail: bit
    exact on integer inputs against OpenBLAS (every partial sum is exact),
    every transpose.  Complex at T = 1024 (the largest complex tile a
    32 MiB cache slot holds).  Integer magnitudes keep every FP32 partial
    sum below 2^24."""
    bx = gpu
    code = bx.dtype_code(dt)
    assert bx.lib().blasx_set_tile_size(code, T) == 0
    cplx = np.dtype(dt).kind == "c"
    try:
        rng = np.random.default_rng(T)
        M, N, K = (2 * T + 333, T + 517, 2 * T + 91)
        pins = []
        for ta, tb in (TRANS_CPLX[:4] + [("C", "C")] if cplx else TRANS_REAL):
            say("tile-rt pinned", np.dtype(dt).name, T, ta, tb, M, N, K)
            ra, ca = (M, K) if ta == "N" else (K, M)
            rb, cb = (K, N) if tb == "N" else (N, K)
            pa, pb, pc = (bx.PinnedArray(ra, ca, dt), bx.PinnedArray(rb, cb, dt),
                          bx.PinnedArray(M, N, dt))
            pins += [pa, pb, pc]

            lim = 8 if dt == np.float32 else 64

            def ints(shape):
                v = rng.integers(-lim, lim + 1, shape).astype(np.float64)
                if cplx:
                    v = v + 1j * rng.integers(-lim, lim + 1, shape)
                return v

            pa.array[:] = ints((ra, ca))
            pb.array[:] = ints((rb, cb))
            C0 = ints((M, N)).astype(dt)
            pc.array[:] = C0
            alpha, beta = (2 - 1j, 1 + 1j) if cplx else (3.0, -2.0)
            ref = np.asfortranarray(C0.copy())
            orc.openblas_gemm(ta, tb, M, N, K, alpha, np.asfortranarray(pa.array), ra,
                              np.asfortranarray(pb.array), rb, beta, ref, M)
            bx.reset_stats()
            bx.cblas_gemm(bx.COL_MAJOR, ta, tb, M, N, K, alpha, pa.array, ra, pb.array, rb, beta,
                          pc.array, M)
            st = bx.stats()
            assert st["calls_gpu"] == 1 and st["cpu_fallbacks"] == 0, st
            assert st["bytes_h2d"] > 0
            assert np.array_equal(pc.array, ref), (T, ta, tb)
            for p in pins:
                p.free()
            pins = []
    finally:
        bx.lib().blasx_set_tile_size(code, 0)
