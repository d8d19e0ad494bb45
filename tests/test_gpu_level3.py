"""GPU parity of SYMM / HEMM / TRMM / SYR2K (one GEMM on an operand
materialised in HBM; SYR2K as the two-term rank-k launch) against the
oracle's naive GEMM on the operand the routine defines.  The reference runs
these on the CPU BLAS (blas/dsymm.c:44-58, blas/dtrmm.c:48-60,
blas/dsyr2k.c:57-70).

Tolerance: ||X_gpu - X_ref||_F / (|alpha| ||op1||_F ||op2||_F + |beta| ||C0||_F)
<= c K eps (c = 1 real, 2 complex; K the inner dimension); SYR2K counts both
terms.  Integer-valued inputs must match bit for bit, and SYR2K's other
triangle must come back unchanged."""
import sys

import numpy as np
import pytest

from level3_common import (COL, CTR, LEFT, LO, NONUNIT, NOTR, RIGHT, ROW, TR, UNIT, UP, call,
                           full_sym, full_tri, ref_her2k, ref_symm, ref_syr2k, ref_trmm,
                           trsm_operand, trsm_residual)

pytestmark = pytest.mark.gpu

EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps,
       np.complex128: np.finfo(np.float64).eps, np.complex64: np.finfo(np.float32).eps}
PFX = {np.float32: "s", np.float64: "d", np.complex128: "z", np.complex64: "c"}


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _mats(orc, dt, seed, *shapes):
    if dt == np.complex64:  # the complex128 generator, rounded
        return [np.asfortranarray(orc.fill_uniform(r, c, seed, i, dtype=np.complex128)
                                  .astype(np.complex64)) for i, (r, c) in enumerate(shapes)]
    return [orc.fill_uniform(r, c, seed, i, dtype=dt) for i, (r, c) in enumerate(shapes)]


def _check(out, ref, num_scale, K, dt, what, mask=None):
    d = (out - ref) if mask is None else (out[mask] - ref[mask])
    err = np.linalg.norm(d.ravel()) / num_scale
    tol = (2 if np.dtype(dt).kind == "c" else 1) * K * EPS[dt]
    assert err <= tol, f"{what}: rel err {err:.3e} > {tol:.3e}"


def _fro(x):
    return np.linalg.norm(np.asarray(x).ravel())


def _sizes(dt):
    # an extent above the GPU threshold (1000 real, 500 complex), ragged against the tiles
    return (1100, 200) if np.dtype(dt).kind != "c" else (600, 150)


def _sc(dt, re, im):
    return complex(re, im) if np.dtype(dt).kind == "c" else re


def _dev(bx, a):
    return bx.DeviceArray.from_host(a)


SYMM_CASES = [(np.float64, "dsymm"), (np.float32, "ssymm"), (np.complex128, "zsymm"),
              (np.complex128, "zhemm"), (np.complex64, "csymm"), (np.complex64, "chemm")]


@pytest.mark.parametrize("dt,r", SYMM_CASES)
@pytest.mark.parametrize("where", ["host", "device"])
def test_gpu_symm_hemm(gpu, orc, dt, r, where):
    bx = gpu
    L = bx.lib()
    M, N = _sizes(dt)
    alpha, beta = _sc(dt, 1.5, -0.5), _sc(dt, 0.5, 0.25)
    for side in (LEFT, RIGHT):
        for uplo in (UP, LO):
            say(r, where, side, uplo, M, N)
            na = M if side == LEFT else N
            A, B, C0 = _mats(orc, dt, 3, (na, na), (M, N), (M, N))
            ref = ref_symm(orc, side, uplo, r.endswith("hemm"), M, N, alpha, A, B, beta, C0)
            bx.reset_stats()
            if where == "host":
                out = C0.copy(order="F")
                call(L, "cblas", r, dt, COL, side, uplo, M, N, alpha, A, na, B, M, beta, out, M)
            else:
                dA, dB, dC = _dev(bx, A), _dev(bx, B), _dev(bx, C0)
                call(L, "cblas", r, dt, COL, side, uplo, M, N, alpha, dA.ptr, na, dB.ptr, M, beta,
                     dC.ptr, M)
                out = dC.download()
                A_check = dA.download()
                assert np.array_equal(np.ascontiguousarray(A_check).view(np.uint8),
                                      np.ascontiguousarray(A).view(np.uint8)), "A modified"
                for d in (dA, dB, dC):
                    d.free()
            assert bx.stats()["calls_gpu"] == 1
            F = full_sym(A, uplo, r.endswith("hemm"))
            _check(out, ref, abs(alpha) * _fro(F) * _fro(B) + abs(beta) * _fro(C0), na, dt,
                   f"{r} {where}")


TRMM_TYPES = [np.float64, np.float32, np.complex128, np.complex64]


@pytest.mark.parametrize("dt", TRMM_TYPES)
@pytest.mark.parametrize("where", ["host", "device"])
def test_gpu_trmm(gpu, orc, dt, where):
    bx = gpu
    L = bx.lib()
    M, N = _sizes(dt)
    alpha = _sc(dt, 1.25, 0.75)
    transes = (NOTR, TR, CTR) if dt != np.float32 else (NOTR, TR)
    for side in (LEFT, RIGHT):
        for uplo in (UP, LO):
            for trans in transes:
                for diag in (NONUNIT, UNIT):
                    say("trmm", np.dtype(dt).name, where, side, uplo, trans, diag)
                    na = M if side == LEFT else N
                    A, B0 = _mats(orc, dt, 4, (na, na), (M, N))
                    ref = ref_trmm(orc, side, uplo, trans, diag, M, N, alpha, A, B0)
                    if where == "host":
                        out = B0.copy(order="F")
                        call(L, "cblas", PFX[dt] + "trmm", dt, COL, side, uplo, trans, diag, M, N,
                             alpha, A, na, out, M)
                    else:
                        dA, dB = _dev(bx, A), _dev(bx, B0)
                        call(L, "cblas", PFX[dt] + "trmm", dt, COL, side, uplo, trans, diag, M,
                             N, alpha, dA.ptr, na, dB.ptr, M)
                        out = dB.download()
                        dA.free()
                        dB.free()
                    T = full_tri(A, uplo, diag)
                    _check(out, ref, abs(alpha) * _fro(T) * _fro(B0), na, dt, "trmm")


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.complex128, np.complex64])
@pytest.mark.parametrize("where", ["host", "device"])
def test_gpu_syr2k(gpu, orc, dt, where):
    bx = gpu
    L = bx.lib()
    N, K = (1100, 320) if np.dtype(dt).kind != "c" else (600, 320)
    alpha, beta = _sc(dt, 0.75, -0.5), _sc(dt, 0.5, 0.5)
    for uplo in (UP, LO):
        for trans in (NOTR, TR):
            say("syr2k", np.dtype(dt).name, where, uplo, trans)
            shape = (N, K) if trans == NOTR else (K, N)
            A, B, C0 = _mats(orc, dt, 5, shape, shape, (N, N))
            ref, m = ref_syr2k(orc, uplo, trans, N, K, alpha, A, B, beta, C0)
            lda = shape[0]
            if where == "host":
                out = C0.copy(order="F")
                call(L, "cblas", PFX[dt] + "syr2k", dt, COL, uplo, trans, N, K, alpha, A, lda, B,
                     lda, beta, out, N)
            else:
                dA, dB, dC = _dev(bx, A), _dev(bx, B), _dev(bx, C0)
                call(L, "cblas", PFX[dt] + "syr2k", dt, COL, uplo, trans, N, K, alpha, dA.ptr,
                     lda, dB.ptr, lda, beta, dC.ptr, N)
                out = dC.download()
                for d in (dA, dB, dC):
                    d.free()
            _check(out, ref, 2 * abs(alpha) * _fro(A) * _fro(B) + abs(beta) * _fro(C0), K, dt,
                   "syr2k", mask=m)
            assert np.array_equal(out[~m].view(np.uint8), C0[~m].view(np.uint8))


def test_gpu_level3_row_major(gpu, orc):
    """RowMajor: the same bytes as the column-major problem on the other
    side / triangle / transposition"""
    bx = gpu
    L = bx.lib()
    dt = np.float64
    M, N = 1100, 200
    A, B, C0 = _mats(orc, dt, 6, (M, M), (M, N), (M, N))
    ref = ref_symm(orc, LEFT, LO, False, M, N, 1.5, A, B, 0.5, C0)
    out = C0.copy(order="F")
    say("row-major dsymm")
    call(L, "cblas", "dsymm", dt, ROW, RIGHT, UP, N, M, 1.5, A, M, B, M, 0.5, out, M)
    _check(out, ref, 1.5 * _fro(full_sym(A, LO, False)) * _fro(B) + 0.5 * _fro(C0), M, dt, "symm")
    ref = ref_trmm(orc, LEFT, UP, TR, NONUNIT, M, N, 2.0, A, B)
    out = B.copy(order="F")
    say("row-major dtrmm")
    call(L, "cblas", "dtrmm", dt, ROW, RIGHT, LO, TR, NONUNIT, N, M, 2.0, A, M, out, M)
    _check(out, ref, 2.0 * _fro(full_tri(A, UP, NONUNIT)) * _fro(B), M, dt, "trmm")
    K = 320
    A2, B2, C2 = _mats(orc, dt, 7, (M, K), (M, K), (M, M))
    ref, m = ref_syr2k(orc, LO, NOTR, M, K, 0.5, A2, B2, 1.0, C2)
    out = C2.copy(order="F")
    say("row-major dsyr2k")
    call(L, "cblas", "dsyr2k", dt, ROW, UP, TR, M, K, 0.5, A2, M, B2, M, 1.0, out, M)
    _check(out, ref, _fro(A2) * _fro(B2) + _fro(C2), K, dt, "syr2k", mask=m)


def test_gpu_level3_exact_integers(gpu, orc):
    """integer-valued operands: every partial sum is exact, so the GPU must
    equal the naive loop bit for bit"""
    bx = gpu
    L = bx.lib()
    rng = np.random.default_rng(9)
    M, N, K = 1100, 96, 320

    def ints(*shape, hi=8):
        return np.asfortranarray(rng.integers(-hi, hi + 1, shape).astype(np.float64))

    A, B, C0 = ints(M, M), ints(M, N), ints(M, N, hi=64)
    for side, uplo in ((LEFT, LO), (LEFT, UP)):
        say("exact dsymm", side, uplo)
        ref = ref_symm(orc, side, uplo, False, M, N, 2.0, A, B, -3.0, C0)
        out = C0.copy(order="F")
        call(L, "cblas", "dsymm", np.float64, COL, side, uplo, M, N, 2.0, A, M, B, M, -3.0, out, M)
        assert np.array_equal(out, ref)
    for trans, diag in ((NOTR, UNIT), (TR, NONUNIT)):
        say("exact dtrmm", trans, diag)
        ref = ref_trmm(orc, LEFT, LO, trans, diag, M, N, 3.0, A, B)
        out = B.copy(order="F")
        call(L, "cblas", "dtrmm", np.float64, COL, LEFT, LO, trans, diag, M, N, 3.0, A, M, out, M)
        assert np.array_equal(out, ref)
    A2, B2, C2 = ints(M, K), ints(M, K), ints(M, M, hi=64)
    say("exact dsyr2k")
    ref, m = ref_syr2k(orc, UP, NOTR, M, K, 2.0, A2, B2, -1.0, C2)
    dA, dB, dC = _dev(bx, A2), _dev(bx, B2), _dev(bx, C2)
    call(L, "cblas", "dsyr2k", np.float64, COL, UP, NOTR, M, K, 2.0, dA.ptr, M, dB.ptr, M, -1.0,
         dC.ptr, M)
    assert np.array_equal(dC.download(), ref)
    for d in (dA, dB, dC):
        d.free()


@pytest.mark.parametrize("dt", TRMM_TYPES)
@pytest.mark.parametrize("where", ["host", "device"])
def test_gpu_trsm(gpu, orc, dt, where):
    """backward error of the solve, c n eps, on a well-conditioned operand"""
    bx = gpu
    L = bx.lib()
    M, N = _sizes(dt)
    alpha = _sc(dt, 1.25, 0.75)
    transes = (NOTR, TR, CTR) if dt != np.float32 else (NOTR, TR)
    for side in (LEFT, RIGHT):
        for uplo in (UP, LO):
            for trans in transes:
                for diag in (NONUNIT, UNIT):
                    say("trsm", np.dtype(dt).name, where, side, uplo, trans, diag)
                    na = M if side == LEFT else N
                    A0, B0 = _mats(orc, dt, 8, (na, na), (M, N))
                    A = trsm_operand(A0, na)
                    if where == "host":
                        X = B0.copy(order="F")
                        call(L, "cblas", PFX[dt] + "trsm", dt, COL, side, uplo, trans, diag, M, N,
                             alpha, A, na, X, M)
                    else:
                        dA, dB = _dev(bx, A), _dev(bx, B0)
                        call(L, "cblas", PFX[dt] + "trsm", dt, COL, side, uplo, trans, diag, M,
                             N, alpha, dA.ptr, na, dB.ptr, M)
                        X = dB.download()
                        dA.free()
                        dB.free()
                    err = trsm_residual(orc, side, uplo, trans, diag, M, N, alpha, A, X, B0)
                    tol = (2 if np.dtype(dt).kind == "c" else 1) * na * EPS[dt]
                    assert err <= tol, f"trsm backward error {err:.3e} > {tol:.3e}"


def test_gpu_trsm_exact_integers_and_row_major(gpu, orc):
    """unit triangular integer T and integer X: B = T X is exact, so the solve
    must give X back bit for bit (ragged leaves, both triangles, both sides,
    RowMajor)"""
    bx = gpu
    L = bx.lib()
    rng = np.random.default_rng(12)
    M, N = 1100, 96
    for side in (LEFT, RIGHT):
        for uplo in (LO, UP):
            na = M if side == LEFT else N
            A = np.asfortranarray(rng.integers(-3, 4, (na, na)).astype(np.float64))
            Xi = np.asfortranarray(rng.integers(-8, 9, (M, N)).astype(np.float64))
            T = full_tri(A, uplo, UNIT)
            B = np.asfortranarray(T @ Xi if side == LEFT else Xi @ T)
            say("exact dtrsm", side, uplo)
            X = B.copy(order="F")
            call(L, "cblas", "dtrsm", np.float64, COL, side, uplo, NOTR, UNIT, M, N, 1.0, A, na,
                 X, M)
            assert np.array_equal(X, Xi)
            X = B.copy(order="F")
            call(L, "cblas", "dtrsm", np.float64, ROW, RIGHT if side == LEFT else LEFT,
                 UP if uplo == LO else LO, NOTR, UNIT, N, M, 1.0, A, na, X, M)
            assert np.array_equal(X, Xi), "row-major"


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("where", ["host", "device"])
def test_gpu_her2k(gpu, orc, dt, where):
    bx = gpu
    L = bx.lib()
    N, K = 600, 320
    alpha, beta = 0.75 - 0.5j, 0.5
    for uplo in (UP, LO):
        for trans in (NOTR, CTR):
            say("her2k", where, uplo, trans)
            shape = (N, K) if trans == NOTR else (K, N)
            A, B, C0 = _mats(orc, dt, 10, shape, shape, (N, N))
            ref, m = ref_her2k(orc, uplo, trans, N, K, alpha, A, B, beta, C0)
            lda = shape[0]
            if where == "host":
                out = C0.copy(order="F")
                call(L, "cblas", PFX[dt] + "her2k", dt, COL, uplo, trans, N, K, alpha, A, lda, B, lda,
                     beta, out, N)
            else:
                dA, dB, dC = _dev(bx, A), _dev(bx, B), _dev(bx, C0)
                call(L, "cblas", PFX[dt] + "her2k", dt, COL, uplo, trans, N, K, alpha, dA.ptr, lda,
                     dB.ptr, lda, beta, dC.ptr, N)
                out = dC.download()
                for d in (dA, dB, dC):
                    d.free()
            _check(out, ref, 2 * abs(alpha) * _fro(A) * _fro(B) + abs(beta) * _fro(C0), K, dt,
                   "her2k", mask=m)
            assert np.array_equal(out[~m].view(np.uint8), C0[~m].view(np.uint8))
            d = np.arange(N)
            assert not np.any(out[d, d].imag), "diagonal not real"
    # RowMajor over the same bytes
    A, B, C0 = _mats(orc, dt, 11, (N, K), (N, K), (N, N))
    ref, m = ref_her2k(orc, LO, NOTR, N, K, alpha, A, B, beta, C0)
    out = C0.copy(order="F")
    say("her2k row-major")
    call(L, "cblas", PFX[dt] + "her2k", dt, ROW, UP, CTR, N, K, np.conj(alpha), A, N, B, N, beta, out, N)
    _check(out, ref, 2 * abs(alpha) * _fro(A) * _fro(B) + abs(beta) * _fro(C0), K, dt,
           "her2k row-major", mask=m)


def test_gpu_level3_edges(gpu, orc):
    """shapes at the edges of the decompositions: a one-column right-hand
    side, a last TRSM leaf of one row (M = 1025), a one-row Right-side solve
    with N = 1100, SYMM sized for the GPU by N alone, alpha = 0 for TRMM / TRSM
    and beta = 0 over NaN for SYMM, SYR2K with K = 0 in HBM"""
    bx = gpu
    L = bx.lib()
    dt = np.float64
    # TRSM: one column, M = 1025 (leaves of 64 ... and a last leaf of 1 row)
    for M, N, side in ((1025, 1, LEFT), (1, 1100, RIGHT)):
        na = M if side == LEFT else N
        A0, B0 = _mats(orc, dt, 20, (na, na), (M, N))
        A = trsm_operand(A0, na)
        for uplo in (LO, UP):
            say("trsm edge", M, N, side, uplo)
            X = B0.copy(order="F")
            call(L, "cblas", "dtrsm", dt, COL, side, uplo, NOTR, NONUNIT, M, N, 1.0, A, na, X, M)
            err = trsm_residual(orc, side, uplo, NOTR, NONUNIT, M, N, 1.0, A, X, B0)
            assert err <= na * EPS[dt], err
    # SYMM: M small, N above the threshold (Side Right: A is N x N)
    M, N = 40, 1100
    A, B, C0 = _mats(orc, dt, 21, (N, N), (M, N), (M, N))
    C0[:, :] = np.nan
    ref = ref_symm(orc, RIGHT, UP, False, M, N, 2.0, A, B, 0.0, np.zeros_like(C0))
    out = C0.copy(order="F")
    say("symm edge beta=0 NaN")
    call(L, "cblas", "dsymm", dt, COL, RIGHT, UP, M, N, 2.0, A, N, B, M, 0.0, out, M)
    _check(out, ref, 2.0 * _fro(full_sym(A, UP, False)) * _fro(B), N, dt, "symm edge")
    # alpha = 0: B becomes zero (TRMM, TRSM), A not needed to be meaningful
    M, N = 1100, 64
    A, B0 = _mats(orc, dt, 22, (M, M), (M, N))
    for r in ("dtrmm", "dtrsm"):
        say(r, "alpha=0")
        dA, dB = _dev(bx, A), _dev(bx, B0)
        call(L, "cblas", r, dt, COL, LEFT, LO, NOTR, UNIT, M, N, 0.0, dA.ptr, M, dB.ptr, M)
        assert not np.any(dB.download()), r
        dA.free()
        dB.free()
    # SYR2K with K = 0 in HBM: C = beta C on the triangle only
    N = 1100
    C0 = orc.fill_uniform(N, N, 23, 0, dtype=dt)
    dC = _dev(bx, C0)
    dummy = _dev(bx, np.zeros((N, 1), order="F"))
    say("syr2k K=0 device")
    call(L, "cblas", "dsyr2k", dt, COL, LO, NOTR, N, 0, 1.0, dummy.ptr, N, dummy.ptr, N, 0.5,
         dC.ptr, N)
    out = dC.download()
    m = np.tril(np.ones((N, N), bool))
    assert np.array_equal(out[m], 0.5 * C0[m])
    assert np.array_equal(out[~m], C0[~m])
    dC.free()
    dummy.free()
