"""SYMM / HEMM / TRMM / SYR2K at the boundary without a GPU: argument errors
(blas/dsymm.c:29-36, blas/dtrmm.c:32-40, blas/dsyr2k.c:43-49, lowest number
wins) and CPU-sized calls (the CPU BLAS runs them, blas/dsymm.c:44-58) in every
precision, side, triangle, transposition and order, against the oracle's
naive GEMM on the operand the routine defines."""
import os
import subprocess
import sys

import numpy as np
import pytest

from level3_common import (COL, CTR, LEFT, LO, NONUNIT, NOTR, RIGHT, ROW, TR, UNIT, UP, call,
                           ref_her2k, ref_symm, ref_syr2k, ref_trmm, trsm_operand,
                           trsm_residual)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
DTS = [np.float32, np.float64, np.complex128, np.complex64]
PFX = {np.float32: "s", np.float64: "d", np.complex128: "z", np.complex64: "c"}


def _rnd(rng, dt, *shape):
    a = rng.uniform(-1, 1, shape)
    if np.dtype(dt).kind == "c":
        a = a + 1j * rng.uniform(-1, 1, shape)
    return np.asfortranarray(a.astype(dt))


def _tol(dt):
    return 2e-4 if dt in (np.float32, np.complex64) else 1e-12


def _run(body):
    code = ("import sys, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
            "import blasx_amd as bx\nfrom level3_common import *\nL = bx.lib()\n"
            "A = np.zeros((8, 8)); B = np.zeros((8, 8)); Cm = np.zeros((8, 8))\n"
            % (ROOT, HERE)) + body + "\nprint('after')"
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return r.stdout.splitlines()


ERRS = [
    ("call(L, 'cblas', 'dsymm', np.float64, COL, 0, UP, 2, 2, 1.0, A, 2, B, 2, 0.0, Cm, 2)",
     "DSYMM ", 1),
    ("call(L, 'cblas', 'dsymm', np.float64, COL, LEFT, 0, 2, 2, 1.0, A, 2, B, 2, 0.0, Cm, 2)",
     "DSYMM ", 2),
    ("call(L, 'cblas', 'dsymm', np.float64, COL, LEFT, UP, -1, 2, 1.0, A, 2, B, 2, 0.0, Cm, 0)",
     "DSYMM ", 3),
    ("call(L, 'cblas', 'dsymm', np.float64, COL, RIGHT, UP, 2, 3, 1.0, A, 2, B, 2, 0.0, Cm, 2)",
     "DSYMM ", 7),
    ("call(L, 'cblas', 'dsymm', np.float64, COL, LEFT, UP, 3, 2, 1.0, A, 3, B, 2, 0.0, Cm, 3)",
     "DSYMM ", 9),
    ("call(L, 'cblas', 'dsymm', np.float64, COL, LEFT, UP, 3, 2, 1.0, A, 3, B, 3, 0.0, Cm, 2)",
     "DSYMM ", 12),
    ("call(L, 'cblas', 'zhemm', np.complex128, COL, 0, UP, 2, 2, 1.0, A, 2, B, 2, 0.0, Cm, 2)",
     "ZHEMM", 1),
    ("call(L, 'cblas', 'dtrmm', np.float64, COL, LEFT, UP, 0, NONUNIT, 2, 2, 1.0, A, 2, B, 2)",
     "DTRMM ", 3),
    ("call(L, 'cblas', 'dtrmm', np.float64, COL, LEFT, UP, NOTR, 0, 2, 2, 1.0, A, 2, B, 2)",
     "DTRMM ", 4),
    ("call(L, 'cblas', 'dtrmm', np.float64, COL, LEFT, UP, NOTR, UNIT, 2, -1, 1.0, A, 2, B, 2)",
     "DTRMM ", 6),
    ("call(L, 'cblas', 'strmm', np.float32, COL, RIGHT, UP, NOTR, UNIT, 2, 3, 1.0, A, 2, B, 2)",
     "STRMM ", 9),
    ("call(L, 'cblas', 'dtrmm', np.float64, COL, LEFT, UP, NOTR, UNIT, 3, 2, 1.0, A, 3, B, 2)",
     "DTRMM ", 11),
    ("call(L, 'cblas', 'dsyr2k', np.float64, COL, UP, NOTR, 3, 2, 1.0, A, 3, B, 2, 0.0, Cm, 3)",
     "DSYR2K", 9),
    ("call(L, 'cblas', 'zsyr2k', np.complex128, COL, UP, CTR, 2, 2, 1.0, A, 2, B, 2, 0.0, Cm, 2)",
     "ZSYR2K", 2),
    ("call(L, 'cblas', 'dsyr2k', np.float64, COL, UP, NOTR, 3, 2, 1.0, A, 3, B, 3, 0.0, Cm, 2)",
     "DSYR2K", 12),
    ("call(L, 'f77', 'dsymm', np.float64, LEFT, 0, 2, 2, 1.0, A, 2, B, 2, 0.0, Cm, 2)",
     "DSYMM ", 2),
    ("call(L, 'f77', 'dtrmm', np.float64, LEFT, UP, NOTR, 0, 2, 2, 1.0, A, 2, B, 2)",
     "DTRMM ", 4),
    ("call(L, 'cblas', 'dtrsm', np.float64, COL, 0, UP, NOTR, UNIT, 2, 2, 1.0, A, 2, B, 2)",
     "DTRSM ", 1),
    ("call(L, 'cblas', 'ztrsm', np.complex128, COL, LEFT, UP, NOTR, UNIT, -1, 2, 1.0, A, 2, B, 2)",
     "ZTRSM ", 5),
    ("call(L, 'cblas', 'dtrsm', np.float64, COL, RIGHT, UP, NOTR, UNIT, 2, 3, 1.0, A, 2, B, 2)",
     "DTRSM ", 9),
    ("call(L, 'f77', 'strsm', np.float32, LEFT, UP, 0, UNIT, 2, 2, 1.0, A, 2, B, 2)",
     "STRSM ", 3),
    ("call(L, 'cblas', 'zher2k', np.complex128, COL, UP, TR, 2, 2, 1.0, A, 2, B, 2, 0.0, Cm, 2)",
     "ZHER2K", 2),
    ("call(L, 'cblas', 'zher2k', np.complex128, COL, UP, NOTR, 3, 2, 1.0, A, 2, B, 3, 0.0, Cm, 3)",
     "ZHER2K", 7),
]


@pytest.mark.parametrize("body,name,info", ERRS)
def test_level3_argument_errors(bx, body, name, info):
    lines = _run(body)
    want = " ** On entry to %6s parameter number %2d had an illegal value" % (name, info)
    assert lines == [want, "after"], lines


def _scal(dt, re, im):
    return complex(re, im) if np.dtype(dt).kind == "c" else re


@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [UP, LO])
def test_symm_hemm_cpu_sized(bx, orc, dt, side, uplo):
    rng = np.random.default_rng(3)
    M, N = 23, 17
    na = M if side == LEFT else N
    A, B, C0 = _rnd(rng, dt, na, na), _rnd(rng, dt, M, N), _rnd(rng, dt, M, N)
    alpha, beta = _scal(dt, 0.75, -0.25), _scal(dt, -0.5, 1.0)
    L = bx.lib()
    routines = [PFX[dt] + "symm"] + ([PFX[dt] + "hemm"] if np.dtype(dt).kind == "c" else [])
    for r in routines:
        ref = ref_symm(orc, side, uplo, r.endswith("hemm"), M, N, alpha, A, B, beta, C0)
        for entry in ("cblas", "f77", "row"):
            out = C0.copy(order="F")
            if entry == "cblas":
                call(L, "cblas", r, dt, COL, side, uplo, M, N, alpha, A, na, B, M, beta, out, M)
            elif entry == "f77":
                call(L, "f77", r, dt, side, uplo, M, N, alpha, A, na, B, M, beta, out, M)
            else:
                # the same bytes read row-major: C^T = B^T A^T, the other side and triangle
                call(L, "cblas", r, dt, ROW, RIGHT if side == LEFT else LEFT,
                     LO if uplo == UP else UP, N, M, alpha, A, na, B, M, beta, out, M)
            assert np.max(np.abs(out - ref)) <= _tol(dt) * 8, (r, entry)


@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [UP, LO])
def test_trmm_cpu_sized(bx, orc, dt, side, uplo):
    rng = np.random.default_rng(4)
    M, N = 19, 13
    na = M if side == LEFT else N
    A, B0 = _rnd(rng, dt, na, na), _rnd(rng, dt, M, N)
    alpha = _scal(dt, 1.5, 0.5)
    L = bx.lib()
    for trans in (NOTR, TR, CTR):
        for diag in (NONUNIT, UNIT):
            ref = ref_trmm(orc, side, uplo, trans, diag, M, N, alpha, A, B0)
            for entry in ("cblas", "f77", "row"):
                B = B0.copy(order="F")
                if entry == "cblas":
                    call(L, "cblas", PFX[dt] + "trmm", dt, COL, side, uplo, trans, diag, M, N,
                         alpha, A, na, B, M)
                elif entry == "f77":
                    call(L, "f77", PFX[dt] + "trmm", dt, side, uplo, trans, diag, M, N, alpha, A,
                         na, B, M)
                else:
                    call(L, "cblas", PFX[dt] + "trmm", dt, ROW, RIGHT if side == LEFT else LEFT,
                         LO if uplo == UP else UP, trans, diag, N, M, alpha, A, na, B, M)
                assert np.max(np.abs(B - ref)) <= _tol(dt) * 8, (trans, diag, entry)


@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("uplo", [UP, LO])
@pytest.mark.parametrize("trans", [NOTR, TR])
def test_syr2k_cpu_sized(bx, orc, dt, uplo, trans):
    rng = np.random.default_rng(5)
    N, K = 21, 15
    shape = (N, K) if trans == NOTR else (K, N)
    A, B, C0 = _rnd(rng, dt, *shape), _rnd(rng, dt, *shape), _rnd(rng, dt, N, N)
    alpha, beta = _scal(dt, 0.5, 0.25), _scal(dt, -1.0, 0.5)
    ref, m = ref_syr2k(orc, uplo, trans, N, K, alpha, A, B, beta, C0)
    L = bx.lib()
    lda = shape[0]
    for entry in ("cblas", "f77", "row"):
        out = C0.copy(order="F")
        if entry == "cblas":
            call(L, "cblas", PFX[dt] + "syr2k", dt, COL, uplo, trans, N, K, alpha, A, lda, B, lda,
                 beta, out, N)
        elif entry == "f77":
            call(L, "f77", PFX[dt] + "syr2k", dt, uplo, trans, N, K, alpha, A, lda, B, lda, beta,
                 out, N)
        else:
            call(L, "cblas", PFX[dt] + "syr2k", dt, ROW, LO if uplo == UP else UP,
                 TR if trans == NOTR else NOTR, N, K, alpha, A, lda, B, lda, beta, out, N)
        assert np.max(np.abs(out[m] - ref[m])) <= _tol(dt) * 8, entry
        assert np.array_equal(out[~m].view(np.uint8), C0[~m].view(np.uint8)), entry


@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [UP, LO])
def test_trsm_cpu_sized(bx, orc, dt, side, uplo):
    rng = np.random.default_rng(6)
    M, N = 19, 13
    na = M if side == LEFT else N
    A = trsm_operand(_rnd(rng, dt, na, na), na)
    B0 = _rnd(rng, dt, M, N)
    alpha = _scal(dt, 1.5, 0.5)
    L = bx.lib()
    for trans in (NOTR, TR, CTR):
        for diag in (NONUNIT, UNIT):
            for entry in ("cblas", "f77", "row"):
                X = B0.copy(order="F")
                if entry == "cblas":
                    call(L, "cblas", PFX[dt] + "trsm", dt, COL, side, uplo, trans, diag, M, N,
                         alpha, A, na, X, M)
                elif entry == "f77":
                    call(L, "f77", PFX[dt] + "trsm", dt, side, uplo, trans, diag, M, N, alpha, A,
                         na, X, M)
                else:
                    call(L, "cblas", PFX[dt] + "trsm", dt, ROW, RIGHT if side == LEFT else LEFT,
                         LO if uplo == UP else UP, trans, diag, N, M, alpha, A, na, X, M)
                err = trsm_residual(orc, side, uplo, trans, diag, M, N, alpha, A, X, B0)
                assert err <= _tol(dt), (trans, diag, entry, err)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("uplo", [UP, LO])
@pytest.mark.parametrize("trans", [NOTR, CTR])
def test_her2k_cpu_sized(bx, orc, dt, uplo, trans):
    rng = np.random.default_rng(7)
    N, K = 21, 15
    shape = (N, K) if trans == NOTR else (K, N)
    A, B, C0 = _rnd(rng, dt, *shape), _rnd(rng, dt, *shape), _rnd(rng, dt, N, N)
    alpha, beta = 0.5 + 0.75j, -0.5
    ref, m = ref_her2k(orc, uplo, trans, N, K, alpha, A, B, beta, C0)
    L = bx.lib()
    lda = shape[0]
    for entry in ("cblas", "f77", "row"):
        out = C0.copy(order="F")
        r = PFX[dt] + "her2k"
        if entry == "cblas":
            call(L, "cblas", r, dt, COL, uplo, trans, N, K, alpha, A, lda, B, lda, beta, out, N)
        elif entry == "f77":
            call(L, "f77", r, dt, uplo, trans, N, K, alpha, A, lda, B, lda, beta, out, N)
        else:
            # the same bytes read row-major: the other triangle and transposition,
            # conj(alpha) (C^T = conj(alpha) B'^H A' + alpha A'^H B')
            call(L, "cblas", r, dt, ROW, LO if uplo == UP else UP,
                 CTR if trans == NOTR else NOTR, N, K, np.conj(alpha), A, lda, B, lda, beta,
                 out, N)
        assert np.max(np.abs(out[m] - ref[m])) <= _tol(dt) * 8, entry
        assert np.array_equal(out[~m].view(np.uint8), C0[~m].view(np.uint8)), entry


@pytest.mark.parametrize("mag", [1e-170, 1e-300, 1.0, 1e170, 1e300])
@pytest.mark.parametrize("conj_ratio", [1, 0])
def test_her2k_unimodular_factor_any_magnitude(bx, mag, conj_ratio):
    """ZHER2K's second term reads B scaled by conj(alpha)/alpha (NoTrans) or
    alpha/conj(alpha) (ConjTrans); the factor must stay finite and of modulus
    one for every finite alpha != 0, however small or large."""
    import ctypes as C

    L = bx.test_lib()
    L.bxtest_her2k_factor.argtypes = [C.c_double, C.c_double, C.c_int,
                                      C.POINTER(C.c_double)]
    for ang in (0.0, 0.3, 1.9, -2.7, np.pi / 2):
        a = mag * complex(np.cos(ang), np.sin(ang))
        out = (C.c_double * 2)()
        L.bxtest_her2k_factor(a.real, a.imag, conj_ratio, out)
        got = complex(out[0], out[1])
        u = complex(np.cos(ang), np.sin(ang))
        want = (u.conjugate() / u) if conj_ratio else (u / u.conjugate())
        assert np.isfinite(got.real) and np.isfinite(got.imag)
        assert abs(got - want) <= 8 * np.finfo(float).eps, (mag, ang, got, want)
