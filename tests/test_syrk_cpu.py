"""Rank-k updates at the boundary without a GPU: argument validation and
xerbla_ output (blas/dsyrk.c:33-38, blas/zherk.c:39-44, blas/dsyrk.c:83-85),
and CPU-sized calls, which go to the CPU BLAS as in the reference
(blas/dsyrk.c:46-58), in every precision, triangle, transposition and
order, checked against the oracle's naive GEMM on the stored triangle."""
import os
import subprocess
import sys

import numpy as np
import pytest

from syrk_common import COL, CTR, LO, NOTR, ROW, TR, UP, cblas_syrk, f77_syrk, mask, reference

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))


def _run(body):
    code = ("import sys, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
            "import blasx_amd as bx\nfrom syrk_common import *\nL = bx.lib()\n"
            "A = np.zeros((8, 8)); Cm = np.zeros((8, 8)); Z = np.zeros((8, 8), complex)\n"
            % (ROOT, HERE)) + body
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return r.stdout


CASES = [
    # (call, expected line)
    ("cblas_syrk(L, np.float64, 0, COL, 120, NOTR, 2, 2, 1.0, A, 2, 0.0, Cm, 2)",
     " ** On entry to DSYRK  parameter number  1 had an illegal value"),
    ("cblas_syrk(L, np.float64, 0, COL, UP, 110, 2, 2, 1.0, A, 2, 0.0, Cm, 2)",
     " ** On entry to DSYRK  parameter number  2 had an illegal value"),
    ("cblas_syrk(L, np.float64, 0, COL, UP, NOTR, -1, 2, 1.0, A, 2, 0.0, Cm, 0)",
     " ** On entry to DSYRK  parameter number  3 had an illegal value"),
    ("cblas_syrk(L, np.float64, 0, COL, UP, NOTR, 2, -1, 1.0, A, 2, 0.0, Cm, 2)",
     " ** On entry to DSYRK  parameter number  4 had an illegal value"),
    ("cblas_syrk(L, np.float64, 0, COL, UP, NOTR, 3, 2, 1.0, A, 2, 0.0, Cm, 3)",
     " ** On entry to DSYRK  parameter number  7 had an illegal value"),
    ("cblas_syrk(L, np.float64, 0, COL, LO, TR, 3, 2, 1.0, A, 1, 0.0, Cm, 3)",
     " ** On entry to DSYRK  parameter number  7 had an illegal value"),
    ("cblas_syrk(L, np.float64, 0, COL, UP, NOTR, 3, 2, 1.0, A, 3, 0.0, Cm, 2)",
     " ** On entry to DSYRK  parameter number 10 had an illegal value"),
    ("cblas_syrk(L, np.float32, 0, COL, UP, NOTR, 3, 2, 1.0, A, 3, 0.0, Cm, 2)",
     " ** On entry to SSYRK  parameter number 10 had an illegal value"),
    ("cblas_syrk(L, np.complex128, 0, COL, UP, CTR, 2, 2, 1.0, Z, 2, 0.0, Z, 2)",
     " ** On entry to ZSYRK  parameter number  2 had an illegal value"),
    ("cblas_syrk(L, np.complex128, 1, COL, UP, TR, 2, 2, 1.0, Z, 2, 0.0, Z, 2)",
     " ** On entry to  ZHERK parameter number  2 had an illegal value"),
    ("f77_syrk(L, np.float64, 0, b'X', b'N', 2, 2, 1.0, A, 2, 0.0, Cm, 2)",
     " ** On entry to DSYRK  parameter number  1 had an illegal value"),
    ("f77_syrk(L, np.float64, 0, b'U', b'Q', 2, 2, 1.0, A, 2, 0.0, Cm, 2)",
     " ** On entry to DSYRK  parameter number  2 had an illegal value"),
]


@pytest.mark.parametrize("call,line", CASES)
def test_syrk_argument_errors(bx, call, line):
    out = _run(call + "\nprint('after')")
    assert out.splitlines() == [line, "after"], out


def test_lowest_info_wins(bx):
    # N < 0 and a bad ldc: parameter 3 is reported
    out = _run("cblas_syrk(L, np.float64, 0, COL, UP, NOTR, -1, 2, 1.0, A, 2, 0.0, Cm, 0)")
    assert "parameter number  3" in out


DTS = [(np.float32, 0), (np.float64, 0), (np.complex128, 0), (np.complex128, 1),
       (np.complex64, 0), (np.complex64, 1)]


@pytest.mark.parametrize("dt,herk", DTS)
@pytest.mark.parametrize("uplo", [UP, LO])
@pytest.mark.parametrize("transname", ["N", "T"])
def test_cpu_sized_calls(bx, orc, dt, herk, uplo, transname):
    """N, K <= 1000 (500 complex): the CPU BLAS runs the call."""
    trans = NOTR if transname == "N" else (CTR if herk else TR)
    N, K = 37, 23
    rng = np.random.default_rng(5)
    shape = (N, K) if trans == NOTR else (K, N)
    A = np.asfortranarray(rng.uniform(-1, 1, shape).astype(dt))
    C0 = np.asfortranarray(rng.uniform(-1, 1, (N, N)).astype(dt))
    if np.dtype(dt).kind == "c":
        A = A + 1j * np.asfortranarray(rng.uniform(-1, 1, shape))
        C0 = C0 + 1j * np.asfortranarray(rng.uniform(-1, 1, (N, N)))
        A, C0 = np.asfortranarray(A.astype(dt)), np.asfortranarray(C0.astype(dt))
    alpha, beta = (0.75, -0.5) if herk or np.dtype(dt).kind != "c" else (0.75 - 0.25j, -0.5 + 1j)
    lda = shape[0]
    ref = reference(orc, dt, herk, uplo, trans, N, K, alpha, A, lda, beta, C0)
    tol = 1e-4 if dt in (np.float32, np.complex64) else 1e-12
    m = mask(N, uplo)
    L = bx.lib()
    for entry in ("cblas", "f77", "row"):
        Cm = C0.copy(order="F")
        if entry == "cblas":
            cblas_syrk(L, dt, herk, COL, uplo, trans, N, K, alpha, A, lda, beta, Cm, N)
        elif entry == "f77":
            f77_syrk(L, dt, herk, uplo, trans, N, K, alpha, A, lda, beta, Cm, N)
        else:
            # RowMajor over the same bytes: C^T with the other triangle, op flipped
            other = LO if uplo == UP else UP
            rtrans = (CTR if herk else TR) if trans == NOTR else NOTR
            cblas_syrk(L, dt, herk, ROW, other, rtrans, N, K, alpha, A, lda, beta, Cm, N)
        assert np.max(np.abs(Cm[m] - ref[m])) <= tol * 8, entry
        assert np.array_equal(Cm[~m].view(np.uint8), C0[~m].view(np.uint8)), entry
