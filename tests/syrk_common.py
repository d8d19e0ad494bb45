"""Shared helpers of the rank-k update tests: ctypes calls of
cblas_{s,d,z}syrk / cblas_zherk / their f77 forms, and the reference result
built from the oracle's naive GEMM with B = A (blas/dsyrk.c:15-95 defines the
routine as the CPU BLAS's SYRK; C = alpha op(A) op(A)^T + beta C on the
stored triangle, the other triangle untouched, and for HERK op(A)^H, real
alpha and beta and a real diagonal)."""
import ctypes as C

import numpy as np

ROW, COL = 101, 102
NOTR, TR, CTR = 111, 112, 113
UP, LO = 121, 122
CH = {NOTR: b"N", TR: b"T", CTR: b"C", UP: b"U", LO: b"L"}


def name_of(dt, herk):
    if herk:
        return "cherk" if dt == np.complex64 else "zherk"
    return {np.float32: "ssyrk", np.float64: "dsyrk", np.complex128: "zsyrk",
            np.complex64: "csyrk"}[dt]


def _scal(dt, herk, v, f77=False):
    if dt == np.float32 or (dt == np.complex64 and herk):
        x = C.c_float(v)
    elif dt == np.complex64:
        x = (C.c_float * 2)(complex(v).real, complex(v).imag)
        return C.cast(x, C.c_void_p), x
    elif dt == np.float64 or herk:
        x = C.c_double(v)
    else:
        x = (C.c_double * 2)(complex(v).real, complex(v).imag)
        return C.cast(x, C.c_void_p), x
    return (C.byref(x) if f77 else x), x


def _p(a):
    return C.c_void_p(a.ctypes.data if isinstance(a, np.ndarray) else a)


def cblas_syrk(L, dt, herk, order, uplo, trans, N, K, alpha, A, lda, beta, Cm, ldc):
    fn = getattr(L, "cblas_" + name_of(dt, herk))
    fn.restype = None
    al, k1 = _scal(dt, herk, alpha)
    be, k2 = _scal(dt, herk, beta)
    fn(order, uplo, trans, N, K, al, _p(A), lda, be, _p(Cm), ldc)


def f77_syrk(L, dt, herk, uplo, trans, N, K, alpha, A, lda, beta, Cm, ldc):
    fn = getattr(L, name_of(dt, herk) + "_")
    fn.restype = None
    al, k1 = _scal(dt, herk, alpha, f77=True)
    be, k2 = _scal(dt, herk, beta, f77=True)
    ints = [C.c_int(v) for v in (N, K, lda, ldc)]
    u = uplo if isinstance(uplo, bytes) else CH[uplo]
    t = trans if isinstance(trans, bytes) else CH[trans]
    fn(C.c_char_p(u), C.c_char_p(t), C.byref(ints[0]), C.byref(ints[1]), al, _p(A),
       C.byref(ints[2]), be, _p(Cm), C.byref(ints[3]))


def mask(N, uplo):
    full = np.ones((N, N), bool)
    return np.tril(full) if uplo == LO else np.triu(full)


def reference(orc, dt, herk, uplo, trans, N, K, alpha, A, lda, beta, C0):
    """The oracle's naive GEMM with B = A, kept on the stored triangle."""
    ta = "N" if trans == NOTR else ("C" if herk else "T")
    tb = ("C" if herk else "T") if trans == NOTR else "N"
    if A.dtype == np.complex64:  # widened exactly for the oracle's complex GEMM
        A, C0 = np.asfortranarray(A.astype(np.complex128)), np.asfortranarray(C0.astype(np.complex128))
    full = np.asfortranarray(C0.copy())
    orc.gemm(ta, tb, N, N, K, alpha, A, lda, A, lda, beta, full, N)
    out = C0.copy(order="F")
    m = mask(N, uplo)
    out[m] = full[m]
    if herk:
        d = np.arange(N)
        out[d, d] = out[d, d].real
    return out
