"""Shared helpers of the SYMM / HEMM / TRMM / SYR2K tests: ctypes calls of
the CBLAS and f77 entry points, and reference results built from the
oracle's naive GEMM on the operand the routine defines (blas/dsymm.c:16,
blas/dtrmm.c:17, blas/dsyr2k.c:16 forward these to the CPU BLAS): the full
symmetric / Hermitian matrix from its stored triangle, or the triangular
matrix with the other triangle zero and, for a unit diagonal, ones."""
import ctypes as C

import numpy as np

ROW, COL = 101, 102
NOTR, TR, CTR = 111, 112, 113
UP, LO = 121, 122
NONUNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142
CH = {NOTR: b"N", TR: b"T", CTR: b"C", UP: b"U", LO: b"L", NONUNIT: b"N", UNIT: b"U",
      LEFT: b"L", RIGHT: b"R"}
PFX = {np.float32: "s", np.float64: "d", np.complex128: "z", np.complex64: "c"}


def up(x):
    """complex64 widened exactly to complex128 for the oracle's complex GEMM"""
    return np.asfortranarray(x.astype(np.complex128)) if x.dtype == np.complex64 else x


def _sc(dt, v, f77=False):
    if dt == np.float32:
        x = C.c_float(v)
    elif dt == np.float64:
        x = C.c_double(v)
    elif dt == np.complex64:
        x = (C.c_float * 2)(complex(v).real, complex(v).imag)
        return C.cast(x, C.c_void_p), x
    else:
        x = (C.c_double * 2)(complex(v).real, complex(v).imag)
        return C.cast(x, C.c_void_p), x
    return (C.byref(x) if f77 else x), x


def _p(a):
    return C.c_void_p(a.ctypes.data if isinstance(a, np.ndarray) else a)


def _i(v):
    x = C.c_int(v)
    return C.byref(x), x


def call(L, entry, routine, dt, *args):
    """entry 'cblas' or 'f77'; args in CBLAS order (with Order first for
    'cblas', without it for 'f77'); scalars are Python numbers, matrices
    numpy arrays or device addresses, enums ints."""
    name = ("cblas_" if entry == "cblas" else "") + routine + ("" if entry == "cblas" else "_")
    fn = getattr(L, name)
    fn.restype = None
    keep, out = [], []
    kinds = KINDS[routine[1:]]
    if entry == "cblas":
        kinds = "e" + kinds
    for kind, a in zip(kinds, args):
        if kind == "e":
            out.append(C.c_int(a) if entry == "cblas" else C.c_char_p(CH.get(a, b"X")))
        elif kind == "i":
            if entry == "cblas":
                out.append(C.c_int(a))
            else:
                r, x = _i(a)
                keep.append(x)
                out.append(r)
        elif kind == "s":
            r, x = _sc(dt, a, f77=entry == "f77")
            keep.append(x)
            out.append(r)
        elif kind == "r":  # a real scalar (HER2K's beta): float for complex64
            x = C.c_float(a) if dt == np.complex64 else C.c_double(a)
            keep.append(x)
            out.append(C.byref(x) if entry == "f77" else x)
        else:
            out.append(_p(a))
    fn(*out)


# argument kinds after Order: e enum, i int, s scalar, m matrix
KINDS = {"symm": "eeiismimismi", "hemm": "eeiismimismi",
         "trmm": "eeeeiismimi", "trsm": "eeeeiismimi", "syr2k": "eeiismimismi",
         "her2k": "eeiismimirmi"}


def full_sym(A, uplo, herm):
    n = A.shape[0]
    t = np.tril(A) if uplo == LO else np.triu(A)
    other = (t.conj().T if herm else t.T)
    f = t + other - np.diag(np.diag(t))
    if herm:
        f[np.arange(n), np.arange(n)] = np.diag(t).real
    return np.asfortranarray(f)


def full_tri(A, uplo, diag):
    t = np.tril(A) if uplo == LO else np.triu(A)
    t = np.array(t, order="F")
    if diag == UNIT:
        t[np.arange(A.shape[0]), np.arange(A.shape[0])] = 1
    return t


def ref_symm(orc, side, uplo, herm, M, N, alpha, A, B, beta, C0):
    A, B, C0 = up(A), up(B), up(C0)
    F = full_sym(A, uplo, herm)
    out = np.asfortranarray(C0.copy())
    if side == LEFT:
        orc.gemm("N", "N", M, N, M, alpha, F, M, B, M, beta, out, M)
    else:
        orc.gemm("N", "N", M, N, N, alpha, B, M, F, N, beta, out, M)
    return out


def ref_trmm(orc, side, uplo, trans, diag, M, N, alpha, A, B0):
    A, B0 = up(A), up(B0)
    T = full_tri(A, uplo, diag)
    t = {NOTR: "N", TR: "T", CTR: "C"}[trans]
    if B0.dtype.kind != "c" and t == "C":
        t = "T"
    out = np.zeros_like(B0, order="F")
    if side == LEFT:
        orc.gemm(t, "N", M, N, M, alpha, T, M, B0, M, 0.0, out, M)
    else:
        orc.gemm("N", t, M, N, N, alpha, B0, M, T, N, 0.0, out, M)
    return out


def trsm_residual(orc, side, uplo, trans, diag, M, N, alpha, A, X, B0):
    """backward error of a TRSM result X: ||op(T) X - alpha B|| (Left) or
    ||X op(T) - alpha B|| (Right) over ||op(T)|| ||X|| + |alpha| ||B||,
    products by the oracle's naive GEMM"""
    A, X, B0 = up(A), up(X), up(B0)
    T = full_tri(A, uplo, diag)
    t = {NOTR: "N", TR: "T", CTR: "C"}[trans]
    if X.dtype.kind != "c" and t == "C":
        t = "T"
    R = np.asfortranarray(B0.copy())
    if side == LEFT:
        orc.gemm(t, "N", M, N, M, 1.0, T, M, X, M, -alpha, R, M)
    else:
        orc.gemm("N", t, M, N, N, 1.0, X, M, T, N, -alpha, R, M)
    den = np.linalg.norm(T.ravel()) * np.linalg.norm(X.ravel()) + \
        abs(alpha) * np.linalg.norm(B0.ravel())
    return np.linalg.norm(R.ravel()) / den


def trsm_operand(A, na):
    """a well-conditioned triangular operand from uniform entries: the
    off-diagonal scaled by 1/n, the diagonal moved near 2"""
    T = np.asfortranarray(A / na)
    T[np.arange(na), np.arange(na)] += 2.0
    return T


def ref_syr2k(orc, uplo, trans, N, K, alpha, A, B, beta, C0):
    A, B, C0 = up(A), up(B), up(C0)
    ta, tb = ("N", "T") if trans == NOTR else ("T", "N")
    lda = A.shape[0]
    full = np.asfortranarray(C0.copy())
    orc.gemm(ta, tb, N, N, K, alpha, A, lda, B, lda, beta, full, N)
    orc.gemm(ta, tb, N, N, K, alpha, B, lda, A, lda, 1.0, full, N)
    out = C0.copy(order="F")
    m = np.tril(np.ones((N, N), bool)) if uplo == LO else np.triu(np.ones((N, N), bool))
    out[m] = full[m]
    return out, m


def ref_her2k(orc, uplo, trans, N, K, alpha, A, B, beta, C0):
    """alpha op(A) op(B)^H + conj(alpha) op(B) op(A)^H + beta C on the stored
    triangle, the diagonal real"""
    A, B, C0 = up(A), up(B), up(C0)
    ta, tb = ("N", "C") if trans == NOTR else ("C", "N")
    lda = A.shape[0]
    full = np.asfortranarray(C0.copy())
    orc.gemm(ta, tb, N, N, K, alpha, A, lda, B, lda, beta, full, N)
    orc.gemm(ta, tb, N, N, K, np.conj(alpha), B, lda, A, lda, 1.0, full, N)
    out = C0.copy(order="F")
    m = np.tril(np.ones((N, N), bool)) if uplo == LO else np.triu(np.ones((N, N), bool))
    out[m] = full[m]
    d = np.arange(N)
    out[d, d] = out[d, d].real
    return out, m
