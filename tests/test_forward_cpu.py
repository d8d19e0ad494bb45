"""The CPU-forwarded BLAS of libblasx (every CBLAS / f77 level-1/2 routine) and
the CPU-sized calls of the level-3 routines, which go to the CPU BLAS as well;
the reference's blas/*.c wrappers (e.g. blas/dgemv.c:9-36, blas/dot.c:74-90,
blas/amax.c:24-40).  Each routine is checked against an
independent numpy formula of its BLAS definition, through both the CBLAS and
the Fortran-77 entry point, including negative increments (which the ILP64
forwarding must sign-extend), packed and banded storage and RowMajor."""
import ctypes as C
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROW, COL = 101, 102
NOTR, TR, CTR = 111, 112, 113
UP, LO = 121, 122
NONUNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142
DT = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}
RT = {"s": np.float32, "d": np.float64, "c": np.float32, "z": np.float64}
TOL = {"s": 2e-4, "d": 1e-12, "c": 2e-4, "z": 1e-12}
CHAR = {NOTR: b"N", TR: b"T", CTR: b"C", UP: b"U", LO: b"L", NONUNIT: b"N", UNIT: b"U",
        LEFT: b"L", RIGHT: b"R"}


class Api:
    """ctypes calls: numpy arrays go as pointers, Python ints as int, real
    scalars as float/double by value (CBLAS) or by reference (f77)."""

    def __init__(self, lib):
        self.lib = lib

    def cblas(self, name, p, *args, restype=None):
        f = getattr(self.lib, "cblas_" + name)
        f.restype = restype
        cargs = []
        for a in args:
            if isinstance(a, np.ndarray):
                cargs.append(a.ctypes.data_as(C.c_void_p))
            elif isinstance(a, int):
                cargs.append(C.c_int(a))
            elif isinstance(a, (np.float32, np.float64, float)):
                cargs.append(C.c_float(a) if RT[p] == np.float32 else C.c_double(a))
            else:
                raise TypeError(type(a))
        f.argtypes = None
        return f(*cargs)

    def f77(self, name, p, *args, restype=None):
        f = getattr(self.lib, name + "_")
        f.restype = restype
        cargs, keep = [], []
        for a in args:
            if isinstance(a, np.ndarray):
                cargs.append(a.ctypes.data_as(C.c_void_p))
            elif isinstance(a, bytes):
                buf = C.create_string_buffer(a)
                keep.append(buf)
                cargs.append(buf)
            elif isinstance(a, int):
                v = C.c_int(a)
                keep.append(v)
                cargs.append(C.byref(v))
            elif isinstance(a, (np.float32, np.float64, float)):
                v = np.array([a], dtype=RT[p])
                keep.append(v)
                cargs.append(v.ctypes.data_as(C.c_void_p))
            else:
                raise TypeError(type(a))
        return f(*cargs)


@pytest.fixture(scope="module")
def api(bx):
    return Api(bx.lib())


def rnd(rng, p, *shape):
    a = rng.uniform(-1, 1, size=shape)
    if p in "cz":
        a = a + 1j * rng.uniform(-1, 1, size=shape)
    return np.asfortranarray(a.astype(DT[p]))


def scal(p, v):
    """alpha/beta as the routine takes it: a 1-element array for complex."""
    return np.array([v], dtype=DT[p]) if p in "cz" else RT[p](v)


def close(p, got, want):
    got, want = np.asarray(got), np.asarray(want)
    scale = max(1.0, float(np.max(np.abs(want))) if want.size else 1.0)
    assert np.max(np.abs(got - want), initial=0) <= TOL[p] * scale * 8, (got, want)


def opm(a, t):
    return a if t == NOTR else (a.T if t == TR else a.conj().T)


# ---------------------------------------------------------------- level 1
@pytest.mark.parametrize("p", "sd")
def test_dot_nrm2_asum_real(api, p):
    rng = np.random.default_rng(1)
    n = 37
    x, y = rnd(rng, p, n), rnd(rng, p, n)
    rt = C.c_float if p == "s" else C.c_double
    close(p, api.cblas(p + "dot", p, n, x, 1, y, 1, restype=rt), x @ y)
    # negative increment: x is traversed from its end
    close(p, api.cblas(p + "dot", p, n, x, -1, y, 1, restype=rt), x[::-1] @ y)
    close(p, api.f77(p + "dot", p, n, x, 1, y, -1, restype=rt), x @ y[::-1])
    close(p, api.f77(p + "dot", p, 19, x, 2, y, 2, restype=rt), x[::2] @ y[::2])
    close(p, api.cblas(p + "nrm2", p, n, x, 1, restype=rt), np.linalg.norm(x))
    close(p, api.f77(p + "nrm2", p, n, x, 1, restype=rt), np.linalg.norm(x))
    close(p, api.cblas(p + "asum", p, n, x, 1, restype=rt), np.abs(x).sum())
    close(p, api.f77(p + "asum", p, 19, x, 2, restype=rt), np.abs(x[::2]).sum())


def test_mixed_precision_dots(api):
    rng = np.random.default_rng(2)
    x, y = rnd(rng, "s", 50), rnd(rng, "s", 50)
    exact = x.astype(np.float64) @ y.astype(np.float64)
    got = api.cblas("dsdot", "s", 50, x, 1, y, 1, restype=C.c_double)
    # OpenBLAS's x86 kernel rounds partial sums through float: bound by that
    assert abs(got - exact) < 1e-6
    assert api.f77("dsdot", "s", 50, x, 1, y, 1, restype=C.c_double) == got
    assert got == _direct(api, "cblas_dsdot", C.c_double, 50, x, 1, y, 1)
    got = api.cblas("sdsdot", "s", 50, np.float32(0.25), x, 1, y, 1, restype=C.c_float)
    assert abs(got - np.float32(0.25 + exact)) <= 1e-6
    got = api.f77("sdsdot", "s", 50, np.float32(0.25), x, 1, y, 1, restype=C.c_float)
    assert abs(got - np.float32(0.25 + exact)) <= 1e-6


def _direct(api, name, restype, *args):
    """the same routine called straight in the CPU BLAS the library loaded"""
    api.lib.blasx_cpublas_path.restype = C.c_char_p
    path = api.lib.blasx_cpublas_path().decode()
    cpu = C.CDLL(path)
    ilp64 = not hasattr(cpu, name)
    f = getattr(cpu, "scipy_%s64_" % name if ilp64 else name)
    f.restype = restype
    it = C.c_int64 if ilp64 else C.c_int
    return f(*[a.ctypes.data_as(C.c_void_p) if isinstance(a, np.ndarray) else it(a)
               for a in args])


class CF(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


class CD(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


@pytest.mark.parametrize("p", "cz")
def test_complex_dots_norms(api, p):
    rng = np.random.default_rng(3)
    n = 29
    x, y = rnd(rng, p, n), rnd(rng, p, n)
    out = np.zeros(1, DT[p])
    api.cblas(p + "dotu_sub", p, n, x, 1, y, 1, out)
    close(p, out[0], x @ y)
    api.cblas(p + "dotc_sub", p, n, x, -1, y, 1, out)
    close(p, out[0], np.conj(x[::-1]) @ y)
    st = CF if p == "c" else CD
    r = api.f77(p + "dotu", p, n, x, 1, y, 1, restype=st)
    close(p, r.re + 1j * r.im, x @ y)
    r = api.f77(p + "dotc", p, n, x, 1, y, 1, restype=st)
    close(p, r.re + 1j * r.im, np.conj(x) @ y)
    rt = C.c_float if p == "c" else C.c_double
    nm = {"c": ("scnrm2", "scasum"), "z": ("dznrm2", "dzasum")}[p]
    close(p, api.cblas(nm[0], p, n, x, 1, restype=rt), np.linalg.norm(x))
    close(p, api.f77(nm[0], p, n, x, 1, restype=rt), np.linalg.norm(x))
    asum = np.abs(x.real).sum() + np.abs(x.imag).sum()
    close(p, api.cblas(nm[1], p, n, x, 1, restype=rt), asum)
    close(p, api.f77(nm[1], p, n, x, 1, restype=rt), asum)


@pytest.mark.parametrize("p", "sdcz")
def test_iamax(api, p):
    rng = np.random.default_rng(4)
    x = rnd(rng, p, 40)
    mag = np.abs(x.real) + np.abs(x.imag) if p in "cz" else np.abs(x)
    k = int(np.argmax(mag))
    assert api.cblas("i%samax" % p, p, 40, x, 1, restype=C.c_size_t) == k
    assert api.f77("i%samax" % p, p, 40, x, 1, restype=C.c_int) == k + 1
    k2 = int(np.argmax(mag[::3]))
    assert api.f77("i%samax" % p, p, 14, x, 3, restype=C.c_int) == k2 + 1
    # n < 1 and incx <= 0 return 0 from the f77 function
    assert api.f77("i%samax" % p, p, 0, x, 1, restype=C.c_int) == 0
    assert api.f77("i%samax" % p, p, 5, x, 0, restype=C.c_int) == 0


@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("entry", ["cblas", "f77"])
def test_swap_copy_axpy_scal(api, p, entry):
    rng = np.random.default_rng(5)
    n = 21
    call = api.cblas if entry == "cblas" else api.f77
    x, y = rnd(rng, p, n), rnd(rng, p, n)
    x0, y0 = x.copy(order="F"), y.copy(order="F")
    call(p + "swap", p, n, x, 1, y, -1)
    close(p, x, y0[::-1])
    close(p, y, x0[::-1])
    x, y = x0.copy(order="F"), y0.copy(order="F")
    call(p + "copy", p, 11, x, 2, y, 1)
    close(p, y[:11], x0[::2])
    close(p, y[11:], y0[11:])
    y = y0.copy(order="F")
    alpha = 0.75 - 0.5j if p in "cz" else 0.75
    call(p + "axpy", p, n, scal(p, alpha), x, 1, y, -1)
    close(p, y, y0 + alpha * x0[::-1])
    x = x0.copy(order="F")
    call(p + "scal", p, 7, scal(p, alpha), x, 3)
    want = x0.copy(order="F")
    want[::3] *= alpha
    close(p, x, want)
    if p in "cz":
        x = x0.copy(order="F")
        call({"c": "csscal", "z": "zdscal"}[p], p, n, RT[p](-1.5), x, 1)
        close(p, x, -1.5 * x0)


@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("entry", ["cblas", "f77"])
def test_rot(api, p, entry):
    rng = np.random.default_rng(6)
    n = 17
    call = api.cblas if entry == "cblas" else api.f77
    x, y = rnd(rng, p, n), rnd(rng, p, n)
    x0, y0 = x.copy(order="F"), y.copy(order="F")
    c, s = RT[p](0.6), RT[p](0.8)
    name = {"s": "srot", "d": "drot", "c": "csrot", "z": "zdrot"}[p]
    call(name, p, n, x, 1, y, 1, c, s)
    close(p, x, c * x0 + s * y0)
    close(p, y, c * y0 - s * x0)


@pytest.mark.parametrize("p", "sd")
@pytest.mark.parametrize("entry", ["cblas", "f77"])
def test_rotg_rotmg_rotm(api, p, entry):
    call = api.cblas if entry == "cblas" else api.f77
    a, b, c, s = (np.array([v], DT[p]) for v in (3.0, 4.0, 0.0, 0.0))
    call(p + "rotg", p, a, b, c, s)
    close(p, [a[0], c[0], s[0]], [5.0, 0.6, 0.8])
    # rotmg builds H with H @ [sqrt(d1) x1, sqrt(d2) y1] zeroing the second
    # component; applying it with rotm to (x1, y1) must give (., 0)
    d1, d2, x1 = (np.array([v], DT[p]) for v in (2.0, 3.0, 1.5))
    y1 = RT[p](0.5)
    param = np.zeros(5, DT[p])
    call(p + "rotmg", p, d1, d2, x1, y1, param)
    xv, yv = np.array([1.5], DT[p]), np.array([0.5], DT[p])
    call(p + "rotm", p, 1, xv, 1, yv, 1, param)
    assert abs(yv[0]) <= TOL[p] * 10
    close(p, xv, x1)
    # rotm with flag -1 applies the full matrix [[h11, h12], [h21, h22]]
    rng = np.random.default_rng(7)
    x, y = rnd(rng, p, 9), rnd(rng, p, 9)
    x0, y0 = x.copy(order="F"), y.copy(order="F")
    P = np.array([-1.0, 0.5, -0.25, 2.0, 1.5], DT[p])
    call(p + "rotm", p, 9, x, 1, y, 1, P)
    close(p, x, 0.5 * x0 + 2.0 * y0)
    close(p, y, -0.25 * x0 + 1.5 * y0)


@pytest.mark.parametrize("p", "cz")
def test_complex_rotg_f77(api, p):
    ca, cb = np.array([3 + 4j], DT[p]), np.array([1 - 2j], DT[p])
    c, s = np.zeros(1, RT[p]), np.zeros(1, DT[p])
    a0, b0 = ca[0], cb[0]
    api.f77(p + "rotg", p, ca, cb, c, s)
    # [c s; -conj(s) c] [a; b] = [r; 0] with c real, |c|^2 + |s|^2 = 1
    close(p, c[0] * a0 + s[0] * b0, ca[0])
    close(p, -np.conj(s[0]) * a0 + c[0] * b0, 0)
    close(p, c[0] ** 2 + abs(s[0]) ** 2, 1)


def test_cabs1_and_lsame(api):
    z = np.array([-3 + 4j], np.complex128)
    api.lib.dcabs1_.restype = C.c_double
    assert api.lib.dcabs1_(z.ctypes.data_as(C.c_void_p)) == 7.0
    zf = np.array([2 - 0.5j], np.complex64)
    api.lib.scabs1_.restype = C.c_float
    assert api.lib.scabs1_(zf.ctypes.data_as(C.c_void_p)) == 2.5
    api.lib.lsame_.restype = C.c_int
    assert api.lib.lsame_(b"n", b"N") == 1
    assert api.lib.lsame_(b"T", b"t") == 1
    assert api.lib.lsame_(b"N", b"T") == 0


# ---------------------------------------------------------------- level 2
def band(a, kl, ku):
    """general band storage: AB[ku + i - j, j] = A[i, j]"""
    m, n = a.shape
    ab = np.zeros((kl + ku + 1, n), a.dtype, order="F")
    for j in range(n):
        for i in range(max(0, j - ku), min(m, j + kl + 1)):
            ab[ku + i - j, j] = a[i, j]
    return ab


def pack(a, uplo):
    n = a.shape[0]
    out = []
    for j in range(n):
        out.extend(a[: j + 1, j] if uplo == UP else a[j:, j])
    return np.array(out, a.dtype)


def sym_band(a, k, uplo):
    n = a.shape[0]
    ab = np.zeros((k + 1, n), a.dtype, order="F")
    for j in range(n):
        if uplo == UP:
            for i in range(max(0, j - k), j + 1):
                ab[k + i - j, j] = a[i, j]
        else:
            for i in range(j, min(n, j + k + 1)):
                ab[i - j, j] = a[i, j]
    return ab


def tri(a, uplo, diag):
    t = np.triu(a) if uplo == UP else np.tril(a)
    if diag == UNIT:
        np.fill_diagonal(t, 1)
    return t


def herm(a, uplo, hermitian):
    """the symmetric/Hermitian matrix the stored triangle of a defines"""
    t = np.triu(a) if uplo == UP else np.tril(a)
    other = t.conj().T if hermitian else t.T
    full = t + other - np.diag(np.diag(t))
    if hermitian:
        np.fill_diagonal(full, np.diag(full).real)
    return full


@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("trans", [NOTR, TR, CTR])
def test_gemv_gbmv(api, p, trans):
    rng = np.random.default_rng(8)
    m, n, kl, ku = 13, 9, 2, 3
    a = rnd(rng, p, m, n)
    alpha, beta = (0.5 + 0.25j, -1.0 + 0.5j) if p in "cz" else (0.5, -1.0)
    ylen = m if trans == NOTR else n
    xlen = n if trans == NOTR else m
    x, y = rnd(rng, p, xlen), rnd(rng, p, ylen)
    want = alpha * opm(a, trans) @ x + beta * y
    yy = y.copy(order="F")
    api.cblas(p + "gemv", p, COL, trans, m, n, scal(p, alpha), a, m, x, 1, scal(p, beta), yy, 1)
    close(p, yy, want)
    yy = y.copy(order="F")
    api.f77(p + "gemv", p, CHAR[trans], m, n, scal(p, alpha), a, m, x, 1, scal(p, beta), yy, 1)
    close(p, yy, want)
    # RowMajor: the same numbers read as the transposed matrix
    ar = np.ascontiguousarray(a)
    yy = y.copy(order="F")
    api.cblas(p + "gemv", p, ROW, trans, m, n, scal(p, alpha), ar, n, x, 1, scal(p, beta), yy, 1)
    close(p, yy, want)
    ab = band(a, kl, ku)
    bandm = np.zeros_like(a)
    for j in range(n):
        lo, hi = max(0, j - ku), min(m, j + kl + 1)
        bandm[lo:hi, j] = a[lo:hi, j]
    want = alpha * opm(bandm, trans) @ x + beta * y
    yy = y.copy(order="F")
    api.cblas(p + "gbmv", p, COL, trans, m, n, kl, ku, scal(p, alpha), ab, kl + ku + 1, x, 1,
              scal(p, beta), yy, 1)
    close(p, yy, want)
    yy = y.copy(order="F")
    api.f77(p + "gbmv", p, CHAR[trans], m, n, kl, ku, scal(p, alpha), ab, kl + ku + 1, x, 1,
            scal(p, beta), yy, 1)
    close(p, yy, want)


@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("uplo", [UP, LO])
@pytest.mark.parametrize("trans", [NOTR, TR, CTR])
@pytest.mark.parametrize("diag", [NONUNIT, UNIT])
def test_triangular_mv_sv(api, p, uplo, trans, diag):
    rng = np.random.default_rng(9)
    n, k = 8, 2
    a = rnd(rng, p, n, n) + 4 * np.eye(n, dtype=DT[p])  # well conditioned
    a = np.asfortranarray(a)
    x0 = rnd(rng, p, n)
    t = tri(a, uplo, diag)
    tb = t.copy(order="F")
    for j in range(n):
        for i in range(n):
            if abs(i - j) > k:
                tb[i, j] = 0
    ab = sym_band(a, k, uplo)
    ap = pack(a, uplo)
    c = CHAR
    for name, mat, stor, extra in (("trmv", t, a, (n,)), ("tbmv", tb, ab, (k + 1,)),
                                   ("tpmv", t, ap, ())):
        kargs = (k,) if name == "tbmv" else ()
        x = x0.copy(order="F")
        api.cblas(p + name, p, COL, uplo, trans, diag, n, *kargs, stor, *extra, x, 1)
        close(p, x, opm(mat, trans) @ x0)
        x = x0.copy(order="F")
        api.f77(p + name, p, c[uplo], c[trans], c[diag], n, *kargs, stor, *extra, x, 1)
        close(p, x, opm(mat, trans) @ x0)
        sv = name[:2] + "sv"
        x = x0.copy(order="F")
        api.cblas(p + sv, p, COL, uplo, trans, diag, n, *kargs, stor, *extra, x, 1)
        close(p, opm(mat, trans) @ x, x0)
        x = x0.copy(order="F")
        api.f77(p + sv, p, c[uplo], c[trans], c[diag], n, *kargs, stor, *extra, x, 1)
        close(p, opm(mat, trans) @ x, x0)


@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("uplo", [UP, LO])
def test_symmetric_hermitian_mv(api, p, uplo):
    rng = np.random.default_rng(10)
    n, k = 9, 2
    hermitian = p in "cz"
    stem = "he" if hermitian else "sy"
    a = rnd(rng, p, n, n)
    full = herm(a, uplo, hermitian)
    fb = full.copy(order="F")
    for j in range(n):
        for i in range(n):
            if abs(i - j) > k:
                fb[i, j] = 0
    x, y = rnd(rng, p, n), rnd(rng, p, n)
    alpha, beta = (0.5 - 0.5j, 2.0 + 0.25j) if hermitian else (0.5, 2.0)
    # the diagonal's imaginary part is not referenced by the Hermitian routines
    names = ((stem + "mv", (a, n)), (("hb" if hermitian else "sb") + "mv",
                                     (sym_band(a, k, uplo), k + 1)),
             (("hp" if hermitian else "sp") + "mv", (pack(a, uplo),)))
    for name, stor in names:
        mat = fb if name[1] == "b" else full
        kargs = (k,) if name[1] == "b" else ()
        want = alpha * mat @ x + beta * y
        yy = y.copy(order="F")
        api.cblas(p + name, p, COL, uplo, n, *kargs, scal(p, alpha), *stor, x, 1, scal(p, beta),
                  yy, 1)
        close(p, yy, want)
        yy = y.copy(order="F")
        api.f77(p + name, p, CHAR[uplo], n, *kargs, scal(p, alpha), *stor, x, 1, scal(p, beta),
                yy, 1)
        close(p, yy, want)


@pytest.mark.parametrize("p", "sdcz")
def test_rank_updates(api, p):
    rng = np.random.default_rng(11)
    m, n = 7, 6
    x, y = rnd(rng, p, m), rnd(rng, p, n)
    a0 = rnd(rng, p, m, n)
    alpha = 0.5 + 1.5j if p in "cz" else 0.5
    gers = ("geru", "gerc") if p in "cz" else ("ger",)
    for g in gers:
        yy = np.conj(y) if g == "gerc" else y
        for entry in (api.cblas, api.f77):
            a = a0.copy(order="F")
            lead = (COL,) if entry == api.cblas else ()
            entry(p + g, p, *lead, m, n, scal(p, alpha), x, 1, y, 1, a, m)
            close(p, a, a0 + alpha * np.outer(x, yy))
    hermitian = p in "cz"
    x = x[:n]
    s0 = rnd(rng, p, n, n)
    for uplo in (UP, LO):
        mask = np.triu(np.ones((n, n), bool)) if uplo == UP else np.tril(np.ones((n, n), bool))
        if hermitian:
            ra = RT[p](0.75)
            upd1 = ra * np.outer(x, np.conj(x))
            upd2 = alpha * np.outer(x, np.conj(y)) + np.conj(alpha) * np.outer(y, np.conj(x))
            r1, r2, a1 = "her", "her2", ra
        else:
            upd1 = alpha * np.outer(x, x)
            upd2 = alpha * (np.outer(x, y) + np.outer(y, x))
            r1, r2, a1 = "syr", "syr2", alpha
        for entry in (api.cblas, api.f77):
            lead = (COL, uplo) if entry == api.cblas else (CHAR[uplo],)
            s = s0.copy(order="F")
            entry(p + r1, p, *lead, n, scal(p, a1) if not hermitian else a1, x, 1, s, n)
            want = s0 + upd1
            if hermitian:
                np.fill_diagonal(want, np.diag(s0).real + np.diag(upd1).real)
            close(p, s[mask], want[mask])
            close(p, s[~mask], s0[~mask])
            s = s0.copy(order="F")
            entry(p + r2, p, *lead, n, scal(p, alpha), x, 1, y, 1, s, n)
            want = s0 + upd2
            if hermitian:
                np.fill_diagonal(want, np.diag(s0).real + np.diag(upd2).real)
            close(p, s[mask], want[mask])
            # packed forms
            ap = pack(s0, uplo)
            entry(p + r1[0] + "pr", p, *lead, n, scal(p, a1) if not hermitian else a1, x, 1, ap)
            want = s0 + upd1
            if hermitian:
                np.fill_diagonal(want, np.diag(s0).real + np.diag(upd1).real)
            close(p, ap, pack(want, uplo))
            ap = pack(s0, uplo)
            entry(p + r2[0] + "pr2", p, *lead, n, scal(p, alpha), x, 1, y, 1, ap)
            want = s0 + upd2
            if hermitian:
                np.fill_diagonal(want, np.diag(s0).real + np.diag(upd2).real)
            close(p, ap, pack(want, uplo))


# ---------------------------------------------------------------- level 3
@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [UP, LO])
def test_symm_hemm(api, p, side, uplo):
    rng = np.random.default_rng(12)
    m, n = 7, 5
    na = m if side == LEFT else n
    a = rnd(rng, p, na, na)
    b, c0 = rnd(rng, p, m, n), rnd(rng, p, m, n)
    alpha, beta = (0.5 + 0.5j, -0.5 + 1j) if p in "cz" else (0.5, -0.5)
    routines = [("symm", herm(a, uplo, False))]
    if p in "cz":
        routines.append(("hemm", herm(a, uplo, True)))
    for name, full in routines:
        want = alpha * (full @ b if side == LEFT else b @ full) + beta * c0
        c = c0.copy(order="F")
        api.cblas(p + name, p, COL, side, uplo, m, n, scal(p, alpha), a, na, b, m, scal(p, beta),
                  c, m)
        close(p, c, want)
        c = c0.copy(order="F")
        api.f77(p + name, p, CHAR[side], CHAR[uplo], m, n, scal(p, alpha), a, na, b, m,
                scal(p, beta), c, m)
        close(p, c, want)


def _rank_k_cases():
    """(precision, routine, trans): syrk/syr2k take N or T; herk/her2k N or C"""
    out = []
    for p in "sdcz":
        for name in ("syrk", "syr2k"):
            out += [(p, name, NOTR), (p, name, TR)]
        if p in "cz":
            for name in ("herk", "her2k"):
                out += [(p, name, NOTR), (p, name, CTR)]
    return out


@pytest.mark.parametrize("p,name,trans", _rank_k_cases())
@pytest.mark.parametrize("uplo", [UP, LO])
def test_rank_k(api, p, name, trans, uplo):
    rng = np.random.default_rng(13)
    n, k = 6, 4
    her = name.startswith("her")
    shape = (n, k) if trans == NOTR else (k, n)
    a, b = rnd(rng, p, *shape), rnd(rng, p, *shape)
    c0 = rnd(rng, p, n, n)
    mask = np.triu(np.ones((n, n), bool)) if uplo == UP else np.tril(np.ones((n, n), bool))
    # op(X) @ opT(X) is X X^T (X X^H for her*) for N, X^T X (X^H X) otherwise
    tr = (lambda x: x.conj().T) if her else (lambda x: x.T)

    def op(x):
        return x if trans == NOTR else tr(x)

    def opT(x):
        return tr(x) if trans == NOTR else x

    alpha = 0.5 + 0.5j if p in "cz" else 0.5
    beta = -0.75 + 0.5j if p in "cz" else -0.75
    if name == "herk":
        alpha, beta = RT[p](0.5), RT[p](-0.75)
    elif name == "her2k":
        beta = RT[p](-0.75)
    if name.endswith("rk"):
        want = alpha * op(a) @ opT(a) + beta * c0
        sargs = (alpha if her else scal(p, alpha), a, n if trans == NOTR else k,
                 beta if her else scal(p, beta))
    else:
        second = np.conj(alpha) if her else alpha
        want = alpha * op(a) @ opT(b) + second * op(b) @ opT(a) + beta * c0
        lda = n if trans == NOTR else k
        sargs = (scal(p, alpha), a, lda, b, lda, beta if her else scal(p, beta))
    if her:
        np.fill_diagonal(want, np.diag(want).real)
    for entry in (api.cblas, api.f77):
        c = c0.copy(order="F")
        lead = (COL, uplo, trans) if entry == api.cblas else (CHAR[uplo], CHAR[trans])
        entry(p + name, p, *lead, n, k, *sargs, c, n)
        close(p, c[mask], want[mask])
        close(p, c[~mask], c0[~mask])


@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [UP, LO])
@pytest.mark.parametrize("trans", [NOTR, TR, CTR])
def test_trmm_trsm(api, p, side, uplo, trans):
    rng = np.random.default_rng(14)
    m, n = 6, 5
    na = m if side == LEFT else n
    a = np.asfortranarray(rnd(rng, p, na, na) + 4 * np.eye(na, dtype=DT[p]))
    b0 = rnd(rng, p, m, n)
    alpha = 0.5 - 0.25j if p in "cz" else 0.5
    for diag in (NONUNIT, UNIT):
        t = opm(tri(a, uplo, diag), trans)
        want = alpha * (t @ b0 if side == LEFT else b0 @ t)
        for entry in (api.cblas, api.f77):
            lead = ((COL, side, uplo, trans, diag) if entry == api.cblas else
                    (CHAR[side], CHAR[uplo], CHAR[trans], CHAR[diag]))
            b = b0.copy(order="F")
            entry(p + "trmm", p, *lead, m, n, scal(p, alpha), a, na, b, m)
            close(p, b, want)
            b = b0.copy(order="F")
            entry(p + "trsm", p, *lead, m, n, scal(p, alpha), a, na, b, m)
            close(p, (t @ b if side == LEFT else b @ t), alpha * b0)


def test_generated_sources_are_current():
    """forward.c, cpu_forward.h and blasx_blas.h are what gen_forward.py writes."""
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([sys.executable, os.path.join(ROOT, "blasx_amd", "csrc", "gen_forward.py"), d],
                       check=True)
        csrc = os.path.join(ROOT, "blasx_amd", "csrc")
        for got, want in (("forward.c", os.path.join(csrc, "forward.c")),
                          ("cpu_forward.h", os.path.join(csrc, "cpu_forward.h")),
                          ("blasx_blas.h", os.path.join(ROOT, "include", "blasx_blas.h"))):
            assert open(os.path.join(d, got)).read() == open(want).read(), want


def test_missing_cpu_routine_ends_the_process_with_a_message(tmp_path):
    """A CPU BLAS without the routine must not turn the call into a silent
    no-op (ddot returning 0, y left unchanged): the process ends with a
    message naming the routine and BLASX_CPUBLAS.  Run in a child process
    whose CPU BLAS is a stub built here that exports cblas_dgemm (so the
    loader accepts it) and nothing else."""
    stub = tmp_path / "libstubblas.so"
    src = tmp_path / "stub.c"
    src.write_text("void cblas_dgemm(void) {}\n")
    subprocess.run(["gcc", "-shared", "-fPIC", "-o", str(stub), str(src)], check=True)
    code = ("import ctypes as C, numpy as np, blasx_amd as bx\n"
            "x = np.ones(4)\n"
            "f = bx.lib().cblas_ddot\n"
            "f.restype = C.c_double\n"
            "v = f(4, x.ctypes.data_as(C.c_void_p), 1, x.ctypes.data_as(C.c_void_p), 1)\n"
            "print('returned', v, flush=True)\n")
    env = dict(os.environ, BLASX_CPUBLAS=str(stub), PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "returned" not in r.stdout, (r.returncode, r.stdout)
    assert "cblas_ddot" in r.stderr and "BLASX_CPUBLAS" in r.stderr, r.stderr[-2000:]
