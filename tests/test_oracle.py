"""Pins the CPU oracle (oracle/) before anything is compared against it.

The reference has no golden vectors or known-answer tests for the GEMM path
(its only driver, testing/gemm.c, checks nothing), so every oracle function
is pinned here by published vectors, exact arithmetic, closed forms, an
independent library (numpy/OpenBLAS), or brute force -- each check chosen so
that a plausible mistake (dropped term, wrong index, transposed operand,
missing conjugate) fails it.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


# ------------------------------------------------------------ generator
def test_splitmix64_published_vectors(orc):
    # Vigna's reference splitmix64.c with state 0 emits these first three
    # values; orc_splitmix64(x) is the output for state x (pre-increment).
    g = 0x9E3779B97F4A7C15
    assert orc.splitmix64(0) == 0xE220A8397B1DCDAF
    assert orc.splitmix64(g) == 0x6E789E6AA1B965F4
    assert orc.splitmix64((2 * g) % 2**64) == 0x06C45D188009454F


def test_uniform_range_and_moments(orc):
    a = orc.fill_uniform(400, 300, 7, 3)
    assert a.min() >= -1.0 and a.max() < 1.0
    assert abs(a.mean()) < 0.01 and abs(a.var() - 1.0 / 3.0) < 0.01
    # element (i,j) uses linear index i + j*rows, independent of ld
    b = orc.fill_uniform(400, 300, 7, 3, ld=411)
    assert np.array_equal(a, b[:400])
    assert a[5, 7] == orc.uniform(7, 3, 5 + 7 * 400)
    # different ids give different streams
    c = orc.fill_uniform(400, 300, 7, 4)
    assert not np.array_equal(a, c)


def test_uniform_mapping_exact(orc):
    # ((x >> 11) * 2^-53) * 2 - 1 for the first published SplitMix64 output
    x = 0xE220A8397B1DCDAF
    expect = float(Fraction(x >> 11, 2**53) * 2 - 1)
    assert orc.uniform(0, 0, 0) == expect


def test_complex_fill_uses_two_counters(orc):
    z = orc.fill_uniform(6, 5, 11, 2, dtype=np.complex128)
    for (i, j) in [(0, 0), (3, 4), (5, 1)]:
        e = i + j * 6
        assert z[i, j].real == orc.uniform(11, 2, 2 * e)
        assert z[i, j].imag == orc.uniform(11, 2, 2 * e + 1)


def test_dynrange_matches_definition(orc):
    import math
    phi = 2.0
    a = orc.fill_dynrange(20, 10, 5, 1, phi)
    for (i, j) in [(0, 0), (7, 3), (19, 9)]:
        idx = i + j * 20
        u = (orc.uniform(5, 1, 3 * idx) + 1.0) * 0.5
        x1 = orc.splitmix64(5 ^ (1 << 56) ^ (3 * idx + 1))
        x2 = orc.splitmix64(5 ^ (1 << 56) ^ (3 * idx + 2))
        u1 = ((x1 >> 11) + 1.0) * 2.0**-53
        u2 = ((x2 >> 11) + 1.0) * 2.0**-53
        g = math.sqrt(-2.0 * math.log(u1)) * math.cos(6.283185307179586 * u2)
        assert a[i, j] == pytest.approx((u - 0.5) * math.exp(phi * g), rel=1e-15)
    big = orc.fill_dynrange(200, 200, 5, 1, 4.0)
    assert np.abs(big).max() / np.median(np.abs(big)) > 1e4  # wide dynamic range


# ------------------------------------------------------------ validation
INFO_CASES = [
    # (transa, transb, M, N, K, lda, ldb, ldc) -> info (blas/dgemm.c:46-53)
    ((110, 111, 1, 1, 1, 1, 1, 1), 1),
    ((111, 114, 1, 1, 1, 1, 1, 1), 2),
    ((111, 111, -1, 1, 1, 1, 1, 1), 3),
    ((111, 111, 1, -1, 1, 1, 1, 1), 4),
    ((111, 111, 1, 1, -1, 1, 1, 1), 5),
    ((111, 111, 5, 1, 1, 4, 1, 5), 8),
    ((112, 111, 5, 1, 7, 6, 7, 5), 8),   # TransA: nrowa = K
    ((111, 111, 5, 3, 7, 5, 6, 5), 10),
    ((111, 112, 5, 3, 7, 5, 2, 5), 10),  # TransB: nrowb = N
    ((111, 111, 5, 3, 7, 5, 7, 4), 13),
    ((111, 111, 0, 0, 0, 1, 1, 1), 0),
    ((111, 111, 0, 0, 0, 0, 1, 1), 8),   # max(1, 0) = 1
    ((113, 113, 5, 3, 7, 7, 3, 5), 0),
    ((110, 114, -1, -1, -1, 0, 0, 0), 1),  # first failing check wins
]


@pytest.mark.parametrize("args,info", INFO_CASES)
def test_gemm_info_numbers(orc, args, info):
    assert orc.gemm_info(*args) == info


def test_f77_trans_chars(orc):
    for ch, v in [("N", 111), ("n", 111), ("T", 112), ("t", 112), ("C", 113), ("c", 113)]:
        assert orc.f77_trans(ch) == v
    for ch in "XxRr ":
        assert orc.f77_trans(ch) == -1


# ------------------------------------------------------------ naive GEMM
def _op(X, t):
    return X if t == "N" else (X.T if t == "T" else X.conj().T)


def _stored(rng, rows, cols, dt, t):
    shape = (rows, cols) if t == "N" else (cols, rows)
    if np.dtype(dt).kind == "c":
        return np.asfortranarray(rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape))
    return np.asfortranarray(rng.uniform(-1, 1, shape).astype(dt))


@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("tb", ["N", "T", "C"])
def test_naive_gemm_matches_numpy(orc, ta, tb):
    rng = np.random.default_rng(1)
    M, N, K = 17, 13, 29
    for dt in (np.float64, np.float32, np.complex128):
        if np.dtype(dt).kind != "c" and "C" in (ta, tb):
            continue
        A = _stored(rng, M, K, dt, ta)
        B = _stored(rng, K, N, dt, tb)
        C0 = np.asfortranarray(_stored(rng, M, N, dt, "N"))
        alpha, beta = (1.5 - 0.25j, 0.5 + 2j) if np.dtype(dt).kind == "c" else (1.5, 0.5)
        ref = alpha * (_op(A.astype(np.complex128) if np.dtype(dt).kind == "c" else A.astype(np.float64), ta)
                       @ _op(B.astype(np.complex128) if np.dtype(dt).kind == "c" else B.astype(np.float64), tb)) + beta * C0
        out = C0.copy(order="F")
        orc.gemm(ta, tb, M, N, K, alpha, A, A.shape[0], B, B.shape[0], beta, out, M)
        eps = EPS32 if dt == np.float32 else EPS64
        assert np.max(np.abs(out - ref)) <= 4 * K * eps * (np.abs(ref).max() + 1)


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_gemm_s_acc64_is_one_rounding_of_the_fp64_product(orc, ta, tb):
    """orc_gemm_s_acc64 (the FP32 reference of the SGEMM and grouped checks):
    float products are exact in FP64, so the result must be the FP64 product
    of the same float32 inputs rounded once to float32, up to K*eps64 of the
    accumulation.  A float32 accumulator (error ~K*eps32) fails this."""
    rng = np.random.default_rng(11)
    M, N, K = 21, 18, 257
    A = _stored(rng, M, K, np.float32, ta)
    B = _stored(rng, K, N, np.float32, tb)
    C0 = np.asfortranarray(_stored(rng, M, N, np.float32, "N"))
    for alpha, beta in ((1.5, 0.5), (-0.75, 0.0)):
        ref = (alpha * (_op(A.astype(np.float64), ta) @ _op(B.astype(np.float64), tb))
               + beta * C0.astype(np.float64))
        mag = (abs(alpha) * (np.abs(_op(A.astype(np.float64), ta)) @ np.abs(_op(B.astype(np.float64), tb)))
               + abs(beta) * np.abs(C0.astype(np.float64)))
        out = C0.copy(order="F")
        orc.gemm(ta, tb, M, N, K, alpha, A, A.shape[0], B, B.shape[0], beta, out, M, acc64=True)
        assert out.dtype == np.float32
        err = np.abs(out.astype(np.float64) - ref)
        assert np.all(err <= 0.5 * EPS32 * np.abs(ref) + 4 * K * EPS64 * mag)
        # the plain float32 loop is measurably worse on the same inputs
        plain = C0.copy(order="F")
        orc.gemm(ta, tb, M, N, K, alpha, A, A.shape[0], B, B.shape[0], beta, plain, M)
        assert np.max(np.abs(plain.astype(np.float64) - ref)) > np.max(err)


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_gemm_s_acc64_integer_exact(orc, ta, tb):
    """Integers up to 2^11 with K = 31: every partial sum is exact in FP64
    (not in FP32: sums reach 2^27), so the result is the exact integer
    rounded once to float32, bit for bit; beta = 0 never reads C (NaN)."""
    rng = np.random.default_rng(12)
    M, N, K = 19, 23, 31
    Ai = rng.integers(-2048, 2049, (M, K) if ta == "N" else (K, M))
    Bi = rng.integers(-2048, 2049, (K, N) if tb == "N" else (N, K))
    Ci = rng.integers(-4096, 4097, (M, N))
    A = np.asfortranarray(Ai, dtype=np.float32)
    B = np.asfortranarray(Bi, dtype=np.float32)
    exact = 3 * (_op(Ai, ta) @ _op(Bi, tb)) - 2 * Ci
    out = np.asfortranarray(Ci, dtype=np.float32)
    orc.gemm(ta, tb, M, N, K, 3.0, A, A.shape[0], B, B.shape[0], -2.0, out, M, acc64=True)
    assert np.array_equal(out, exact.astype(np.float32))
    assert np.abs(exact).max() > 2 ** 24   # beyond float32's exact integers
    out = np.full((M, N), np.nan, dtype=np.float32, order="F")
    orc.gemm(ta, tb, M, N, K, 1.0, A, A.shape[0], B, B.shape[0], 0.0, out, M, acc64=True)
    assert np.array_equal(out, (_op(Ai, ta) @ _op(Bi, tb)).astype(np.float32))


def test_naive_gemm_integer_exact(orc):
    rng = np.random.default_rng(2)
    M, N, K = 23, 19, 31
    for ta in "NT":
        for tb in "NT":
            Ai = rng.integers(-256, 257, (M, K) if ta == "N" else (K, M))
            Bi = rng.integers(-256, 257, (K, N) if tb == "N" else (N, K))
            Ci = rng.integers(-1000, 1001, (M, N))
            exact = 3 * (_op(Ai, ta) @ _op(Bi, tb)) - 2 * Ci
            A, B = np.asfortranarray(Ai, dtype=np.float64), np.asfortranarray(Bi, dtype=np.float64)
            out = np.asfortranarray(Ci, dtype=np.float64)
            orc.gemm(ta, tb, M, N, K, 3.0, A, A.shape[0], B, B.shape[0], -2.0, out, M)
            assert np.array_equal(out, exact.astype(np.float64))


def test_naive_gemm_closed_forms(orc):
    rng = np.random.default_rng(3)
    M, N, K = 9, 7, 5
    A = np.asfortranarray(rng.uniform(-1, 1, (M, K)))
    B = np.asfortranarray(rng.uniform(-1, 1, (K, N)))
    C0 = np.asfortranarray(rng.uniform(-1, 1, (M, N)))
    # alpha = 0: C = beta C, A and B never read (NaN inputs do not propagate)
    An = np.full_like(A, np.nan)
    out = C0.copy(order="F")
    orc.gemm("N", "N", M, N, K, 0.0, An, M, An, K, 0.5, out, M)
    assert np.array_equal(out, 0.5 * C0)
    # beta = 0: C not read
    out = np.full((M, N), np.nan, order="F")
    orc.gemm("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, out, M)
    assert np.all(np.isfinite(out))
    # K = 0: C = beta C
    out = C0.copy(order="F")
    orc.gemm("N", "N", M, N, 0, 1.0, A, M, B, 1, 2.0, out, M)
    assert np.array_equal(out, 2.0 * C0)
    # A = I: C = alpha B + beta C (exact)
    I = np.asfortranarray(np.eye(K))
    Bs = np.asfortranarray(rng.uniform(-1, 1, (K, N)))
    C1 = np.asfortranarray(rng.uniform(-1, 1, (K, N)))
    out = C1.copy(order="F")
    orc.gemm("N", "N", K, N, K, 1.0, I, K, Bs, K, 1.0, out, K)
    assert np.array_equal(out, Bs + C1)


def test_naive_gemm_ld_padding_untouched(orc):
    rng = np.random.default_rng(4)
    M, N, K, ld = 6, 5, 4, 9
    A = np.asfortranarray(rng.uniform(-1, 1, (ld, K)))
    B = np.asfortranarray(rng.uniform(-1, 1, (ld, N)))
    C0 = np.asfortranarray(rng.uniform(-1, 1, (ld, N)))
    out = C0.copy(order="F")
    orc.gemm("N", "N", M, N, K, 1.0, A, ld, B, ld, 1.0, out, ld)
    assert np.array_equal(out[M:], C0[M:])
    ref = A[:M] @ B[:K] + C0[:M]
    assert np.allclose(out[:M], ref, rtol=0, atol=1e-14)


def test_complex_conjugation_is_applied(orc):
    # a dropped conjugate changes the answer: pin C-ops against numpy
    rng = np.random.default_rng(5)
    A = _stored(rng, 3, 4, np.complex128, "C")
    B = _stored(rng, 4, 2, np.complex128, "N")
    out = np.zeros((3, 2), dtype=np.complex128, order="F")
    orc.gemm("C", "N", 3, 2, 4, 1.0, A, 4, B, 4, 0.0, out, 3)
    assert np.allclose(out, A.conj().T @ B, atol=1e-14)
    assert not np.allclose(out, A.T @ B, atol=1e-3)


# ------------------------------------------------------------ tiled restatement
def test_tiled_equals_naive_bruteforce(orc):
    """Tile grid, edge extents and transposed stored-tile addressing
    (blas/blasx_dgemm.c:28-94, blas/blasx_mem_control.c:24-42): bit-exact on
    integer inputs for every tile edge 1..7 and ragged sizes."""
    rng = np.random.default_rng(6)
    for T in range(1, 8):
        for (M, N, K) in [(3 * T + 1, 2 * T, T + 1), (T, 3 * T + 1, 2 * T - 1 if T > 1 else 1),
                          (2 * T + 1, T + 2, 3 * T + 1)]:
            for ta in "NT":
                for tb in "NT":
                    A = np.asfortranarray(rng.integers(-8, 9, (M, K) if ta == "N" else (K, M)), dtype=np.float64)
                    B = np.asfortranarray(rng.integers(-8, 9, (K, N) if tb == "N" else (N, K)), dtype=np.float64)
                    C0 = np.asfortranarray(rng.integers(-8, 9, (M, N)), dtype=np.float64)
                    r1, r2 = C0.copy(order="F"), C0.copy(order="F")
                    orc.gemm(ta, tb, M, N, K, 2.0, A, A.shape[0], B, B.shape[0], 3.0, r1, M)
                    orc.gemm_tiled(ta, tb, M, N, K, T, 2.0, A, A.shape[0], B, B.shape[0], 3.0, r2, M)
                    assert np.array_equal(r1, r2), (T, M, N, K, ta, tb)


def test_flop_formulas(orc):
    assert orc.flops(False, 2, 3, 5) == 60.0
    assert orc.flops(True, 2, 3, 5) == 240.0


# ------------------------------------------------------------ double-double
def test_two_sum_two_prod_exact(orc):
    rng = np.random.default_rng(7)
    for _ in range(2000):
        a, b = rng.standard_normal(2) * np.exp(rng.uniform(-30, 30, 2))
        s, e = orc.two_sum(a, b)
        assert Fraction(s) + Fraction(e) == Fraction(a) + Fraction(b)
        p, f = orc.two_prod(a, b)
        assert Fraction(p) + Fraction(f) == Fraction(a) * Fraction(b)


def test_dd_dot_vs_exact_rational(orc):
    rng = np.random.default_rng(8)
    for n in (1, 2, 7, 33, 100):
        x = rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))
        y = rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))
        exact = sum(Fraction(a) * Fraction(b) for a, b in zip(x, y))
        hi, lo = orc.dd_dot(x, y)
        err = abs(Fraction(hi) + Fraction(lo) - exact)
        scale = sum(abs(Fraction(a) * Fraction(b)) for a, b in zip(x, y))
        assert err <= Fraction(4 * n) * Fraction(2) ** -104 * scale
        if exact != 0:
            # hi is the correctly rounded value except in exact-tie-adjacent cases
            assert abs(Fraction(hi) - exact) <= Fraction(1, 2**52) * abs(exact)


def test_dd_gemm_golden_mpmath(orc):
    g = np.load(os.path.join(HERE, "golden", "dd_golden.npz"), allow_pickle=False)
    checked = 0
    for pre in ("r0", "r1", "r2"):
        M, N, K, al, be, ta, tb = g[pre + "_meta"]
        M, N, K, ta, tb = int(M), int(N), int(K), chr(int(ta)), chr(int(tb))
        A, B, C0 = (np.asfortranarray(g[pre + s]) for s in ("_A", "_B", "_C"))
        hi, lo = orc.gemm_dd(ta, tb, M, N, K, float(al), A, A.shape[0], B, B.shape[0],
                             float(be), C0, M)
        ex_hi, ex_lo = g[pre + "_hi"], g[pre + "_lo"]
        scale = abs(al) * (np.abs(_op(A, ta)) @ np.abs(_op(B, tb))) + abs(be) * np.abs(C0)
        err = np.abs((hi - ex_hi) + (lo - ex_lo))
        assert np.all(err <= 8 * K * 2.0**-104 * scale + 1e-300), pre
        assert np.mean(hi == ex_hi) >= 0.9, pre
        checked += 1
    for pre in ("z0", "z1"):
        meta = g[pre + "_meta"]
        M, N, K = int(meta[0].real), int(meta[1].real), int(meta[2].real)
        al, be = complex(meta[3]), complex(meta[4])
        ta, tb = chr(int(meta[5].real)), chr(int(meta[6].real))
        A, B, C0 = (np.asfortranarray(g[pre + s]) for s in ("_A", "_B", "_C"))
        hi, lo = orc.gemm_dd(ta, tb, M, N, K, al, A, A.shape[0], B, B.shape[0], be, C0, M)
        ex = g[pre + "_hi"]
        scale = abs(al) * (np.abs(_op(A, ta)) @ np.abs(_op(B, tb))) + abs(be) * np.abs(C0)
        assert np.all(np.abs(hi - ex) <= 2 * EPS64 * np.abs(ex) + 32 * K * 2.0**-104 * scale), pre
        checked += 1
    assert checked == 5


# ------------------------------------------------------------ CPU BLAS (reference CPU path)
def test_openblas_loader_matches_numpy(orc):
    path = orc.openblas_load()
    assert path
    rng = np.random.default_rng(9)
    for dt in (np.float64, np.float32, np.complex128):
        for ta, tb in [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")]:
            M, N, K = 31, 17, 23
            A = _stored(rng, M, K, dt, ta)
            B = _stored(rng, K, N, dt, tb)
            C0 = np.asfortranarray(_stored(rng, M, N, dt, "N"))
            out = C0.copy(order="F")
            orc.openblas_gemm(ta, tb, M, N, K, 1.25, A, A.shape[0], B, B.shape[0], -0.5, out, M)
            ref = 1.25 * (_op(A, ta) @ _op(B, tb)) - 0.5 * C0
            eps = EPS32 if dt == np.float32 else EPS64
            assert np.max(np.abs(out - ref)) <= 8 * K * eps * np.abs(ref).max()


def test_relerr_metric_catches_transposed_operand(orc):
    """The tolerance rule used by every GPU parity test rejects the classic
    mistake of a transposed operand, even at the size of the smallest test."""
    rng = np.random.default_rng(10)
    M = N = K = 64
    A = np.asfortranarray(rng.uniform(-1, 1, (M, K)))
    B = np.asfortranarray(rng.uniform(-1, 1, (K, N)))
    C0 = np.zeros((M, N), order="F")
    good = A @ B
    bad = A.T @ B
    assert orc.relerr_fro(good, good, 1.0, A, B, 0.0, C0) == 0.0
    assert orc.relerr_fro(bad, good, 1.0, A, B, 0.0, C0) > 1e3 * K * EPS64


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128])
def test_sample_uniform_equals_fill_at_indices(orc, dtype):
    """The sampled generator (used to check full-size results on sampled
    rows and columns) returns exactly what the full fill holds there,
    including the last row and column and repeated indices."""
    rows, cols = 53, 41
    F = orc.fill_uniform(rows, cols, 11, 2, dtype=dtype)
    ri = [0, 7, 52, 7, 31]
    ci = [40, 0, 19, 19]
    S = orc.sample_uniform(rows, ri, ci, 11, 2, dtype)
    assert S.dtype == F.dtype
    assert np.array_equal(S, F[np.ix_(ri, ci)])
