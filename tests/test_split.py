"""Split-precision DGEMM (Ozaki scheme on INT8 MFMA).

No reference counterpart: the algorithm is pinned from its mathematics.
  * CPU: the oracle's digit split reconstructs its input exactly (rational
    arithmetic); on integer inputs the whole oracle pipeline equals the exact
    product; its error against the double-double oracle falls with the
    number of slices.
  * GPU: the library's result equals the oracle restatement bit for bit
    (every split step and every integer product is exact and the FP64 fold
    order is fixed), and on integer inputs equals the exact product.
"""
import sys
from fractions import Fraction

import numpy as np
import pytest

EPS = np.finfo(np.float64).eps


def _op(X, t):
    return X if t == "N" else X.T


def _exact_int(A, B, ta, tb):
    opA = _op(A, ta).astype(np.int64)
    opB = _op(B, tb).astype(np.int64)
    return opA @ opB


# ------------------------------------------------------------------ CPU
def test_digit_split_is_exact(orc):
    rng = np.random.default_rng(30)
    for _ in range(3000):
        x = float(rng.standard_normal() * np.exp(rng.uniform(-40, 40)))
        e = int(np.frexp(abs(x) * rng.uniform(1.0, 1e6))[1])  # any e with |x| 2^-e < 1
        for s in (1, 3, 7, 8):
            d = orc.ozaki_digits(x, e, s)
            assert all(-127 <= v <= 127 for v in d)
            assert all(v == 0 or np.sign(v) == np.sign(x) for v in d)
            rec = sum(Fraction(v) * Fraction(2) ** (-7 * (t + 1)) for t, v in enumerate(d))
            rem = Fraction(x) * Fraction(2) ** (-e) - rec
            assert abs(rem) < Fraction(2) ** (-7 * s), (x, e, s)
            if s == 8 and abs(x) * 2.0 ** -e >= 2.0 ** -3:  # all 53 bits fit in 56
                assert rem == 0


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_oracle_split_exact_on_14bit_integers(orc, ta, tb):
    """|a| < 2^14 with row maxima >= 2^13 splits into exactly two digits;
    with s >= 3 every digit product is kept, so the result is exact."""
    rng = np.random.default_rng(31)
    M, N, K = 23, 17, 41
    A = rng.integers(-(2**14) + 1, 2**14, (M, K) if ta == "N" else (K, M)).astype(np.float64)
    B = rng.integers(-(2**14) + 1, 2**14, (K, N) if tb == "N" else (N, K)).astype(np.float64)
    # different scale per row / column exercises the exponents
    if ta == "N":
        A *= 2.0 ** rng.integers(-20, 20, (M, 1))
    else:
        A *= 2.0 ** rng.integers(-20, 20, (1, M))
    A, B = np.asfortranarray(A), np.asfortranarray(B)
    C0 = np.asfortranarray(rng.integers(-50, 50, (M, N)).astype(np.float64))
    out = C0.copy(order="F")
    orc.dgemm_ozaki(ta, tb, M, N, K, 1.0, A, A.shape[0], B, B.shape[0], 1.0, out, M, 3)
    opA = [[Fraction(v) for v in row] for row in _op(A, ta)]
    opB = [[Fraction(v) for v in row] for row in _op(B, tb)]
    for i in range(M):
        for j in range(N):
            ex = sum(opA[i][p] * opB[p][j] for p in range(K)) + Fraction(C0[i, j])
            assert out[i, j] == float(ex), (i, j)


def test_oracle_split_error_vs_dd_falls_with_slices(orc):
    M, N, K = 40, 30, 300
    A = orc.fill_dynrange(M, K, 3, 0, 0.5)
    B = orc.fill_dynrange(K, N, 3, 1, 0.5)
    C0 = np.zeros((M, N), order="F")
    hi, lo = orc.gemm_dd("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, C0, M)
    scale = np.abs(A) @ np.abs(B)
    errs = []
    for s in (3, 4, 5, 6, 7, 8, 9):
        out = np.zeros((M, N), order="F")
        orc.dgemm_ozaki("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, out, M, s)
        errs.append(np.max(np.abs(out - hi) / scale))
    assert all(errs[i + 1] < errs[i] for i in range(4)), errs
    # with 8 and 9 slices the split product is as accurate as FP64 GEMM allows
    plain = np.zeros((M, N), order="F")
    orc.gemm("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, plain, M)
    assert max(errs[-2:]) <= 2 * np.max(np.abs(plain - hi) / scale) + 4 * EPS


# ------------------------------------------------------------------ GPU
def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
@pytest.mark.parametrize("s", [3, 5, 7, 9])
def test_gpu_split_bit_exact_vs_oracle(gpu, orc, s):
    bx = gpu
    M, N, K = 203, 151, 333
    for ta, tb in [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")]:
        say("split", s, ta, tb, M, N, K)
        A = orc.fill_dynrange(M if ta == "N" else K, K if ta == "N" else M, 5, 0, 2.0)
        B = orc.fill_dynrange(K if tb == "N" else N, N if tb == "N" else K, 5, 1, 2.0)
        C0 = orc.fill_uniform(M, N, 5, 2)
        ref = C0.copy(order="F")
        orc.dgemm_ozaki(ta, tb, M, N, K, 1.25, A, A.shape[0], B, B.shape[0], -0.5, ref, M, s)
        dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
        bx.dgemm_split_device(ta, tb, M, N, K, 1.25, dA.ptr, A.shape[0], dB.ptr, B.shape[0],
                              -0.5, dC.ptr, M, s)
        bx.synchronize()
        out = dC.download()
        nbad = int(np.sum(out != ref))
        assert nbad == 0, f"{nbad} of {M * N} entries differ (max {np.max(np.abs(out - ref))})"


@pytest.mark.gpu
def test_gpu_split_exact_integers_and_beta_zero(gpu, orc):
    bx = gpu
    rng = np.random.default_rng(32)
    M, N, K = 300, 260, 1000
    A = np.asfortranarray(rng.integers(-(2**14) + 1, 2**14, (M, K)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-(2**14) + 1, 2**14, (K, N)).astype(np.float64))
    exact = A.astype(np.int64) @ B.astype(np.int64)
    dA, dB = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(B)
    dC = bx.DeviceArray.from_host(np.full((M, N), np.nan, order="F"))
    say("split exact integers, beta=0 over NaN")
    bx.dgemm_split_device("N", "N", M, N, K, 1.0, dA.ptr, M, dB.ptr, K, 0.0, dC.ptr, M, 3)
    bx.synchronize()
    assert np.array_equal(dC.download(), exact.astype(np.float64))


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "T")])
def test_gpu_split_big_tile_path_sampled(gpu, orc, ta, tb):
    """M, N >= 4096 pads to 256 and runs the 256x256 INT8 kernel; ragged M, N
    and K exercise the padding.  Sampled 48x40 blocks (including the last
    rows/columns) must equal the oracle restatement bit for bit."""
    bx = gpu
    M, N, K = 4100, 4097, 777
    A = orc.fill_dynrange(M if ta == "N" else K, K if ta == "N" else M, 8, 0, 1.0)
    B = orc.fill_dynrange(K if tb == "N" else N, N if tb == "N" else K, 8, 1, 1.0)
    C0 = orc.fill_uniform(M, N, 8, 2)
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
    say("split big tiles", ta, tb, M, N, K)
    bx.dgemm_split_device(ta, tb, M, N, K, 0.75, dA.ptr, A.shape[0], dB.ptr, B.shape[0], 2.0,
                          dC.ptr, M, 6)
    bx.synchronize()
    out = dC.download()
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    for (r0, c0) in [(0, 0), (M - 48, N - 40), (1999, 3001)]:
        As = np.asfortranarray(opA[r0:r0 + 48])
        Bs = np.asfortranarray(opB[:, c0:c0 + 40])
        ref = np.asfortranarray(C0[r0:r0 + 48, c0:c0 + 40])
        orc.dgemm_ozaki("N", "N", 48, 40, K, 0.75, As, 48, Bs, K, 2.0, ref, 48, 6)
        assert np.array_equal(out[r0:r0 + 48, c0:c0 + 40], ref), (r0, c0)


@pytest.mark.gpu
def test_gpu_split_large_sampled_blocks(gpu, orc):
    """2048^3 with a wide dynamic range: sampled 64x64 blocks equal the
    oracle restatement bit for bit, and the error against the double-double
    oracle stays inside the scheme's bound.  With a = 2^e (sum_p d_p 2^-7p +
    r), |r| < 2^-7s, and the pairs p+q > s+1 dropped, every product is off by
    less than 2^(e_i+f_j-7s) (s+2), so
        |C - C_exact|_ij <= K 2^(e_i+f_j-7s) (s+2) + 16 u (|A||B|)_ij
    (the last term covers the FP64 folding of the digit groups)."""
    bx = gpu
    n = 2048
    A = orc.fill_dynrange(n, n, 7, 0, 2.0)
    B = orc.fill_dynrange(n, n, 7, 1, 2.0)
    dA, dB = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(B)
    dC = bx.DeviceArray(n, n, np.float64)
    say("split 2048^3 s=7")
    bx.dgemm_split_device("N", "N", n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0, dC.ptr, n, 7)
    bx.synchronize()
    out = dC.download()
    rng = np.random.default_rng(33)
    for _ in range(2):
        r0, c0 = (int(v) for v in rng.integers(0, n - 64, 2))
        As = np.asfortranarray(A[r0:r0 + 64])
        Bs = np.asfortranarray(B[:, c0:c0 + 64])
        ref = np.zeros((64, 64), order="F")
        orc.dgemm_ozaki("N", "N", 64, 64, n, 1.0, As, 64, Bs, n, 0.0, ref, 64, 7)
        blk = out[r0:r0 + 64, c0:c0 + 64]
        assert np.array_equal(blk, ref)
        hi, _ = orc.gemm_dd("N", "N", 64, 64, n, 1.0, As, 64, Bs, n, 0.0, ref * 0, 64)
        s = 7
        ea = np.frexp(np.abs(As).max(axis=1))[1].astype(np.float64)  # e_i of each row
        eb = np.frexp(np.abs(Bs).max(axis=0))[1].astype(np.float64)  # f_j of each column
        bound = n * (s + 2) * np.exp2(ea[:, None] + eb[None, :] - 7 * s)
        bound += 16 * EPS * (np.abs(As) @ np.abs(Bs))
        assert np.all(np.abs(blk - hi) <= bound)
        # and the bound is not vacuous: it is far below the entries themselves
        assert np.median(bound / (np.abs(As) @ np.abs(Bs))) < 1e-9


def test_oracle_split_error_within_bound(orc):
    """The per-entry bound K 2^(e_i+f_j-7s) (s+2) + 16u(|A||B|)_ij holds for
    the oracle restatement against the double-double oracle, for every s."""
    n = 300
    A = orc.fill_dynrange(40, n, 4, 0, 2.0)
    B = orc.fill_dynrange(n, 30, 4, 1, 2.0)
    hi, _ = orc.gemm_dd("N", "N", 40, 30, n, 1.0, A, 40, B, n, 0.0, np.zeros((40, 30), order="F"), 40)
    ea = np.frexp(np.abs(A).max(axis=1))[1].astype(np.float64)
    eb = np.frexp(np.abs(B).max(axis=0))[1].astype(np.float64)
    for s in range(2, 9):
        out = np.zeros((40, 30), order="F")
        orc.dgemm_ozaki("N", "N", 40, 30, n, 1.0, A, 40, B, n, 0.0, out, 40, s)
        bound = n * (s + 2) * np.exp2(ea[:, None] + eb[None, :] - 7 * s)
        bound += 16 * EPS * (np.abs(A) @ np.abs(B))
        assert np.all(np.abs(out - hi) <= bound), s


def _cplx_int(rng, shape):
    return np.asfortranarray(rng.integers(-(2**14) + 1, 2**14, shape)
                             + 1j * rng.integers(-(2**14) + 1, 2**14, shape))


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "N"), ("T", "C")])
def test_oracle_zsplit_exact_on_integers(orc, ta, tb):
    rng = np.random.default_rng(34)
    M, N, K = 9, 7, 13
    A = _cplx_int(rng, (M, K) if ta == "N" else (K, M))
    B = _cplx_int(rng, (K, N) if tb == "N" else (N, K))
    C0 = _cplx_int(rng, (M, N))

    def op(X, t):
        return X if t == "N" else (X.T if t == "T" else X.conj().T)
    exact = (2 - 1j) * (op(A, ta) @ op(B, tb)) + (1 + 1j) * C0
    out = C0.copy(order="F")
    orc.zgemm_ozaki(ta, tb, M, N, K, 2 - 1j, A, A.shape[0], B, B.shape[0], 1 + 1j, out, M, 3)
    assert np.array_equal(out, exact)


def test_oracle_zsplit_close_to_dd(orc):
    M, N, K = 20, 16, 200
    A = orc.fill_uniform(M, K, 6, 0, dtype=np.complex128)
    B = orc.fill_uniform(K, N, 6, 1, dtype=np.complex128)
    C0 = np.zeros((M, N), dtype=np.complex128, order="F")
    hi, _ = orc.gemm_dd("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, C0, M)
    out = C0.copy(order="F")
    orc.zgemm_ozaki("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, out, M, 8)
    scale = np.abs(A) @ np.abs(B)
    assert np.max(np.abs(out - hi) / scale) <= 8 * EPS


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "T"), ("T", "C")])
def test_gpu_zsplit_bit_exact_vs_oracle(gpu, orc, ta, tb):
    bx = gpu
    M, N, K = 150, 131, 270
    A = orc.fill_uniform(M if ta == "N" else K, K if ta == "N" else M, 9, 0, dtype=np.complex128)
    B = orc.fill_uniform(K if tb == "N" else N, N if tb == "N" else K, 9, 1, dtype=np.complex128)
    C0 = orc.fill_uniform(M, N, 9, 2, dtype=np.complex128)
    ref = C0.copy(order="F")
    orc.zgemm_ozaki(ta, tb, M, N, K, 0.5 + 1j, A, A.shape[0], B, B.shape[0], -1 + 0.25j, ref, M, 6)
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
    say("zsplit", ta, tb, M, N, K)
    bx.zgemm_split_device(ta, tb, M, N, K, 0.5 + 1j, dA.ptr, A.shape[0], dB.ptr, B.shape[0],
                          -1 + 0.25j, dC.ptr, M, 6)
    bx.synchronize()
    assert np.array_equal(dC.download(), ref)
