"""BF16x3 split DGEMM (blasx_dgemm_bf16x3_device): the BF16 variant of the
split-precision GEMMs.  The piece split is pinned on the CPU (the library's
host build of the kernels' split, the oracle's independent restatement, and
exact rational checks); on the GPU the product is bit-exact where every FP32
partial sum is exact (small integers, K <= 256) and otherwise within the FP32
summation bound of the oracle's restatement, whose own error against the
double-double oracle is the pieces' truncation."""
import ctypes as C
import sys
from fractions import Fraction

import numpy as np
import pytest

U32 = 2.0 ** -24


def _pieces_lib(bx, x):
    L = bx.test_lib()
    out = (C.c_uint16 * 3)()
    L.bxtest_bf16x3_pieces(C.c_double(x), out)
    return [float(np.frombuffer(np.uint32(int(v) << 16).tobytes(), dtype=np.float32)[0])
            for v in out]


def _is_bf16(v):
    if v == 0.0:
        return True
    m, e = np.frexp(v)
    return float(m * 256) == int(m * 256) and e - 1 >= -126


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bf16x3_pieces_library_equals_oracle(bx, orc, seed):
    """The kernels' split built for the host (bit manipulation of the
    double) equals the oracle's frexp / nearbyint restatement; every piece is
    a bf16 value; the three pieces rebuild x to 2^-24 |x|; s1 is x rounded
    to nearest (ties to even) at 8 significant bits, checked exactly."""
    rng = np.random.default_rng(seed)
    xs = list(rng.uniform(-1, 1, 400)) + list(rng.uniform(-1, 1, 100) * 2.0 ** -60)
    xs += [0.0, 0.5, -0.5, 1 - 2 ** -53, 2 ** -126, 2 ** -130, 1 + 2 ** -8 - 1,
           float(Fraction(257, 512)), float(Fraction(259, 512))]  # ties at 8 bits
    for x in xs:
        lib_p = _pieces_lib(bx, x)
        orc_p = orc.bf16x3_pieces(x)
        assert lib_p == orc_p, (x, lib_p, orc_p)
        assert all(_is_bf16(v) for v in orc_p), (x, orc_p)
        if abs(x) >= 2.0 ** -100:
            assert abs(x - sum(orc_p)) <= U32 * abs(x), x
        fx = Fraction(x)
        if x != 0 and abs(x) >= 2.0 ** -126:
            m, e = np.frexp(x)
            t = Fraction(float(m)) * 256
            r = round(t)  # Python rounds half to even
            want = float(r * Fraction(2) ** (int(e) - 8))
            if abs(want) < 2.0 ** -126:
                want = 0.0
            assert orc_p[0] == want, (x, orc_p[0], want)
        del fx


def test_bf16x3_oracle_small_integers_exact(orc):
    rng = np.random.default_rng(5)
    M, N, K = 23, 17, 64
    A = np.asfortranarray(rng.integers(-16, 17, (M, K)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-16, 17, (K, N)).astype(np.float64))
    C0 = np.asfortranarray(rng.integers(-50, 50, (M, N)).astype(np.float64))
    out = C0.copy(order="F")
    orc.dgemm_bf16x3("N", "N", M, N, K, 2.0, A, M, B, K, -1.0, out, M)
    assert np.array_equal(out, 2.0 * (A @ B) - C0)


def test_bf16x3_oracle_error_is_the_pieces_truncation(orc):
    """Against the double-double oracle the restatement's error is the
    pieces' truncation: below 2^-23 of sum_k |a_ik||b_kj| (the FP64 sums of
    the restatement add only 2^-52 K of that)."""
    M, N, K = 30, 20, 200
    A = orc.fill_dynrange(M, K, 4, 0, 2.0)
    B = orc.fill_dynrange(K, N, 4, 1, 2.0)
    hi, _ = orc.gemm_dd("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, np.zeros((M, N), order="F"), M)
    out = np.zeros((M, N), order="F")
    orc.dgemm_bf16x3("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, out, M)
    scale = np.abs(A) @ np.abs(B)
    err = np.max(np.abs(out - hi) / scale)
    assert 0 < err <= 2.0 ** -23, err


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _run(bx, ta, tb, M, N, K, alpha, A, B, beta, C0):
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
    rc = bx.lib().blasx_dgemm_bf16x3_device(bx.TRANS[ta], bx.TRANS[tb], M, N, K, alpha, dA.ptr,
                                            A.shape[0], dB.ptr, B.shape[0], beta, dC.ptr, M,
                                            None)
    assert rc == 0, rc
    bx.synchronize()
    out = dC.download()
    for d in (dA, dB, dC):
        d.free()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_gpu_bf16x3_integer_exact(gpu, orc, ta, tb):
    """Small integers with K <= 256: every piece product and FP32 partial
    sum is exact, so the GPU equals the exact product bit for bit; ragged
    shapes cross the padded edges of both tile engines."""
    bx = gpu
    rng = np.random.default_rng(7)
    for (M, N, K) in ((300, 270, 250), (4100, 4099, 192)):
        say("bf16x3 int", ta, tb, M, N, K)
        A = np.asfortranarray(rng.integers(-16, 17, (M, K) if ta == "N" else (K, M)),
                              dtype=np.float64)
        B = np.asfortranarray(rng.integers(-16, 17, (K, N) if tb == "N" else (N, K)),
                              dtype=np.float64)
        C0 = np.asfortranarray(rng.integers(-50, 50, (M, N)), dtype=np.float64)
        ref = C0.copy(order="F")
        orc.gemm(ta, tb, M, N, K, 2.0, A, A.shape[0], B, B.shape[0], -1.0, ref, M)
        out = _run(bx, ta, tb, M, N, K, 2.0, A, B, -1.0, C0)
        assert np.array_equal(out, ref), (ta, tb, M, N, K)


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "T")])
def test_gpu_bf16x3_within_fp32_sum_bound(gpu, orc, ta, tb):
    """Dynamic-range inputs: the GPU differs from the oracle's restatement
    (FP64 sums of the same pieces) only by the FP32 sums of the three
    groups, |dC| <= 3 K 2^-24 |alpha| (|A||B|)_ij; beta = 0 ignores NaN."""
    bx = gpu
    M, N, K = 700, 650, 1500
    say("bf16x3 dynrange", ta, tb, M, N, K)
    A = orc.fill_dynrange(M if ta == "N" else K, K if ta == "N" else M, 8, 0, 2.0)
    B = orc.fill_dynrange(K if tb == "N" else N, N if tb == "N" else K, 8, 1, 2.0)
    C0 = np.full((M, N), np.nan, order="F")
    ref = np.zeros((M, N), order="F")
    orc.dgemm_bf16x3(ta, tb, M, N, K, 1.5, A, A.shape[0], B, B.shape[0], 0.0, ref, M)
    out = _run(bx, ta, tb, M, N, K, 1.5, A, B, 0.0, C0)
    assert np.isfinite(out).all()
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    bound = 3 * K * U32 * 1.5 * (np.abs(opA) @ np.abs(opB))
    assert np.all(np.abs(out - ref) <= bound)
