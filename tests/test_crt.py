"""CRT split DGEMM (Ozaki scheme II on INT8 MFMA).

No reference counterpart: the method is pinned from its mathematics.
  * CPU: the reconstruction constants of the library and of the oracle
    equal the exact big-integer values; the residue helpers the kernels use
    (compiled for the host from the same source) and the oracle's equal
    Python's exact arithmetic; the oracle's floating CRT recovers integers
    to 2u at every magnitude; Garner's digits rebuild the exact integer; on
    integer inputs the oracle pipeline equals the exact product; its error
    obeys the scheme's bound and falls as moduli are added; the INT8 GEMM's
    block -> tile traversal (host build) takes every tile once.
  * GPU: the library's result equals the oracle restatement bit for bit
    (residues and integer products are exact, the FP64 reconstruction order
    is fixed) and equals the exact product on integer inputs.
"""
import ctypes as C
import math
from fractions import Fraction
import random
import sys

import numpy as np
import pytest

EPS = np.finfo(np.float64).eps
MODS = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
        223, 217, 211, 199, 197, 193, 191, 181, 179, 173]


def say(*a):
    print("[crt]", *a)
    sys.stdout.flush()


def _smod(v, m):
    r = v % m
    return r - m if r >= (m + 1) // 2 else r


def _exact_consts(L):
    M = math.prod(MODS[:L])
    E = M.bit_length()
    ws = [(M // m) * pow(M // m, -1, m) for m in MODS[:L]]

    def chunks(x):
        out = []
        for k in range(4):
            hi, lo = E - 40 * k, max(0, E - 40 * (k + 1))
            out.append(((x >> lo) & ((1 << (hi - lo)) - 1)) << lo if hi > 0 else 0)
        return out

    return M, [chunks(w) for w in ws], chunks(M)


# ------------------------------------------------------------------ CPU

def test_moduli_pairwise_coprime_and_beta(orc, bx):
    assert orc.crt_moduli(20) == MODS
    for i in range(20):
        for j in range(i):
            assert math.gcd(MODS[i], MODS[j]) == 1
    L_ = bx.lib()
    L_.blasx_crt_bits.argtypes = [C.c_int, C.c_int]
    for K in (1, 7, 1000, 16384, 131071):
        for L in range(2, 21):
            b = orc.crt_beta(K, L)
            assert L_.blasx_crt_bits(K, L) == b
            if b >= 1:  # worst case: |A'B'| < K 2^(2b) < M / 2 exactly
                assert 2 * K * 2 ** (2 * b) < math.prod(MODS[:L])
    for L in range(2, 21):
        # scaling by row 2-norms: |A'B'| <= ||A'_i|| ||B'_j|| < 2^(2g) < M / 2
        # with room for the floating quotient (2^(2g+1) < M (1 - 2^-20)), and
        # g is the largest such integer
        g = orc.crt_gamma(L)
        M = math.prod(MODS[:L])
        assert 2 ** (2 * g + 1) * (2 ** 20) < M * (2 ** 20 - 1)
        assert M < 2 * 2 ** (2 * g + 2)


@pytest.mark.parametrize("L", [2, 5, 8, 13, 16, 20])
def test_reconstruction_constants_exact(orc, bx, L):
    M, W, Mk = _exact_consts(L)
    Lb = bx.test_lib()
    Lb.bxtest_crt_consts.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    Wl = np.zeros(80)
    Ml = np.zeros(4)
    inv = C.c_double()
    Lb.bxtest_crt_consts(L, Wl.ctypes.data, Ml.ctypes.data, C.byref(inv))
    Wo, Mo, invo = orc.crt_consts(L)
    for k in range(4):
        assert int(Ml[k]) == Mk[k] and int(Mo[k]) == Mk[k]
        for l in range(L):
            assert int(Wl[k * 20 + l]) == W[l][k] and int(Wo[k][l]) == W[l][k]
    assert sum(Mk) == M
    md = ((float(Mk[3]) + float(Mk[2])) + float(Mk[1])) + float(Mk[0])
    assert inv.value == invo == 1.0 / md


def test_residue_helpers_match_exact_arithmetic(orc, bx):
    """The kernels' residue of an integer-valued double (signed 7-bit digits
    in three 32-bit words, reduced on the int8 dot-product unit and by one
    float quotient) on every magnitude below 2^84, and below 2^56 with two
    words (the split's |A'| < 2^beta), including the digit and word
    boundaries, and the residue of any int32 product sum, against Python's
    exact integers: the library's helpers (host build of the kernel code)
    and the oracle's independent int64 version.  Non-finite values and
    magnitudes beyond the digits give residue 0."""
    Lb = bx.test_lib()
    Lb.bxtest_crt_resid.argtypes = [C.c_double, C.c_int]
    Lb.bxtest_crt_resid2.argtypes = [C.c_double, C.c_int]
    Lb.bxtest_crt_smod_i32.argtypes = [C.c_int, C.c_int]
    rng = random.Random(11)
    vals = [0.0, -0.0, 1.0, -1.0, 127.0, -127.0, 128.0, -128.0, 2.0 ** 53, -(2.0 ** 53),
            2.0 ** 53 - 1, 2.0 ** 26, 2.0 ** 26 - 1, -(2.0 ** 26), 2.0 ** 78,
            -(2.0 ** 78) + 2.0 ** 30, 2.0 ** 83, -(2.0 ** 84 - 2.0 ** 31), 2.0 ** 56 - 2.0 ** 3,
            -(2.0 ** 56 - 2.0 ** 4)]
    for b in range(7, 84, 7):  # digit boundaries (28, 56: word boundaries)
        for v in (2.0 ** b, 2.0 ** b - 1, 2.0 ** b + 1, 2.0 ** b - 2.0 ** max(b - 52, 0)):
            if v < 2.0 ** 53 or v == math.trunc(v):
                vals += [float(v), -float(v)]
    for _ in range(3000):
        e = rng.randrange(0, 84)
        vals.append(float(math.trunc(rng.uniform(-1, 1) * 2.0 ** e)))
    for _ in range(300):  # all-ones digits (127 in every byte) and their negatives
        e = rng.randrange(1, 84)
        v = float(2 ** e - 2 ** max(e - 53, 0))
        vals += [v, -v]
    for v in vals:
        iv = int(v)
        for m in MODS:
            want = _smod(iv, m)
            assert Lb.bxtest_crt_resid(v, m) == want, (v, m)
            assert Lb.bxtest_crt_resid2(v, m) == (want if abs(v) < 2.0 ** 56 else 0), (v, m)
            if abs(v) < 2.0 ** 79:
                assert orc.crt_resid(v, m) == want, (v, m)
    for v in (float("inf"), -float("inf"), float("nan"), 2.0 ** 84, -(2.0 ** 90)):
        for m in MODS:
            assert Lb.bxtest_crt_resid(v, m) == 0 and Lb.bxtest_crt_resid2(v, m) == 0, (v, m)


def _norm_exp_exact(ss):
    """Smallest integer e with fl(ss * (1 + 2^-30)) < 4^e, exactly."""
    u = ss * (1.0 + 2.0 ** -30)
    if u == 0.0 or not math.isfinite(u):
        return 0
    q = Fraction(u)
    e = -600
    while Fraction(4) ** e <= q:
        e += 1
    while e > -600 and Fraction(4) ** (e - 1) > q:
        e -= 1
    return e


def test_norm_exponent_matches_definition(orc, bx):
    """The row exponent of the CRT scaling (smallest e with ss (1 + 2^-30) <
    4^e, from the binary exponent alone): the library's host build of the
    kernel code and the oracle against exact rational arithmetic, on powers
    of 4 and 2 and their neighbours, tiny and huge sums, zero and inf."""
    Lb = bx.test_lib()
    Lb.bxtest_crt_norm_exp.argtypes = [C.c_double]
    rng = random.Random(5)
    vals = [0.0, float("inf"), 5e-324, 1e-300, 1e300]
    for k in range(-60, 61):
        for b in (2.0 ** k, 4.0 ** (k // 2)):
            vals += [b, b * (1 - 2 ** -52), b * (1 + 2 ** -52), b * (1 - 2 ** -31), b * (1 - 2 ** -29)]
    vals += [rng.uniform(0, 1) * 2.0 ** rng.randrange(-200, 200) for _ in range(2000)]
    for ss in vals:
        want = _norm_exp_exact(ss)
        assert Lb.bxtest_crt_norm_exp(ss) == want, ss
        assert orc.norm_exp(ss) == want, ss
        if 0 < ss < float("inf"):  # the row norm is below 2^e
            assert math.sqrt(ss) < 2.0 ** want
    for _ in range(20000):
        x = rng.randrange(-2 ** 31, 2 ** 31)
        m = rng.choice(MODS)
        assert Lb.bxtest_crt_smod_i32(x, m) == _smod(x, m), (x, m)
    for m in MODS:  # quotient ties and range ends
        for x in (m // 2, -(m // 2), (m + 1) // 2, 3 * m // 2, -3 * m // 2, 2 ** 31 - 1, -2 ** 31):
            assert Lb.bxtest_crt_smod_i32(x, m) == _smod(x, m), (x, m)


@pytest.mark.parametrize("mp,np_", [(8192, 4096), (16384, 16384), (4096, 4096),
                                     (12288, 4096), (6144, 2048), (2560, 2304), (256, 256),
                                     (8192, 12288)])
def test_crt_tile_traversal_covers_every_tile_once(bx, mp, np_):
    """Block -> tile origin of the INT8 CRT GEMM (the kernel's own code, host
    build), one block per 256x256 tile: every tile is taken exactly once, and
    where the grid divides into 32x16 super-tiles (16x16 when 32 rows do not
    divide) each window of 8 x 64 consecutive blocks covers one whole
    super-tile, and the 32 blocks of one XCD in each 256 of them (b % 8
    equal) a 4 x 8 patch whose panels they share in L2."""
    Lb = bx.test_lib()
    f = Lb.bxtest_crt_tile_origin
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    T = 256
    tm, tn = mp // T, np_ // T
    nb = tm * tn
    m0, n0 = C.c_int(), C.c_int()
    tiles = []
    for b in range(nb):
        f(b, nb, np_, C.byref(m0), C.byref(n0))
        assert m0.value % T == 0 and n0.value % T == 0, b
        assert 0 <= m0.value < mp and 0 <= n0.value < np_, (b, m0.value, n0.value)
        tiles.append((m0.value // T, n0.value // T))
    assert sorted(tiles) == [(i, j) for i in range(tm) for j in range(tn)]
    if nb % 8 or tn % 16 or tm % 16:
        return
    sr = 32 if tm % 32 == 0 else 16
    win = 8 * 2 * sr  # blocks per super-tile of sr x 16 tiles
    for w0 in range(0, nb, win):
        sup = tiles[w0:w0 + win]
        rows = {t[0] for t in sup}
        cols = {t[1] for t in sup}
        assert len(rows) == sr and len(cols) == 16 and len(set(sup)) == win, w0
        assert max(rows) - min(rows) == sr - 1 and max(cols) - min(cols) == 15, w0
        for x in range(8):
            for h in range(win // 256):
                patch = [tiles[w0 + 256 * h + 8 * c + x] for c in range(32)]
                pr = {t[0] for t in patch}
                pc = {t[1] for t in patch}
                assert len(pr) * len(pc) == 32 and len(pr) == 4, (w0, x, h, pr, pc)


@pytest.mark.parametrize("L", [8, 12, 16, 20])
def test_floating_crt_and_garner(orc, L):
    """Every integer |P| < 2^(2 gamma) (< M/2) is recovered to 2u by the
    floating CRT, and exactly by Garner's digits: the range ends, values
    within 2^-19 of M/2, and random magnitudes."""
    ms = MODS[:L]
    M = math.prod(ms)
    lim = 2 ** (2 * orc.crt_gamma(L))  # the scaling's bound on |A'B'|
    consts = orc.crt_consts(L)
    rng = random.Random(L)
    ends = [lim - 1, -(lim - 1), lim // 2, -(lim // 3), 1, -1, 0]
    for i in range(1500):
        if i < len(ends):
            P = ends[i]
        elif rng.random() < 0.2:
            P = rng.randrange(-2 ** rng.randrange(0, M.bit_length() - 1), 2 ** 60)
            P = max(-(lim - 1), min(lim - 1, P))
        else:
            P = rng.randrange(-(lim - 1), lim) >> rng.randrange(0, 120)
        r = [_smod(P, m) for m in ms]
        S = orc.crt_float(r, consts)
        assert abs(S - P) <= 2 * EPS / 2 * abs(P), (P, S)
        v = orc.crt_digits(r)
        acc, mult = 0, 1
        for d, m in zip(v, ms):
            acc += d * mult
            mult *= m
        assert acc == P


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_oracle_exact_on_integer_inputs(orc, ta, tb):
    """Integers whose row maxima reach 2^20 are kept exactly when beta >= 21
    (A'_ik = a_ik 2^(beta - e_i) is then an integer), so the result is the
    exact integer product."""
    rng = np.random.default_rng(5)
    M, N, K = 37, 29, 61
    A = rng.integers(-(2 ** 20) + 1, 2 ** 20, (M, K) if ta == "N" else (K, M)).astype(np.float64)
    B = rng.integers(-(2 ** 20) + 1, 2 ** 20, (K, N) if tb == "N" else (N, K)).astype(np.float64)
    A, B = np.asfortranarray(A), np.asfortranarray(B)
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    exact = opA.astype(np.int64) @ opB.astype(np.int64)
    emax = max(max(orc.norm_exp(float(np.sum(opA * opA, axis=1).max())), 0),
               max(orc.norm_exp(float(np.sum(opB * opB, axis=0).max())), 0))
    for L in (8, 12, 20):
        assert orc.crt_gamma(L) >= emax  # A' = a 2^(gamma - e) is an integer
        Cc = np.full((M, N), np.nan, order="F")
        orc.dgemm_crt(ta, tb, M, N, K, 1.0, A, A.shape[0], B, B.shape[0], 0.0, Cc, M, L)
        assert np.array_equal(Cc, exact.astype(np.float64)), L


def _crt_bound(orc, As, Bs, beta, alpha=1.0):
    """|C - C_exact| for C = alpha op(A) op(B) with gamma = beta scaling bits:
    truncation of both operands (|a - A' 2^(e-beta)| < 2^(e-beta), e the
    row's 2-norm exponent, taken one higher here so that a different
    summation order of the sum of squares cannot tighten it) plus the
    reconstruction and scaling roundings."""
    ea = 1.0 + np.array([orc.norm_exp(float(v)) for v in np.sum(As * As, axis=1)])
    eb = 1.0 + np.array([orc.norm_exp(float(v)) for v in np.sum(Bs * Bs, axis=0)])
    K = As.shape[1]
    t = (np.exp2(ea - beta)[:, None] * np.abs(Bs).sum(axis=0)[None, :]
         + np.exp2(eb - beta)[None, :] * np.abs(As).sum(axis=1)[:, None]
         + K * np.exp2(ea[:, None] + eb[None, :] - 2 * beta))
    return abs(alpha) * (t * (1 + 1e-12) + 4 * EPS * (np.abs(As) @ np.abs(Bs)))


def test_oracle_error_within_bound_and_falls_with_moduli(orc):
    M, N, K = 48, 40, 300
    A = orc.fill_dynrange(M, K, 3, 0, 2.0)
    B = orc.fill_dynrange(K, N, 3, 1, 2.0)
    hi, lo = orc.gemm_dd("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, np.zeros((M, N), order="F"), M)
    errs = []
    for L in (8, 11, 14, 17, 20):
        Cc = np.zeros((M, N), order="F")
        orc.dgemm_crt("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, Cc, M, L)
        err = np.abs(Cc - hi)
        assert np.all(err <= _crt_bound(orc, A, B, orc.crt_gamma(L))), L
        errs.append(np.max(err / (np.abs(A) @ np.abs(B))))
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 4 * EPS


def test_oracle_alpha_beta_and_zero_rows(orc):
    """alpha/beta follow C = alpha*P + fma(beta, C); a zero row of A and a
    zero column of B give exact zeros (times alpha) plus beta*C."""
    M, N, K = 20, 18, 50
    A = orc.fill_uniform(M, K, 9, 0)
    B = orc.fill_uniform(K, N, 9, 1)
    A[3, :] = 0.0
    B[:, 5] = 0.0
    C0 = orc.fill_uniform(M, N, 9, 2)
    Cc = C0.copy(order="F")
    orc.dgemm_crt("N", "N", M, N, K, -2.0, A, M, B, K, 0.5, Cc, M, 16)
    assert np.array_equal(Cc[3, :], 0.5 * C0[3, :])
    assert np.array_equal(Cc[:, 5], 0.5 * C0[:, 5])
    ref = -2.0 * (A @ B) + 0.5 * C0
    assert np.max(np.abs(Cc - ref)) < 1e-13


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
@pytest.mark.parametrize("L", [8, 14, 20])
def test_gpu_crt_bit_exact_vs_oracle(gpu, orc, L):
    bx = gpu
    M, N, K = 203, 151, 333
    for ta, tb in [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")]:
        say("crt", L, ta, tb, M, N, K)
        A = orc.fill_dynrange(M if ta == "N" else K, K if ta == "N" else M, 5, 0, 2.0)
        B = orc.fill_dynrange(K if tb == "N" else N, N if tb == "N" else K, 5, 1, 2.0)
        C0 = orc.fill_uniform(M, N, 5, 2)
        ref = C0.copy(order="F")
        orc.dgemm_crt(ta, tb, M, N, K, 1.25, A, A.shape[0], B, B.shape[0], -0.5, ref, M, L)
        dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
        bx.dgemm_crt_device(ta, tb, M, N, K, 1.25, dA.ptr, A.shape[0], dB.ptr, B.shape[0],
                            -0.5, dC.ptr, M, L)
        bx.synchronize()
        out = dC.download()
        nbad = int(np.sum(out != ref))
        assert nbad == 0, f"{nbad} of {M * N} entries differ (max {np.max(np.abs(out - ref))})"


@pytest.mark.gpu
def test_gpu_crt_exact_integers_and_beta_zero(gpu, orc):
    bx = gpu
    rng = np.random.default_rng(32)
    M, N, K = 300, 260, 1000
    A = np.asfortranarray(rng.integers(-(2 ** 20) + 1, 2 ** 20, (M, K)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-(2 ** 20) + 1, 2 ** 20, (K, N)).astype(np.float64))
    exact = A.astype(np.int64) @ B.astype(np.int64)
    dA, dB = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(B)
    dC = bx.DeviceArray.from_host(np.full((M, N), np.nan, order="F"))
    say("crt exact integers, beta=0 over NaN")
    bx.dgemm_crt_device("N", "N", M, N, K, 1.0, dA.ptr, M, dB.ptr, K, 0.0, dC.ptr, M, 12)
    bx.synchronize()
    assert np.array_equal(dC.download(), exact.astype(np.float64))


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "T")])
def test_gpu_crt_big_tile_path_sampled(gpu, orc, ta, tb):
    """M, N >= 4096 pads to 256 and runs the 256x256 INT8 kernel; ragged M,
    N and K exercise the padding.  Sampled blocks (including the last
    rows/columns) equal the oracle restatement bit for bit."""
    bx = gpu
    M, N, K = 4100, 4097, 777
    A = orc.fill_dynrange(M if ta == "N" else K, K if ta == "N" else M, 8, 0, 1.0)
    B = orc.fill_dynrange(K if tb == "N" else N, N if tb == "N" else K, 8, 1, 1.0)
    C0 = orc.fill_uniform(M, N, 8, 2)
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
    say("crt big tiles", ta, tb, M, N, K)
    bx.dgemm_crt_device(ta, tb, M, N, K, 0.75, dA.ptr, A.shape[0], dB.ptr, B.shape[0], 2.0,
                        dC.ptr, M, 16)
    bx.synchronize()
    out = dC.download()
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    for (r0, c0) in [(0, 0), (M - 48, N - 40), (1999, 3001)]:
        As = np.asfortranarray(opA[r0:r0 + 48])
        Bs = np.asfortranarray(opB[:, c0:c0 + 40])
        ref = np.asfortranarray(C0[r0:r0 + 48, c0:c0 + 40])
        orc.dgemm_crt("N", "N", 48, 40, K, 0.75, As, 48, Bs, K, 2.0, ref, 48, 16)
        assert np.array_equal(out[r0:r0 + 48, c0:c0 + 40], ref), (r0, c0)


@pytest.mark.gpu
@pytest.mark.parametrize("M,N,K", [(6100, 5889, 300), (8100, 8050, 200), (12200, 4000, 64)])
def test_gpu_crt_many_tiles_sampled(gpu, orc, M, N, K):
    """More 256x256 tiles than CUs, ragged M and N, each tile traversal:
    24x24 tiles (M, N pad to 6144: XCD ranges of 8 tile rows), 32x32 (two
    32x16 super-tiles) and 48x16 (three 16x16 super-tiles).  Sampled blocks
    from the first, middle and last tiles equal the oracle restatement bit
    for bit."""
    bx = gpu
    A = orc.fill_dynrange(M, K, 9, 0, 1.0)
    B = orc.fill_dynrange(K, N, 9, 1, 1.0)
    C0 = orc.fill_uniform(M, N, 9, 2)
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
    say("crt many tiles", M, N, K)
    bx.dgemm_crt_device("N", "N", M, N, K, 1.25, dA.ptr, M, dB.ptr, K, -0.5, dC.ptr, M, 15)
    bx.synchronize()
    out = dC.download()
    for (r0, c0) in [(0, 0), (M - 48, N - 40), (3071, 2999), (5000, 200), (255, N - 300),
                     (M - 300, N // 2), (M // 2, 1000)]:
        As = np.asfortranarray(A[r0:r0 + 48])
        Bs = np.asfortranarray(B[:, c0:c0 + 40])
        ref = np.asfortranarray(C0[r0:r0 + 48, c0:c0 + 40])
        orc.dgemm_crt("N", "N", 48, 40, K, 1.25, As, 48, Bs, K, -0.5, ref, 48, 15)
        assert np.array_equal(out[r0:r0 + 48, c0:c0 + 40], ref), (r0, c0)


@pytest.mark.gpu
def test_gpu_crt_large_within_bound(gpu, orc):
    """2048^3 with a wide dynamic range and the default moduli: sampled
    blocks equal the oracle bit for bit and stay inside the scheme's error
    bound against the double-double oracle."""
    bx = gpu
    n = 2048
    A = orc.fill_dynrange(n, n, 7, 0, 2.0)
    B = orc.fill_dynrange(n, n, 7, 1, 2.0)
    dA, dB = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(B)
    dC = bx.DeviceArray(n, n, np.float64)
    say("crt 2048^3 default moduli")
    bx.dgemm_crt_device("N", "N", n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0, dC.ptr, n, 0)
    bx.synchronize()
    out = dC.download()
    L = next(L for L in range(2, 21) if orc.crt_beta(n, L) >= 56)
    rng = np.random.default_rng(33)
    for _ in range(2):
        r0, c0 = (int(v) for v in rng.integers(0, n - 64, 2))
        As = np.asfortranarray(A[r0:r0 + 64])
        Bs = np.asfortranarray(B[:, c0:c0 + 64])
        ref = np.zeros((64, 64), order="F")
        orc.dgemm_crt("N", "N", 64, 64, n, 1.0, As, 64, Bs, n, 0.0, ref, 64, L)
        blk = out[r0:r0 + 64, c0:c0 + 64]
        assert np.array_equal(blk, ref)
        hi, _ = orc.gemm_dd("N", "N", 64, 64, n, 1.0, As, 64, Bs, n, 0.0, ref * 0, 64)
        assert np.all(np.abs(blk - hi) <= _crt_bound(orc, As, Bs, orc.crt_gamma(L)))


# ------------------------------------------------------------------ complex

def _cplx_int(rng, shape, bits):
    return np.asfortranarray(rng.integers(-(2 ** bits) + 1, 2 ** bits, shape)
                             + 1j * rng.integers(-(2 ** bits) + 1, 2 ** bits, shape))


def _zop(X, t):
    return X if t == "N" else (X.T if t == "T" else X.conj().T)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "N"), ("T", "C"), ("N", "T")])
def test_oracle_zcrt_exact_on_integers(orc, ta, tb):
    """Gaussian integers whose row maxima reach 2^18 are kept exactly
    (beta(2K, L) >= 19), and the three-product Karatsuba form is exact in
    residues, so the result is the exact complex product."""
    rng = np.random.default_rng(35)
    M, N, K = 23, 19, 41
    A = _cplx_int(rng, (M, K) if ta == "N" else (K, M), 18)
    B = _cplx_int(rng, (K, N) if tb == "N" else (N, K), 18)
    C0 = _cplx_int(rng, (M, N), 18)
    exact = (2 - 1j) * (_zop(A, ta) @ _zop(B, tb)) + (1 + 1j) * C0
    oa, ob = _zop(A, ta), _zop(B, tb)
    emax = max(orc.norm_exp(float(np.sum(np.abs(oa) ** 2, axis=1).max())),
               orc.norm_exp(float(np.sum(np.abs(ob) ** 2, axis=0).max())))
    for L in (10, 16):
        assert orc.crt_gamma(L) >= emax  # both parts scale to integers
        out = C0.copy(order="F")
        orc.zgemm_crt(ta, tb, M, N, K, 2 - 1j, A, A.shape[0], B, B.shape[0], 1 + 1j, out, M, L)
        assert np.array_equal(out, exact), L


def test_oracle_zcrt_close_to_dd(orc):
    M, N, K = 20, 16, 200
    A = orc.fill_uniform(M, K, 6, 0, dtype=np.complex128)
    B = orc.fill_uniform(K, N, 6, 1, dtype=np.complex128)
    C0 = np.zeros((M, N), dtype=np.complex128, order="F")
    hi, _ = orc.gemm_dd("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, C0, M)
    scale = np.abs(A) @ np.abs(B)
    errs = []
    for L in (10, 14, 18):
        out = C0.copy(order="F")
        orc.zgemm_crt("N", "N", M, N, K, 1.0, A, M, B, K, 0.0, out, M, L)
        errs.append(np.max(np.abs(out - hi) / scale))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] <= 4 * EPS


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "T"), ("T", "C")])
def test_gpu_zcrt_bit_exact_vs_oracle(gpu, orc, ta, tb):
    bx = gpu
    M, N, K = 150, 131, 270
    A = orc.fill_uniform(M if ta == "N" else K, K if ta == "N" else M, 9, 0, dtype=np.complex128)
    B = orc.fill_uniform(K if tb == "N" else N, N if tb == "N" else K, 9, 1, dtype=np.complex128)
    C0 = orc.fill_uniform(M, N, 9, 2, dtype=np.complex128)
    ref = C0.copy(order="F")
    orc.zgemm_crt(ta, tb, M, N, K, 0.5 + 1j, A, A.shape[0], B, B.shape[0], -1 + 0.25j, ref, M, 15)
    dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (A, B, C0))
    say("zcrt", ta, tb, M, N, K)
    bx.zgemm_crt_device(ta, tb, M, N, K, 0.5 + 1j, dA.ptr, A.shape[0], dB.ptr, B.shape[0],
                        -1 + 0.25j, dC.ptr, M, 15)
    bx.synchronize()
    assert np.array_equal(dC.download(), ref)


@pytest.mark.gpu
def test_gpu_zcrt_big_tile_path_sampled(gpu, orc):
    bx = gpu
    M, N, K = 4097, 4100, 300
    A = orc.fill_uniform(M, K, 10, 0, dtype=np.complex128)
    B = orc.fill_uniform(N, K, 10, 1, dtype=np.complex128)  # op(B) = B^H
    dA, dB = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(B)
    dC = bx.DeviceArray(M, N, np.complex128)
    say("zcrt big tiles N C", M, N, K)
    bx.zgemm_crt_device("N", "C", M, N, K, 1.0, dA.ptr, M, dB.ptr, N, 0.0, dC.ptr, M, 0)
    bx.synchronize()
    out = dC.download()
    L = next(L for L in range(2, 21) if orc.crt_beta(2 * K, L) >= 56)
    for (r0, c0) in [(0, 0), (M - 40, N - 36), (2048, 1111)]:
        As = np.asfortranarray(A[r0:r0 + 40])
        Bs = np.asfortranarray(B[c0:c0 + 36])
        ref = np.zeros((40, 36), dtype=np.complex128, order="F")
        orc.zgemm_crt("N", "C", 40, 36, K, 1.0, As, 40, Bs, 36, 0.0, ref, 40, L)
        assert np.array_equal(out[r0:r0 + 40, c0:c0 + 36], ref), (r0, c0)
