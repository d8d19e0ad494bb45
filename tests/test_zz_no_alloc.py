"""The allocation-free grouped GEMM entry point
(blasx_gemm_grouped_device_noalloc) through the C ABI.

It shares its kernels with blasx_gemm_grouped_device and keeps its problem
tables in a static per-GPU store, so on the same inputs its results must be
bit-identical to that entry point's (and within c*K*eps of the oracle).  Two
calls in a row, the second with fewer and smaller problems, check that the
store's leftover tables from the first call never leak into the second.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps,
       np.complex128: np.finfo(np.float64).eps}


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _problems(orc, seed, count, max_dim):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        dt = [np.float64, np.float32, np.complex128][i % 3]
        cplx = np.dtype(dt).kind == "c"
        ta = str(rng.choice(["N", "T", "C"] if cplx else ["N", "T"]))
        tb = str(rng.choice(["N", "T", "C"] if cplx else ["N", "T"]))
        M, N, K = (int(v) for v in rng.integers(1, max_dim, 3))
        if i == 4:
            K = 0          # C = beta C
        alpha = complex(1.25, -0.5) if cplx else 1.25
        beta = 0.0 if i % 5 == 2 else (complex(0.5, 0.25) if cplx else -0.5)
        A = orc.fill_uniform(M if ta == "N" else K, K if ta == "N" else M, seed + i, 0, dtype=dt)
        B = orc.fill_uniform(K if tb == "N" else N, N if tb == "N" else K, seed + i, 1, dtype=dt)
        C0 = orc.fill_uniform(M, N, seed + i, 2, dtype=dt)
        if beta == 0.0:
            C0[:] = np.nan   # must not be read
        out.append(dict(dt=dt, ta=ta, tb=tb, M=M, N=N, K=K, alpha=alpha, beta=beta, A=A, B=B,
                        C0=C0))
    return out


def _run(bx, probs, noalloc):
    dev = []
    for p in probs:
        dev.append(tuple(bx.DeviceArray.from_host(x) for x in (p["A"], p["B"], p["C0"])))
    descs = [bx.make_desc(p["dt"], p["ta"], p["tb"], p["M"], p["N"], p["K"], p["alpha"],
                          d[0].ptr, max(p["A"].shape[0], 1), d[1].ptr, max(p["B"].shape[0], 1),
                          p["beta"], d[2].ptr, p["M"]) for p, d in zip(probs, dev)]
    bx.gemm_grouped_device(descs, noalloc=noalloc)
    bx.synchronize()
    outs = [d[2].download() for d in dev]
    for d in dev:
        for x in d:
            x.free()
    return outs


def _check_oracle(orc, out, p):
    ref = np.nan_to_num(p["C0"]).copy(order="F") if p["beta"] == 0.0 else p["C0"].copy(order="F")
    orc.gemm(p["ta"], p["tb"], p["M"], p["N"], p["K"], p["alpha"], p["A"],
             max(p["A"].shape[0], 1), p["B"], max(p["B"].shape[0], 1), p["beta"], ref, p["M"],
             acc64=(p["dt"] == np.float32))
    c0 = np.nan_to_num(p["C0"])
    err = orc.relerr_fro(out, ref, p["alpha"], p["A"], p["B"], p["beta"], c0)
    tol = (2 if np.dtype(p["dt"]).kind == "c" else 1) * max(p["K"], 1) * EPS[p["dt"]]
    assert np.isfinite(out).all() and err <= tol, (err, tol)


def test_noalloc_grouped_equals_grouped_twice(gpu, orc):
    bx = gpu
    for call, (seed, count, max_dim) in enumerate(((300, 24, 700), (400, 7, 200))):
        say("noalloc call", call, "problems", count, "max dim", max_dim)
        probs = _problems(orc, seed, count, max_dim)
        ref = _run(bx, probs, noalloc=False)
        got = _run(bx, probs, noalloc=True)
        for p, r, g in zip(probs, ref, got):
            assert np.array_equal(r, g, equal_nan=True), {k: p[k] for k in ("dt", "ta", "tb",
                                                                             "M", "N", "K")}
            _check_oracle(orc, g, p)


def test_noalloc_grouped_rejects_more_than_the_table(gpu):
    bx = gpu
    L = bx.lib()
    n = 1025
    arr = (bx.GemmDesc * n)()
    for d in arr:
        d.dtype, d.transa, d.transb, d.M, d.N, d.K = 1, 111, 111, 1, 1, 1
        d.lda = d.ldb = d.ldc = 1
    say("noalloc over the limit")
    assert L.blasx_gemm_grouped_device_noalloc(arr, n, None) == -7   # BLASX_ERANGE
    assert L.blasx_gemm_grouped_device_noalloc(arr, -1, None) == -2  # BLASX_EINVAL
