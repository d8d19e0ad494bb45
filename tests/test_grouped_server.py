"""Grouped GEMM and the GEMM request server (continuous batching).

Every problem is checked against the oracle within the stated tolerance, and
against the same problem run alone: a tile's arithmetic does not depend on
what else shares the launch, so "alone" and "in the mix" must agree bit for
bit.
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


def _problem(orc, rng, i):
    dt = [np.float64, np.float32, np.complex128][i % 3]
    cplx = np.dtype(dt).kind == "c"
    ta = rng.choice(["N", "T", "C"] if cplx else ["N", "T"])
    tb = rng.choice(["N", "T", "C"] if cplx else ["N", "T"])
    M, N, K = (int(v) for v in np.exp(rng.uniform(0, np.log(700), 3)).astype(int) + 1)
    if i % 11 == 5:
        K = 0
    alpha = complex(rng.uniform(-2, 2), rng.uniform(-1, 1)) if cplx else float(rng.uniform(-2, 2))
    beta = complex(rng.uniform(-1, 1), rng.uniform(-1, 1)) if cplx else float(rng.uniform(-1, 1))
    if i % 7 == 3:
        beta = 0.0
    if i % 13 == 4:
        alpha = 0.0
    A = orc.fill_uniform(M if ta == "N" else K, K if ta == "N" else M, 40 + i, 0, dtype=dt)
    B = orc.fill_uniform(K if tb == "N" else N, N if tb == "N" else K, 40 + i, 1, dtype=dt)
    C0 = orc.fill_uniform(M, N, 40 + i, 2, dtype=dt)
    return dict(dt=dt, ta=ta, tb=tb, M=M, N=N, K=K, alpha=alpha, beta=beta, A=A, B=B, C0=C0)


def ld(x):
    """Leading dimension of a column-major array or column-major view."""
    if x.ndim == 2 and x.shape[1] > 1:
        return max(x.strides[1] // x.itemsize, 1)
    return max(x.shape[0], 1)


def _ref(orc, p):
    ref = p["C0"].copy(order="F")
    orc.gemm(p["ta"], p["tb"], p["M"], p["N"], p["K"], p["alpha"], p["A"], ld(p["A"]),
             p["B"], ld(p["B"]), p["beta"], ref, p["M"], acc64=(p["dt"] == np.float32))
    return ref


def _check(orc, out, p):
    ref = _ref(orc, p)
    err = orc.relerr_fro(out, ref, p["alpha"], p["A"], p["B"], p["beta"], p["C0"])
    tol = (2 if np.dtype(p["dt"]).kind == "c" else 1) * max(p["K"], 1) * EPS[p["dt"]]
    assert err <= tol, (err, tol, {k: p[k] for k in ("dt", "ta", "tb", "M", "N", "K")})


def test_grouped_uneven_mixed(gpu, orc):
    bx = gpu
    rng = np.random.default_rng(50)
    probs = [_problem(orc, rng, i) for i in range(30)]
    dev = []
    for p in probs:
        dA, dB, dC = (bx.DeviceArray.from_host(x) for x in (p["A"], p["B"], p["C0"]))
        dev.append((dA, dB, dC))
    descs = [bx.make_desc(p["dt"], p["ta"], p["tb"], p["M"], p["N"], p["K"], p["alpha"],
                          d[0].ptr, ld(p["A"]), d[1].ptr, ld(p["B"]), p["beta"],
                          d[2].ptr, p["M"]) for p, d in zip(probs, dev)]
    say("grouped launch of", len(descs), "problems")
    bx.reset_stats()
    bx.gemm_grouped_device(descs)
    bx.synchronize()
    assert bx.stats()["kernel_launches"] <= 3  # one launch per element type
    for p, d in zip(probs, dev):
        out = d[2].download()
        _check(orc, out, p)
        # the same problem alone through the plain device GEMM
        solo = bx.DeviceArray.from_host(p["C0"])
        bx.device_gemm(p["ta"], p["tb"], p["M"], p["N"], p["K"], p["alpha"], d[0].ptr,
                       ld(p["A"]), d[1].ptr, ld(p["B"]), p["beta"], solo.ptr, p["M"],
                       p["dt"])
        bx.synchronize()
        assert np.array_equal(solo.download(), out)


def _fixed(orc, dt, ta, tb, M, N, K, seed, alpha=1.25, beta=-0.5):
    A = orc.fill_uniform(M if ta == "N" else K, K if ta == "N" else M, seed, 0, dtype=dt)
    B = orc.fill_uniform(K if tb == "N" else N, N if tb == "N" else K, seed, 1, dtype=dt)
    C0 = orc.fill_uniform(M, N, seed, 2, dtype=dt)
    return dict(dt=dt, ta=ta, tb=tb, M=M, N=N, K=K, alpha=alpha, beta=beta, A=A, B=B, C0=C0)


@pytest.mark.parametrize("shape", [0, 1, 2])
def test_grouped_block_shapes_equal_plain(gpu, orc, shape):
    """Each block tile shape a grouped launch may pick for a real problem
    (0: 128x128, 1: 128x64, 2: 64x64; forced here for S and D), on a mix of
    ragged and aligned problems of every element type: results within the
    oracle's tolerance, and bit-identical to the plain device GEMM of the same
    problem, since every kernel and block shape sums a C element's products
    in one K order."""
    bx = gpu
    L = bx.lib()
    rng = np.random.default_rng(60 + shape)
    probs = [_problem(orc, rng, i) for i in range(24)]
    probs += [_fixed(orc, np.float64, "N", "N", 256, 384, 64, 90),
              _fixed(orc, np.float64, "T", "N", 640, 130, 77, 91),
              _fixed(orc, np.float32, "N", "T", 256, 256, 96, 92),
              _fixed(orc, np.float32, "T", "T", 65, 200, 300, 93),
              _fixed(orc, np.float64, "N", "T", 64, 64, 16, 94, beta=0.0)]
    dev = [tuple(bx.DeviceArray.from_host(x) for x in (p["A"], p["B"], p["C0"])) for p in probs]
    descs = [bx.make_desc(p["dt"], p["ta"], p["tb"], p["M"], p["N"], p["K"], p["alpha"],
                          d[0].ptr, ld(p["A"]), d[1].ptr, ld(p["B"]), p["beta"],
                          d[2].ptr, p["M"]) for p, d in zip(probs, dev)]
    say("grouped launch, block shape", shape, len(descs), "problems")
    try:
        for dt in (bx.S, bx.D):
            assert L.blasx_set_grouped_shape(dt, shape) == 0
        bx.gemm_grouped_device(descs)
        bx.synchronize()
    finally:
        for dt in (bx.S, bx.D):
            L.blasx_set_grouped_shape(dt, -1)
    for p, d in zip(probs, dev):
        out = d[2].download()
        _check(orc, out, p)
        solo = bx.DeviceArray.from_host(p["C0"])
        bx.device_gemm(p["ta"], p["tb"], p["M"], p["N"], p["K"], p["alpha"], d[0].ptr,
                       ld(p["A"]), d[1].ptr, ld(p["B"]), p["beta"], solo.ptr, p["M"], p["dt"])
        bx.synchronize()
        assert np.array_equal(solo.download(), out), (shape, p["dt"], p["ta"], p["tb"],
                                                      p["M"], p["N"], p["K"])
        solo.free()
    for d in dev:
        for x in d:
            x.free()


def test_grouped_mixed_block_shapes_in_one_launch(gpu):
    """The measured table picks different block tile shapes for different
    problems of one launch (skinny 4096x64x4096 problems and 2048^3 cubes
    of one element type share the launch's load): every problem's result is
    bit-identical to the plain device GEMM of the same problem."""
    bx = gpu
    L = bx.lib()
    specs = [(4096, 64, 4096)] * 12 + [(2048, 2048, 2048)] * 2 + [(256, 256, 4096)] * 4
    load = sum(-(-M // 128) * -(-N // 128) for M, N, _ in specs)
    shapes = {L.blasx_grouped_shape(bx.D, M, N, K, load) for M, N, K in specs}
    say("grouped mixed shapes", len(specs), "problems, load", load, "shapes", sorted(shapes))
    assert len(shapes) >= 2, shapes
    arrs = []
    for i, (M, N, K) in enumerate(specs):
        dA, dB, dC = (bx.DeviceArray(r, c, np.float64) for r, c in ((M, K), (K, N), (M, N)))
        for j, d in enumerate((dA, dB, dC)):
            d.fill_uniform(70 + i, j)
        arrs.append((dA, dB, dC))
    bx.synchronize()
    C0 = [d[2].download() for d in arrs]
    # odd problems use B transposed: the K x N buffer read as N x K (ld = N)
    ops = [("N", "T" if i % 2 else "N", N if i % 2 else K) for i, (M, N, K) in enumerate(specs)]
    descs = [bx.make_desc(np.float64, ta, tb, M, N, K, 1.25, d[0].ptr, M, d[1].ptr, ldb, -0.5,
                          d[2].ptr, M)
             for (M, N, K), d, (ta, tb, ldb) in zip(specs, arrs, ops)]
    bx.gemm_grouped_device(descs)
    bx.synchronize()
    for i, ((M, N, K), d, c0, (ta, tb, ldb)) in enumerate(zip(specs, arrs, C0, ops)):
        out = d[2].download()
        solo = bx.DeviceArray.from_host(c0)
        bx.device_gemm(ta, tb, M, N, K, 1.25, d[0].ptr, M, d[1].ptr, ldb, -0.5, solo.ptr, M,
                       np.float64)
        bx.synchronize()
        assert np.array_equal(solo.download(), out), (i, M, N, K)
        assert np.isfinite(out).all()
        solo.free()
    for d in arrs:
        for x in d:
            x.free()


def _run_server(bx, srv, probs, resident_B=None, one_at_a_time=False):
    outs, ids = [], []
    for p in probs:
        Cw = p["C0"].copy(order="F")
        B = p["B"]
        d = bx.make_desc(p["dt"], p["ta"], p["tb"], p["M"], p["N"], p["K"], p["alpha"], p["A"],
                         ld(p["A"]), B, ld(B), p["beta"], Cw, p["M"])
        rid = srv.submit(d, keep=(p["A"], B, Cw))
        if one_at_a_time:
            srv.wait(rid)
        outs.append(Cw)
        ids.append(rid)
    lat = [srv.wait(r) for r in ids]
    return outs, lat


def test_server_alone_equals_in_mix(gpu, orc):
    bx = gpu
    rng = np.random.default_rng(51)
    probs = [_problem(orc, rng, i) for i in range(40)]
    # a shared resident B: 4096 x 600 FP64, a few requests point into it
    shared = orc.fill_uniform(900, 600, 77, 1)
    for i in (2, 9, 17, 26):
        p = probs[i]
        p.update(dt=np.float64, tb="N", K=min(p["K"] or 1, 900), N=min(p["N"], 600))
        p["A"] = orc.fill_uniform(p["M"] if p["ta"] == "N" else p["K"],
                                  p["K"] if p["ta"] == "N" else p["M"], 60 + i, 0)
        p["B"] = shared[:p["K"], :p["N"]]  # a view with ld = 900 inside the resident operand
        p["C0"] = orc.fill_uniform(p["M"], p["N"], 60 + i, 2)
        p["alpha"], p["beta"] = 1.5, 0.5
    srv = bx.Server(4 << 30, max_batch=64)
    try:
        srv.register_operand(shared)
        say("server: one at a time")
        alone, _ = _run_server(bx, srv, probs, one_at_a_time=True)
        say("server: all in flight together")
        mixed, lat = _run_server(bx, srv, probs)
        for p, a, m in zip(probs, alone, mixed):
            assert np.array_equal(a, m)
            if p["B"].flags["F_CONTIGUOUS"]:
                _check(orc, m, p)
            else:
                q = dict(p)
                q["B"] = np.asfortranarray(p["B"])
                _check(orc, m, q)
        c = srv.counters()
        assert c["done"] == 2 * len(probs)
        assert c["arena_used"] == 600 * 900 * 8 + (-(600 * 900 * 8)) % 256  # only the resident B remains
        assert all(x >= 0 for x in lat)
        # continuous batching: small requests that arrive while a large one
        # computes are admitted together into few batches, not one each
        n = 4096
        big = dict(dt=np.float64, ta="N", tb="N", M=n, N=n, K=n, alpha=1.0, beta=0.0,
                   A=orc.fill_uniform(n, n, 3, 0), B=orc.fill_uniform(n, n, 3, 1),
                   C0=np.zeros((n, n), order="F"))
        smalls = [_problem(orc, rng, 100 + i) for i in range(24)]
        before = srv.counters()["batches"]
        say("server: burst behind a large request")
        outs, _ = _run_server(bx, srv, [big] + smalls)
        used = srv.counters()["batches"] - before
        assert used < len(smalls), used
        for p, o in zip(smalls, outs[1:]):
            _check(orc, o, p)
    finally:
        srv.close()


def test_server_tiny_arena_admission(gpu, orc):
    """Arena smaller than the whole workload: requests wait for memory in
    arrival order and all complete; one that can never fit is rejected."""
    bx = gpu
    rng = np.random.default_rng(52)
    n = 512
    srv = bx.Server(3 * n * n * 8 + 4096, max_batch=8)  # exactly one request at a time
    try:
        probs = []
        for i in range(6):
            p = dict(dt=np.float64, ta="N", tb="N", M=n, N=n, K=n, alpha=1.0, beta=1.0)
            p["A"] = orc.fill_uniform(n, n, 90 + i, 0)
            p["B"] = orc.fill_uniform(n, n, 90 + i, 1)
            p["C0"] = orc.fill_uniform(n, n, 90 + i, 2)
            probs.append(p)
        say("server: tiny arena")
        outs, _ = _run_server(bx, srv, probs)
        for p, o in zip(probs, outs):
            _check(orc, o, p)
        big = orc.fill_uniform(2 * n, 2 * n, 1, 1)
        Cb = np.zeros((2 * n, 2 * n), order="F")
        d = bx.make_desc(np.float64, "N", "N", 2 * n, 2 * n, 2 * n, 1.0, big, 2 * n, big, 2 * n,
                         0.0, Cb, 2 * n)
        with pytest.raises(RuntimeError):
            srv.submit(d)
    finally:
        srv.close()
