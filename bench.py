#!/usr/bin/env python3
"""bench.py -- headline benchmark of the GEMM hot path on one MI355X.

Workload (the in-core tiled GEMM config): DGEMM, M = N = K = 32768, NN,
alpha = 1, beta = 1, entries uniform[-1,1) from the seeded SplitMix64
generator (seed 0, matrix ids 0/1/2), column-major.  A "step" is one full
C = A*B + C through libblasx's FP64-MFMA kernel with the operands already
resident in HBM.  value = 2*M*N*K*steps / (timed wall time) in TFLOP/s.

Also reported (not `value`):
  * roofline of the dominant kernel from HIP events on its stream,
  * the CPU baseline: OpenBLAS cblas_dgemm -- the reference's CPU BLAS path --
    timed on this host on a bounded sample,
  * extras: the same DGEMM through the cblas_dgemm ABI with host-resident
    pinned operands (cold and resident-warm), bytes moved vs the lower bound;
    split-precision (INT8 Ozaki) vs plain FP64 DGEMM at M=N=K=16384 on inputs
    (u-0.5)*exp(phi*g), phi in {0.5, 2, 4}, with errors against a
    double-double reference on sampled 256x256 blocks.

Progress goes to stderr; the single JSON result line goes to stdout.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 matrix, AMD published
FP32_PEAK_TFLOPS = 157.3
INT8_PEAK_TOPS = 5000.0
METRIC = "FP64-accurate GEMM TFLOP/s, plain vs split-MFMA; GEMM-stream req/s & p99 ms"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def load_traffic(key):
    """HBM bytes per launch measured offline with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), scaled by a copy-kernel calibration; see profiles/."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        e = t[key]
        return float(e["fetch_bytes"]) + float(e["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def headline(bx, n, steps, warmup):
    import numpy as np

    A = bx.DeviceArray(n, n, np.float64)
    B = bx.DeviceArray(n, n, np.float64)
    Cm = bx.DeviceArray(n, n, np.float64)
    A.fill_uniform(0, 0)
    B.fill_uniform(0, 1)
    Cm.fill_uniform(0, 2)
    bx.synchronize()
    log(f"headline: DGEMM {n}^3 NN, {warmup} warmup + {steps} timed steps")
    for _ in range(warmup):
        bx.device_gemm("N", "N", n, n, n, 1.0, A.ptr, n, B.ptr, n, 1.0, Cm.ptr, n, np.float64)
    bx.synchronize()
    evs = [bx.Event() for _ in range(steps + 1)]
    bx.synchronize()
    t0 = time.perf_counter()
    evs[0].record()
    for s in range(steps):
        bx.device_gemm("N", "N", n, n, n, 1.0, A.ptr, n, B.ptr, n, 1.0, Cm.ptr, n, np.float64)
        evs[s + 1].record()
    bx.synchronize()
    wall = time.perf_counter() - t0
    per_launch = [evs[i].elapsed_ms(evs[i + 1]) for i in range(steps)]
    flops = 2.0 * n * n * n
    avg_ms = sum(per_launch) / steps
    log(f"headline: wall {wall:.3f}s, kernel avg {avg_ms:.2f} ms "
        f"({flops / avg_ms / 1e9:.2f} TFLOP/s)")
    # spot check: one element against a dd-free FP64 dot product on the host
    for X in (A, B, Cm):
        X.free()
    return wall, avg_ms, flops


def cpu_baseline():
    """OpenBLAS cblas_dgemm (the reference's CPU path, loaded the way the
    reference loads it) on a bounded sample of the workload."""
    import numpy as np

    import oracle

    path = oracle.openblas_load()
    threads = os.cpu_count() or 1
    env_thr = os.environ.get("OMP_NUM_THREADS")
    if env_thr and env_thr.isdigit():
        threads = min(threads, int(env_thr))
    oracle.openblas_set_threads(threads)
    n = 8192
    A = oracle.fill_uniform(n, n, 0, 0)
    B = oracle.fill_uniform(n, n, 0, 1)
    C0 = oracle.fill_uniform(n, n, 0, 2)
    log(f"cpu baseline: OpenBLAS DGEMM {n}^3 on {threads} threads ({path})")
    Cw = C0.copy(order="F")
    oracle.openblas_gemm("N", "N", n, n, n, 1.0, A, n, B, n, 1.0, Cw, n)  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_start < 10.0 and len(times) < 30):
        Cw = C0.copy(order="F")
        t = time.perf_counter()
        oracle.openblas_gemm("N", "N", n, n, n, 1.0, A, n, B, n, 1.0, Cw, n)
        times.append(time.perf_counter() - t)
    med = statistics.median(times)
    tfl = 2.0 * n ** 3 / med / 1e12
    info = oracle.openblas_info()
    log(f"cpu baseline: median {med:.3f}s over {len(times)} calls -> {tfl:.3f} TFLOP/s")
    return {"value": round(tfl, 4), "unit": "TFLOP/s", "cores": threads, "kind": "reference",
            "sample": (f"OpenBLAS cblas_dgemm (the reference's CPU BLAS path, dlopen+dlsym as in "
                       f"blas/blasx_internal.c:26-36) DGEMM M=N=K={n} NN alpha=1 beta=1, same "
                       f"seeded generator; median of {len(times)} calls after 1 warm-up; "
                       f"{os.path.basename(info['path'])} {info['config'][:40]}"),
            "cfg0": cfg0_cpu_and_parity(oracle, threads)}


def cfg0_cpu_and_parity(oracle, threads):
    """The reference's CPU-runnable case: DGEMM M=N=K=1024 NN, alpha=1.5,
    beta=0.5, uniform[-1,1) seeds 0 (ids 0, 1, 2), ld=1024.  OpenBLAS
    GFLOP/s (median of >= 3 calls after a warm-up), and the same inputs
    through this library's cblas_dgemm (1024 > 1000 routes to the GPU, as in
    the reference), both checked against the oracle's naive triple loop:
    relative Frobenius error ||C - C_naive|| / (|alpha| ||A|| ||B|| + |beta|
    ||C0||)."""
    import numpy as np

    import blasx_amd as bx

    n, alpha, beta = 1024, 1.5, 0.5
    A = oracle.fill_uniform(n, n, 0, 0)
    B = oracle.fill_uniform(n, n, 0, 1)
    C0 = oracle.fill_uniform(n, n, 0, 2)
    ref = C0.copy(order="F")
    oracle.gemm("N", "N", n, n, n, alpha, A, n, B, n, beta, ref, n)
    scale = abs(alpha) * np.linalg.norm(A) * np.linalg.norm(B) + abs(beta) * np.linalg.norm(C0)
    times, Cc = [], None
    for r in range(6):
        Cc = C0.copy(order="F")
        t = time.perf_counter()
        oracle.openblas_gemm("N", "N", n, n, n, alpha, A, n, B, n, beta, Cc, n)
        if r:
            times.append(time.perf_counter() - t)
    cpu_err = float(np.linalg.norm(Cc - ref) / scale)
    gt = []
    for r in range(4):
        Cg = C0.copy(order="F")
        t = time.perf_counter()
        bx.gemm("N", "N", alpha, A, B, beta, Cg)
        if r:
            gt.append(time.perf_counter() - t)
    gpu_err = float(np.linalg.norm(Cg - ref) / scale)
    flops = 2.0 * n ** 3
    res = {"n": n, "cpu_gflops": round(flops / statistics.median(times) / 1e9, 1),
           "cpu_threads": threads, "cpu_rel_fro_err": cpu_err,
           "abi_gflops": round(flops / statistics.median(gt) / 1e9, 1),
           "abi_path": "GPU (1024 > 1000, host operands, pageable)",
           "abi_rel_fro_err": gpu_err, "tolerance": n * 2.0 ** -53}
    log(f"cfg0: {res}")
    return res


def host_resident(bx, n):
    """The same DGEMM through cblas_dgemm with host-resident pinned operands:
    cold (empty tile cache) and warm (A and B registered resident)."""
    import numpy as np

    L = bx.lib()
    log(f"host-resident: allocating 3 x {n}^2 FP64 pinned host matrices")
    mats = [bx.PinnedArray(n, n, np.float64) for _ in range(3)]
    for i, m in enumerate(mats):
        bx.fill_uniform_host(m.array, 0, i)
    A, B, Cm = (m.array for m in mats)
    L.blasx_cache_flush()
    bx.reset_stats()
    flops = 2.0 * n ** 3
    t = time.perf_counter()
    bx.gemm("N", "N", 1.0, A, B, 1.0, Cm)
    cold = time.perf_counter() - t
    st = bx.stats()
    log(f"host-resident cold: {cold:.3f}s ({flops / cold / 1e12:.2f} TFLOP/s), "
        f"H2D {st['bytes_h2d'] / 1e9:.1f} GB, D2H {st['bytes_d2h'] / 1e9:.1f} GB")
    L.blasx_resident_register(A.ctypes.data, A.nbytes)
    L.blasx_resident_register(B.ctypes.data, B.nbytes)
    bx.gemm("N", "N", 1.0, A, B, 1.0, Cm)  # loads A and B once more, now kept
    bx.reset_stats()
    t = time.perf_counter()
    bx.gemm("N", "N", 1.0, A, B, 1.0, Cm)
    warm = time.perf_counter() - t
    st2 = bx.stats()
    log(f"host-resident warm: {warm:.3f}s ({flops / warm / 1e12:.2f} TFLOP/s), "
        f"H2D {st2['bytes_h2d'] / 1e9:.1f} GB")
    L.blasx_resident_unregister(A.ctypes.data)
    L.blasx_resident_unregister(B.ctypes.data)
    L.blasx_cache_flush()
    for m in mats:
        m.free()
    # ordinary (pageable) host memory, as most BLAS callers pass it
    log("host-resident: pageable operands")
    Ap = bx.fill_uniform_host(np.empty((n, n), order="F"), 0, 0)
    Bp = bx.fill_uniform_host(np.empty((n, n), order="F"), 0, 1)
    Cp = bx.fill_uniform_host(np.empty((n, n), order="F"), 0, 2)
    t = time.perf_counter()
    bx.gemm("N", "N", 1.0, Ap, Bp, 1.0, Cp)
    pageable = time.perf_counter() - t
    log(f"host-resident pageable: {pageable:.3f}s ({flops / pageable / 1e12:.2f} TFLOP/s)")
    del Ap, Bp, Cp
    return {
        "cold_tflops": round(flops / cold / 1e12, 3),
        "warm_tflops": round(flops / warm / 1e12, 3),
        "pageable_tflops": round(flops / pageable / 1e12, 3),
        "cold_h2d_GBps": round(st["bytes_h2d"] / cold / 1e9, 2),
        "cold_bytes_over_lower_bound": round((st["bytes_h2d"] + st["bytes_d2h"]) /
                                             max(st["bytes_lower_bound"], 1), 3),
        "warm_h2d_bytes": st2["bytes_h2d"],
    }


def _syrk_call(L, name, uplo, trans, n, k, alpha, A, lda, beta, Cm, ldc):
    import ctypes as C

    fn = getattr(L, "cblas_" + name)
    fn.restype = None
    if name == "zsyrk":
        al = (C.c_double * 2)(alpha, 0.0)
        be = (C.c_double * 2)(beta, 0.0)
        a_arg, b_arg = C.cast(al, C.c_void_p), C.cast(be, C.c_void_p)
    else:
        a_arg, b_arg = C.c_double(alpha), C.c_double(beta)
    fn(102, uplo, trans, n, k, a_arg, C.c_void_p(A), lda, b_arg, C.c_void_p(Cm), ldc)


def rank_k(bx, n=32768):
    """Rank-k updates through the ABI: DSYRK (lower, NoTrans, alpha = beta = 1,
    N = K = n) with operands in HBM (one grouped launch) and host-resident
    (tile runtime over the triangle's C tiles, cold and warm), ZHERK in HBM,
    and OpenBLAS DSYRK on the host for comparison.  FLOPs: N(N+1)K real,
    4N(N+1)K complex."""
    import ctypes as C

    import numpy as np

    L = bx.lib()
    out = {}
    flops = float(n) * (n + 1) * n
    dA = bx.DeviceArray(n, n, np.float64)
    dC = bx.DeviceArray(n, n, np.float64)
    dA.fill_uniform(0, 0)
    dC.fill_uniform(0, 2)
    bx.synchronize()
    _syrk_call(L, "dsyrk", 122, 111, n, n, 1.0, dA.ptr, n, 1.0, dC.ptr, n)  # warm-up
    times = []
    for _ in range(3):
        t = time.perf_counter()
        _syrk_call(L, "dsyrk", 122, 111, n, n, 1.0, dA.ptr, n, 1.0, dC.ptr, n)
        times.append(time.perf_counter() - t)
    best = min(times)
    log(f"dsyrk in HBM N=K={n}: {best:.3f}s ({flops / best / 1e12:.2f} TFLOP/s, "
        f"{flops / best / 1e12 / FP64_PEAK_TFLOPS:.3f} of FP64 peak)")
    out["dsyrk_inhbm_tflops"] = round(flops / best / 1e12, 3)
    dA.free()
    dC.free()

    nz, kz = n // 2, n
    zfl = 4.0 * nz * (nz + 1) * kz
    dA = bx.DeviceArray(nz, kz, np.complex128)
    dC = bx.DeviceArray(nz, nz, np.complex128)
    dA.fill_uniform(0, 0)
    dC.fill_uniform(0, 2)
    bx.synchronize()
    _syrk_call(L, "zherk", 122, 111, nz, kz, 1.0, dA.ptr, nz, 1.0, dC.ptr, nz)
    times = []
    for _ in range(3):
        t = time.perf_counter()
        _syrk_call(L, "zherk", 122, 111, nz, kz, 1.0, dA.ptr, nz, 1.0, dC.ptr, nz)
        times.append(time.perf_counter() - t)
    best = min(times)
    log(f"zherk in HBM N={nz} K={kz}: {best:.3f}s ({zfl / best / 1e12:.2f} TFLOP/s)")
    out["zherk_inhbm_tflops"] = round(zfl / best / 1e12, 3)
    out["zherk_shape"] = [nz, kz]
    dA.free()
    dC.free()

    log(f"dsyrk host-resident: allocating 2 x {n}^2 FP64 pinned host matrices")
    mats = [bx.PinnedArray(n, n, np.float64) for _ in range(2)]
    bx.fill_uniform_host(mats[0].array, 0, 0)
    bx.fill_uniform_host(mats[1].array, 0, 2)
    A, Cm = mats[0].array, mats[1].array
    L.blasx_cache_flush()
    bx.reset_stats()
    t = time.perf_counter()
    _syrk_call(L, "dsyrk", 122, 111, n, n, 1.0, A.ctypes.data, n, 1.0, Cm.ctypes.data, n)
    cold = time.perf_counter() - t
    st = bx.stats()
    L.blasx_resident_register(A.ctypes.data, A.nbytes)
    _syrk_call(L, "dsyrk", 122, 111, n, n, 1.0, A.ctypes.data, n, 1.0, Cm.ctypes.data, n)
    t = time.perf_counter()
    _syrk_call(L, "dsyrk", 122, 111, n, n, 1.0, A.ctypes.data, n, 1.0, Cm.ctypes.data, n)
    warm = time.perf_counter() - t
    L.blasx_resident_unregister(A.ctypes.data)
    L.blasx_cache_flush()
    for m in mats:
        m.free()
    log(f"dsyrk host-resident: cold {cold:.3f}s ({flops / cold / 1e12:.2f} TFLOP/s), "
        f"warm {warm:.3f}s ({flops / warm / 1e12:.2f} TFLOP/s), cold H2D "
        f"{st['bytes_h2d'] / 1e9:.1f} GB D2H {st['bytes_d2h'] / 1e9:.1f} GB")
    out["dsyrk_host_cold_tflops"] = round(flops / cold / 1e12, 3)
    out["dsyrk_host_warm_tflops"] = round(flops / warm / 1e12, 3)
    out["dsyrk_host_cold_bytes_over_lower_bound"] = round(
        (st["bytes_h2d"] + st["bytes_d2h"]) / max(st["bytes_lower_bound"], 1), 3)

    # OpenBLAS DSYRK on the host cores, a bounded sample (the reference's path)
    L.blasx_cpublas_path.restype = C.c_char_p
    path = L.blasx_cpublas_path().decode()
    cpu = C.CDLL(path)
    threads = os.cpu_count() or 1
    env_thr = os.environ.get("OMP_NUM_THREADS")
    if env_thr and env_thr.isdigit():
        threads = min(threads, int(env_thr))
    ilp = not hasattr(cpu, "cblas_dsyrk")
    fn = cpu.scipy_cblas_dsyrk64_ if ilp else cpu.cblas_dsyrk
    it = C.c_int64 if ilp else C.c_int
    if ilp and hasattr(cpu, "scipy_openblas_set_num_threads64_"):
        cpu.scipy_openblas_set_num_threads64_(C.c_int(threads))
    ns = 8192
    As = bx.fill_uniform_host(np.empty((ns, ns), order="F"), 0, 0)
    Cs = bx.fill_uniform_host(np.empty((ns, ns), order="F"), 0, 2)
    args = (C.c_int(102), C.c_int(122), C.c_int(111), it(ns), it(ns), C.c_double(1.0),
            C.c_void_p(As.ctypes.data), it(ns), C.c_double(1.0), C.c_void_p(Cs.ctypes.data),
            it(ns))
    fn(*args)
    ct = []
    for _ in range(3):
        t = time.perf_counter()
        fn(*args)
        ct.append(time.perf_counter() - t)
    cflops = float(ns) * (ns + 1) * ns
    med = statistics.median(ct)
    log(f"OpenBLAS dsyrk N=K={ns} on {threads} threads: {med:.3f}s "
        f"({cflops / med / 1e12:.3f} TFLOP/s)")
    out["cpu_dsyrk"] = {"value": round(cflops / med / 1e12, 4), "unit": "TFLOP/s",
                        "cores": threads, "sample": f"OpenBLAS cblas_dsyrk N=K={ns} lower NoTrans, "
                        "median of 3 after 1 warm-up"}
    return out


def level3_inhbm(bx, n=32768):
    """DSYMM (Left, Lower), DTRMM (Left, Lower, NoTrans, NonUnit), DSYR2K
    (Lower, NoTrans), DTRSM (Left, Lower, NoTrans, Unit) and CGEMM (NN)
    through the BLAS ABI on operands in HBM, M = N (= K) = n.  FLOPs: SYMM
    2M^2N, TRMM and TRSM M^2N, SYR2K 2N(N+1)K, CGEMM 8MNK."""
    import ctypes as C

    import numpy as np

    L = bx.lib()
    out = {}
    mats = [bx.DeviceArray(n, n, np.float64) for _ in range(3)]
    for i, m in enumerate(mats):
        m.fill_uniform(1, i)
    bx.synchronize()
    dA, dB, dC = (m.ptr for m in mats)
    V, D = C.c_void_p, C.c_double
    calls = {
        "dsymm": (2.0 * n ** 3, lambda: L.cblas_dsymm(102, 141, 122, n, n, D(1.0), V(dA), n, V(dB),
                                                     n, D(1.0), V(dC), n)),
        "dtrmm": (float(n) ** 3, lambda: L.cblas_dtrmm(102, 141, 122, 111, 131, n, n, D(1.0),
                                                      V(dA), n, V(dB), n)),
        "dsyr2k": (2.0 * n * (n + 1) * n, lambda: L.cblas_dsyr2k(102, 122, 111, n, n, D(1.0),
                                                               V(dA), n, V(dB), n, D(1.0),
                                                               V(dC), n)),
        # the solve's values are not inspected (A is not conditioned for it); the
        # time does not depend on them
        "dtrsm": (float(n) ** 3, lambda: L.cblas_dtrsm(102, 141, 122, 111, 132, n, n, D(1.0),
                                                      V(dA), n, V(dB), n)),
    }
    for name, (flops, fn) in calls.items():
        fn()  # warm-up
        times = []
        for _ in range(2):
            t = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t)
        best = min(times)
        log(f"{name} in HBM n={n}: {best:.3f}s ({flops / best / 1e12:.2f} TFLOP/s)")
        out[name + "_tflops"] = round(flops / best / 1e12, 3)
    for m in mats:
        m.free()
    # CGEMM (complex64) on the FP32 matrix cores: 8 n^3 real FLOPs
    cm = [bx.DeviceArray(n, n, np.complex64) for _ in range(3)]
    for i, m in enumerate(cm):
        # an n x n complex64 matrix is a 2n x n float32 matrix: (re, im) pairs
        rc = L.blasx_fill_uniform_device(bx.S, m.ptr, 2 * n, n, 2 * n, 2, i, None)
        if rc:
            raise RuntimeError(f"blasx_fill_uniform_device failed ({rc})")
    bx.synchronize()
    al = (C.c_float * 2)(1.0, 0.0)
    fn = lambda: L.cblas_cgemm(102, 111, 111, n, n, n, al, V(cm[0].ptr), n, V(cm[1].ptr), n,
                               al, V(cm[2].ptr), n)
    fn()
    times = []
    for _ in range(2):
        t = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t)
    best = min(times)
    log(f"cgemm in HBM n={n}: {best:.3f}s ({8.0 * n ** 3 / best / 1e12:.2f} TFLOP/s)")
    out["cgemm_tflops"] = round(8.0 * n ** 3 / best / 1e12, 3)
    for m in cm:
        m.free()
    return out


def matrix_inhbm(bx, n=32768):
    """The kernel on operands resident in HBM for every element type and
    transpose the ABI serves: SGEMM and DGEMM NN/NT/TN/TT, ZGEMM NN/NT/TN/TT/
    CN/NC/CC, M=N=K=n, alpha=beta=1, seeded uniform operands generated on
    the device.  One warm-up per type, one timed call per variant (HIP
    events on the compute stream); TFLOP/s and the fraction of the
    published FP32 (157.3) / FP64 (78.6) matrix peak."""
    import numpy as np

    peaks = {np.float32: 157.3, np.float64: 78.6, np.complex128: 78.6}
    variants = {np.float32: ("NN", "NT", "TN", "TT"), np.float64: ("NN", "NT", "TN", "TT"),
                np.complex128: ("NN", "NT", "TN", "TT", "CN", "NC", "CC")}
    res = {}
    ev0, ev1 = bx.Event(), bx.Event()
    for dt in (np.float32, np.float64, np.complex128):
        name = {np.float32: "sgemm", np.float64: "dgemm", np.complex128: "zgemm"}[dt]
        X = [bx.DeviceArray(n, n, dt) for _ in range(3)]
        for i, x in enumerate(X):
            x.fill_uniform(0, i)
        bx.synchronize()
        flops = (8.0 if dt == np.complex128 else 2.0) * n ** 3

        def call(t):
            bx.device_gemm(t[0], t[1], n, n, n, 1.0, X[0].ptr, n, X[1].ptr, n, 1.0, X[2].ptr, n,
                           dt)

        call("NN")
        bx.synchronize()
        for t in variants[dt]:
            ev0.record()
            call(t)
            ev1.record()
            ev1.synchronize()
            ms = ev0.elapsed_ms(ev1)
            tf = flops / (ms * 1e-3) / 1e12
            res[f"{name}_{t}"] = {"tflops": round(tf, 2), "frac_of_peak": round(tf / peaks[dt], 4)}
            log(f"matrix: {name} {t} {n}^3: {ms:.1f} ms, {tf:.2f} TFLOP/s")
        for x in X:
            x.free()
    return res


def host_resident_sz(bx, n=32768):
    """SGEMM and ZGEMM NN through the BLAS ABI with host operands (the DGEMM
    case is host_resident): cold call (empty tile cache) and warm call (A
    and B registered resident), bytes moved vs the lower bound
    e (MK + KN + 2MN).  Pinned memory where the driver grants it, else
    pageable (reported)."""
    import numpy as np

    L = bx.lib()
    res = {}
    for dt in (np.float32, np.complex128):
        name = "sgemm" if dt == np.float32 else "zgemm"
        flops = (8.0 if dt == np.complex128 else 2.0) * n ** 3
        pinned = []
        try:
            pinned = [bx.PinnedArray(n, n, dt) for _ in range(3)]
            mats = [p.array for p in pinned]
            kind = "pinned"
        except MemoryError:
            for p in pinned:
                p.free()
            pinned = []
            mats = [np.empty((n, n), dtype=dt, order="F") for _ in range(3)]
            kind = "pageable"
        log(f"host-resident {name}: {kind} 3 x {n}^2")
        for i, m in enumerate(mats):
            bx.fill_uniform_host(m, 0, i)
        A, B, Cm = mats
        L.blasx_cache_flush()
        bx.reset_stats()
        t = time.perf_counter()
        bx.gemm("N", "N", 1.0, A, B, 1.0, Cm)
        cold = time.perf_counter() - t
        st = bx.stats()
        L.blasx_resident_register(A.ctypes.data, A.nbytes)
        L.blasx_resident_register(B.ctypes.data, B.nbytes)
        bx.gemm("N", "N", 1.0, A, B, 1.0, Cm)
        bx.reset_stats()
        t = time.perf_counter()
        bx.gemm("N", "N", 1.0, A, B, 1.0, Cm)
        warm = time.perf_counter() - t
        st2 = bx.stats()
        L.blasx_resident_unregister(A.ctypes.data)
        L.blasx_resident_unregister(B.ctypes.data)
        L.blasx_cache_flush()
        res[name] = {"memory": kind, "cold_tflops": round(flops / cold / 1e12, 3),
                     "warm_tflops": round(flops / warm / 1e12, 3),
                     "cold_h2d_GBps": round(st["bytes_h2d"] / cold / 1e9, 2),
                     "cold_bytes_over_lower_bound": round(
                         (st["bytes_h2d"] + st["bytes_d2h"]) / max(st["bytes_lower_bound"], 1), 3),
                     "warm_h2d_bytes": st2["bytes_h2d"]}
        log(f"host-resident {name}: {res[name]}")
        for p in pinned:
            p.free()
        del A, B, Cm, mats
    return res


def out_of_core(bx, n=65536, budget_frac=0.75, host_limit=160 << 30):
    """DGEMM through cblas_dgemm with pageable host operands whose working set
    exceeds the tile-cache budget (budget = budget_frac * operand bytes),
    against the same call with a budget that holds everything.  Reports
    throughput, the ratio, bytes moved vs the lower bound e(MK + KN + 2MN) and
    the host-link rates.

    When three n x n matrices exceed `host_limit` bytes of host memory (n =
    131072: 3 x 128 GiB), B is the same host matrix as A (C = A A + C, a
    legal BLAS call: both operands are read-only), so the run needs 2 x 128
    GiB; the tile cache then keys both operands' tiles on one matrix, and the
    budgets are fractions of that one matrix."""
    import numpy as np

    L = bx.lib()
    alias = 3 * n * n * 8 > host_limit
    # ordinary host memory: 3 x 32 GiB of pinned memory is more than the
    # driver grants one process here; the runtime stages pageable operands
    log(f"out-of-core: allocating {2 if alias else 3} x {n}^2 FP64 host matrices"
        + (" (B aliases A)" if alias else ""))
    A = bx.fill_uniform_host(np.empty((n, n), order="F"), 3, 0)
    B = A if alias else bx.fill_uniform_host(np.empty((n, n), order="F"), 3, 1)
    Cm = bx.fill_uniform_host(np.empty((n, n), order="F"), 3, 2)
    operand_bytes = A.nbytes * (1 if alias else 2)
    flops = 2.0 * n ** 3
    res = {"n": n, "b_aliases_a": alias, "operand_bytes": operand_bytes}
    old = L.blasx_get_hbm_budget()
    try:
        cases = (("in_core", old), ("out_of_core", int(budget_frac * operand_bytes)),
                 ("out_of_core_below_B", int(0.4 * operand_bytes)))
        for name, budget in cases:
            L.blasx_finalize()
            L.blasx_set_hbm_budget(budget)
            bx.reset_stats()
            t = time.perf_counter()
            bx.gemm("N", "N", 1.0, A, B, 1.0, Cm)
            dt = time.perf_counter() - t
            st = bx.stats()
            res[name] = {"budget_bytes": budget, "tflops": round(flops / dt / 1e12, 3),
                         "seconds": round(dt, 3),
                         "h2d_GBps": round(st["bytes_h2d"] / dt / 1e9, 2),
                         "d2h_GBps": round(st["bytes_d2h"] / dt / 1e9, 2),
                         "bytes_over_lower_bound": round((st["bytes_h2d"] + st["bytes_d2h"]) /
                                                         max(st["bytes_lower_bound"], 1), 3),
                         "tile_evictions": st["tile_evictions"]}
            log(f"out-of-core: {name}: {res[name]}")
    finally:
        L.blasx_finalize()
        L.blasx_set_hbm_budget(old)
        del A, B, Cm
    res["out_of_core_over_in_core"] = round(res["out_of_core"]["tflops"] /
                                            res["in_core"]["tflops"], 3)
    res["below_B_over_in_core"] = round(res["out_of_core_below_B"]["tflops"] /
                                        res["in_core"]["tflops"], 3)
    return res


def _errors(c, ref, scale):
    d = np.abs(c - ref)
    nz = np.abs(ref) > 0
    return {"max_rel": float(np.max(d[nz] / np.abs(ref[nz]))),
            "max_normwise": float(np.max(d / scale)),
            "fro_rel": float(np.linalg.norm(d) / np.linalg.norm(ref))}


def split_precision(bx, n=16384, phis=(0.5, 2.0, 4.0), slices=(6, 7, 8),
                    moduli=(14, 15, 16, 17, 18, 20), blocks=2):
    """Split-precision vs plain FP64 DGEMM on inputs a = (u-0.5) exp(phi g):
    effective TFLOP/s (2n^3 / time) and errors against a double-double
    reference on sampled 256x256 blocks of C, for the digit split (s slices,
    s(s+1)/2 INT8 products) and the CRT split (L moduli, L INT8 products)."""
    global np
    import numpy as np

    dA = bx.DeviceArray(n, n, np.float64)
    dB = bx.DeviceArray(n, n, np.float64)
    dC = bx.DeviceArray(n, n, np.float64)
    flops = 2.0 * n ** 3
    rng = np.random.default_rng(2024)
    corners = [tuple(int(v) for v in rng.integers(0, n - 256, 2)) for _ in range(blocks)]
    ev0, ev1 = bx.Event(), bx.Event()

    def timed(fn):
        fn()
        bx.synchronize()
        ev0.record()
        fn()
        ev1.record()
        ev1.synchronize()
        return ev0.elapsed_ms(ev1) * 1e-3

    def sample():
        out = []
        for (r0, c0) in corners:
            full = np.empty((256, 256), order="F")
            # column c of the block lives at dC + (r0 + (c0+c)*n)*8
            for c in range(256):
                src = dC.ptr + (r0 + (c0 + c) * n) * 8
                bx.lib().blasx_memcpy_d2h(full[:, c].ctypes.data, src, 256 * 8)
            out.append(full)
        return out

    res = {"n": n, "blocks": f"{blocks} x 256x256, corners {corners}", "per_phi": {}}
    A = np.zeros((n, n), order="F")
    B = np.zeros((n, n), order="F")
    for phi in phis:
        log(f"split: phi={phi}: generating inputs")
        bx.fill_dynrange_host(A, 0, 0, phi)
        bx.fill_dynrange_host(B, 0, 1, phi)
        dA.upload(A)
        dB.upload(B)
        refs, scales = [], []
        for (r0, c0) in corners:
            As = np.asfortranarray(A[r0:r0 + 256])
            Bs = np.asfortranarray(B[:, c0:c0 + 256])
            refs.append(bx.dd_gemm_block("N", "N", 1.0, As, Bs))
            scales.append(np.abs(As) @ np.abs(Bs))
        entry = {}
        t = timed(lambda: bx.device_gemm("N", "N", n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0,
                                         dC.ptr, n, np.float64))
        errs = [_errors(c, r, s) for c, r, s in zip(sample(), refs, scales)]
        entry["plain"] = {"tflops": round(flops / t / 1e12, 3),
                          **{k: max(e[k] for e in errs) for k in errs[0]}}
        log(f"split: phi={phi} plain {entry['plain']}")
        for s in slices:
            t = timed(lambda: bx.dgemm_split_device("N", "N", n, n, n, 1.0, dA.ptr, n, dB.ptr,
                                                    n, 0.0, dC.ptr, n, s))
            errs = [_errors(c, r, sc) for c, r, sc in zip(sample(), refs, scales)]
            entry[f"s{s}"] = {"tflops_effective": round(flops / t / 1e12, 3),
                              "int8_pair_products": s * (s + 1) // 2,
                              **{k: max(e[k] for e in errs) for k in errs[0]}}
            log(f"split: phi={phi} s={s} {entry[f's{s}']}")
        for L in moduli:
            t = timed(lambda: bx.dgemm_crt_device("N", "N", n, n, n, 1.0, dA.ptr, n, dB.ptr, n,
                                                  0.0, dC.ptr, n, L))
            errs = [_errors(c, r, sc) for c, r, sc in zip(sample(), refs, scales)]
            entry[f"crt{L}"] = {"tflops_effective": round(flops / t / 1e12, 3),
                                "int8_products": L, "bits": bx.crt_bits(n, L),
                                **{k: max(e[k] for e in errs) for k in errs[0]}}
            log(f"split: phi={phi} crt L={L} {entry[f'crt{L}']}")
        variants = [k for k in entry if k != "plain"]
        ok = [k for k in variants if entry[k]["max_rel"] <= entry["plain"]["max_rel"]]
        if ok:
            best = max(ok, key=lambda k: entry[k]["tflops_effective"])
            entry["fastest_at_or_below_plain_max_rel"] = best
            entry["speedup_at_equal_or_lower_error"] = round(
                entry[best]["tflops_effective"] / entry["plain"]["tflops"], 3)
        else:
            entry["fastest_at_or_below_plain_max_rel"] = None
        res["per_phi"][str(phi)] = entry
    for X in (dA, dB, dC):
        X.free()
    return res


def zsplit_precision(bx, n=8192, phi=0.5, slices=(6, 7), blocks=2):
    """Complex128: plain ZGEMM (four FP64 MFMAs per complex MAC) vs the digit
    split (4 real split products) and the CRT split (3 INT8 products per
    modulus); effective TFLOP/s = 8n^3 / time, errors on sampled 128x128
    blocks against double-double references of the real and imaginary parts
    (Re C = [Ar, -Ai] [Br; Bi], Im C = [Ar, Ai] [Bi; Br], one rounding each)."""
    import numpy as np

    flops = 8.0 * n ** 3
    A = np.empty((n, n), dtype=np.complex128, order="F")
    B = np.empty((n, n), dtype=np.complex128, order="F")
    plane = np.empty((n, n), order="F")
    for X, ids in ((A, (0, 1)), (B, (2, 3))):
        bx.fill_dynrange_host(plane, 11, ids[0], phi)
        X.real = plane
        bx.fill_dynrange_host(plane, 11, ids[1], phi)
        X.imag = plane
    del plane
    dA, dB = bx.DeviceArray.from_host(A), bx.DeviceArray.from_host(B)
    dC = bx.DeviceArray(n, n, np.complex128)
    rng = np.random.default_rng(2025)
    corners = [tuple(int(v) for v in rng.integers(0, n - 128, 2)) for _ in range(blocks)]
    refs, scales = [], []
    for (r0, c0) in corners:
        As, Bs = A[r0:r0 + 128], B[:, c0:c0 + 128]
        Ar, Ai = np.asfortranarray(As.real), np.asfortranarray(As.imag)
        Br, Bi = np.asfortranarray(Bs.real), np.asfortranarray(Bs.imag)
        re = bx.dd_gemm_block("N", "N", 1.0, np.asfortranarray(np.hstack([Ar, -Ai])),
                              np.asfortranarray(np.vstack([Br, Bi])))
        im = bx.dd_gemm_block("N", "N", 1.0, np.asfortranarray(np.hstack([Ar, Ai])),
                              np.asfortranarray(np.vstack([Bi, Br])))
        refs.append(re + 1j * im)
        scales.append(np.abs(As) @ np.abs(Bs))
    del A, B
    ev0, ev1 = bx.Event(), bx.Event()

    def timed(fn):
        fn()
        bx.synchronize()
        ev0.record()
        fn()
        ev1.record()
        ev1.synchronize()
        return ev0.elapsed_ms(ev1) * 1e-3

    def errors():
        out = []
        for (r0, c0), ref, sc in zip(corners, refs, scales):
            blk = np.empty((128, 128), dtype=np.complex128, order="F")
            for c in range(128):
                src = dC.ptr + (r0 + (c0 + c) * n) * 16
                bx.lib().blasx_memcpy_d2h(blk[:, c].ctypes.data, src, 128 * 16)
            out.append(_errors(blk, ref, sc))
        return {k: max(e[k] for e in out) for k in out[0]}

    res = {"n": n, "phi": phi, "blocks": f"{blocks} x 128x128, corners {corners}"}
    t = timed(lambda: bx.device_gemm("N", "N", n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0, dC.ptr,
                                     n, np.complex128))
    res["plain"] = {"tflops": round(flops / t / 1e12, 3), **errors()}
    log(f"zsplit: plain {res['plain']}")
    for s in slices:
        t = timed(lambda: bx.zgemm_split_device("N", "N", n, n, n, 1.0, dA.ptr, n, dB.ptr, n,
                                                0.0, dC.ptr, n, s))
        res[f"s{s}"] = {"tflops_effective": round(flops / t / 1e12, 3),
                        "int8_products": 4 * s * (s + 1) // 2, **errors()}
        log(f"zsplit: s={s} {res[f's{s}']}")
    L0 = next(L for L in range(2, 21) if bx.crt_bits(2 * n, L) >= 56)
    for L in range(L0 - 3, L0 + 1):
        t = timed(lambda: bx.zgemm_crt_device("N", "N", n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0,
                                              dC.ptr, n, L))
        res[f"crt{L}"] = {"tflops_effective": round(flops / t / 1e12, 3), "int8_products": 3 * L,
                          "bits": bx.crt_bits(2 * n, L), **errors()}
        log(f"zsplit: crt L={L} {res[f'crt{L}']}")
    ok = [k for k in res if k.startswith(("s", "crt")) and isinstance(res[k], dict)
          and res[k]["max_rel"] <= res["plain"]["max_rel"]]
    if ok:
        best = max(ok, key=lambda k: res[k]["tflops_effective"])
        res["fastest_at_or_below_plain_max_rel"] = best
        res["speedup_at_equal_or_lower_error"] = round(
            res[best]["tflops_effective"] / res["plain"]["tflops"], 3)
    for X in (dA, dB, dC):
        X.free()
    return res


def serving(bx, rates=(50, 100, 200, 500, 1000, 2000), nreq=200, budget=200 << 30, seed=7):
    """Mixed GEMM request stream through the continuous-batching server vs the
    same trace through the BLAS ABI one call at a time.

    Requests: Poisson arrivals at `rate`/s; M, N, K log-uniform in [64, 8192];
    60% FP64 / 30% FP32 / 10% complex128; random transposes; 30% use one of 8
    resident 4096x4096 FP64 B operands (K, N <= 4096, a view into it).
    Operands are views into pinned host pools (their values do not matter
    for timing; results are checked by the test-suite), so C outputs of
    concurrent requests may overlap in host memory."""
    import numpy as np

    L = bx.lib()
    L.blasx_finalize()  # release the tile cache of the sections before
    rng = np.random.default_rng(seed)
    dts = [np.float64, np.float32, np.complex128]
    pools = {}
    for dt in dts:
        pools[dt] = [bx.PinnedArray(8192, 8192, dt) for _ in range(3)]
        for i, p in enumerate(pools[dt]):
            bx.fill_uniform_host(p.array, 11, i)
    shared = [bx.PinnedArray(4096, 4096, np.float64) for _ in range(8)]
    for i, p in enumerate(shared):
        bx.fill_uniform_host(p.array, 12, i)

    def lu():
        return int(np.exp(rng.uniform(np.log(64), np.log(8192))))

    def make_trace(rate):
        t, reqs = 0.0, []
        for _ in range(nreq):
            t += rng.exponential(1.0 / rate)
            dt = dts[rng.choice(3, p=[0.6, 0.3, 0.1])]
            M, N, K = lu(), lu(), lu()
            cplx = dt == np.complex128
            ta = str(rng.choice(["N", "T", "C"] if cplx else ["N", "T"]))
            tb = str(rng.choice(["N", "T", "C"] if cplx else ["N", "T"]))
            Ap, Bp, Cp = (p.array for p in pools[dt])
            sh = None
            if rng.random() < 0.3:
                dt, Ap, Bp, Cp = np.float64, *(p.array for p in pools[np.float64])
                N, K, tb = min(N, 4096), min(K, 4096), "N"
                sh = int(rng.integers(8))
                Bp = shared[sh].array
                ta = "N" if ta == "C" else ta
            reqs.append(dict(t=t, dt=dt, ta=ta, tb=tb, M=M, N=N, K=K, A=Ap, B=Bp, C=Cp,
                             shared=sh))
        return reqs

    def desc(r):
        ld = r["A"].shape[0]
        return bx.make_desc(r["dt"], r["ta"], r["tb"], r["M"], r["N"], r["K"], 1.0, r["A"], ld,
                            r["B"], r["B"].shape[0], 1.0, r["C"], r["C"].shape[0])

    def flops(r):
        return (8.0 if r["dt"] == np.complex128 else 2.0) * r["M"] * r["N"] * r["K"]

    out = {"requests_per_rate": nreq, "budget_bytes": budget, "rates": {}}
    srv = bx.Server(budget, max_batch=256)
    try:
        for p in shared:
            srv.register_operand(p.array)
        for rate in rates:
            trace = make_trace(rate)
            log(f"serving: rate {rate}/s, {nreq} requests")
            ids = []
            t0 = time.perf_counter()
            for r in trace:
                while True:
                    dtw = t0 + r["t"] - time.perf_counter()
                    if dtw <= 0:
                        break
                    time.sleep(min(dtw, 0.001))
                ids.append(srv.submit(desc(r)))
            lat = [srv.wait(i) for i in ids]
            wall = time.perf_counter() - t0
            fl = sum(flops(r) for r in trace)
            lat_s = sorted(lat)
            out["rates"][str(rate)] = {
                "req_per_s": round(nreq / wall, 2),
                "tflops": round(fl / wall / 1e12, 3),
                "p50_ms": round(lat_s[len(lat_s) // 2], 3),
                "p99_ms": round(lat_s[min(len(lat_s) - 1, int(0.99 * len(lat_s)))], 3),
            }
            log(f"serving: rate {rate}: {out['rates'][str(rate)]}")
        out["server_counters"] = srv.counters()
    finally:
        srv.close()
    # the same request mix, one blocking ABI call at a time (default dispatch)
    trace = make_trace(1e9)
    log("serving: serial ABI baseline")
    t0 = time.perf_counter()
    lat = []
    for r in trace:
        t = time.perf_counter()
        bx.cblas_gemm(bx.COL_MAJOR, r["ta"], r["tb"], r["M"], r["N"], r["K"], 1.0, r["A"],
                      r["A"].shape[0], r["B"], r["B"].shape[0], 1.0, r["C"], r["C"].shape[0])
        lat.append((time.perf_counter() - t) * 1e3)
    wall = time.perf_counter() - t0
    lat.sort()
    out["serial_abi"] = {"req_per_s": round(nreq / wall, 2),
                         "tflops": round(sum(flops(r) for r in trace) / wall / 1e12, 3),
                         "p50_ms_service": round(lat[len(lat) // 2], 3),
                         "p99_ms_service": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))], 3)}
    log(f"serving: serial ABI {out['serial_abi']}")
    best = max(v["req_per_s"] for v in out["rates"].values())
    out["best_req_per_s_vs_serial"] = round(best / out["serial_abi"]["req_per_s"], 3)
    for dt in dts:
        for p in pools[dt]:
            p.free()
    for p in shared:
        p.free()
    return out


def traffic_probe(bx, n):
    """One DGEMM launch at n^3 and one 1 GiB copy-kernel launch, for the
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes: the copy kernel moves a
    known 1 GiB each way and calibrates the counters."""
    import numpy as np

    A = bx.DeviceArray(n, n, np.float64)
    B = bx.DeviceArray(n, n, np.float64)
    Cm = bx.DeviceArray(n, n, np.float64)
    for i, X in enumerate((A, B, Cm)):
        X.fill_uniform(0, i)
    bx.synchronize()
    bx.device_gemm("N", "N", n, n, n, 1.0, A.ptr, n, B.ptr, n, 1.0, Cm.ptr, n, np.float64)
    bx.synchronize()
    src = bx.DeviceArray(1 << 17, 1024, np.float64)  # 1 GiB
    dst = bx.DeviceArray(1 << 17, 1024, np.float64)
    rc = bx.lib().blasx_device_copy(dst.ptr, src.ptr, 1 << 30, None)
    bx.synchronize()
    log(f"traffic probe: one DGEMM {n}^3 and one 1 GiB copy (rc {rc})")
    print(json.dumps({"traffic_probe": {"n": n, "copy_bytes": 1 << 30}}), flush=True)


def crt_probe(bx, n, moduli):
    """One CRT split DGEMM launch sequence at n^3 (NN, operands generated in
    HBM), for rocprofv3 kernel-trace and counter passes."""
    import numpy as np

    A = bx.DeviceArray(n, n, np.float64)
    B = bx.DeviceArray(n, n, np.float64)
    Cm = bx.DeviceArray(n, n, np.float64)
    A.fill_uniform(0, 0)
    B.fill_uniform(0, 1)
    bx.synchronize()
    bx.dgemm_crt_device("N", "N", n, n, n, 1.0, A.ptr, n, B.ptr, n, 0.0, Cm.ptr, n, moduli)
    bx.synchronize()
    log(f"crt probe: one CRT DGEMM {n}^3 with {moduli} moduli")
    print(json.dumps({"crt_probe": {"n": n, "moduli": moduli,
                                    "bits": bx.crt_bits(n, moduli)}}), flush=True)


def peaks(bx):
    """Measured MFMA and HBM ceilings of this GPU (blasx_measure_peaks)."""
    import ctypes as C

    out = (C.c_double * 4)()
    rc = bx.lib().blasx_measure_peaks(out)
    if rc:
        raise RuntimeError(f"blasx_measure_peaks failed ({rc})")
    res = {"fp64_mfma_tflops": round(out[0], 2), "fp32_mfma_tflops": round(out[1], 2),
           "int8_mfma_tops": round(out[2], 1), "hbm_copy_GBps": round(out[3], 1)}
    log(f"peaks: {res}")
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--no-extras", action="store_true", help="headline line only")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--split-n", type=int, default=16384, help="size of the split-precision run")
    ap.add_argument("--serve-requests", type=int, default=200, help="requests per arrival rate")
    ap.add_argument("--ooc-n", type=int, default=65536, help="size of the out-of-core run")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="one DGEMM + one 1 GiB copy kernel, for rocprofv3 --pmc passes")
    ap.add_argument("--peaks", action="store_true",
                    help="measure the MFMA and HBM ceilings of this GPU and exit")
    ap.add_argument("--crt-probe", type=int, default=0, metavar="MODULI",
                    help="one CRT split DGEMM at --n with MODULI moduli, for rocprofv3 passes")
    ap.add_argument("--only", default="", help="comma list of extras to run "
                    "(matrix,host,hostsz,ooc,split,zsplit,serve,syrk,level3); default all")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build targets one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    import blasx_amd as bx

    if not bx.gpu_available():
        print("bench.py: libblasx reports no usable GPU", file=sys.stderr)
        sys.exit(1)
    n = args.n
    if args.traffic_probe:
        traffic_probe(bx, n)
        return
    if args.peaks:
        print(json.dumps({"peaks": peaks(bx)}), flush=True)
        return
    if args.crt_probe:
        crt_probe(bx, n, args.crt_probe)
        return
    wall, avg_ms, flops = headline(bx, n, args.steps, args.warmup)
    value = flops * args.steps / wall / 1e12
    achieved = flops / (avg_ms * 1e-3) / 1e12
    traffic = load_traffic(f"dgemm_nn_{n}")
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 uniform[-1,1), generated on the device)",
        "config": {"workload": f"in-core DGEMM M=N=K={n} NN alpha=1 beta=1, operands resident "
                               "in HBM (plain FP64 MFMA)",
                   "M": n, "N": n, "K": n, "transa": "N", "transb": "N",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
                     "kernel": "k_dgemm_dma<MN,K> (FP64 MFMA 16x16x4, global->LDS DMA staging)",
                     "launch_ms": round(avg_ms, 3)},
    }
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # reported, never silently dropped
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    if not args.no_extras:
        only = set(x for x in args.only.split(",") if x)
        sections = [("matrix", "matrix_inhbm", lambda: matrix_inhbm(bx, n)),
                    ("host", "host_resident", lambda: host_resident(bx, n)),
                    ("hostsz", "host_resident_sz", lambda: host_resident_sz(bx, n)),
                    ("ooc", "out_of_core", lambda: out_of_core(bx, n=args.ooc_n)),
                    ("split", "split_precision", lambda: split_precision(bx, n=args.split_n)),
                    ("zsplit", "zsplit_precision", lambda: zsplit_precision(bx)),
                    ("serve", "serving", lambda: serving(bx, nreq=args.serve_requests)),
                    ("syrk", "rank_k", lambda: rank_k(bx, n)),
                    ("level3", "level3_inhbm", lambda: level3_inhbm(bx, n))]
        for key, name, fn in sections:
            if only and key not in only:
                continue
            try:
                out[name] = fn()
            except Exception as e:  # reported in the line, never dropped silently
                out[name] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
