"""The C-ABI boundary without a GPU: exported symbols, argument validation
and xerbla_ messages (blas/dgemm.c:46-57, blas/error.c:5-8), the CPU-BLAS
dispatch of small problems (blas/dgemm.c:61-74), quick returns, RowMajor,
and the no-GPU behaviour of GPU-sized calls."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EPS = np.finfo(np.float64).eps


def _header_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", src, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "sizeof", "return"):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol(bx):
    declared = _header_functions()
    assert {"cblas_dgemm", "dgemm_", "zgemm_", "xerbla_", "blasx_dgemm_device"} <= declared
    out = subprocess.run(["nm", "-D", "--defined-only", bx.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(declared - exported)
    assert not missing, f"declared in include/ but not exported: {missing}"
    # the test-suite's hooks live in libblasx_test.so, not in the drop-in library
    assert not [n for n in exported if n.startswith("bxtest_")]


def _run_c_snippet(body, env_extra=None, expect_ok=True):
    """Run a tiny Python child that calls the library, capturing C stdout."""
    code = ("import sys, ctypes as C, numpy as np; sys.path.insert(0, %r)\n"
            "import blasx_amd as bx\nL = bx.lib()\n" % ROOT) + body
    env = dict(os.environ)
    for k, v in (env_extra or {}).items():
        if v is None:
            env.pop(k, None)
        else:
            env[k] = v
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                       timeout=300)
    if expect_ok:
        assert r.returncode == 0, r.stderr[-3000:]
    return r


XERBLA_CASES = [
    # CBLAS_ORDER, TransA, TransB, M, N, K, lda, ldb, ldc -> info
    ((102, 110, 111, 2, 2, 2, 2, 2, 2), 1),
    ((102, 111, 0, 2, 2, 2, 2, 2, 2), 2),
    ((102, 111, 111, -1, 2, 2, 2, 2, 2), 3),
    ((102, 111, 111, 2, -1, 2, 2, 2, 2), 4),
    ((102, 111, 111, 2, 2, -1, 2, 2, 2), 5),
    ((102, 111, 111, 3, 2, 2, 2, 2, 3), 8),
    ((102, 112, 111, 3, 2, 4, 3, 4, 3), 8),
    ((102, 111, 111, 3, 2, 4, 3, 3, 3), 10),
    ((102, 111, 112, 3, 5, 4, 3, 4, 3), 10),
    ((102, 111, 111, 3, 2, 4, 3, 4, 2), 13),
]


@pytest.mark.parametrize("args,info", XERBLA_CASES)
def test_xerbla_info_and_message(bx, orc, args, info):
    order, ta, tb, M, N, K, lda, ldb, ldc = args
    # the oracle's restatement of the reference validation agrees
    assert orc.gemm_info(ta, tb, M, N, K, lda, ldb, ldc) == info
    body = ("buf = np.zeros(64)\n"
            "L.cblas_dgemm(%d,%d,%d,%d,%d,%d,1.0,buf.ctypes.data,%d,buf.ctypes.data,%d,0.0,"
            "buf.ctypes.data,%d)\nassert np.all(buf == 0)\n" % (order, ta, tb, M, N, K, lda, ldb, ldc))
    r = _run_c_snippet(body)
    expect = " ** On entry to DGEMM  parameter number %2d had an illegal value" % info
    assert [x for x in r.stdout.splitlines() if x.strip()] == [expect]


@pytest.mark.parametrize("prec,name", [("s", "SGEMM "), ("d", "DGEMM "), ("z", "ZGEMM "),
                                       ("c", "CGEMM ")])
def test_f77_bad_trans_char(bx, prec, name):
    body = ("one = C.c_int(1); buf = np.zeros(16)\n"
            "L.%sgemm_(b'X', b'N', C.byref(one), C.byref(one), C.byref(one), buf.ctypes.data,"
            " buf.ctypes.data, C.byref(one), buf.ctypes.data, C.byref(one), buf.ctypes.data,"
            " buf.ctypes.data, C.byref(one))\n"
            "L.%sgemm_(b'n', b'q', C.byref(one), C.byref(one), C.byref(one), buf.ctypes.data,"
            " buf.ctypes.data, C.byref(one), buf.ctypes.data, C.byref(one), buf.ctypes.data,"
            " buf.ctypes.data, C.byref(one))\n" % (prec, prec))
    r = _run_c_snippet(body)
    lines = [x for x in r.stdout.splitlines() if x.strip()]
    assert lines == [" ** On entry to %6s parameter number  1 had an illegal value" % name,
                     " ** On entry to %6s parameter number  2 had an illegal value" % name]


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.complex128, np.complex64])
def test_small_calls_go_to_cpu_blas(bx, orc, dt):
    rng = np.random.default_rng(20)
    M, N, K = 300, 200, 250  # below both thresholds
    for ta, tb in [("N", "N"), ("T", "N"), ("N", "T"), ("C", "T")]:
        shapeA = (M, K) if ta == "N" else (K, M)
        shapeB = (K, N) if tb == "N" else (N, K)
        if np.dtype(dt).kind == "c":
            A = np.asfortranarray(rng.uniform(-1, 1, shapeA) + 1j * rng.uniform(-1, 1, shapeA)).astype(dt)
            B = np.asfortranarray(rng.uniform(-1, 1, shapeB) + 1j * rng.uniform(-1, 1, shapeB)).astype(dt)
            C0 = np.asfortranarray(rng.uniform(-1, 1, (M, N)) + 1j * rng.uniform(-1, 1, (M, N))).astype(dt)
            al, be = 1 - 1j, 0.5j
        else:
            A = np.asfortranarray(rng.uniform(-1, 1, shapeA)).astype(dt)
            B = np.asfortranarray(rng.uniform(-1, 1, shapeB)).astype(dt)
            C0 = np.asfortranarray(rng.uniform(-1, 1, (M, N))).astype(dt)
            al, be = 1.5, -0.5

        def op(X, t):
            return X if t == "N" else (X.T if t == "T" else (X.conj().T if np.dtype(dt).kind == "c" else X.T))
        ref = al * (op(A.astype(np.complex128 if np.dtype(dt).kind == "c" else np.float64), ta)
                    @ op(B.astype(np.complex128 if np.dtype(dt).kind == "c" else np.float64), tb)) + be * C0
        out = C0.copy(order="F")
        bx.reset_stats()
        bx.cblas_gemm(bx.COL_MAJOR, ta, tb, M, N, K, al, A, A.shape[0], B, B.shape[0], be, out, M)
        st = bx.stats()
        assert st["calls_cpu"] == 1 and st["calls_gpu"] == 0
        eps = np.finfo(dt).eps
        assert np.max(np.abs(out - ref)) <= 8 * K * eps * np.abs(ref).max()


def test_quick_returns_and_alpha_zero(bx):
    M, N, K = 3000, 2000, 1500  # GPU-sized, but these cases never reach the GPU
    C0 = np.asfortranarray(np.random.default_rng(1).uniform(-1, 1, (M, N)))
    An = np.full((1, 1), np.nan)
    out = C0.copy(order="F")
    bx.reset_stats()
    bx.cblas_gemm(bx.COL_MAJOR, "N", "N", M, N, K, 0.0, An, M, An, K, 0.5, out, M)
    assert np.array_equal(out, 0.5 * C0)
    out = C0.copy(order="F")
    bx.cblas_gemm(bx.COL_MAJOR, "N", "N", M, N, 0, 2.0, An, M, An, 1, 0.0, out, M)
    assert np.all(out == 0.0)
    out = C0.copy(order="F")
    bx.cblas_gemm(bx.COL_MAJOR, "N", "N", 0, N, K, 1.0, An, 1, An, K, 3.0, out, 1)
    assert np.array_equal(out, C0)
    st = bx.stats()
    assert st["calls_gpu"] == 0 and st["calls_cpu"] == 0 and st["cpu_fallbacks"] == 0


def test_row_major_cpu_path(bx):
    rng = np.random.default_rng(21)
    M, N, K = 120, 90, 70
    A = rng.uniform(-1, 1, (M, K))  # C-ordered = row-major
    B = rng.uniform(-1, 1, (N, K))
    C0 = rng.uniform(-1, 1, (M, N))
    out = C0.copy()
    bx.cblas_gemm(bx.ROW_MAJOR, "N", "T", M, N, K, 2.0, A, K, B, K, -1.0, out, N)
    assert np.allclose(out, 2.0 * A @ B.T - C0, atol=1e-12)


def test_gpu_sized_call_without_gpu_falls_back_loudly(bx):
    """The failure policy: a GPU-sized call whose GPU path fails is logged and
    rerun on the CPU BLAS by default; BLASX_CPU_FALLBACK=0 makes it fatal."""
    if bx.gpu_available():
        pytest.skip("a GPU is present; this checks the no-GPU behaviour")
    body = ("n = 1100; A = np.ones((n, n), order='F'); C_ = np.zeros((n, n), order='F')\n"
            "bx.gemm('N', 'N', 1.0, A, A, 0.0, C_)\n")
    check = ("st = bx.stats(); assert st['cpu_fallbacks'] == 1 and st['calls_gpu'] == 0, st\n"
             "assert np.all(C_ == 1100.0)\n")
    r = _run_c_snippet(body + check, env_extra={"BLASX_CPU_FALLBACK": None})
    assert "DGEMM : GPU path failed" in r.stderr and "running on the CPU BLAS" in r.stderr
    r = _run_c_snippet(body, env_extra={"BLASX_CPU_FALLBACK": "0"}, expect_ok=False)
    assert r.returncode != 0
    assert "BLASX_CPU_FALLBACK=0 forbids the CPU BLAS" in r.stderr


def test_f77_two_bad_characters_report_the_later_check(bx):
    """Both checks run and the later wins: blas/dgemm.c:141-147 (GEMM),
    blas/dsyrk.c:83-85 (SYRK/HERK), blas/dsyr2k.c:96-98 (SYR2K/HER2K)."""
    body = ("one = C.c_int(1); buf = np.zeros(16)\n"
            "for p in 'sdcz':\n"
            "    getattr(L, p + 'gemm_')(b'X', b'Q', C.byref(one), C.byref(one), C.byref(one),"
            " buf.ctypes.data, buf.ctypes.data, C.byref(one), buf.ctypes.data, C.byref(one),"
            " buf.ctypes.data, buf.ctypes.data, C.byref(one))\n"
            "for r in ('dsyrk_', 'zherk_'):\n"
            "    getattr(L, r)(b'X', b'Q', C.byref(one), C.byref(one), buf.ctypes.data,"
            " buf.ctypes.data, C.byref(one), buf.ctypes.data, buf.ctypes.data, C.byref(one))\n"
            "for r in ('dsyr2k_', 'zher2k_'):\n"
            "    getattr(L, r)(b'X', b'Q', C.byref(one), C.byref(one), buf.ctypes.data,"
            " buf.ctypes.data, C.byref(one), buf.ctypes.data, C.byref(one), buf.ctypes.data,"
            " buf.ctypes.data, C.byref(one))\n"
            "assert np.all(buf == 0)\n")
    r = _run_c_snippet(body)
    lines = [x for x in r.stdout.splitlines() if x.strip()]
    names = ["SGEMM ", "DGEMM ", "CGEMM ", "ZGEMM ", "DSYRK ", " ZHERK", "DSYR2K", "ZHER2K"]
    assert lines == [" ** On entry to %6s parameter number  2 had an illegal value" % n
                     for n in names]


def test_host_generator_matches_oracle(bx, orc):
    for dt in (np.float64, np.float32, np.complex128):
        a = np.zeros((1500, 900), dtype=dt, order="F")  # large enough for the threaded path
        bx.fill_uniform_host(a, 42, 7)
        assert np.array_equal(a, orc.fill_uniform(1500, 900, 42, 7, dtype=dt))


def test_tile_size_heuristic(bx):
    """The heuristic behind the measured table (shapes far from the table's
    cubes): a problem inside one tile gets the smallest edge holding it;
    otherwise 2048 (1024 complex), halved while the C grid has < 16 tiles."""
    L = bx.test_lib()
    assert L.bxtest_autotune_tile(bx.D, 100000, 50000, 3000) == 2048
    assert L.bxtest_autotune_tile(bx.Z, 100000, 50000, 3000) == 1024
    assert L.bxtest_autotune_tile(bx.D, 4000, 8000, 100000) == 1024  # 2 x 4 tiles, then 4 x 8
    assert L.bxtest_autotune_tile(bx.D, 300, 3000, 50000) == 512     # 1 x 2, 1 x 3, stops at 512
    # inside one tile: the smallest edge that holds it
    assert L.bxtest_autotune_tile(bx.D, 1500, 800, 1900) == 2048
    assert L.bxtest_autotune_tile(bx.D, 600, 700, 900) == 1024
    assert L.bxtest_autotune_tile(bx.D, 100, 200, 300) == 512
    assert L.bxtest_autotune_tile(bx.Z, 1500, 800, 1000) == 1024
    assert bx.lib().blasx_set_tile_size(bx.D, 4096) != 0   # larger than a cache slot
    assert bx.lib().blasx_set_tile_size(bx.D, 300) != 0    # not a supported edge


def test_dynrange_host_generator_matches_oracle(bx, orc):
    a = np.zeros((700, 600), order="F")
    bx.fill_dynrange_host(a, 9, 4, 2.0)
    assert np.array_equal(a, orc.fill_dynrange(700, 600, 9, 4, 2.0))


def test_dd_block_reference_matches_oracle(bx, orc):
    A = orc.fill_dynrange(37, 500, 1, 0, 4.0)
    B = orc.fill_dynrange(29, 500, 1, 1, 4.0)   # used transposed
    C0 = orc.fill_uniform(37, 29, 1, 2)
    out = bx.dd_gemm_block("N", "T", 1.5, A, B, -0.25, C0)
    hi, _ = orc.gemm_dd("N", "T", 37, 29, 500, 1.5, A, 37, B, 29, -0.25, C0, 37)
    assert np.array_equal(out, hi)


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128])
def test_sampled_generator_matches_oracle(bx, orc, dtype):
    """The library's sampled generator (the benchmark's full-size checks
    regenerate sampled operand rows and columns with it) equals the
    oracle's fill at the same indices, past 2^31 linear indices too."""
    rows = 131072
    ri = [0, 1, 2047, 65536, 131071]
    ci = [0, 16383, 16384, 131071]          # i + j*rows up to 1.7e10
    S = bx.sample_uniform_host(rows, ri, ci, 3, 1, dtype)
    assert np.array_equal(S, orc.sample_uniform(rows, ri, ci, 3, 1, dtype))
    F = orc.fill_uniform(300, 200, 3, 1, dtype=dtype)
    S = bx.sample_uniform_host(300, [0, 299, 17], [199, 0], 3, 1, dtype)
    assert np.array_equal(S, F[np.ix_([0, 299, 17], [199, 0])])
