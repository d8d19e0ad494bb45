"""CPU oracle for the GEMM hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker or the CPU baseline.  The product
(blasx_amd/, libblasx.so) never loads it.

liboracle.so (oracle/oracle.c, oracle/openblas.c) restates:
  * the BLAS GEMM definition the reference hands to cuBLAS / OpenBLAS
    (blas/dgemm.c:18-128, blas/zgemm.c:16-130, blas/blasx_dgemm.c:82-94),
  * argument validation and info numbers (blas/dgemm.c:29-57),
  * the f77 trans-character mapping (blas/blasx_internal.c:161-174),
  * the tile decomposition (blas/blasx_dgemm.c:28-94,284-296,
    blas/blasx_mem_control.c:24-42),
  * FLOP counts (include/flops.h:218-221),
  * the seeded SplitMix64 input generator,
  * a double-double GEMM (Ogita-Rump-Oishi TwoSum/TwoProd),
  * the reference's CPU BLAS call (dlopen + dlsym of cblas_?gemm,
    blas/blasx_internal.c:26-36).

_ref/libref_rt.so is compiled from the reference's own
blas/red_black_tree.c and blas/blasx_task_queue.c (see oracle/Makefile) and is
used for differential tests of the tile index and task queue.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_rt.so")

_lib = None

TRANS = {"N": 111, "T": 112, "C": 113}

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = C.CDLL(_LIB_PATH)
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_uniform.restype = C.c_double
        L.orc_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64]
        L.orc_dynrange.restype = C.c_double
        L.orc_dynrange.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_double]
        for nm, p in (("orc_fill_uniform_d", _dp), ("orc_fill_uniform_s", _fp),
                      ("orc_fill_uniform_z", _dp)):
            getattr(L, nm).argtypes = [p, C.c_int64, C.c_int64, C.c_int64, C.c_uint64, C.c_uint32]
        L.orc_bf16_rn8.restype = C.c_double
        L.orc_bf16_rn8.argtypes = [C.c_double]
        L.orc_bf16x3_pieces.argtypes = [C.c_double, _dp]
        L.orc_dgemm_bf16x3.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double,
                                       _dp, C.c_int, _dp, C.c_int, C.c_double, _dp, C.c_int]
        L.orc_sample_uniform.argtypes = [_dp, C.c_int, C.c_int64, C.c_int, _i64p, C.c_int,
                                         _i64p, C.c_uint64, C.c_uint32]
        L.orc_fill_dynrange_d.argtypes = [_dp, C.c_int64, C.c_int64, C.c_int64, C.c_uint64,
                                          C.c_uint32, C.c_double]
        L.orc_gemm_info.argtypes = [C.c_int] * 8
        L.orc_gemm_info.restype = C.c_int
        L.orc_f77_trans.argtypes = [C.c_char]
        L.orc_f77_trans.restype = C.c_int
        L.orc_gemm_d.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double,
                                 _dp, C.c_int, _dp, C.c_int, C.c_double, _dp, C.c_int]
        L.orc_gemm_tiled_d.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_double, _dp, C.c_int, _dp, C.c_int, C.c_double, _dp,
                                       C.c_int]
        for nm in ("orc_gemm_s", "orc_gemm_s_acc64"):
            getattr(L, nm).argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_float,
                                       _fp, C.c_int, _fp, C.c_int, C.c_float, _fp, C.c_int]
        L.orc_gemm_z.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, _dp,
                                 _dp, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int]
        L.orc_gemm_dd_d.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double,
                                    _dp, C.c_int, _dp, C.c_int, C.c_double, _dp, C.c_int,
                                    _dp, _dp, C.c_int]
        L.orc_gemm_dd_z.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, _dp,
                                    _dp, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int,
                                    _dp, _dp, C.c_int]
        L.orc_dd_two_sum.argtypes = [C.c_double, C.c_double, _dp, _dp]
        L.orc_dd_two_prod.argtypes = [C.c_double, C.c_double, _dp, _dp]
        L.orc_dd_dot.argtypes = [_dp, _dp, C.c_int, _dp, _dp]
        L.orc_ozaki_digits.argtypes = [C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int8)]
        L.orc_dgemm_ozaki.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double,
                                      _dp, C.c_int, _dp, C.c_int, C.c_double, _dp, C.c_int,
                                      C.c_int, C.c_int]
        L.orc_zgemm_ozaki.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                      C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                      C.c_int]
        L.orc_crt_modulus.argtypes = [C.c_int]
        L.orc_crt_beta.argtypes = [C.c_int, C.c_int]
        L.orc_crt_gamma.argtypes = [C.c_int]
        L.orc_norm_exp.argtypes = [C.c_double]
        L.orc_smod.restype = C.c_int64
        L.orc_smod.argtypes = [C.c_int64, C.c_int64]
        L.orc_crt_resid.argtypes = [C.c_double, C.c_int]
        L.orc_crt_digits.argtypes = [C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
        L.orc_crt_horner.restype = C.c_double
        L.orc_crt_horner.argtypes = [C.POINTER(C.c_int), C.c_int]
        L.orc_crt_consts.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_crt_float.restype = C.c_double
        L.orc_crt_float.argtypes = [C.POINTER(C.c_int), C.c_int, _dp, _dp, C.c_double]
        L.orc_zgemm_crt.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                    C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, C.c_int]
        L.orc_dgemm_crt.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double,
                                    _dp, C.c_int, _dp, C.c_int, C.c_double, _dp, C.c_int,
                                    C.c_int]
        L.orc_flops_gemm.restype = C.c_double
        L.orc_flops_gemm.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_openblas_load.argtypes = [C.c_char_p]
        L.orc_openblas_load.restype = C.c_int
        L.orc_openblas_path.restype = C.c_char_p
        L.orc_openblas_config.restype = C.c_char_p
        L.orc_openblas_is_ilp64.restype = C.c_int
        L.orc_openblas_set_threads.argtypes = [C.c_int]
        L.orc_openblas_dgemm.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int64,
                                         C.c_double, _dp, C.c_int64, _dp, C.c_int64, C.c_double,
                                         _dp, C.c_int64]
        L.orc_openblas_sgemm.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int64,
                                         C.c_float, _fp, C.c_int64, _fp, C.c_int64, C.c_float,
                                         _fp, C.c_int64]
        L.orc_openblas_zgemm.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int64,
                                         _dp, _dp, C.c_int64, _dp, C.c_int64, _dp, _dp, C.c_int64]
        _lib = L
    return _lib


def _ptr(a, t):
    assert a.flags["F_CONTIGUOUS"] or a.ndim == 1 or a.size == 0
    return a.ctypes.data_as(t)


def _fortran_buffer(a):
    """Raw column-major storage: the caller's (ld x cols) Fortran array."""
    return a


# ------------------------------------------------------------------ generator
def splitmix64(x):
    return lib().orc_splitmix64(x)


def uniform(seed, mid, idx):
    return lib().orc_uniform(seed, mid, idx)


def fill_uniform(rows, cols, seed, mid, dtype=np.float64, ld=None):
    """Column-major (ld x cols) array with element (i,j) = U(seed, mid, i + j*rows)."""
    ld = rows if ld is None else ld
    if dtype == np.float64:
        out = np.zeros((ld, cols), dtype=np.float64, order="F")
        lib().orc_fill_uniform_d(_ptr(out, _dp), rows, cols, ld, seed, mid)
    elif dtype == np.float32:
        out = np.zeros((ld, cols), dtype=np.float32, order="F")
        lib().orc_fill_uniform_s(_ptr(out, _fp), rows, cols, ld, seed, mid)
    elif dtype == np.complex128:
        out = np.zeros((ld, cols), dtype=np.complex128, order="F")
        lib().orc_fill_uniform_z(out.ctypes.data_as(_dp), rows, cols, ld, seed, mid)
    else:
        raise TypeError(dtype)
    return out


def sample_uniform(rows, ri, ci, seed, mid, dtype=np.float64):
    """Entries (ri[a], ci[b]) of the seeded matrix with `rows` rows, as the
    (len(ri), len(ci)) array fill_uniform would hold there."""
    ri = np.ascontiguousarray(ri, dtype=np.int64)
    ci = np.ascontiguousarray(ci, dtype=np.int64)
    kind = {np.float64: 0, np.float32: 1, np.complex128: 2}[np.dtype(dtype).type]
    out = np.zeros((len(ri), len(ci)), dtype=np.complex128 if kind == 2 else np.float64,
                   order="F")
    lib().orc_sample_uniform(out.ctypes.data_as(_dp), kind, rows, len(ri), _ptr(ri, _i64p),
                             len(ci), _ptr(ci, _i64p), seed, mid)
    return out.astype(np.float32, order="F") if kind == 1 else out


def fill_dynrange(rows, cols, seed, mid, phi, ld=None):
    ld = rows if ld is None else ld
    out = np.zeros((ld, cols), dtype=np.float64, order="F")
    lib().orc_fill_dynrange_d(_ptr(out, _dp), rows, cols, ld, seed, mid, phi)
    return out


# ------------------------------------------------------------------ validation
def gemm_info(transa, transb, M, N, K, lda, ldb, ldc):
    return lib().orc_gemm_info(transa, transb, M, N, K, lda, ldb, ldc)


def f77_trans(ch):
    return lib().orc_f77_trans(ch.encode() if isinstance(ch, str) else ch)


# ------------------------------------------------------------------ GEMM
def gemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C_, ldc, acc64=False):
    """Naive triple loop on column-major buffers; C_ is updated in place."""
    L = lib()
    ta_b, tb_b = ta.encode(), tb.encode()
    if C_.dtype == np.float64:
        L.orc_gemm_d(ta_b, tb_b, M, N, K, alpha, _ptr(A, _dp), lda, _ptr(B, _dp), ldb, beta,
                     _ptr(C_, _dp), ldc)
    elif C_.dtype == np.float32:
        fn = L.orc_gemm_s_acc64 if acc64 else L.orc_gemm_s
        fn(ta_b, tb_b, M, N, K, alpha, _ptr(A, _fp), lda, _ptr(B, _fp), ldb, beta,
           _ptr(C_, _fp), ldc)
    elif C_.dtype == np.complex128:
        al = np.array([complex(alpha).real, complex(alpha).imag])
        be = np.array([complex(beta).real, complex(beta).imag])
        L.orc_gemm_z(ta_b, tb_b, M, N, K, _ptr(al, _dp), A.ctypes.data_as(_dp), lda,
                     B.ctypes.data_as(_dp), ldb, _ptr(be, _dp), C_.ctypes.data_as(_dp), ldc)
    else:
        raise TypeError(C_.dtype)
    return C_


def gemm_tiled(ta, tb, M, N, K, T, alpha, A, lda, B, ldb, beta, C_, ldc):
    lib().orc_gemm_tiled_d(ta.encode(), tb.encode(), M, N, K, T, alpha, _ptr(A, _dp), lda,
                           _ptr(B, _dp), ldb, beta, _ptr(C_, _dp), ldc)
    return C_


def gemm_dd(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C_, ldc):
    """Double-double GEMM.  Returns (hi, lo) arrays of shape (M, N); C_ is read only."""
    L = lib()
    if C_.dtype == np.float64:
        hi = np.zeros((M, N), order="F")
        lo = np.zeros((M, N), order="F")
        L.orc_gemm_dd_d(ta.encode(), tb.encode(), M, N, K, alpha, _ptr(A, _dp), lda,
                        _ptr(B, _dp), ldb, beta, _ptr(C_, _dp), ldc, _ptr(hi, _dp),
                        _ptr(lo, _dp), max(M, 1))
        return hi, lo
    if C_.dtype == np.complex128:
        hi = np.zeros((M, N), dtype=np.complex128, order="F")
        lo = np.zeros((M, N), dtype=np.complex128, order="F")
        al = np.array([complex(alpha).real, complex(alpha).imag])
        be = np.array([complex(beta).real, complex(beta).imag])
        L.orc_gemm_dd_z(ta.encode(), tb.encode(), M, N, K, _ptr(al, _dp),
                        A.ctypes.data_as(_dp), lda, B.ctypes.data_as(_dp), ldb, _ptr(be, _dp),
                        C_.ctypes.data_as(_dp), ldc, hi.ctypes.data_as(_dp),
                        lo.ctypes.data_as(_dp), max(M, 1))
        return hi, lo
    raise TypeError(C_.dtype)


def ozaki_digits(x, e, s):
    d = (C.c_int8 * 9)()
    lib().orc_ozaki_digits(x, e, s, d)
    return [int(d[t]) for t in range(s)]


def ozaki_segmax(K, k_align=128):
    """Slice pairs per exact int32 launch: nseg * Kp * 127^2 < 2^31, with Kp
    the K extent the library pads to (a multiple of 128: pairs of 64-byte
    K tiles)."""
    kp = (K + k_align - 1) // k_align * k_align
    return (2**31 - 1) // (kp * 127 * 127)


def dgemm_ozaki(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C_, ldc, s, segmax=None):
    """CPU restatement of the split-precision DGEMM; C_ updated in place."""
    segmax = ozaki_segmax(K) if segmax is None else segmax
    lib().orc_dgemm_ozaki(ta.encode(), tb.encode(), M, N, K, alpha, _ptr(A, _dp), lda,
                          _ptr(B, _dp), ldb, beta, _ptr(C_, _dp), ldc, s, segmax)
    return C_


def zgemm_ozaki(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C_, ldc, s, segmax=None):
    """CPU restatement of the complex split-precision GEMM; C_ updated in place."""
    segmax = ozaki_segmax(K) if segmax is None else segmax
    al = np.array([complex(alpha).real, complex(alpha).imag])
    be = np.array([complex(beta).real, complex(beta).imag])
    lib().orc_zgemm_ozaki(ta.encode(), tb.encode(), M, N, K, _ptr(al, _dp),
                          A.ctypes.data_as(_dp), lda, B.ctypes.data_as(_dp), ldb, _ptr(be, _dp),
                          C_.ctypes.data_as(_dp), ldc, s, segmax)
    return C_


def bf16x3_pieces(x):
    out = (C.c_double * 3)()
    lib().orc_bf16x3_pieces(float(x), out)
    return [out[0], out[1], out[2]]


def bf16_rn8(x):
    return lib().orc_bf16_rn8(float(x))


def dgemm_bf16x3(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C_, ldc):
    lib().orc_dgemm_bf16x3(ta.encode(), tb.encode(), M, N, K, alpha, _ptr(A, _dp), lda,
                           _ptr(B, _dp), ldb, beta, _ptr(C_, _dp), ldc)
    return C_


def crt_moduli(L):
    """The first L moduli of the CRT split (pairwise coprime, <= 256)."""
    return [lib().orc_crt_modulus(l) for l in range(L)]


def crt_beta(K, L):
    """Worst-case bits kept per operand entry, relative to its row's
    maximum, by the CRT split with L moduli (all entries of a row equal in
    magnitude)."""
    return lib().orc_crt_beta(K, L)


def crt_gamma(L):
    """Scaling bits of the CRT split with L moduli, relative to the row's
    2-norm: 2^(2 gamma + 3) <= prod m_l."""
    return lib().orc_crt_gamma(L)


def norm_exp(ss):
    """Smallest e with ss (1 + 2^-30) < 4^e (the row exponent from a sum of
    squares)."""
    return lib().orc_norm_exp(ss)


def smod(v, m):
    return lib().orc_smod(v, m)


def crt_resid(v, m):
    """Symmetric residue of the integer-valued double v modulo m."""
    return lib().orc_crt_resid(v, m)


def crt_digits(r):
    """Balanced mixed-radix (Garner) digits of residues r[l] mod m_l."""
    L = len(r)
    ri = (C.c_int * L)(*r)
    v = (C.c_int * L)()
    lib().orc_crt_digits(ri, L, v)
    return [int(v[l]) for l in range(L)]


def crt_horner(v):
    vi = (C.c_int * len(v))(*v)
    return lib().orc_crt_horner(vi, len(v))


def crt_consts(L):
    """(Wk[4][L], Mk[4], 1/Md): the 40-bit chunks of the CRT weights and of M."""
    W = np.zeros(4 * L)
    Mk = np.zeros(4)
    inv = C.c_double()
    lib().orc_crt_consts(L, W.ctypes.data_as(_dp), Mk.ctypes.data_as(_dp), C.byref(inv))
    return W.reshape(4, L), Mk, inv.value


def crt_float(r, consts):
    """Floating CRT reconstruction of residues r with crt_consts(len(r))."""
    W, Mk, inv = consts
    ri = (C.c_int * len(r))(*r)
    Wc = np.ascontiguousarray(W, dtype=np.float64)
    return lib().orc_crt_float(ri, len(r), Wc.ctypes.data_as(_dp), Mk.ctypes.data_as(_dp), inv)


def dgemm_crt(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C_, ldc, L):
    """CPU restatement of the CRT split DGEMM; C_ updated in place."""
    lib().orc_dgemm_crt(ta.encode(), tb.encode(), M, N, K, alpha, _ptr(A, _dp), lda,
                        _ptr(B, _dp), ldb, beta, _ptr(C_, _dp), ldc, L)
    return C_


def zgemm_crt(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C_, ldc, L):
    """CPU restatement of the complex CRT split GEMM; C_ updated in place."""
    al = np.array([complex(alpha).real, complex(alpha).imag])
    be = np.array([complex(beta).real, complex(beta).imag])
    lib().orc_zgemm_crt(ta.encode(), tb.encode(), M, N, K, _ptr(al, _dp),
                        A.ctypes.data_as(_dp), lda, B.ctypes.data_as(_dp), ldb, _ptr(be, _dp),
                        C_.ctypes.data_as(_dp), ldc, L)
    return C_


def two_sum(a, b):
    s, e = C.c_double(), C.c_double()
    lib().orc_dd_two_sum(a, b, C.byref(s), C.byref(e))
    return s.value, e.value


def two_prod(a, b):
    p, e = C.c_double(), C.c_double()
    lib().orc_dd_two_prod(a, b, C.byref(p), C.byref(e))
    return p.value, e.value


def dd_dot(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    hi, lo = C.c_double(), C.c_double()
    lib().orc_dd_dot(_ptr(x, _dp), _ptr(y, _dp), len(x), C.byref(hi), C.byref(lo))
    return hi.value, lo.value


def flops(complex_, m, n, k):
    return lib().orc_flops_gemm(1 if complex_ else 0, m, n, k)


# ------------------------------------------------------------------ OpenBLAS
def openblas_load(path=None):
    ok = lib().orc_openblas_load(path.encode() if path else None)
    if not ok:
        raise RuntimeError("no CPU BLAS (OpenBLAS) library found")
    return lib().orc_openblas_path().decode()


def openblas_info():
    return {"path": lib().orc_openblas_path().decode(),
            "config": lib().orc_openblas_config().decode(),
            "ilp64": bool(lib().orc_openblas_is_ilp64())}


def openblas_set_threads(n):
    lib().orc_openblas_set_threads(n)


def openblas_gemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C_, ldc):
    openblas_load()
    L = lib()
    if C_.dtype == np.float64:
        L.orc_openblas_dgemm(TRANS[ta], TRANS[tb], M, N, K, alpha, _ptr(A, _dp), lda,
                             _ptr(B, _dp), ldb, beta, _ptr(C_, _dp), ldc)
    elif C_.dtype == np.float32:
        L.orc_openblas_sgemm(TRANS[ta], TRANS[tb], M, N, K, alpha, _ptr(A, _fp), lda,
                             _ptr(B, _fp), ldb, beta, _ptr(C_, _fp), ldc)
    elif C_.dtype == np.complex128:
        al = np.array([complex(alpha).real, complex(alpha).imag])
        be = np.array([complex(beta).real, complex(beta).imag])
        L.orc_openblas_zgemm(TRANS[ta], TRANS[tb], M, N, K, _ptr(al, _dp),
                             A.ctypes.data_as(_dp), lda, B.ctypes.data_as(_dp), ldb,
                             _ptr(be, _dp), C_.ctypes.data_as(_dp), ldc)
    else:
        raise TypeError(C_.dtype)
    return C_


# ------------------------------------------------------------------ reference pieces
def ref_lib():
    """The reference's own red-black tree + task queue (oracle/_ref), or None."""
    if not os.path.exists(_REF_PATH):
        return None
    L = C.CDLL(_REF_PATH)
    L.rbt_init.restype = C.c_void_p
    L.rbt_find.restype = C.c_void_p
    L.rbt_find.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    L.rbt_insert.restype = C.c_void_p
    L.rbt_insert.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]
    L.rbt_delete.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    L.rbt_free.argtypes = [C.POINTER(C.c_void_p)]
    L.init_queue.argtypes = [C.c_void_p, C.c_int]
    L.dequeue.argtypes = [C.c_void_p]
    L.dequeue.restype = C.c_int
    return L


def relerr_fro(Cg, Cr, alpha, A, B, beta, C0):
    """||Cg - Cr||_F / (|alpha| ||A||_F ||B||_F + |beta| ||C0||_F)."""
    num = np.linalg.norm((Cg - Cr).ravel())
    den = abs(alpha) * np.linalg.norm(A.ravel()) * np.linalg.norm(B.ravel())
    den += abs(beta) * np.linalg.norm(C0.ravel())
    return num / den if den > 0 else num
