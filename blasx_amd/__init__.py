"""blasx_amd -- Python view of libblasx.so, the MI355X-native tiled GEMM library.

The product is the C-ABI shared library ``blasx_amd/lib/libblasx.so`` (header
``include/blasx.h``).  This module only binds it with ctypes so tests and
benchmarks can call the same entry points a C, Fortran or MATLAB program
would:

* ``cblas_gemm(...)`` -> ``cblas_{s,d,c,z}gemm`` (reference blas/dgemm.c:18-128)
* ``f77_gemm(...)``   -> ``{s,d,c,z}gemm_``     (reference blas/dgemm.c:131-154)
* ``device_gemm(...)`` -> ``blasx_{s,d,z}gemm_device`` (operands in HBM)

There is no CPU fallback here: if the library is missing, ``lib()`` raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libblasx.so")

S, D, Z, CPLX64 = 0, 1, 2, 3
ROW_MAJOR, COL_MAJOR = 101, 102
TRANS = {"N": 111, "T": 112, "C": 113}

_DTYPES = {np.dtype(np.float32): S, np.dtype(np.float64): D,
           np.dtype(np.complex128): Z, np.dtype(np.complex64): CPLX64}
_PREFIX = {S: "s", D: "d", Z: "z", CPLX64: "c"}

_vp = C.c_void_p
_lib = None


class Stats(C.Structure):
    _fields_ = [("calls_gpu", C.c_uint64), ("calls_cpu", C.c_uint64),
                ("cpu_fallbacks", C.c_uint64), ("kernel_launches", C.c_uint64),
                ("bytes_h2d", C.c_uint64), ("bytes_d2h", C.c_uint64),
                ("bytes_lower_bound", C.c_uint64), ("tile_hits", C.c_uint64),
                ("tile_misses", C.c_uint64), ("tile_evictions", C.c_uint64),
                ("tiles_generated", C.c_uint64), ("flops", C.c_double)]


class GemmDesc(C.Structure):
    """blasx_gemm_desc_t (include/blasx_ext.h)."""
    _fields_ = [("dtype", C.c_int), ("transa", C.c_int), ("transb", C.c_int),
                ("M", C.c_int), ("N", C.c_int), ("K", C.c_int),
                ("alpha", C.c_double * 2), ("beta", C.c_double * 2),
                ("A", C.c_void_p), ("lda", C.c_int64), ("B", C.c_void_p), ("ldb", C.c_int64),
                ("C", C.c_void_p), ("ldc", C.c_int64)]


def make_desc(dtype, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C_, ldc):
    d = GemmDesc()
    d.dtype = dtype_code(dtype)
    d.transa, d.transb = TRANS[ta], TRANS[tb]
    d.M, d.N, d.K = M, N, K
    d.alpha[0], d.alpha[1] = complex(alpha).real, complex(alpha).imag
    d.beta[0], d.beta[1] = complex(beta).real, complex(beta).imag
    d.A, d.lda, d.B, d.ldb, d.C, d.ldc = _addr(A), lda, _addr(B), ldb, _addr(C_), ldc
    return d


def gemm_grouped_device(descs, stream=None, noalloc=False):
    """blasx_gemm_grouped_device, or its allocation-free variant"""
    arr = (GemmDesc * len(descs))(*descs)
    fn = lib().blasx_gemm_grouped_device_noalloc if noalloc else lib().blasx_gemm_grouped_device
    rc = fn(arr, len(descs), stream)
    if rc:
        raise RuntimeError(f"grouped gemm{' (noalloc)' if noalloc else ''} failed ({rc})")


class Server:
    """GEMM request server (continuous batching); requests carry host arrays."""

    def __init__(self, hbm_bytes, max_batch=256):
        self.s = lib().blasx_server_create(hbm_bytes, max_batch)
        if not self.s:
            raise RuntimeError("blasx_server_create failed")
        self._keep = {}

    def register_operand(self, a):
        rc = lib().blasx_server_register_operand(self.s, a.ctypes.data, a.nbytes)
        if rc:
            raise RuntimeError(f"register_operand failed ({rc})")
        self._keep[id(a)] = a

    def submit(self, desc, keep=()):
        rid = lib().blasx_server_submit(self.s, C.byref(desc))
        if rid < 0:
            raise RuntimeError(f"submit failed ({rid})")
        self._keep[("r", rid)] = keep
        return rid

    def wait(self, rid):
        lat = C.c_double()
        rc = lib().blasx_server_wait(self.s, rid, C.byref(lat))
        self._keep.pop(("r", rid), None)
        if rc:
            raise RuntimeError(f"request {rid} failed ({rc})")
        return lat.value

    def poll(self, rid):
        return lib().blasx_server_poll(self.s, rid)

    def counters(self):
        b, dn, used = C.c_uint64(), C.c_uint64(), C.c_size_t()
        lib().blasx_server_counters(self.s, C.byref(b), C.byref(dn), C.byref(used))
        return {"batches": b.value, "done": dn.value, "arena_used": used.value}

    def close(self):
        if self.s:
            lib().blasx_server_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lib():
    """Load libblasx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                           "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    i, i64, d, f = C.c_int, C.c_int64, C.c_double, C.c_float
    L.cblas_sgemm.argtypes = [i, i, i, i, i, i, f, _vp, i, _vp, i, f, _vp, i]
    L.cblas_dgemm.argtypes = [i, i, i, i, i, i, d, _vp, i, _vp, i, d, _vp, i]
    L.cblas_zgemm.argtypes = [i, i, i, i, i, i, _vp, _vp, i, _vp, i, _vp, _vp, i]
    L.cblas_cgemm.argtypes = [i, i, i, i, i, i, _vp, _vp, i, _vp, i, _vp, _vp, i]
    for p in "sdcz":
        getattr(L, p + "gemm_").argtypes = [C.c_char_p, C.c_char_p] + [_vp] * 11
    L.xerbla_.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
    L.blasx_init.restype = i
    L.blasx_gpu_available.restype = i
    L.blasx_version.restype = C.c_char_p
    L.blasx_cpublas_path.restype = C.c_char_p
    L.blasx_get_stats.argtypes = [C.POINTER(Stats)]
    L.blasx_set_hbm_budget.argtypes = [C.c_size_t]
    L.blasx_get_hbm_budget.restype = C.c_size_t
    L.blasx_set_tile_size.argtypes = [i, i]
    L.blasx_get_tile_size.argtypes = [i, i, i, i]
    L.blasx_host_alloc.restype = _vp
    L.blasx_host_alloc.argtypes = [C.c_size_t]
    L.blasx_host_free.argtypes = [_vp]
    L.blasx_dev_alloc.restype = _vp
    L.blasx_dev_alloc.argtypes = [C.c_size_t]
    L.blasx_dev_free.argtypes = [_vp]
    L.blasx_memcpy_h2d.argtypes = [_vp, _vp, C.c_size_t]
    L.blasx_memcpy_d2h.argtypes = [_vp, _vp, C.c_size_t]
    L.blasx_memset_device.argtypes = [_vp, i, C.c_size_t]
    L.blasx_compute_stream.restype = _vp
    L.blasx_device_copy.argtypes = [_vp, _vp, C.c_size_t, _vp]
    L.blasx_event_create.restype = _vp
    L.blasx_event_destroy.argtypes = [_vp]
    L.blasx_event_record.argtypes = [_vp, _vp]
    L.blasx_event_synchronize.argtypes = [_vp]
    L.blasx_event_elapsed_ms.argtypes = [C.POINTER(C.c_float), _vp, _vp]
    L.blasx_dgemm_device.argtypes = [i, i, i, i, i, d, _vp, i64, _vp, i64, d, _vp, i64, _vp]
    L.blasx_sgemm_device.argtypes = [i, i, i, i, i, f, _vp, i64, _vp, i64, f, _vp, i64, _vp]
    L.blasx_zgemm_device.argtypes = [i, i, i, i, i, _vp, _vp, i64, _vp, i64, _vp, _vp, i64,
                                     _vp]
    L.blasx_dgemm_split_device.argtypes = [i, i, i, i, i, d, _vp, i64, _vp, i64, d, _vp, i64,
                                           i, _vp]
    L.blasx_dgemm_crt_device.argtypes = [i, i, i, i, i, d, _vp, i64, _vp, i64, d, _vp, i64, i,
                                         _vp]
    L.blasx_crt_bits.argtypes = [i, i]
    L.blasx_measure_peaks.argtypes = [C.POINTER(C.c_double)]
    L.blasx_zgemm_crt_device.argtypes = [i, i, i, i, i, _vp, _vp, i64, _vp, i64, _vp, _vp, i64,
                                         i, _vp]
    L.blasx_gemm_grouped_device.argtypes = [C.POINTER(GemmDesc), i, _vp]
    L.blasx_gemm_grouped_device_noalloc.argtypes = [C.POINTER(GemmDesc), i, _vp]
    L.blasx_set_grouped_shape.argtypes = [i, i]
    L.blasx_grouped_shape.argtypes = [i, i, i, i, i64]
    L.blasx_grouped_shape_add.argtypes = [i, i, i, i, i64, i]
    L.blasx_server_create.restype = _vp
    L.blasx_server_create.argtypes = [C.c_size_t, i]
    L.blasx_server_register_operand.argtypes = [_vp, _vp, C.c_size_t]
    L.blasx_server_submit.restype = i64
    L.blasx_server_submit.argtypes = [_vp, C.POINTER(GemmDesc)]
    L.blasx_server_poll.argtypes = [_vp, i64]
    L.blasx_server_wait.argtypes = [_vp, i64, C.POINTER(C.c_double)]
    L.blasx_server_counters.argtypes = [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_size_t)]
    L.blasx_server_destroy.argtypes = [_vp]
    L.blasx_zgemm_split_device.argtypes = [i, i, i, i, i, _vp, _vp, i64, _vp, i64, _vp, _vp,
                                           i64, i, _vp]
    L.blasx_resident_register.argtypes = [_vp, C.c_size_t]
    L.blasx_resident_unregister.argtypes = [_vp]
    L.blasx_fill_uniform_device.argtypes = [i, _vp, i64, i64, i64, C.c_uint64, C.c_uint32, _vp]
    L.blasx_fill_uniform_host.argtypes = [i, _vp, i64, i64, i64, C.c_uint64, C.c_uint32]
    L.blasx_tune_table_add.argtypes = [i, i, i, i, i]
    L.blasx_fill_dynrange_host.argtypes = [_vp, i64, i64, i64, C.c_uint64, C.c_uint32, d]
    L.blasx_dd_gemm_block_host.argtypes = [i, i, i, i, i, d, _vp, i64, _vp, i64, d, _vp, i64,
                                           _vp, i64]
    L.blasx_sample_uniform_host.argtypes = [i, _vp, i64, i, _vp, i, _vp, C.c_uint64, C.c_uint32]
    L.blasx_dgemm_bf16x3_device.argtypes = [i, i, i, i, i, d, _vp, i64, _vp, i64, d, _vp, i64,
                                            _vp]
    _lib = L
    return L


TEST_LIB_PATH = os.path.join(_HERE, "lib", "libblasx_test.so")
_test_lib = None


def test_lib():
    """Load libblasx_test.so: the same library plus the CPU-side hooks
    (tile cache over host memory, task queue, tile heuristic, CRT helpers)
    that only the test-suite calls.  It has its own library state."""
    global _test_lib
    if _test_lib is not None:
        return _test_lib
    if not os.path.exists(TEST_LIB_PATH):
        raise RuntimeError(f"{TEST_LIB_PATH} is missing: build it with "
                           "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(TEST_LIB_PATH)
    i, i64 = C.c_int, C.c_int64
    L.bxtest_cache_create.restype = _vp
    L.bxtest_cache_create.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t]
    L.bxtest_cache_destroy.argtypes = [_vp]
    L.bxtest_cache_acquire.argtypes = [_vp, C.c_uint64, i64, i, i, i, i64, C.POINTER(i),
                                       C.POINTER(i64)]
    L.bxtest_cache_lookup.argtypes = [_vp, C.c_uint64, i64, i, i, i]
    L.bxtest_cache_capacity.argtypes = [_vp]
    L.bxtest_cache_live.argtypes = [_vp]
    L.bxtest_cache_take_scratch.argtypes = [_vp, i64]
    L.bxtest_cache_return_scratch.argtypes = [_vp, i]
    L.bxtest_cache_invalidate_all.argtypes = [_vp]
    L.bxtest_cache_counters.argtypes = [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64)]
    L.bxtest_cache_slot_ptr.argtypes = [_vp, i]
    L.bxtest_cache_slot_ptr.restype = _vp
    L.bxtest_queue_create.argtypes = [i64]
    L.bxtest_queue_create.restype = _vp
    L.bxtest_queue_next.argtypes = [_vp]
    L.bxtest_queue_next.restype = i64
    L.bxtest_queue_destroy.argtypes = [_vp]
    L.bxtest_autotune_tile.argtypes = [i, i, i, i]
    L.blasx_tune_table_add.argtypes = [i, i, i, i, i]
    _test_lib = L
    return L


def gpu_available():
    return bool(lib().blasx_gpu_available())


def dtype_code(dt):
    return _DTYPES[np.dtype(dt)]


def _scalar(code, v):
    if code == S:
        return C.c_float(v)
    if code == D:
        return C.c_double(v)
    if code == Z:
        return (C.c_double * 2)(complex(v).real, complex(v).imag)
    return (C.c_float * 2)(complex(v).real, complex(v).imag)


def _addr(a):
    """The address a BLAS entry point receives: a numpy buffer (host), a
    DeviceArray (HBM) or a raw integer address."""
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if isinstance(a, DeviceArray):
        return a.ptr
    return a


def cblas_gemm(order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C_, ldc):
    """Call cblas_?gemm on numpy buffers (dtype taken from C_)."""
    code = dtype_code(C_.dtype)
    L = lib()
    fn = getattr(L, "cblas_" + _PREFIX[code] + "gemm")
    al, be = _scalar(code, alpha), _scalar(code, beta)
    if code in (Z, CPLX64):
        al, be = C.cast(al, _vp), C.cast(be, _vp)
    o = order if isinstance(order, int) else (COL_MAJOR if order == "F" else ROW_MAJOR)
    t_a = ta if isinstance(ta, int) else TRANS[ta]
    t_b = tb if isinstance(tb, int) else TRANS[tb]
    fn(o, t_a, t_b, M, N, K, al, _addr(A), lda, _addr(B), ldb, be, _addr(C_), ldc)
    return C_


def f77_gemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C_, ldc):
    """Call {s,d,c,z}gemm_ with Fortran by-reference arguments."""
    code = dtype_code(C_.dtype)
    fn = getattr(lib(), _PREFIX[code] + "gemm_")
    ints = [C.c_int(v) for v in (M, N, K, lda, ldb, ldc)]
    al, be = _scalar(code, alpha), _scalar(code, beta)
    fn(ta.encode(), tb.encode(), C.byref(ints[0]), C.byref(ints[1]), C.byref(ints[2]),
       C.byref(al), _vp(_addr(A)), C.byref(ints[3]), _vp(_addr(B)), C.byref(ints[4]),
       C.byref(be), _vp(_addr(C_)), C.byref(ints[5]))
    return C_


def gemm(ta, tb, alpha, A, B, beta, C_):
    """Column-major convenience wrapper: C_ <- alpha op(A) op(B) + beta C_ (in place)."""
    M, N = C_.shape
    K = A.shape[1] if ta == "N" else A.shape[0]
    for a in (A, B, C_):
        assert a.flags["F_CONTIGUOUS"], "column-major (Fortran-ordered) arrays expected"
    return cblas_gemm(COL_MAJOR, ta, tb, M, N, K, alpha, A, max(A.shape[0], 1), B,
                      max(B.shape[0], 1), beta, C_, max(M, 1))


def stats():
    s = Stats()
    lib().blasx_get_stats(C.byref(s))
    return {k: getattr(s, k) for k, _ in Stats._fields_}


def reset_stats():
    lib().blasx_reset_stats()


# ----------------------------------------------------------------- device memory
class DeviceArray:
    """A column-major matrix in HBM owned by libblasx (blasx_dev_alloc)."""

    def __init__(self, rows, cols, dtype, ld=None):
        self.rows, self.cols = rows, cols
        self.ld = rows if ld is None else ld
        self.dtype = np.dtype(dtype)
        self.nbytes = self.ld * cols * self.dtype.itemsize
        self.ptr = lib().blasx_dev_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError(f"blasx_dev_alloc({self.nbytes}) failed")

    @classmethod
    def from_host(cls, a):
        a = np.asfortranarray(a)
        d = cls(a.shape[0], a.shape[1], a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.asfortranarray(a, dtype=self.dtype)
        assert a.shape == (self.ld, self.cols) or (self.ld == self.rows and a.shape == (self.rows, self.cols))
        rc = lib().blasx_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes)
        if rc:
            raise RuntimeError(f"blasx_memcpy_h2d failed ({rc})")

    def download(self):
        out = np.empty((self.ld, self.cols), dtype=self.dtype, order="F")
        rc = lib().blasx_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes)
        if rc:
            raise RuntimeError(f"blasx_memcpy_d2h failed ({rc})")
        return out[: self.rows]

    def fill_uniform(self, seed, mid, stream=None):
        rc = lib().blasx_fill_uniform_device(dtype_code(self.dtype), self.ptr, self.rows,
                                             self.cols, self.ld, seed, mid, stream)
        if rc:
            raise RuntimeError(f"blasx_fill_uniform_device failed ({rc})")

    def free(self):
        if self.ptr:
            lib().blasx_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_gemm(ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, dtype, stream=None):
    """blasx_?gemm_device on HBM pointers (asynchronous on `stream`)."""
    code = dtype_code(dtype)
    L = lib()
    t_a, t_b = TRANS[ta], TRANS[tb]
    if code == D:
        rc = L.blasx_dgemm_device(t_a, t_b, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, stream)
    elif code == S:
        rc = L.blasx_sgemm_device(t_a, t_b, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, stream)
    elif code == Z:
        al, be = _scalar(Z, alpha), _scalar(Z, beta)
        rc = L.blasx_zgemm_device(t_a, t_b, M, N, K, C.cast(al, _vp), dA, lda, dB, ldb,
                                  C.cast(be, _vp), dC, ldc, stream)
    else:
        raise TypeError(dtype)
    if rc:
        raise RuntimeError(f"device gemm failed ({rc})")


class Event:
    """A HIP event recorded on a library stream (NULL = compute stream)."""

    def __init__(self):
        self.ev = lib().blasx_event_create()
        if not self.ev:
            raise RuntimeError("blasx_event_create failed")

    def record(self, stream=None):
        rc = lib().blasx_event_record(self.ev, stream)
        if rc:
            raise RuntimeError(f"blasx_event_record failed ({rc})")

    def synchronize(self):
        lib().blasx_event_synchronize(self.ev)

    def elapsed_ms(self, later):
        ms = C.c_float()
        rc = lib().blasx_event_elapsed_ms(C.byref(ms), self.ev, later.ev)
        if rc:
            raise RuntimeError(f"blasx_event_elapsed_ms failed ({rc})")
        return ms.value

    def __del__(self):
        try:
            lib().blasx_event_destroy(self.ev)
        except Exception:
            pass


def dgemm_split_device(ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, slices=0,
                       stream=None):
    """Split-precision (Ozaki, INT8 MFMA) DGEMM on HBM pointers."""
    rc = lib().blasx_dgemm_split_device(TRANS[ta], TRANS[tb], M, N, K, alpha, dA, lda, dB, ldb,
                                        beta, dC, ldc, slices, stream)
    if rc:
        raise RuntimeError(f"split dgemm failed ({rc})")


def dgemm_crt_device(ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, moduli=0,
                     stream=None):
    """CRT split (Ozaki scheme II, INT8 MFMA) DGEMM on HBM pointers."""
    rc = lib().blasx_dgemm_crt_device(TRANS[ta], TRANS[tb], M, N, K, alpha, dA, lda, dB, ldb,
                                      beta, dC, ldc, moduli, stream)
    if rc:
        raise RuntimeError(f"crt dgemm failed ({rc})")


def zgemm_crt_device(ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, moduli=0,
                     stream=None):
    """Complex128 CRT split GEMM (3 INT8 products per modulus) on HBM pointers."""
    al, be = _scalar(Z, alpha), _scalar(Z, beta)
    rc = lib().blasx_zgemm_crt_device(TRANS[ta], TRANS[tb], M, N, K, C.cast(al, _vp), dA, lda,
                                      dB, ldb, C.cast(be, _vp), dC, ldc, moduli, stream)
    if rc:
        raise RuntimeError(f"crt zgemm failed ({rc})")


def crt_bits(K, moduli=0):
    """Bits kept per operand entry by the CRT split at this K."""
    return lib().blasx_crt_bits(K, moduli)


def zgemm_split_device(ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, slices=0,
                       stream=None):
    """Split-precision complex128 GEMM on HBM pointers."""
    al, be = _scalar(Z, alpha), _scalar(Z, beta)
    rc = lib().blasx_zgemm_split_device(TRANS[ta], TRANS[tb], M, N, K, C.cast(al, _vp), dA, lda,
                                        dB, ldb, C.cast(be, _vp), dC, ldc, slices, stream)
    if rc:
        raise RuntimeError(f"split zgemm failed ({rc})")


def synchronize():
    rc = lib().blasx_device_synchronize()
    if rc:
        raise RuntimeError(f"blasx_device_synchronize failed ({rc})")


class PinnedArray:
    """Column-major host matrix in pinned memory (blasx_host_alloc)."""

    def __init__(self, rows, cols, dtype):
        self.dtype = np.dtype(dtype)
        nbytes = rows * cols * self.dtype.itemsize
        self.ptr = lib().blasx_host_alloc(max(nbytes, 1))
        if not self.ptr:
            raise MemoryError(f"blasx_host_alloc({nbytes}) failed")
        buf = (C.c_char * max(nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=rows * cols).reshape(
            (rows, cols), order="F")

    def free(self):
        if self.ptr:
            self.array = None
            lib().blasx_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def fill_dynrange_host(a, seed, mid, phi):
    """(u - 0.5) * exp(phi * g) inputs of the split-precision workload."""
    assert a.dtype == np.float64 and a.flags["F_CONTIGUOUS"]
    rows, cols = a.shape
    rc = lib().blasx_fill_dynrange_host(a.ctypes.data, rows, cols, rows, seed, mid, phi)
    if rc:
        raise RuntimeError(f"blasx_fill_dynrange_host failed ({rc})")
    return a


def dd_gemm_block(ta, tb, alpha, A, B, beta=0.0, C0=None):
    """Double-double reference of alpha op(A) op(B) + beta C0 (host arrays)."""
    M = A.shape[0] if ta == "N" else A.shape[1]
    K = A.shape[1] if ta == "N" else A.shape[0]
    N = B.shape[1] if tb == "N" else B.shape[0]
    out = np.zeros((M, N), order="F")
    c0 = 0 if C0 is None else C0.ctypes.data
    rc = lib().blasx_dd_gemm_block_host(TRANS[ta], TRANS[tb], M, N, K, alpha, A.ctypes.data,
                                        A.shape[0], B.ctypes.data, B.shape[0], beta, c0,
                                        max(M, 1), out.ctypes.data, max(M, 1))
    if rc:
        raise RuntimeError(f"blasx_dd_gemm_block_host failed ({rc})")
    return out


def fill_uniform_host(a, seed, mid):
    """Seeded generator on the host, same values as the device fill."""
    rows, cols = a.shape
    rc = lib().blasx_fill_uniform_host(dtype_code(a.dtype), a.ctypes.data, rows, cols, rows,
                                       seed, mid)
    if rc:
        raise RuntimeError(f"blasx_fill_uniform_host failed ({rc})")
    return a


def sample_uniform_host(rows, ri, ci, seed, mid, dtype=np.float64):
    """Entries (ri[a], ci[b]) of the seeded matrix with `rows` rows, as a
    (len(ri), len(ci)) Fortran array: what the fills hold there."""
    ri = np.ascontiguousarray(ri, dtype=np.int64)
    ci = np.ascontiguousarray(ci, dtype=np.int64)
    out = np.zeros((len(ri), len(ci)), dtype=dtype, order="F")
    rc = lib().blasx_sample_uniform_host(dtype_code(dtype), out.ctypes.data, rows, len(ri),
                                         ri.ctypes.data, len(ci), ci.ctypes.data, seed, mid)
    if rc:
        raise RuntimeError(f"blasx_sample_uniform_host failed ({rc})")
    return out
