#!/usr/bin/env python3
"""Generate the CPU-BLAS forwarders of libblasx: every CBLAS and Fortran-77
level-1/2/3 routine other than the GEMMs and the GPU level-3 routines (ON_GPU
below), each forwarded to the dlopen'ed CPU BLAS through the library handle,
as the reference forwards them (blas/blasx_globalpointers.c, one wrapper per
routine in blas/*.c, e.g. blas/dgemv.c:9-36).

The CPU BLAS may be LP64 (libopenblas.so: cblas_dgemv, dgemv_) or the ILP64
build numpy bundles (scipy_cblas_dgemv64_, scipy_dgemv_64_); the forwarders
widen every integer argument to int64_t for the latter.  Argument checking,
xerbla numbering and RowMajor handling are the CPU BLAS's own (netlib
semantics).  Writes csrc/forward.c, csrc/cpu_forward.h (the CPU routines
behind the GPU entry points) and ../include/blasx_blas.h.
Run from anywhere: python3 gen_forward.py [outdir-for-tests]
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

# ---------------------------------------------------------------------------
# Signatures.  Argument tokens:
#   ORDER TRANS UPLO DIAG SIDE  CBLAS enums (f77: const char *)
#   INT      int (ILP64 target: int64_t)
#   SCAL     scalar of the element type; by value for s/d, const void * for c/z
#   RSCAL    real scalar of the precision (float for s/c, double for d/z)
#   CARR ARR const / mutable element array (void * for complex)
#   CRARR RARR  const / mutable real array
# Return tokens: VOID, REAL (real type of the precision), DOUBLE, INDEX.
# f77 modes: "call" forwards to the CPU BLAS's f77 symbol; "value" returns the
# CBLAS result; "index" is 1-based i?amax; "cdot" returns a complex value;
# None: no f77 routine.
# ---------------------------------------------------------------------------
L1 = [
    # names per precision, return, args, f77 mode
    ({"s": "sdot", "d": "ddot"}, "REAL", "INT N, CARR X, INT incX, CARR Y, INT incY", "value"),
    ({"s": "dsdot"}, "DOUBLE", "INT N, CARR X, INT incX, CARR Y, INT incY", "value"),
    ({"s": "sdsdot"}, "REAL", "INT N, SCAL alpha, CARR X, INT incX, CARR Y, INT incY", "value"),
    ({"c": "cdotu_sub", "z": "zdotu_sub"}, "VOID",
     "INT N, CARR X, INT incX, CARR Y, INT incY, ARR dotu", "cdot"),
    ({"c": "cdotc_sub", "z": "zdotc_sub"}, "VOID",
     "INT N, CARR X, INT incX, CARR Y, INT incY, ARR dotc", "cdot"),
    ({"s": "snrm2", "d": "dnrm2", "c": "scnrm2", "z": "dznrm2"}, "REAL", "INT N, CARR X, INT incX",
     "value"),
    ({"s": "sasum", "d": "dasum", "c": "scasum", "z": "dzasum"}, "REAL", "INT N, CARR X, INT incX",
     "value"),
    ({p: "i%samax" % p for p in "sdcz"}, "INDEX", "INT N, CARR X, INT incX", "index"),
    ({p: p + "swap" for p in "sdcz"}, "VOID", "INT N, ARR X, INT incX, ARR Y, INT incY", "call"),
    ({p: p + "copy" for p in "sdcz"}, "VOID", "INT N, CARR X, INT incX, ARR Y, INT incY", "call"),
    ({p: p + "axpy" for p in "sdcz"}, "VOID", "INT N, SCAL alpha, CARR X, INT incX, ARR Y, INT incY",
     "call"),
    ({"s": "srotg", "d": "drotg"}, "VOID", "RARR a, RARR b, RARR c, RARR s", "call"),
    ({"s": "srotmg", "d": "drotmg"}, "VOID", "RARR d1, RARR d2, RARR b1, RSCAL b2, RARR P", "call"),
    ({"s": "srot", "d": "drot", "c": "csrot", "z": "zdrot"}, "VOID",
     "INT N, ARR X, INT incX, ARR Y, INT incY, RSCAL c, RSCAL s", "call"),
    ({"s": "srotm", "d": "drotm"}, "VOID", "INT N, ARR X, INT incX, ARR Y, INT incY, CRARR P", "call"),
    ({p: p + "scal" for p in "sdcz"}, "VOID", "INT N, SCAL alpha, ARR X, INT incX", "call"),
    ({"c": "csscal", "z": "zdscal"}, "VOID", "INT N, RSCAL alpha, ARR X, INT incX", "call"),
]

_TRI = "ORDER Order, UPLO Uplo, TRANS TransA, DIAG Diag, INT N"
L2 = [
    ("sdcz", "gemv", "ORDER Order, TRANS TransA, INT M, INT N, SCAL alpha, CARR A, INT lda, "
     "CARR X, INT incX, SCAL beta, ARR Y, INT incY"),
    ("sdcz", "gbmv", "ORDER Order, TRANS TransA, INT M, INT N, INT KL, INT KU, SCAL alpha, "
     "CARR A, INT lda, CARR X, INT incX, SCAL beta, ARR Y, INT incY"),
    ("sdcz", "trmv", _TRI + ", CARR A, INT lda, ARR X, INT incX"),
    ("sdcz", "tbmv", _TRI + ", INT K, CARR A, INT lda, ARR X, INT incX"),
    ("sdcz", "tpmv", _TRI + ", CARR Ap, ARR X, INT incX"),
    ("sdcz", "trsv", _TRI + ", CARR A, INT lda, ARR X, INT incX"),
    ("sdcz", "tbsv", _TRI + ", INT K, CARR A, INT lda, ARR X, INT incX"),
    ("sdcz", "tpsv", _TRI + ", CARR Ap, ARR X, INT incX"),
    ("sd", "symv", "ORDER Order, UPLO Uplo, INT N, SCAL alpha, CARR A, INT lda, CARR X, INT incX, "
     "SCAL beta, ARR Y, INT incY"),
    ("sd", "sbmv", "ORDER Order, UPLO Uplo, INT N, INT K, SCAL alpha, CARR A, INT lda, CARR X, "
     "INT incX, SCAL beta, ARR Y, INT incY"),
    ("sd", "spmv", "ORDER Order, UPLO Uplo, INT N, SCAL alpha, CARR Ap, CARR X, INT incX, "
     "SCAL beta, ARR Y, INT incY"),
    ("sd", "ger", "ORDER Order, INT M, INT N, SCAL alpha, CARR X, INT incX, CARR Y, INT incY, "
     "ARR A, INT lda"),
    ("sd", "syr", "ORDER Order, UPLO Uplo, INT N, SCAL alpha, CARR X, INT incX, ARR A, INT lda"),
    ("sd", "spr", "ORDER Order, UPLO Uplo, INT N, SCAL alpha, CARR X, INT incX, ARR Ap"),
    ("sd", "syr2", "ORDER Order, UPLO Uplo, INT N, SCAL alpha, CARR X, INT incX, CARR Y, "
     "INT incY, ARR A, INT lda"),
    ("sd", "spr2", "ORDER Order, UPLO Uplo, INT N, SCAL alpha, CARR X, INT incX, CARR Y, "
     "INT incY, ARR A"),
    ("cz", "hemv", "ORDER Order, UPLO Uplo, INT N, SCAL alpha, CARR A, INT lda, CARR X, INT incX, "
     "SCAL beta, ARR Y, INT incY"),
    ("cz", "hbmv", "ORDER Order, UPLO Uplo, INT N, INT K, SCAL alpha, CARR A, INT lda, CARR X, "
     "INT incX, SCAL beta, ARR Y, INT incY"),
    ("cz", "hpmv", "ORDER Order, UPLO Uplo, INT N, SCAL alpha, CARR Ap, CARR X, INT incX, "
     "SCAL beta, ARR Y, INT incY"),
    ("cz", "geru", "ORDER Order, INT M, INT N, SCAL alpha, CARR X, INT incX, CARR Y, INT incY, "
     "ARR A, INT lda"),
    ("cz", "gerc", "ORDER Order, INT M, INT N, SCAL alpha, CARR X, INT incX, CARR Y, INT incY, "
     "ARR A, INT lda"),
    ("cz", "her", "ORDER Order, UPLO Uplo, INT N, RSCAL alpha, CARR X, INT incX, ARR A, INT lda"),
    ("cz", "hpr", "ORDER Order, UPLO Uplo, INT N, RSCAL alpha, CARR X, INT incX, ARR A"),
    ("cz", "her2", "ORDER Order, UPLO Uplo, INT N, SCAL alpha, CARR X, INT incX, CARR Y, "
     "INT incY, ARR A, INT lda"),
    ("cz", "hpr2", "ORDER Order, UPLO Uplo, INT N, SCAL alpha, CARR X, INT incX, CARR Y, "
     "INT incY, ARR Ap"),
]

L3 = [
    ("sdcz", "symm", "ORDER Order, SIDE Side, UPLO Uplo, INT M, INT N, SCAL alpha, CARR A, "
     "INT lda, CARR B, INT ldb, SCAL beta, ARR C, INT ldc"),
    ("sdcz", "syrk", "ORDER Order, UPLO Uplo, TRANS Trans, INT N, INT K, SCAL alpha, CARR A, "
     "INT lda, SCAL beta, ARR C, INT ldc"),
    ("sdcz", "syr2k", "ORDER Order, UPLO Uplo, TRANS Trans, INT N, INT K, SCAL alpha, CARR A, "
     "INT lda, CARR B, INT ldb, SCAL beta, ARR C, INT ldc"),
    ("sdcz", "trmm", "ORDER Order, SIDE Side, UPLO Uplo, TRANS TransA, DIAG Diag, INT M, INT N, "
     "SCAL alpha, CARR A, INT lda, ARR B, INT ldb"),
    ("sdcz", "trsm", "ORDER Order, SIDE Side, UPLO Uplo, TRANS TransA, DIAG Diag, INT M, INT N, "
     "SCAL alpha, CARR A, INT lda, ARR B, INT ldb"),
    ("cz", "hemm", "ORDER Order, SIDE Side, UPLO Uplo, INT M, INT N, SCAL alpha, CARR A, "
     "INT lda, CARR B, INT ldb, SCAL beta, ARR C, INT ldc"),
    ("cz", "herk", "ORDER Order, UPLO Uplo, TRANS Trans, INT N, INT K, RSCAL alpha, CARR A, "
     "INT lda, RSCAL beta, ARR C, INT ldc"),
    ("cz", "her2k", "ORDER Order, UPLO Uplo, TRANS Trans, INT N, INT K, SCAL alpha, CARR A, "
     "INT lda, CARR B, INT ldb, RSCAL beta, ARR C, INT ldc"),
]

# complex Givens rotation: f77 only (netlib CROTG / ZROTG; no CBLAS form)
F77_ONLY = [("crotg", "c", "ARR ca, CARR cb, RARR c, ARR s"),
            ("zrotg", "z", "ARR ca, CARR cb, RARR c, ARR s")]

# Where the reference defines each routine (its wrappers forward to the CPU
# BLAS the same way); routines absent there are netlib BLAS names it lacks.
CITES = {
    "caxpy_": "blas/axpy.c:36", "cblas_caxpy": "blas/axpy.c:28", "cblas_ccopy": "blas/copy.c:30",
    "cblas_cdotc_sub": "blas/dot.c:74", "cblas_cdotu_sub": "blas/dot.c:91",
    "cblas_cgbmv": "blas/cgbmv.c:8", "cblas_cgemm": "blas/cgemm.c:16",
    "cblas_cgemv": "blas/cgemv.c:9", "cblas_cgerc": "blas/cgerc.c:8",
    "cblas_cgeru": "blas/cgeru.c:8", "cblas_chbmv": "blas/chbmv.c:8",
    "cblas_chemm": "blas/chemm.c:16", "cblas_chemv": "blas/chemv.c:8",
    "cblas_cher": "blas/cher.c:8", "cblas_cher2": "blas/cher2.c:8",
    "cblas_cher2k": "blas/cher2k.c:15", "cblas_cherk": "blas/cherk.c:15",
    "cblas_chpmv": "blas/chpmv.c:8", "cblas_chpr": "blas/chpr.c:8",
    "cblas_chpr2": "blas/chpr2.c:8", "cblas_cscal": "blas/scal.c:29",
    "cblas_csrot": "blas/rot.c:37", "cblas_csscal": "blas/scal.c:55",
    "cblas_cswap": "blas/swap.c:29", "cblas_csymm": "blas/csymm.c:16",
    "cblas_csyr2k": "blas/csyr2k.c:16", "cblas_csyrk": "blas/csyrk.c:15",
    "cblas_ctbmv": "blas/ctbmv.c:9", "cblas_ctbsv": "blas/ctbsv.c:9",
    "cblas_ctpmv": "blas/ctpmv.c:9", "cblas_ctpsv": "blas/ctpsv.c:9",
    "cblas_ctrmm": "blas/ctrmm.c:17", "cblas_ctrmv": "blas/ctrmv.c:9",
    "cblas_ctrsm": "blas/ctrsm.c:16", "cblas_ctrsv": "blas/ctrsv.c:9",
    "cblas_dasum": "blas/asum.c:18", "cblas_daxpy": "blas/axpy.c:16",
    "cblas_dcopy": "blas/copy.c:16", "cblas_ddot": "blas/dot.c:17",
    "cblas_dgbmv": "blas/dgbmv.c:8", "cblas_dgemm": "blas/dgemm.c:18",
    "cblas_dgemv": "blas/dgemv.c:9", "cblas_dger": "blas/dger.c:8",
    "cblas_dnrm2": "blas/nrm2.c:18", "cblas_drot": "blas/rot.c:22",
    "cblas_drotg": "blas/rot.c:115", "cblas_drotm": "blas/rot.c:147",
    "cblas_drotmg": "blas/rot.c:174", "cblas_dsbmv": "blas/dsbmv.c:9",
    "cblas_dscal": "blas/scal.c:16", "cblas_dsdot": "blas/dot.c:30",
    "cblas_dspmv": "blas/dspmv.c:9", "cblas_dspr": "blas/dspr.c:8",
    "cblas_dspr2": "blas/dspr2.c:8", "cblas_dswap": "blas/swap.c:16",
    "cblas_dsymm": "blas/dsymm.c:16", "cblas_dsymv": "blas/dsymv.c:8",
    "cblas_dsyr": "blas/dsyr.c:8", "cblas_dsyr2": "blas/dsyr2.c:8",
    "cblas_dsyr2k": "blas/dsyr2k.c:16", "cblas_dsyrk": "blas/dsyrk.c:15",
    "cblas_dtbmv": "blas/dtbmv.c:9", "cblas_dtbsv": "blas/dtbsv.c:9",
    "cblas_dtpmv": "blas/dtpmv.c:9", "cblas_dtpsv": "blas/dtpsv.c:9",
    "cblas_dtrmm": "blas/dtrmm.c:17", "cblas_dtrmv": "blas/dtrmv.c:9",
    "cblas_dtrsm": "blas/dtrsm.c:16", "cblas_dtrsv": "blas/dtrsv.c:9",
    "cblas_dzasum": "blas/asum.c:43", "cblas_dznrm2": "blas/nrm2.c:47",
    "cblas_icamax": "blas/amax.c:44", "cblas_idamax": "blas/amax.c:24",
    "cblas_isamax": "blas/amax.c:4", "cblas_izamax": "blas/amax.c:58",
    "cblas_sasum": "blas/asum.c:5", "cblas_saxpy": "blas/axpy.c:3",
    "cblas_scasum": "blas/asum.c:31", "cblas_scnrm2": "blas/nrm2.c:33",
    "cblas_scopy": "blas/copy.c:3", "cblas_sdot": "blas/dot.c:3", "cblas_sdsdot": "blas/dot.c:44",
    "cblas_sgbmv": "blas/sgbmv.c:8", "cblas_sgemm": "blas/sgemm.c:16",
    "cblas_sgemv": "blas/sgemv.c:9", "cblas_sger": "blas/sger.c:8",
    "cblas_snrm2": "blas/nrm2.c:3", "cblas_srot": "blas/rot.c:3", "cblas_srotg": "blas/rot.c:97",
    "cblas_srotm": "blas/rot.c:133", "cblas_srotmg": "blas/rot.c:161",
    "cblas_ssbmv": "blas/ssbmv.c:9", "cblas_sscal": "blas/scal.c:3",
    "cblas_sspmv": "blas/sspmv.c:9", "cblas_sspr": "blas/sspr.c:8",
    "cblas_sspr2": "blas/sspr2.c:8", "cblas_sswap": "blas/swap.c:3",
    "cblas_ssymm": "blas/ssymm.c:16", "cblas_ssymv": "blas/ssymv.c:8",
    "cblas_ssyr": "blas/ssyr.c:8", "cblas_ssyr2": "blas/ssyr2.c:8",
    "cblas_ssyr2k": "blas/ssyr2k.c:16", "cblas_ssyrk": "blas/ssyrk.c:15",
    "cblas_stbmv": "blas/stbmv.c:9", "cblas_stbsv": "blas/stbsv.c:9",
    "cblas_stpmv": "blas/stpmv.c:9", "cblas_stpsv": "blas/stpsv.c:9",
    "cblas_strmm": "blas/strmm.c:17", "cblas_strmv": "blas/strmv.c:9",
    "cblas_strsm": "blas/strsm.c:16", "cblas_strsv": "blas/strsv.c:9",
    "cblas_zaxpy": "blas/axpy.c:41", "cblas_zcopy": "blas/copy.c:43",
    "cblas_zdotc_sub": "blas/dot.c:57", "cblas_zdotu_sub": "blas/dot.c:106",
    "cblas_zdrot": "blas/rot.c:67", "cblas_zdscal": "blas/scal.c:68",
    "cblas_zgbmv": "blas/zgbmv.c:8", "cblas_zgemm": "blas/zgemm.c:16",
    "cblas_zgemv": "blas/zgemv.c:9", "cblas_zgerc": "blas/zgerc.c:8",
    "cblas_zgeru": "blas/zgeru.c:8", "cblas_zhbmv": "blas/zhbmv.c:8",
    "cblas_zhemm": "blas/zhemm.c:16", "cblas_zhemv": "blas/zhemv.c:8",
    "cblas_zher": "blas/zher.c:8", "cblas_zher2": "blas/zher2.c:8",
    "cblas_zher2k": "blas/zher2k.c:15", "cblas_zherk": "blas/zherk.c:15",
    "cblas_zhpmv": "blas/zhpmv.c:8", "cblas_zhpr": "blas/zhpr.c:8",
    "cblas_zhpr2": "blas/zhpr2.c:8", "cblas_zscal": "blas/scal.c:42",
    "cblas_zswap": "blas/swap.c:42", "cblas_zsymm": "blas/zsymm.c:16",
    "cblas_zsyr2k": "blas/zsyr2k.c:16", "cblas_zsyrk": "blas/zsyrk.c:15",
    "cblas_ztbmv": "blas/ztbmv.c:9", "cblas_ztbsv": "blas/ztbsv.c:9",
    "cblas_ztpmv": "blas/ztpmv.c:9", "cblas_ztpsv": "blas/ztpsv.c:9",
    "cblas_ztrmm": "blas/ztrmm.c:17", "cblas_ztrmv": "blas/ztrmv.c:9",
    "cblas_ztrsm": "blas/ztrsm.c:16", "cblas_ztrsv": "blas/ztrsv.c:9", "ccopy_": "blas/copy.c:37",
    "cdotc_": "blas/dot.c:83", "cdotu_": "blas/dot.c:99", "cgbmv_": "blas/cgbmv.c:39",
    "cgemm_": "blas/cgemm.c:119", "cgemv_": "blas/cgemv.c:38", "cgerc_": "blas/cgerc.c:33",
    "cgeru_": "blas/cgeru.c:33", "chbmv_": "blas/chbmv.c:36", "chemm_": "blas/chemm.c:80",
    "chemv_": "blas/chemv.c:33", "cher2_": "blas/cher2.c:30", "cher2k_": "blas/cher2k.c:89",
    "cher_": "blas/cher.c:31", "cherk_": "blas/cherk.c:82", "chpmv_": "blas/chpmv.c:33",
    "chpr2_": "blas/chpr2.c:31", "chpr_": "blas/chpr.c:30", "crotg_": "blas/rot.c:110",
    "cscal_": "blas/scal.c:36", "csrot_": "blas/rot.c:61", "csscal_": "blas/scal.c:62",
    "cswap_": "blas/swap.c:36", "csymm_": "blas/csymm.c:74", "csyr2k_": "blas/csyr2k.c:88",
    "csyrk_": "blas/csyrk.c:76", "ctbmv_": "blas/ctbmv.c:38", "ctbsv_": "blas/ctbsv.c:38",
    "ctpmv_": "blas/ctpmv.c:35", "ctpsv_": "blas/ctpsv.c:35", "ctrmm_": "blas/ctrmm.c:67",
    "ctrmv_": "blas/ctrmv.c:37", "ctrsm_": "blas/ctrsm.c:78", "ctrsv_": "blas/ctrsv.c:37",
    "dasum_": "blas/asum.c:25", "daxpy_": "blas/axpy.c:23", "dcabs1_": "blas/rot.c:12",
    "dcopy_": "blas/copy.c:24", "ddot_": "blas/dot.c:25", "dgbmv_": "blas/dgbmv.c:39",
    "dgemm_": "blas/dgemm.c:131", "dgemv_": "blas/dgemv.c:38", "dger_": "blas/dger.c:35",
    "dnrm2_": "blas/nrm2.c:26", "drot_": "blas/rot.c:31", "drotg_": "blas/rot.c:127",
    "drotm_": "blas/rot.c:155", "drotmg_": "blas/rot.c:181", "dsbmv_": "blas/dsbmv.c:36",
    "dscal_": "blas/scal.c:23", "dsdot_": "blas/dot.c:38", "dspmv_": "blas/dspmv.c:33",
    "dspr2_": "blas/dspr2.c:31", "dspr_": "blas/dspr.c:30", "dswap_": "blas/swap.c:23",
    "dsymm_": "blas/dsymm.c:74", "dsymv_": "blas/dsymv.c:33", "dsyr2_": "blas/dsyr2.c:33",
    "dsyr2k_": "blas/dsyr2k.c:88", "dsyr_": "blas/dsyr.c:31", "dsyrk_": "blas/dsyrk.c:76",
    "dtbmv_": "blas/dtbmv.c:38", "dtbsv_": "blas/dtbsv.c:38", "dtpmv_": "blas/dtpmv.c:35",
    "dtpsv_": "blas/dtpsv.c:35", "dtrmm_": "blas/dtrmm.c:67", "dtrmv_": "blas/dtrmv.c:37",
    "dtrsm_": "blas/dtrsm.c:78", "dtrsv_": "blas/dtrsv.c:37", "dzasum_": "blas/asum.c:50",
    "dznrm2_": "blas/nrm2.c:55", "icamax_": "blas/amax.c:51", "idamax_": "blas/amax.c:37",
    "isamax_": "blas/amax.c:17", "izamax_": "blas/amax.c:65", "lsame_": "blas/lsame.c:6",
    "sasum_": "blas/asum.c:12", "saxpy_": "blas/axpy.c:11", "scasum_": "blas/asum.c:38",
    "scnrm2_": "blas/nrm2.c:41", "scopy_": "blas/copy.c:10", "sdot_": "blas/dot.c:11",
    "sdsdot_": "blas/dot.c:52", "sgbmv_": "blas/sgbmv.c:39", "sgemm_": "blas/sgemm.c:130",
    "sgemv_": "blas/sgemv.c:38", "sger_": "blas/sger.c:34", "snrm2_": "blas/nrm2.c:11",
    "srot_": "blas/rot.c:16", "srotg_": "blas/rot.c:105", "srotm_": "blas/rot.c:141",
    "srotmg_": "blas/rot.c:168", "ssbmv_": "blas/ssbmv.c:36", "sscal_": "blas/scal.c:10",
    "sspmv_": "blas/sspmv.c:33", "sspr2_": "blas/sspr2.c:31", "sspr_": "blas/sspr.c:30",
    "sswap_": "blas/swap.c:10", "ssymm_": "blas/ssymm.c:74", "ssymv_": "blas/ssymv.c:33",
    "ssyr2_": "blas/ssyr2.c:33", "ssyr2k_": "blas/ssyr2k.c:88", "ssyr_": "blas/ssyr.c:31",
    "ssyrk_": "blas/ssyrk.c:76", "stbmv_": "blas/stbmv.c:38", "stbsv_": "blas/stbsv.c:38",
    "stpmv_": "blas/stpmv.c:35", "stpsv_": "blas/stpsv.c:35", "strmm_": "blas/strmm.c:67",
    "strmv_": "blas/strmv.c:37", "strsm_": "blas/strsm.c:78", "strsv_": "blas/strsv.c:37",
    "xerbla_": "blas/error.c:5", "zaxpy_": "blas/axpy.c:49", "zcopy_": "blas/copy.c:50",
    "zdotc_": "blas/dot.c:66", "zdotu_": "blas/dot.c:115", "zdrot_": "blas/rot.c:89",
    "zdscal_": "blas/scal.c:75", "zgbmv_": "blas/zgbmv.c:39", "zgemm_": "blas/zgemm.c:133",
    "zgemv_": "blas/zgemv.c:38", "zgerc_": "blas/zgerc.c:33", "zgeru_": "blas/zgeru.c:33",
    "zhbmv_": "blas/zhbmv.c:36", "zhemm_": "blas/zhemm.c:80", "zhemv_": "blas/zhemv.c:33",
    "zher2_": "blas/zher2.c:30", "zher2k_": "blas/zher2k.c:89", "zher_": "blas/zher.c:31",
    "zherk_": "blas/zherk.c:82", "zhpmv_": "blas/zhpmv.c:33", "zhpr2_": "blas/zhpr2.c:31",
    "zhpr_": "blas/zhpr.c:30", "zrotg_": "blas/rot.c:122", "zscal_": "blas/scal.c:49",
    "zswap_": "blas/swap.c:49", "zsymm_": "blas/zsymm.c:74", "zsyr2k_": "blas/zsyr2k.c:88",
    "zsyrk_": "blas/zsyrk.c:76", "ztbmv_": "blas/ztbmv.c:38", "ztbsv_": "blas/ztbsv.c:38",
    "ztpmv_": "blas/ztpmv.c:35", "ztpsv_": "blas/ztpsv.c:35", "ztrmm_": "blas/ztrmm.c:67",
    "ztrmv_": "blas/ztrmv.c:37", "ztrsm_": "blas/ztrsm.c:78", "ztrsv_": "blas/ztrsv.c:37",
}

ELEM = {"s": "float", "d": "double", "c": "void", "z": "void"}
REAL = {"s": "float", "d": "double", "c": "float", "z": "double"}
ENUMS = {"ORDER": "CBLAS_ORDER", "TRANS": "CBLAS_TRANSPOSE", "UPLO": "CBLAS_UPLO",
         "DIAG": "CBLAS_DIAG", "SIDE": "CBLAS_SIDE"}


class Arg:
    def __init__(self, tok, name, p):
        self.tok, self.name, self.p = tok, name, p
        self.f77name = name.lower()

    @property
    def is_enum(self):
        return self.tok in ENUMS

    def cblas_type(self):
        t, p = self.tok, self.p
        if self.is_enum:
            return "const enum %s" % ENUMS[t]
        if t == "INT":
            return "const int"
        if t == "SCAL":
            return "const void *" if p in "cz" else "const %s" % ELEM[p]
        if t == "RSCAL":
            return "const %s" % REAL[p]
        if t == "CARR":
            return "const %s *" % ELEM[p]
        if t == "ARR":
            return "%s *" % ELEM[p]
        if t == "CRARR":
            return "const %s *" % REAL[p]
        if t == "RARR":
            return "%s *" % REAL[p]
        raise ValueError(t)

    def cblas_call_type(self, ilp):
        if self.is_enum:
            return "int"
        if self.tok == "INT":
            return "int64_t" if ilp else "int"
        return self.cblas_type()

    def f77_type(self):
        t, p = self.tok, self.p
        if self.is_enum:
            return "const char *"
        if t == "INT":
            return "const int *"
        if t == "SCAL":
            return "const %s *" % ELEM[p]
        if t == "RSCAL":
            return "const %s *" % REAL[p]
        return self.cblas_type()

    def f77_call_type(self, ilp):
        if self.tok == "INT":
            return "const int64_t *" if ilp else "const int *"
        return self.f77_type()


def parse(sig, p):
    out = []
    for part in sig.split(","):
        tok, name = part.split()
        out.append(Arg(tok, name, p))
    return out


def decl(typ, name):
    return typ + name if typ.endswith("*") else "%s %s" % (typ, name)


def ret_type(ret, p):
    return {"VOID": "void", "REAL": REAL[p], "DOUBLE": "double", "INDEX": "CBLAS_INDEX"}[ret]


def cite(name):
    c = CITES.get(name)
    return ("reference %s" % c) if c else "no reference counterpart (netlib BLAS name)"


class Routine:
    def __init__(self, base, p, ret, sig, f77):
        self.base, self.p, self.ret, self.f77 = base, p, ret, f77
        self.args = parse(sig, p)

    @property
    def cname(self):
        return "cblas_" + self.base

    @property
    def fname(self):
        base = self.base
        if self.f77 == "cdot":
            base = base[: -len("_sub")]
        return base + "_"

    def cblas_proto(self):
        params = ", ".join(decl(a.cblas_type(), a.name) for a in self.args)
        return "%s %s(%s)" % (ret_type(self.ret, self.p), self.cname, params)

    def f77_args(self):
        return [a for a in self.args if a.tok != "ORDER"]

    def f77_proto(self):
        if self.f77 == "cdot":
            args = self.f77_args()[:-1]
            rt = "bx_f77_complex_float" if self.p == "c" else "bx_f77_complex_double"
        elif self.f77 == "index":
            args, rt = self.f77_args(), "int"
        else:
            args, rt = self.f77_args(), ret_type(self.ret, self.p)
            if self.f77 == "call":
                rt = "void"
        params = ", ".join(decl(a.f77_type(), a.f77name) for a in args)
        return "%s %s(%s)" % (rt, self.fname, params)


# run on the GPU by syrk.c and level3.c (declared in blasx.h): no public
# forwarder, but an internal bx_cpu_<name> for their CPU-sized calls
ON_GPU = {"ssyrk", "dsyrk", "zsyrk", "zherk", "ssymm", "dsymm", "zsymm", "zhemm", "strmm",
          "dtrmm", "ztrmm", "strsm", "dtrsm", "ztrsm", "ssyr2k", "dsyr2k", "zsyr2k", "zher2k",
          "csyrk", "cherk", "csymm", "chemm", "ctrmm", "ctrsm", "csyr2k", "cher2k"}


def routines():
    out = []
    for names, ret, sig, f77 in L1:
        for p, base in names.items():
            out.append(Routine(base, p, ret, sig, f77))
    for group in (L2, L3):
        for precs, stem, sig in group:
            for p in precs:
                out.append(Routine(p + stem, p, "VOID", sig, "call"))
    return out


def fn_type(ret, params):
    return "%s (*)(%s)" % (ret, ", ".join(params))


def emit_cblas(r, w):
    rt = ret_type(r.ret, r.p)
    names = ", ".join(a.name for a in r.args)
    internal = r.base in ON_GPU
    w("/* %s: %s */" % (r.cname, cite(r.cname)))
    params = ", ".join(decl(a.cblas_type(), a.name) for a in r.args)
    if internal:
        # the CPU BLAS's own routine, for the GPU entry point's CPU-sized calls
        w("%s bx_cpu_%s(%s)" % (rt, r.base, params))
    else:
        w("static %s fw_%s(%s)" % (rt, r.base, params))
    w("{")
    w("    static void *slot;")
    w("    void *f = fwd_sym(&slot, \"%s\");" % r.cname)
    lp = [a.cblas_call_type(False) for a in r.args]
    ilp = [a.cblas_call_type(True) for a in r.args]
    ret = "" if rt == "void" else "return "
    if lp != ilp:
        w("    if (bx_cpublas_ilp64())")
        w("        %s((%s)f)(%s);" % (ret, fn_type(rt, ilp), names))
        w("    else")
        w("        %s((%s)f)(%s);" % (ret, fn_type(rt, lp), names))
    else:
        w("    %s((%s)f)(%s);" % (ret, fn_type(rt, lp), names))
    w("}")
    if internal:
        w("")
        return
    w("BX_EXPORT %s" % r.cblas_proto())
    w("{")
    w("    %sfw_%s(%s);" % (ret, r.base, names))
    w("}")
    w("")


def emit_f77(r, w):
    args = r.f77_args()
    w("/* %s: %s */" % (r.fname, cite(r.fname)))
    w("BX_EXPORT %s" % r.f77_proto())
    w("{")
    if r.f77 == "value":
        vals = ", ".join(("*" + a.f77name) if a.tok in ("INT", "SCAL", "RSCAL") else a.f77name
                         for a in args)
        w("    return fw_%s(%s);" % (r.base, vals))
    elif r.f77 == "index":
        n, x, inc = (a.f77name for a in args)
        w("    if (*%s < 1 || *%s <= 0) return 0;" % (n, inc))
        w("    return (int)fw_%s(*%s, %s, *%s) + 1;" % (r.base, n, x, inc))
    elif r.f77 == "cdot":
        a = args[:-1]
        rt = "bx_f77_complex_float" if r.p == "c" else "bx_f77_complex_double"
        w("    %s res = 0;" % rt)
        vals = ", ".join(("*" + x.f77name) if x.tok == "INT" else x.f77name for x in a)
        w("    fw_%s(%s, &res);" % (r.base, vals))
        w("    return res;")
    else:
        emit_f77_call(r.fname, args, w)
    w("}")
    w("")


def emit_f77_call(fname, args, w):
    nchar = sum(a.is_enum for a in args)
    names_lp = [a.f77name for a in args] + ["1"] * nchar
    names_ilp = [("&%s64" % a.f77name) if a.tok == "INT" else a.f77name for a in args] + \
        ["1"] * nchar
    lp = [a.f77_call_type(False) for a in args] + ["size_t"] * nchar
    ilp = [a.f77_call_type(True) for a in args] + ["size_t"] * nchar
    w("    static void *slot;")
    w("    void *f = fwd_sym(&slot, \"%s\");" % fname)
    ints = [a.f77name for a in args if a.tok == "INT"]
    if ints:
        w("    if (bx_cpublas_ilp64()) {")
        w("        const int64_t %s;" % ", ".join("%s64 = *%s" % (n, n) for n in ints))
        w("        ((%s)f)(%s);" % (fn_type("void", ilp), ", ".join(names_ilp)))
        w("    } else")
        w("        ((%s)f)(%s);" % (fn_type("void", lp), ", ".join(names_lp)))
    else:
        w("    ((%s)f)(%s);" % (fn_type("void", lp), ", ".join(names_lp)))


def generate():
    rs = routines()
    c, h = [], []
    cw, hw = c.append, h.append
    cw("/* forward.c -- generated by gen_forward.py; do not edit.")
    cw(" *")
    cw(" * CPU-BLAS forwarders for every CBLAS and Fortran-77 level-1/2/3 routine")
    cw(" * except the GEMMs and the GPU level-3 routines (syrk.c, level3.c; for those")
    cw(" * only an internal bx_cpu_* form): each resolves its symbol in the dlopen'ed")
    cw(" * CPU BLAS once")
    cw(" * (through the library handle, never RTLD_NEXT) and calls it, widening")
    cw(" * integers to int64_t when that BLAS is ILP64.  The reference forwards the")
    cw(" * same routines the same way (blas/blasx_globalpointers.c, blas/dgemv.c:9-36).")
    cw(" * Argument errors are reported by the CPU BLAS's own checks (netlib numbering).")
    cw(" * Fortran character arguments are passed on with hidden lengths of 1. */")
    cw("#include <ctype.h>")
    cw("#include <math.h>")
    cw("#include <stddef.h>")
    cw("#include <stdint.h>")
    cw("")
    cw("#include \"blasx.h\"")
    cw("#include \"cpu_forward.h\"")
    cw("#include \"internal.h\"")
    cw("")
    cw("/* The routine's symbol in the CPU BLAS, looked up once.  A CPU BLAS that")
    cw(" * lacks it ends the process with a message: returning without computing")
    cw(" * would hand the caller a wrong result (a dot product of 0, an unchanged y). */")
    cw("static void *fwd_sym(void **slot, const char *name)")
    cw("{")
    cw("    void *f = __atomic_load_n(slot, __ATOMIC_ACQUIRE);")
    cw("    if (f) return f;")
    cw("    f = bx_cpublas_sym(name);")
    cw("    if (!f)")
    cw("        bx_fatal(\"%s: the CPU BLAS (%s) has no such routine; set BLASX_CPUBLAS to a \"")
    cw("                 \"complete BLAS\",")
    cw("                 name, blasx_cpublas_path());")
    cw("    __atomic_store_n(slot, f, __ATOMIC_RELEASE);")
    cw("    return f;")
    cw("}")
    cw("")
    for r in rs:
        emit_cblas(r, cw)
        if r.f77 and r.base not in ON_GPU:
            emit_f77(r, cw)
    for name, p, sig in F77_ONLY:
        args = parse(sig, p)
        w = cw
        params = ", ".join(decl(a.f77_type(), a.f77name) for a in args)
        w("/* %s_: %s */" % (name, cite(name + "_")))
        w("BX_EXPORT void %s_(%s)" % (name, params))
        w("{")
        emit_f77_call(name + "_", args, w)
        w("}")
        w("")
    cw("/* scabs1_, dcabs1_: |re| + |im| of one complex value (netlib SCABS1 and")
    cw(" * DCABS1; reference dcabs1_ at %s returns the real part only) */" %
       CITES.get("dcabs1_", "?"))
    cw("BX_EXPORT float scabs1_(const void *z)")
    cw("{")
    cw("    const float *v = (const float *)z;")
    cw("    return fabsf(v[0]) + fabsf(v[1]);")
    cw("}")
    cw("BX_EXPORT double dcabs1_(const void *z)")
    cw("{")
    cw("    const double *v = (const double *)z;")
    cw("    return fabs(v[0]) + fabs(v[1]);")
    cw("}")
    cw("")
    cw("/* lsame_: case-insensitive equality of two characters (netlib LSAME; the")
    cw(" * reference's lsame_ at %s returns the negation) */" % CITES.get("lsame_", "?"))
    cw("BX_EXPORT int lsame_(const char *ca, const char *cb)")
    cw("{")
    cw("    return toupper((unsigned char)*ca) == toupper((unsigned char)*cb);")
    cw("}")

    hw("/* blasx_blas.h -- generated by blasx_amd/csrc/gen_forward.py; do not edit.")
    hw(" *")
    hw(" * The CBLAS and Fortran-77 level-1/2/3 routines libblasx exports besides the")
    hw(" * GEMMs and the GPU level-3 routines (declared in blasx.h).  Each forwards to")
    hw(" * the CPU BLAS, as the reference's wrappers do; the comment before each")
    hw(" * names the reference routine it replaces.  Integers are LP64 int, matrices")
    hw(" * column-major unless an Order argument says otherwise, complex values")
    hw(" * interleaved (re, im). */")
    hw("#ifndef BLASX_BLAS_H")
    hw("#define BLASX_BLAS_H")
    hw("")
    hw("#include <stddef.h>")
    hw("")
    hw("#ifdef __cplusplus")
    hw("extern \"C\" {")
    hw("#endif")
    hw("")
    hw("#ifndef CBLAS_INDEX")
    hw("#define CBLAS_INDEX size_t")
    hw("#endif")
    hw("")
    hw("/* complex function results of the f77 routines, returned in registers as")
    hw(" * gfortran returns COMPLEX and COMPLEX*16 */")
    hw("#ifdef __cplusplus")
    hw("typedef struct { float re, im; } bx_f77_complex_float;")
    hw("typedef struct { double re, im; } bx_f77_complex_double;")
    hw("#else")
    hw("typedef float _Complex bx_f77_complex_float;")
    hw("typedef double _Complex bx_f77_complex_double;")
    hw("#endif")
    hw("")
    for r in rs:
        if r.base in ON_GPU:
            continue
        hw("/* %s */" % cite(r.cname))
        hw(r.cblas_proto() + ";")
        if r.f77:
            hw("/* %s */" % cite(r.fname))
            hw(r.f77_proto() + ";")
    for name, p, sig in F77_ONLY:
        args = parse(sig, p)
        hw("/* %s */" % cite(name + "_"))
        hw("void %s_(%s);" % (name, ", ".join(decl(a.f77_type(), a.f77name) for a in args)))
    hw("/* netlib SCABS1 (no reference counterpart) */")
    hw("float scabs1_(const void *z);")
    hw("/* %s (returns |re| + |im|, netlib DCABS1) */" % cite("dcabs1_"))
    hw("double dcabs1_(const void *z);")
    hw("/* %s (netlib LSAME semantics) */" % cite("lsame_"))
    hw("int lsame_(const char *ca, const char *cb);")
    hw("")
    hw("#ifdef __cplusplus")
    hw("}")
    hw("#endif")
    hw("")
    hw("#endif /* BLASX_BLAS_H */")
    g = ["/* cpu_forward.h -- generated by gen_forward.py; do not edit.",
         " *",
         " * The CPU BLAS's own routines behind the GPU entry points (syrk.c,",
         " * level3.c), for their CPU-sized calls: LP64 CBLAS signatures, forwarded",
         " * like every other routine in forward.c.  Internal to libblasx. */",
         "#ifndef BX_CPU_FORWARD_H",
         "#define BX_CPU_FORWARD_H",
         "",
         "#include \"blasx.h\"",
         ""]
    for r in rs:
        if r.base in ON_GPU:
            params = ", ".join(decl(a.cblas_type(), a.name) for a in r.args)
            g.append("%s bx_cpu_%s(%s);" % (ret_type(r.ret, r.p), r.base, params))
    g += ["", "#endif /* BX_CPU_FORWARD_H */"]
    return "\n".join(c) + "\n", "\n".join(h) + "\n", "\n".join(g) + "\n"


def main():
    csrc, hdr, internal = generate()
    if len(sys.argv) > 1:
        outc = os.path.join(sys.argv[1], "forward.c")
        outh = os.path.join(sys.argv[1], "blasx_blas.h")
        outi = os.path.join(sys.argv[1], "cpu_forward.h")
    else:
        outc = os.path.join(HERE, "forward.c")
        outh = os.path.join(HERE, "..", "..", "include", "blasx_blas.h")
        outi = os.path.join(HERE, "cpu_forward.h")
    for path, text in ((outc, csrc), (outh, hdr), (outi, internal)):
        with open(path, "w") as f:
            f.write(text)


if __name__ == "__main__":
    main()
