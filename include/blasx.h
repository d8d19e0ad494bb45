/*
 * blasx.h -- public C ABI of libblasx.so, the MI355X-native tiled GEMM
 * library.  Drop-in for the GEMM path of a system BLAS (LD_PRELOAD,
 * LD_LIBRARY_PATH substitution, MATLAB BLAS_VERSION).
 *
 * Every entry point below replaces the reference interface cited beside it.
 * Integers are LP64 `int` (the reference's blasint, include/f77blas_auxiliary.h:70-80).
 * Matrices are column-major; complex values are interleaved (re, im) doubles.
 */
#ifndef BLASX_H
#define BLASX_H

#ifdef __cplusplus
extern "C" {
#endif

/* CBLAS enums, same values as the reference's include/cblas.h:10-14 */
#ifndef CBLAS_H
enum CBLAS_ORDER { CblasRowMajor = 101, CblasColMajor = 102 };
enum CBLAS_TRANSPOSE { CblasNoTrans = 111, CblasTrans = 112, CblasConjTrans = 113 };
enum CBLAS_UPLO { CblasUpper = 121, CblasLower = 122 };
enum CBLAS_DIAG { CblasNonUnit = 131, CblasUnit = 132 };
enum CBLAS_SIDE { CblasLeft = 141, CblasRight = 142 };
#endif

/* ---- CBLAS entry points -------------------------------------------------
 * replaces cblas_sgemm, blas/sgemm.c:16-127 (prototype include/cblas.h:424-428) */
void cblas_sgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_TRANSPOSE TransB, const int M, const int N, const int K,
                 const float alpha, const float *A, const int lda, const float *B,
                 const int ldb, const float beta, float *C, const int ldc);

/* replaces cblas_dgemm, blas/dgemm.c:18-128 (prototype include/cblas.h:429-433) */
void cblas_dgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_TRANSPOSE TransB, const int M, const int N, const int K,
                 const double alpha, const double *A, const int lda, const double *B,
                 const int ldb, const double beta, double *C, const int ldc);

/* replaces cblas_cgemm, blas/cgemm.c:16-142 (the reference keeps it on the CPU;
 * here above 500 it runs as one real GEMM on the FP32 matrix cores) */
void cblas_cgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_TRANSPOSE TransB, const int M, const int N, const int K,
                 const void *alpha, const void *A, const int lda, const void *B,
                 const int ldb, const void *beta, void *C, const int ldc);

/* replaces cblas_zgemm, blas/zgemm.c:16-130 (prototype include/cblas.h:514-518) */
void cblas_zgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_TRANSPOSE TransB, const int M, const int N, const int K,
                 const void *alpha, const void *A, const int lda, const void *B,
                 const int ldb, const void *beta, void *C, const int ldc);

/* ---- Fortran-77 entry points (hidden string lengths ignored) --------------
 * replaces sgemm_, blas/sgemm.c:130-153 */
void sgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const float *alpha, const float *A, const int *lda, const float *B, const int *ldb,
            const float *beta, float *C, const int *ldc);
/* replaces dgemm_, blas/dgemm.c:131-154 */
void dgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const double *alpha, const double *A, const int *lda, const double *B,
            const int *ldb, const double *beta, double *C, const int *ldc);
/* replaces cgemm_, blas/cgemm.c:119 */
void cgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const void *alpha, const void *A, const int *lda, const void *B, const int *ldb,
            const void *beta, void *C, const int *ldc);
/* replaces zgemm_, blas/zgemm.c:133-156 */
void zgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const void *alpha, const void *A, const int *lda, const void *B, const int *ldb,
            const void *beta, void *C, const int *ldc);

/* ---- rank-k updates on the GPU (C = alpha op(A) op(A)^T + beta C on one
 * triangle; op(A)^H and real alpha, beta for HERK) ---------------------------
 * replaces cblas_ssyrk, blas/ssyrk.c:15 */
void cblas_ssyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const float alpha,
                 const float *A, const int lda, const float beta, float *C, const int ldc);
/* replaces cblas_dsyrk, blas/dsyrk.c:15 */
void cblas_dsyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const double alpha,
                 const double *A, const int lda, const double beta, double *C, const int ldc);
/* replaces cblas_zsyrk, blas/zsyrk.c:15 */
void cblas_zsyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                 const void *A, const int lda, const void *beta, void *C, const int ldc);
/* replaces cblas_zherk, blas/zherk.c:15 */
void cblas_zherk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const double alpha,
                 const void *A, const int lda, const double beta, void *C, const int ldc);
/* replaces ssyrk_, blas/ssyrk.c:76 */
void ssyrk_(const char *uplo, const char *trans, const int *n, const int *k, const float *alpha,
            const float *A, const int *lda, const float *beta, float *C, const int *ldc);
/* replaces dsyrk_, blas/dsyrk.c:76 */
void dsyrk_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha,
            const double *A, const int *lda, const double *beta, double *C, const int *ldc);
/* replaces zsyrk_, blas/zsyrk.c:76 */
void zsyrk_(const char *uplo, const char *trans, const int *n, const int *k, const void *alpha,
            const void *A, const int *lda, const void *beta, void *C, const int *ldc);
/* replaces zherk_, blas/zherk.c:82 */
void zherk_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha,
            const void *A, const int *lda, const double *beta, void *C, const int *ldc);

/* ---- symmetric / Hermitian / triangular products and rank-2k updates on
 * the GPU (one GEMM on an operand materialised in HBM; SYR2K as a two-term
 * rank-k launch) ------------------------------------------------------------
 * replace cblas_ssymm blas/ssymm.c:16, cblas_dsymm blas/dsymm.c:16,
 * cblas_zsymm blas/zsymm.c:16, cblas_zhemm blas/zhemm.c:16 */
void cblas_ssymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N, const float alpha,
                 const float *A, const int lda, const float *B, const int ldb, const float beta,
                 float *C, const int ldc);
void cblas_dsymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N, const double alpha,
                 const double *A, const int lda, const double *B, const int ldb,
                 const double beta, double *C, const int ldc);
void cblas_zsymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                 const void *A, const int lda, const void *B, const int ldb, const void *beta,
                 void *C, const int ldc);
void cblas_zhemm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                 const void *A, const int lda, const void *B, const int ldb, const void *beta,
                 void *C, const int ldc);
/* replace cblas_strmm blas/strmm.c:17, cblas_dtrmm blas/dtrmm.c:17,
 * cblas_ztrmm blas/ztrmm.c:17 */
void cblas_strmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const float alpha,
                 const float *A, const int lda, float *B, const int ldb);
void cblas_dtrmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const double alpha,
                 const double *A, const int lda, double *B, const int ldb);
void cblas_ztrmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                 const void *A, const int lda, void *B, const int ldb);
/* replace cblas_strsm blas/strsm.c:16, cblas_dtrsm blas/dtrsm.c:16,
 * cblas_ztrsm blas/ztrsm.c:16 */
void cblas_strsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const float alpha,
                 const float *A, const int lda, float *B, const int ldb);
void cblas_dtrsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const double alpha,
                 const double *A, const int lda, double *B, const int ldb);
void cblas_ztrsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                 const void *A, const int lda, void *B, const int ldb);
/* replace cblas_ssyr2k blas/ssyr2k.c:16, cblas_dsyr2k blas/dsyr2k.c:16,
 * cblas_zsyr2k blas/zsyr2k.c:16 */
void cblas_ssyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const float alpha,
                  const float *A, const int lda, const float *B, const int ldb, const float beta,
                  float *C, const int ldc);
void cblas_dsyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const double alpha,
                  const double *A, const int lda, const double *B, const int ldb,
                  const double beta, double *C, const int ldc);
void cblas_zsyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                  const void *A, const int lda, const void *B, const int ldb, const void *beta,
                  void *C, const int ldc);
/* replaces cblas_zher2k, blas/zher2k.c:15 */
void cblas_zher2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                  const void *A, const int lda, const void *B, const int ldb, const double beta,
                  void *C, const int ldc);
/* replaces zher2k_, blas/zher2k.c:89 */
void zher2k_(const char *uplo, const char *trans, const int *n, const int *k, const void *alpha,
             const void *A, const int *lda, const void *B, const int *ldb, const double *beta,
             void *C, const int *ldc);
/* replace ssymm_ blas/ssymm.c:74, dsymm_ blas/dsymm.c:74, zsymm_ blas/zsymm.c:74,
 * zhemm_ blas/zhemm.c:80 */
void ssymm_(const char *side, const char *uplo, const int *m, const int *n, const float *alpha,
            const float *A, const int *lda, const float *B, const int *ldb, const float *beta,
            float *C, const int *ldc);
void dsymm_(const char *side, const char *uplo, const int *m, const int *n, const double *alpha,
            const double *A, const int *lda, const double *B, const int *ldb, const double *beta,
            double *C, const int *ldc);
void zsymm_(const char *side, const char *uplo, const int *m, const int *n, const void *alpha,
            const void *A, const int *lda, const void *B, const int *ldb, const void *beta,
            void *C, const int *ldc);
void zhemm_(const char *side, const char *uplo, const int *m, const int *n, const void *alpha,
            const void *A, const int *lda, const void *B, const int *ldb, const void *beta,
            void *C, const int *ldc);
/* replace strmm_ blas/strmm.c:67, dtrmm_ blas/dtrmm.c:67, ztrmm_ blas/ztrmm.c:67 */
void strmm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const float *alpha, const float *A, const int *lda,
            float *B, const int *ldb);
void dtrmm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const double *alpha, const double *A, const int *lda,
            double *B, const int *ldb);
void ztrmm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const void *alpha, const void *A, const int *lda,
            void *B, const int *ldb);
/* replace strsm_ blas/strsm.c:78, dtrsm_ blas/dtrsm.c:78, ztrsm_ blas/ztrsm.c:78 */
void strsm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const float *alpha, const float *A, const int *lda,
            float *B, const int *ldb);
void dtrsm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const double *alpha, const double *A, const int *lda,
            double *B, const int *ldb);
void ztrsm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const void *alpha, const void *A, const int *lda,
            void *B, const int *ldb);
/* replace ssyr2k_ blas/ssyr2k.c:88, dsyr2k_ blas/dsyr2k.c:88, zsyr2k_ blas/zsyr2k.c:88 */
void ssyr2k_(const char *uplo, const char *trans, const int *n, const int *k, const float *alpha,
             const float *A, const int *lda, const float *B, const int *ldb, const float *beta,
             float *C, const int *ldc);
void dsyr2k_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha,
             const double *A, const int *lda, const double *B, const int *ldb,
             const double *beta, double *C, const int *ldc);
void zsyr2k_(const char *uplo, const char *trans, const int *n, const int *k, const void *alpha,
             const void *A, const int *lda, const void *B, const int *ldb, const void *beta,
             void *C, const int *ldc);

/* ---- complex64 forms of the above, on the FP32 matrix cores (one real GEMM
 * over packed operands; rank updates as a full product whose stored triangle
 * is copied back) ------------------------------------------------------------
 * replace cblas_csyrk blas/csyrk.c:15, cblas_cherk blas/cherk.c:15,
 * cblas_csymm blas/csymm.c:16, cblas_chemm blas/chemm.c:16,
 * cblas_ctrmm blas/ctrmm.c:17, cblas_ctrsm blas/ctrsm.c:16,
 * cblas_csyr2k blas/csyr2k.c:16, cblas_cher2k blas/cher2k.c:15 */
void cblas_csyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                 const void *A, const int lda, const void *beta, void *C, const int ldc);
void cblas_cherk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const float alpha,
                 const void *A, const int lda, const float beta, void *C, const int ldc);
void cblas_csymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                 const void *A, const int lda, const void *B, const int ldb, const void *beta,
                 void *C, const int ldc);
void cblas_chemm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                 const void *A, const int lda, const void *B, const int ldb, const void *beta,
                 void *C, const int ldc);
void cblas_ctrmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                 const void *A, const int lda, void *B, const int ldb);
void cblas_ctrsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                 const void *A, const int lda, void *B, const int ldb);
void cblas_csyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                  const void *A, const int lda, const void *B, const int ldb, const void *beta,
                  void *C, const int ldc);
void cblas_cher2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                  const void *A, const int lda, const void *B, const int ldb, const float beta,
                  void *C, const int ldc);
/* replace csyrk_ blas/csyrk.c:76, cherk_ blas/cherk.c:82, csymm_ blas/csymm.c:74,
 * chemm_ blas/chemm.c:80, ctrmm_ blas/ctrmm.c:67, ctrsm_ blas/ctrsm.c:78,
 * csyr2k_ blas/csyr2k.c:88, cher2k_ blas/cher2k.c:89 */
void csyrk_(const char *uplo, const char *trans, const int *n, const int *k, const void *alpha,
            const void *A, const int *lda, const void *beta, void *C, const int *ldc);
void cherk_(const char *uplo, const char *trans, const int *n, const int *k, const float *alpha,
            const void *A, const int *lda, const float *beta, void *C, const int *ldc);
void csymm_(const char *side, const char *uplo, const int *m, const int *n, const void *alpha,
            const void *A, const int *lda, const void *B, const int *ldb, const void *beta,
            void *C, const int *ldc);
void chemm_(const char *side, const char *uplo, const int *m, const int *n, const void *alpha,
            const void *A, const int *lda, const void *B, const int *ldb, const void *beta,
            void *C, const int *ldc);
void ctrmm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const void *alpha, const void *A, const int *lda,
            void *B, const int *ldb);
void ctrsm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const void *alpha, const void *A, const int *lda,
            void *B, const int *ldb);
void csyr2k_(const char *uplo, const char *trans, const int *n, const int *k, const void *alpha,
             const void *A, const int *lda, const void *B, const int *ldb, const void *beta,
             void *C, const int *ldc);
void cher2k_(const char *uplo, const char *trans, const int *n, const int *k, const void *alpha,
             const void *A, const int *lda, const void *B, const int *ldb, const float *beta,
             void *C, const int *ldc);

/* replaces xerbla_, blas/error.c:5-8.  Weak: a caller (LAPACK) may override it. */
void xerbla_(const char *srname, const int *info);

#ifdef __cplusplus
}
#endif

#include "blasx_blas.h"
#include "blasx_ext.h"

#endif /* BLASX_H */
