/*
 * hiplayer.h -- the thin C ABI between the C host runtime (abi.c, tile_rt.c,
 * ...) and the HIP side (hip/ kernels).  Plain C types only: the C files never
 * include HIP headers.  Streams and events travel as void*.
 *
 * Every bxh_* function returns 0 on success or a negative BLASX_E* code;
 * bxh_last_error() names the failing HIP call.
 */
#ifndef BLASX_HIPLAYER_H
#define BLASX_HIPLAYER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* one K-segment of a GEMM problem: op(A) columns / op(B) rows [0, k) of
 * the segment live at A (leading dim lda) and B (ldb), stored as the
 * problem's ops say.  A plain GEMM has one segment; the tile runtime gives
 * one segment per cached K-tile. */
typedef struct bxh_seg {
    const void *A;
    const void *B;
    int64_t lda;
    int64_t ldb;
    int64_t k;
} bxh_seg_t;

/* one GEMM problem of a grouped launch: C[m x n] = alpha * sum_s op(A_s) op(B_s) + beta * C */
typedef struct bxh_problem {
    void *C;
    int64_t ldc;
    int32_t m, n;
    int32_t nseg, seg_off;       /* segments [seg_off, seg_off + nseg) of the segment table */
    int32_t opa, opb;            /* 0 = N, 1 = T, 2 = C (conjugate transpose) */
    int32_t tiles_m, tiles_n;    /* block tiles of this problem */
    int32_t tile_start;          /* prefix sum of tiles_m*tiles_n over earlier problems */
    int32_t beta_zero;           /* C is not read when set */
    /* rank-k updates (SYRK/HERK): with BX_TRI_LOWER only C elements with
     * row + tri_off >= col are written, with BX_TRI_UPPER those with
     * row + tri_off <= col (row, col relative to C); BX_TRI_HERM also zeroes
     * the imaginary part where row + tri_off == col.  0: every element. */
    int32_t tri, tri_off;
    /* block tile shape (real types): 0 = 128 x 128, 1 = 128 x 64, 2 = 64 x 64
     * (BXH_SHAPE_*); complex problems use 0 (128 x 64 complex) */
    int32_t shape, pad0;
    double alpha_re, alpha_im;
    double beta_re, beta_im;
} bxh_problem_t;

#define BX_TRI_LOWER 1
#define BX_TRI_UPPER 2
#define BX_TRI_HERM 4

/* block-tile edges per element type (for computing tiles_m/tiles_n) */
int bxh_tile_m(int dtype);
int bxh_tile_n(int dtype);
#define BXH_SHAPE_128x128 0
#define BXH_SHAPE_128x64 1
#define BXH_SHAPE_64x64 2
#define BXH_NSHAPES 3
/* block-tile edges of a grouped problem of this element type and shape */
void bxh_tile_dims(int dtype, int shape, int *bm, int *bn);

/* ---- runtime ---- */
const char *bxh_last_error(void);
int bxh_device_count(int *n);
int bxh_set_device(int dev);
int bxh_device_info(int dev, char *name, size_t name_len, int *cus, size_t *total_mem);
int bxh_mem_get_info(size_t *free_b, size_t *total_b);
int bxh_malloc(void **p, size_t bytes);
int bxh_free(void *p);
int bxh_host_malloc(void **p, size_t bytes);
int bxh_host_free(void *p);
int bxh_pointer_is_pinned(const void *p);      /* 1 pinned host, 2 device, 0 pageable/unknown */
int bxh_stream_create(void **s);
int bxh_stream_destroy(void *s);
int bxh_stream_sync(void *s);
int bxh_device_sync(void);
int bxh_event_create(void **e, int timing);
int bxh_event_destroy(void *e);
int bxh_event_record(void *e, void *s);
int bxh_event_query(void *e);                  /* 1 done, 0 pending, <0 error */
int bxh_event_sync(void *e);
int bxh_event_elapsed_ms(float *ms, void *start, void *stop);
int bxh_stream_wait_event(void *s, void *e);
int bxh_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *s);
int bxh_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *s);
int bxh_memcpy_d2d_async(void *dst, const void *src, size_t bytes, void *s);
int bxh_memcpy_h2d(void *dst, const void *src, size_t bytes);
int bxh_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* strided 2-D copies: `height` columns of `width` bytes */
int bxh_memcpy2d_h2d_async(void *dst, size_t dpitch, const void *src, size_t spitch,
                           size_t width, size_t height, void *s);
int bxh_memcpy2d_d2h_async(void *dst, size_t dpitch, const void *src, size_t spitch,
                           size_t width, size_t height, void *s);
int bxh_memcpy2d_d2d_async(void *dst, size_t dpitch, const void *src, size_t spitch,
                           size_t width, size_t height, void *s);
int bxh_memset_async(void *dst, int v, size_t bytes, void *s);
/* HIP graphs: capture what is enqueued on s between begin and end */
int bxh_graph_begin(void *s);
int bxh_graph_end(void *s, void **exec);
int bxh_graph_launch(void *exec, void *s);
int bxh_graph_destroy(void *exec);

/* ---- kernels ---- */
/* C = alpha op(A) op(B) + beta C on device pointers; alpha/beta are (re, im) */
int bxh_gemm(int dtype, int opa, int opb, int m, int n, int k, const double *alpha,
             const void *A, int64_t lda, const void *B, int64_t ldb, const double *beta,
             void *C, int64_t ldc, void *stream);
/* grouped launch over a device-resident problem and segment table; all
 * problems share dtype.  grid = 0 picks one block per tile. */
int bxh_gemm_grouped(int dtype, const bxh_problem_t *probs_dev, int nprob,
                     const bxh_seg_t *segs_dev, int total_tiles, int grid, void *stream);
/* persistent grouped launch for graph replay: counts_dev[0] = nprob,
 * counts_dev[1] = total tiles, both read by the kernel */
int bxh_gemm_grouped_dyn(int dtype, const bxh_problem_t *probs_dev, const bxh_seg_t *segs_dev,
                         const int *counts_dev, int grid, void *stream);
/* C = beta * C (beta == 0 writes zeros without reading C) */
int bxh_scale(int dtype, int m, int n, const double *beta, void *C, int64_t ldc, void *stream);
/* the same on the triangle tri names (BX_TRI_* bits), square n x n C */
int bxh_scale_tri(int dtype, int m, int n, int tri, const double *beta, void *C, int64_t ldc,
                  void *stream);
int bxh_fill_uniform(int dtype, void *p, int64_t rows, int64_t cols, int64_t ld,
                     uint64_t seed, uint32_t id, void *stream);
/* materialise an n x n operand in place from its stored triangle (lower != 0:
 * the lower one): BX_FILL_SYM mirrors it, BX_FILL_HERM mirrors it conjugated
 * and zeroes the diagonal's imaginary part, BX_FILL_TRI zeroes the other
 * triangle, BX_FILL_UNIT also sets the diagonal to one */
#define BX_FILL_SYM 0
#define BX_FILL_HERM 1
#define BX_FILL_TRI 2
#define BX_FILL_UNIT 3
int bxh_tri_fill(int dtype, int n, int lower, int mode, void *A, int64_t lda, void *stream);
/* out (cols x rows) = in^T (in^H with conj, complex only) */
int bxh_transpose(int dtype, int rows, int cols, const void *in, int64_t ldi, void *out, int64_t ldo,
                  int conj, void *stream);
/* dst = src on the triangle tri names (BX_TRI_* bits), n x n; BX_TRI_HERM also
 * zeroes the diagonal's imaginary part */
int bxh_tri_copy(int dtype, int n, int tri, const void *src, int64_t lds, void *dst, int64_t ldd,
                 void *stream);
/* solve T X = X in place for an m x m triangular T (lower or upper) with
 * m <= bxh_trsm_leaf_nb(dtype) */
int bxh_trsm_leaf_nb(int dtype);
int bxh_trsm_leaf(int dtype, int lower, int m, int n, const void *T, int64_t ldt, void *X,
                  int64_t ldx, void *stream);
/* complex64 GEMM as one real GEMM: Ap = [Re op(A), Im op(A)] (M x 2K),
 * Bp = [[Re, Im], [-Im, Re]] of op(B) (2K x 2N), op 0 N / 1 T / 2 C; then
 * C = alpha (P[:, :N] + i P[:, N:]) + beta C (P NULL: C = beta C) */
int bxh_cpack_a(int M, int K, int op, const void *A, int64_t lda, float *Ap, void *stream);
int bxh_cpack_b(int K, int N, int op, const void *B, int64_t ldb, float *Bp, void *stream);
int bxh_ccombine(int M, int N, const float *P, int64_t ldp, const double *alpha,
                 const double *beta, void *C, int64_t ldc, void *stream);
/* streaming device copy kernel (16-byte aligned, multiple of 16 bytes) */
int bxh_copy_kernel(void *dst, const void *src, size_t bytes, void *stream);

/* ---- allocation-free grouped GEMM (hip/noalloc.hip) ----
 * The problem and segment tables of up to BXH_NOALLOC_MAX_PROBLEMS
 * problems per element type live in a static device array, one per GPU:
 * three slots (S, D, Z) of BXH_NOALLOC_SLOT_BYTES, problems first, then
 * segments. */
#define BXH_NOALLOC_MAX_PROBLEMS 1024
#define BXH_NOALLOC_SEG_OFFSET ((size_t)BXH_NOALLOC_MAX_PROBLEMS * sizeof(bxh_problem_t))
#define BXH_NOALLOC_SLOT_BYTES \
    ((BXH_NOALLOC_SEG_OFFSET + (size_t)BXH_NOALLOC_MAX_PROBLEMS * sizeof(bxh_seg_t) + 255) & \
     ~(size_t)255)
#define BXH_NOALLOC_SCRATCH_BYTES (3 * BXH_NOALLOC_SLOT_BYTES)
int bxh_noalloc_setup(void);
/* the current device's copy of the table store and the event recorded after
 * the last call's kernels (set up on first use) */
int bxh_noalloc_scratch(void **ptr, void **done_event);
const char *bxh_noalloc_error(void);

#ifdef __cplusplus
}
#endif

#endif
