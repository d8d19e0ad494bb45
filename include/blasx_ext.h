/*
 * blasx_ext.h -- extension API of libblasx.so beyond the BLAS ABI.
 *
 * The reference has no counterpart for most of these: it exposes only the
 * BLAS symbols and keeps device memory, streams and its tile cache private
 * and per call (blas/dgemm.c:104-114).  The functions here give callers
 * device-resident GEMMs, cross-call operand residency, pinned host memory,
 * statistics, split-precision GEMM, grouped GEMM and the GEMM server.
 *
 * All functions return 0 on success and a negative BLASX_E* code on error
 * unless stated otherwise.  Device pointers are HBM addresses on the current
 * GPU; `stream` arguments are hipStream_t values passed as void* (NULL = the
 * library's own compute stream).
 */
#ifndef BLASX_EXT_H
#define BLASX_EXT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types */
enum { BLASX_S = 0, BLASX_D = 1, BLASX_Z = 2, BLASX_C = 3 };

/* error codes */
enum {
    BLASX_OK = 0,
    BLASX_ENOGPU = -1,     /* no usable GPU */
    BLASX_EINVAL = -2,     /* invalid argument */
    BLASX_ENOMEM = -3,     /* device or pinned allocation failed */
    BLASX_EHIP = -4,       /* HIP runtime error */
    BLASX_ECPU = -5,       /* CPU BLAS could not be loaded */
    BLASX_EBUSY = -6,      /* all cache slots pinned / server full */
    BLASX_ERANGE = -7      /* problem outside a documented limit */
};

/* ---- library lifetime ----------------------------------------------------
 * Lazy by default (first GEMM call), as the reference's blasx_init
 * (blas/blasx_internal.c:79-100).  Explicit init is optional. */
int blasx_init(void);
void blasx_finalize(void);
int blasx_gpu_available(void);          /* 1 if a GPU path is usable */
const char *blasx_version(void);
const char *blasx_cpublas_path(void);   /* the dlopen'ed CPU BLAS, "" if none */

/* ---- statistics ---------------------------------------------------------- */
typedef struct blasx_stats {
    uint64_t calls_gpu;          /* ABI calls that ran on the GPU */
    uint64_t calls_cpu;          /* ABI calls forwarded to the CPU BLAS by size rule */
    uint64_t cpu_fallbacks;      /* GPU-sized calls run on the CPU (only with BLASX_CPU_FALLBACK=1) */
    uint64_t kernel_launches;    /* GEMM kernel launches */
    uint64_t bytes_h2d;          /* host -> device bytes moved by the tile runtime */
    uint64_t bytes_d2h;          /* device -> host bytes */
    uint64_t bytes_lower_bound;  /* sum over calls of e*(|A|+|B|+|C| in (+ out)) */
    uint64_t tile_hits;
    uint64_t tile_misses;
    uint64_t tile_evictions;
    uint64_t tiles_generated;    /* unused by the ABI path */
    double flops;                /* useful FLOPs executed on the GPU (2MNK / 8MNK) */
} blasx_stats_t;

void blasx_get_stats(blasx_stats_t *out);
void blasx_reset_stats(void);

/* ---- configuration (also read from the environment at init) --------------
 * BLASX_HBM_BUDGET bytes of tile cache (default: 80% of free HBM)
 * BLASX_TILE       tile edge for host-resident GEMMs (default: autotuned)
 * BLASX_GPU_MIN_DIM size threshold (default 1000 for S/D, 500 for Z,
 *                  the reference's blas/dgemm.c:62 and blas/zgemm.c:61)
 * BLASX_CPUBLAS    path of the CPU BLAS to dlopen
 * BLASX_CPU_FALLBACK  1 (default): a GPU-sized call on host operands whose
 *                  GPU path fails is logged, counted in cpu_fallbacks and run
 *                  on the CPU BLAS; 0: the process ends with a message.
 *                  Operands in HBM cannot fall back (the CPU BLAS cannot
 *                  read them): such a failure always ends with a message. */
int blasx_set_hbm_budget(size_t bytes);
size_t blasx_get_hbm_budget(void);
int blasx_set_tile_size(int dtype, int tile);   /* 0 = autotune */
int blasx_get_tile_size(int dtype, int M, int N, int K);
/* Autotuner for the tile edge of host-resident GEMMs (the reference fixes it
 * at compile time, include/blasx.h:5-7): times every candidate edge
 * (256, 512, 1024, 2048 within a cache slot) on synthetic pinned operands of
 * the given shape, records the fastest in the table consulted per call (and
 * appends it to $BLASX_TUNE_FILE).  times_out[i] = seconds for edge 256<<i. */
int blasx_autotune_tile_size(int dtype, int M, int N, int K, int reps, int *best_T,
                             double times_out[4]);
int blasx_tune_table_add(int dtype, int M, int N, int K, int T);

/* ---- memory ---------------------------------------------------------------- */
void *blasx_host_alloc(size_t bytes);    /* pinned host memory (hipHostMalloc) */
void blasx_host_free(void *p);
void *blasx_dev_alloc(size_t bytes);     /* HBM */
void blasx_dev_free(void *p);
int blasx_memcpy_h2d(void *dst, const void *src, size_t bytes);
int blasx_memcpy_d2h(void *dst, const void *src, size_t bytes);
int blasx_memset_device(void *dst, int value, size_t bytes);
int blasx_device_synchronize(void);
/* HBM -> HBM streaming copy by a kernel (16-byte aligned, multiple of 16 bytes) */
int blasx_device_copy(void *dst, const void *src, size_t bytes, void *stream);
/* Measured ceilings of this GPU: out[0] FP64 MFMA TFLOP/s, out[1] FP32 MFMA
 * TFLOP/s, out[2] INT8 MFMA TOP/s (register-only MFMA loops, 16 independent
 * accumulators per wave, best of 4/8/16 waves per CU), out[3] HBM GB/s
 * (read + write of a 4 GiB device copy).  Takes a few hundred ms. */
int blasx_measure_peaks(double *out);
void *blasx_compute_stream(void);        /* hipStream_t of the library's compute stream */
/* HIP events, for timing work on a stream (NULL = the compute stream) */
void *blasx_event_create(void);
void blasx_event_destroy(void *ev);
int blasx_event_record(void *ev, void *stream);
int blasx_event_synchronize(void *ev);
int blasx_event_elapsed_ms(float *ms, void *start, void *stop);

/* ---- device-resident GEMM (operands already in HBM) ----------------------
 * Same argument meaning as cblas_?gemm with Order = ColMajor; trans values
 * are the CBLAS enums.  Asynchronous on `stream`.  Returns BLASX_EINVAL with
 * the BLAS info number written to *info (if non-NULL) on bad arguments. */
int blasx_sgemm_device(int transa, int transb, int M, int N, int K, float alpha,
                       const float *dA, int64_t lda, const float *dB, int64_t ldb, float beta,
                       float *dC, int64_t ldc, void *stream);
int blasx_dgemm_device(int transa, int transb, int M, int N, int K, double alpha,
                       const double *dA, int64_t lda, const double *dB, int64_t ldb,
                       double beta, double *dC, int64_t ldc, void *stream);
int blasx_zgemm_device(int transa, int transb, int M, int N, int K, const double *alpha,
                       const double *dA, int64_t lda, const double *dB, int64_t ldb,
                       const double *beta, double *dC, int64_t ldc, void *stream);

/* ---- split-precision DGEMM (Ozaki scheme on INT8 MFMA) --------------------
 * No counterpart in the reference.  C = alpha op(A) op(B) + beta C with
 * op(A) rows and op(B) columns scaled by powers of two and cut into
 * `slices` exact 7-bit int8 digits (2..9; 0 = default 7); the products of
 * digit pairs with p + q <= slices + 1 run on the INT8 matrix cores with
 * exact int32 accumulation and are summed in FP64.  K <= 133143.
 * Accuracy: each entry of op(A) keeps 7*slices bits relative to the
 * largest magnitude in its row (op(B): column), so inputs with a wide
 * dynamic range within a row need more slices for FP64-level error.
 * Inputs must be finite: a NaN or infinity in a row (column) makes that
 * row's (column's) scaling meaningless and its results unspecified, as the
 * digits of a non-finite value do not exist. */
int blasx_dgemm_split_device(int transa, int transb, int M, int N, int K, double alpha,
                             const double *dA, int64_t lda, const double *dB, int64_t ldb,
                             double beta, double *dC, int64_t ldc, int slices, void *stream);
/* complex128 version: four real split products on the real/imaginary
 * planes (Pr = ArBr - AiBi, Pi = ArBi + AiBr), then C = alpha P + beta C;
 * alpha/beta are (re, im); trans 113 conjugates */
int blasx_zgemm_split_device(int transa, int transb, int M, int N, int K, const double *alpha,
                             const double *dA, int64_t lda, const double *dB, int64_t ldb,
                             const double *beta, double *dC, int64_t ldc, int slices,
                             void *stream);

/* ---- CRT split DGEMM (Ozaki scheme II on INT8 MFMA) ------------------------
 * No counterpart in the reference.  Entries of op(A) rows / op(B) columns
 * are scaled by powers of two to integers of `bits` bits, reduced modulo
 * `moduli` pairwise coprime moduli <= 256 (2..20; 0 = the fewest reaching
 * 56 bits at this K), multiplied with one exact INT8 GEMM per modulus and
 * recombined by the Chinese remainder theorem in FP64.  bits =
 * floor((log2 prod(moduli) - log2 K - 3) / 2), reported by blasx_crt_bits.
 * K <= 131071.  Each entry keeps `bits` bits relative to the largest
 * magnitude in its row (op(B): column).  Inputs must be finite. */
int blasx_dgemm_crt_device(int transa, int transb, int M, int N, int K, double alpha,
                           const double *dA, int64_t lda, const double *dB, int64_t ldb,
                           double beta, double *dC, int64_t ldc, int moduli, void *stream);
int blasx_crt_bits(int K, int moduli);
/* complex128 version: one exponent per op(A) row / op(B) column for both
 * parts, three exact INT8 products per modulus on the residues of Re, Im
 * and Re + Im (Karatsuba, exact in modular arithmetic), so 3 * moduli INT8
 * GEMMs; bits = blasx_crt_bits(2K, moduli); alpha/beta are (re, im);
 * trans 113 conjugates */
int blasx_zgemm_crt_device(int transa, int transb, int M, int N, int K, const double *alpha,
                           const double *dA, int64_t lda, const double *dB, int64_t ldb,
                           const double *beta, double *dC, int64_t ldc, int moduli,
                           void *stream);

/* ---- BF16x3 split DGEMM (BF16 matrix cores) ---------------------------------
 * No counterpart in the reference.  op(A) rows / op(B) columns are scaled
 * by powers of two into (-1, 1) and each entry is cut into three bf16
 * pieces (about 24 significant bits); the six piece products with
 * p + q <= 4 run on the BF16 matrix cores with FP32 sums over K, added into
 * an FP64 workspace; C = alpha 2^(e_i + f_j) W + beta C in FP64.  Accuracy
 * is FP32-like (about sqrt(K) 2^-24 relative to sum_k |a_ik b_kj|), not
 * FP64: a fast variant for callers that accept it.  Inputs must be finite. */
int blasx_dgemm_bf16x3_device(int transa, int transb, int M, int N, int K, double alpha,
                              const double *dA, int64_t lda, const double *dB, int64_t ldb,
                              double beta, double *dC, int64_t ldc, void *stream);

/* ---- grouped GEMM -----------------------------------------------------------
 * No counterpart in the reference (which issues one vendor call per tile
 * product, blas/blasx_dgemm.c:87-94).  Many independent GEMMs of mixed
 * shape, transposes and element type (S, D, Z) in one launch per element
 * type; pointers are HBM addresses.  Asynchronous on `stream`. */
typedef struct blasx_gemm_desc {
    int dtype;              /* BLASX_S, BLASX_D or BLASX_Z */
    int transa, transb;     /* CBLAS enums (111/112/113) */
    int M, N, K;
    double alpha[2];        /* (re, im); im ignored for real types */
    double beta[2];
    const void *A;
    int64_t lda;
    const void *B;
    int64_t ldb;
    void *C;
    int64_t ldc;
} blasx_gemm_desc_t;

int blasx_gemm_grouped_device(const blasx_gemm_desc_t *descs, int n, void *stream);

/* Block tile shape of a real problem in a grouped launch (or a server
 * batch): 0 = 128 x 128, 1 = 128 x 64, 2 = 64 x 64; complex problems keep
 * 128 x 64.  No counterpart in the reference (its tile sizes are compile-time
 * constants, include/blasx.h:5-7).  The shape is picked per problem from a
 * table of launch times measured on an MI355X (built in, written by
 * python -m blasx_amd.tune --grouped; entries added at run time carry a
 * winner only, come later and win ties): the workload nearest in M, N, K,
 * then the fastest shape at the launch's load (its tile count at the large
 * tile for that element type), interpolated between measured loads; with
 * no entry, the large tile once the load reaches two tiles per CU.  Every
 * shape gives bit-identical results.  blasx_set_grouped_shape forces one
 * shape for S or D (-1: back to the table); blasx_grouped_shape reports the
 * choice; blasx_grouped_shape_add adds a measured winner. */
int blasx_set_grouped_shape(int dtype, int shape);
int blasx_grouped_shape(int dtype, int M, int N, int K, int64_t load);
int blasx_grouped_shape_add(int dtype, int M, int N, int K, int64_t load, int shape);

/* Allocation-free variant of blasx_gemm_grouped_device.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and kernels; the problem tables live
 * in a static per-GPU table store (about 430 KiB, held until the process
 * ends) instead of a buffer the library allocates, so at most
 * BLASX_NOALLOC_MAX_PROBLEMS descriptors per call (more: BLASX_ERANGE).  The
 * call allocates nothing, on the device or the host.  Calls share the
 * store: each waits on the device for the previous call's kernels before
 * rewriting it, so calls on different streams are serialised on the GPU;
 * their outputs must not overlap. */
#define BLASX_NOALLOC_MAX_PROBLEMS 1024
int blasx_gemm_grouped_device_noalloc(const blasx_gemm_desc_t *descs, int n, void *stream);

/* ---- GEMM request server (continuous batching) ------------------------------
 * No counterpart in the reference (its ABI is one blocking, non-reentrant
 * call at a time).  Requests carry HOST pointers (pinned memory gives full
 * copy overlap); each is admitted when its operands fit in the server's HBM
 * arena of `hbm_bytes`, joins the next batch, runs in a persistent grouped
 * launch replayed from a hipGraph, and has C copied back before it is
 * reported done.  A request that can never fit is rejected (BLASX_ERANGE).
 * Operands registered with blasx_server_register_operand are uploaded once
 * and shared by every request whose A or B lies inside them; the caller
 * must not modify them while the server lives. */
typedef struct blasx_server blasx_server_t;
blasx_server_t *blasx_server_create(size_t hbm_bytes, int max_batch);
int blasx_server_register_operand(blasx_server_t *s, const void *host, size_t bytes);
int64_t blasx_server_submit(blasx_server_t *s, const blasx_gemm_desc_t *d); /* id or <0 */
int blasx_server_poll(blasx_server_t *s, int64_t id);       /* 1 done, 0 pending, <0 error */
int blasx_server_wait(blasx_server_t *s, int64_t id, double *latency_ms);
int blasx_server_counters(blasx_server_t *s, uint64_t *batches, uint64_t *done,
                          size_t *arena_used);
void blasx_server_destroy(blasx_server_t *s);

/* ---- cross-call residency ---------------------------------------------------
 * The reference frees its tile cache after every call (blas/dgemm.c:114).
 * A caller that promises not to modify [ptr, ptr+bytes) until it
 * unregisters the range lets cached tiles of it survive between calls. */
int blasx_resident_register(const void *host_ptr, size_t bytes);
int blasx_resident_unregister(const void *host_ptr);
void blasx_cache_flush(void);

/* ---- seeded inputs --------------------------------------------------------
 * Element (i,j) of a rows x cols matrix = SplitMix64(seed ^ id<<56 ^ (i+j*rows))
 * mapped to uniform[-1,1); complex uses counters 2e (re) and 2e+1 (im). */
int blasx_fill_uniform_device(int dtype, void *dptr, int64_t rows, int64_t cols, int64_t ld,
                              uint64_t seed, uint32_t id, void *stream);
int blasx_fill_uniform_host(int dtype, void *hptr, int64_t rows, int64_t cols, int64_t ld,
                            uint64_t seed, uint32_t id);
/* entries (ri[a], ci[b]) of the matrix the fills above produce, column-major
 * into out (nr x nc): sampled rows/columns of a full-size operand */
int blasx_sample_uniform_host(int dtype, void *out, int64_t rows, int nr, const int64_t *ri,
                              int nc, const int64_t *ci, uint64_t seed, uint32_t id);
/* split-precision workload inputs: a = (u - 0.5) exp(phi g), u ~ U(0,1),
 * g ~ N(0,1) by Box-Muller on glibc log/cos/exp, FP64 */
int blasx_fill_dynrange_host(double *h, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                             uint32_t id, double phi);
/* out = alpha op(A) op(B) + beta C0 evaluated in double-double and rounded
 * once (TwoSum/TwoProd, Ogita, Rump & Oishi 2005): an accuracy reference
 * for one block of C (C0 may be NULL when beta == 0) */
int blasx_dd_gemm_block_host(int transa, int transb, int m, int n, int K, double alpha,
                             const double *A, int64_t lda, const double *B, int64_t ldb,
                             double beta, const double *C0, int64_t ldc, double *out,
                             int64_t ldo);

#ifdef __cplusplus
}
#endif

#endif /* BLASX_EXT_H */
