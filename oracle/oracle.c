/*
 * oracle.c -- CPU restatement of the reference GEMM path, used ONLY as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Nothing in the product (blasx_amd/) links, loads or calls this file.
 *
 * What it restates (reference file:line):
 *   - the BLAS GEMM definition the reference delegates to cuBLAS/OpenBLAS:
 *     C <- alpha*op(A)*op(B) + beta*C, column-major, op in {N,T,C}
 *     (blas/dgemm.c:18-128, blas/sgemm.c:16-127, blas/zgemm.c:16-130; per
 *     tile blas/blasx_dgemm.c:82-94, blas/blasx_zgemm.c:82)
 *   - argument validation order and Fortran info numbers
 *     (blas/dgemm.c:29-57, blas/zgemm.c:26-56, f77 shim blas/dgemm.c:131-154)
 *   - the tile decomposition: tile grid (blas/blasx_dgemm.c:284-296), C tile
 *     (i,k) = task/(y+1), task%(y+1) (blas/blasx_dgemm.c:28-29), transposed
 *     stored-tile addressing (blas/blasx_dgemm.c:42-53), edge extents
 *     (blas/blasx_mem_control.c:24-42), beta_inner = beta at j==0 else 1
 *     (blas/blasx_dgemm.c:82)
 *   - the FLOP formulas (include/flops.h:218-221)
 *   - the seeded counter-based generator that defines every benchmark input
 *     (no counterpart in the reference, whose driver uses unseeded rand(),
 *     testing/gemm.c:10-24)
 *   - a double-double (Ogita-Rump-Oishi TwoSum/TwoProd) GEMM used as the
 *     high-accuracy oracle for the split-precision GEMM.
 *   - the split-precision schemes themselves (no reference counterpart; from
 *     their published methods): the 7-bit digit split (Ozaki, Ogita, Oishi,
 *     Rump 2012), the CRT split and floating reconstruction (Ozaki scheme
 *     II, 2025) and the BF16x3 piece split, each pinned by exact rational
 *     or closed-form checks in tests/test_split.py, tests/test_crt.py and
 *     tests/test_bf16x3.py.
 *
 * Compiled with -ffp-contract=off so that every a*b+c below rounds twice
 * unless fma() is written explicitly.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>

/* ------------------------------------------------------------------ */
/* Generator: SplitMix64 (Steele, Lea, Flood 2014) of                  */
/*   seed ^ (id << 56) ^ index                                         */
/* uniform[-1,1)  = ((x >> 11) * 2^-53) * 2 - 1                        */
/* ------------------------------------------------------------------ */
uint64_t orc_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t orc_key(uint64_t seed, uint32_t id, uint64_t idx)
{
    return seed ^ ((uint64_t)id << 56) ^ idx;
}

double orc_uniform(uint64_t seed, uint32_t id, uint64_t idx)
{
    uint64_t x = orc_splitmix64(orc_key(seed, id, idx));
    return ((double)(x >> 11) * 0x1.0p-53) * 2.0 - 1.0;
}

/* open unit interval (0,1] draw used by Box-Muller */
static double orc_unit_open(uint64_t seed, uint32_t id, uint64_t idx)
{
    uint64_t x = orc_splitmix64(orc_key(seed, id, idx));
    return ((double)(x >> 11) + 1.0) * 0x1.0p-53;
}

/* Element (i,j) of a rows x cols matrix uses linear index i + j*rows
 * (independent of the storage leading dimension). */
void orc_fill_uniform_d(double *out, int64_t rows, int64_t cols, int64_t ld,
                        uint64_t seed, uint32_t id)
{
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i)
            out[i + j * ld] = orc_uniform(seed, id, (uint64_t)(i + j * rows));
}

void orc_fill_uniform_s(float *out, int64_t rows, int64_t cols, int64_t ld,
                        uint64_t seed, uint32_t id)
{
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i)
            out[i + j * ld] = (float)orc_uniform(seed, id, (uint64_t)(i + j * rows));
}

/* complex: real part from counter 2*idx, imaginary from 2*idx+1 */
void orc_fill_uniform_z(double *out, int64_t rows, int64_t cols, int64_t ld,
                        uint64_t seed, uint32_t id)
{
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i) {
            uint64_t e = (uint64_t)(i + j * rows);
            out[2 * (i + j * ld)] = orc_uniform(seed, id, 2 * e);
            out[2 * (i + j * ld) + 1] = orc_uniform(seed, id, 2 * e + 1);
        }
}

/* Entries (ri[a], ci[b]) of the seeded matrix with `rows` rows (element
 * (i,j) = U(seed, id, i + j*rows), as the fills above), written column-major
 * to out (nr x nc, ld nr).  kind: 0 FP64, 1 FP32 (rounded as the FP32 fill),
 * 2 complex128 (out holds nr*nc (re, im) pairs).  Lets a checker regenerate
 * sampled rows and columns of a matrix too large to hold twice. */
void orc_sample_uniform(double *out, int kind, int64_t rows, int nr, const int64_t *ri, int nc,
                        const int64_t *ci, uint64_t seed, uint32_t id)
{
    for (int b = 0; b < nc; ++b)
        for (int a = 0; a < nr; ++a) {
            const uint64_t e = (uint64_t)(ri[a] + ci[b] * rows);
            const int64_t o = a + (int64_t)b * nr;
            if (kind == 2) {
                out[2 * o] = orc_uniform(seed, id, 2 * e);
                out[2 * o + 1] = orc_uniform(seed, id, 2 * e + 1);
            } else if (kind == 1) {
                out[o] = (double)(float)orc_uniform(seed, id, e);
            } else {
                out[o] = orc_uniform(seed, id, e);
            }
        }
}

/* Dynamic-range matrix a = (u - 0.5) * exp(phi * g), u ~ U(0,1), g ~ N(0,1).
 * u uses matrix id `id` (counter 3*idx), the Box-Muller pair uses counters
 * 3*idx+1 and 3*idx+2. */
double orc_dynrange(uint64_t seed, uint32_t id, uint64_t idx, double phi)
{
    double u = (orc_uniform(seed, id, 3 * idx) + 1.0) * 0.5;
    double u1 = orc_unit_open(seed, id, 3 * idx + 1);
    double u2 = orc_unit_open(seed, id, 3 * idx + 2);
    double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    return (u - 0.5) * exp(phi * g);
}

void orc_fill_dynrange_d(double *out, int64_t rows, int64_t cols, int64_t ld,
                         uint64_t seed, uint32_t id, double phi)
{
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i)
            out[i + j * ld] = orc_dynrange(seed, id, (uint64_t)(i + j * rows), phi);
}

/* ------------------------------------------------------------------ */
/* Argument validation, restating blas/dgemm.c:29-57 (ColMajor).       */
/* trans codes: 111 NoTrans, 112 Trans, 113 ConjTrans (include/cblas.h:12) */
/* Returns the Fortran info number or 0.                               */
/* ------------------------------------------------------------------ */
int orc_gemm_info(int transa, int transb, int M, int N, int K, int lda, int ldb, int ldc)
{
    int ok_a = (transa == 111 || transa == 112 || transa == 113);
    int ok_b = (transb == 111 || transb == 112 || transb == 113);
    int nrowa = (transa == 111) ? M : K;
    int nrowb = (transb == 111) ? K : N;
    if (!ok_a) return 1;
    if (!ok_b) return 2;
    if (M < 0) return 3;
    if (N < 0) return 4;
    if (K < 0) return 5;
    if (lda < (nrowa > 1 ? nrowa : 1)) return 8;
    if (ldb < (nrowb > 1 ? nrowb : 1)) return 10;
    if (ldc < (M > 1 ? M : 1)) return 13;
    return 0;
}

/* f77 character -> CBLAS enum, restating blas/blasx_internal.c:161-174 */
int orc_f77_trans(char c)
{
    if (c == 'N' || c == 'n') return 111;
    if (c == 'T' || c == 't') return 112;
    if (c == 'C' || c == 'c') return 113;
    return -1;
}

/* ------------------------------------------------------------------ */
/* Naive triple-loop GEMM in the working precision.                   */
/* ta, tb: 'N','T','C'.  p runs 0..K-1 in order; acc starts at 0.      */
/* ------------------------------------------------------------------ */
#define OPIDX(t, r, c, ld) (((t) == 'N') ? ((r) + (int64_t)(c) * (ld)) : ((c) + (int64_t)(r) * (ld)))

void orc_gemm_d(char ta, char tb, int M, int N, int K, double alpha,
                const double *A, int lda, const double *B, int ldb,
                double beta, double *C, int ldc)
{
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            double acc = 0.0;
            if (alpha != 0.0)
                for (int p = 0; p < K; ++p)
                    acc += A[OPIDX(ta, i, p, lda)] * B[OPIDX(tb, p, j, ldb)];
            double *c = &C[i + (int64_t)j * ldc];
            if (beta == 0.0)
                *c = alpha * acc;
            else
                *c = alpha * acc + beta * *c;
        }
}

void orc_gemm_s(char ta, char tb, int M, int N, int K, float alpha,
                const float *A, int lda, const float *B, int ldb,
                float beta, float *C, int ldc)
{
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            float acc = 0.0f;
            if (alpha != 0.0f)
                for (int p = 0; p < K; ++p)
                    acc += A[OPIDX(ta, i, p, lda)] * B[OPIDX(tb, p, j, ldb)];
            float *c = &C[i + (int64_t)j * ldc];
            if (beta == 0.0f)
                *c = alpha * acc;
            else
                *c = alpha * acc + beta * *c;
        }
}

/* Same loop accumulated in double (a more accurate FP32 reference). */
void orc_gemm_s_acc64(char ta, char tb, int M, int N, int K, float alpha,
                      const float *A, int lda, const float *B, int ldb,
                      float beta, float *C, int ldc)
{
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            double acc = 0.0;
            if (alpha != 0.0f)
                for (int p = 0; p < K; ++p)
                    acc += (double)A[OPIDX(ta, i, p, lda)] * (double)B[OPIDX(tb, p, j, ldb)];
            float *c = &C[i + (int64_t)j * ldc];
            if (beta == 0.0f)
                *c = (float)((double)alpha * acc);
            else
                *c = (float)((double)alpha * acc + (double)beta * (double)*c);
        }
}

/* complex128, interleaved (re,im); alpha/beta given as (re,im) pairs */
void orc_gemm_z(char ta, char tb, int M, int N, int K, const double *alpha,
                const double *A, int lda, const double *B, int ldb,
                const double *beta, double *C, int ldc)
{
    double ar = alpha[0], ai = alpha[1], br = beta[0], bi = beta[1];
    int alpha_zero = (ar == 0.0 && ai == 0.0), beta_zero = (br == 0.0 && bi == 0.0);
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            double sr = 0.0, si = 0.0;
            if (!alpha_zero)
                for (int p = 0; p < K; ++p) {
                    int64_t ia = OPIDX(ta, i, p, lda), ib = OPIDX(tb, p, j, ldb);
                    double xr = A[2 * ia], xi = A[2 * ia + 1];
                    double yr = B[2 * ib], yi = B[2 * ib + 1];
                    if (ta == 'C') xi = -xi;
                    if (tb == 'C') yi = -yi;
                    sr += xr * yr - xi * yi;
                    si += xr * yi + xi * yr;
                }
            double *c = &C[2 * (i + (int64_t)j * ldc)];
            double outr = ar * sr - ai * si, outi = ar * si + ai * sr;
            if (!beta_zero) {
                outr += br * c[0] - bi * c[1];
                outi += br * c[1] + bi * c[0];
            }
            c[0] = outr;
            c[1] = outi;
        }
}

/* ------------------------------------------------------------------ */
/* Double-double arithmetic (Dekker 1971; Ogita, Rump, Oishi 2005).    */
/* ------------------------------------------------------------------ */
typedef struct { double hi, lo; } dd_t;

static inline dd_t two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double err = (a - (s - bb)) + (b - bb);
    dd_t r = {s, err};
    return r;
}

static inline dd_t fast_two_sum(double a, double b)
{
    double s = a + b;
    dd_t r = {s, b - (s - a)};
    return r;
}

static inline dd_t two_prod(double a, double b)
{
    double p = a * b;
    dd_t r = {p, fma(a, b, -p)};
    return r;
}

/* accurate dd + dd (two TwoSums, then renormalise) */
static inline dd_t dd_add(dd_t x, dd_t y)
{
    dd_t s = two_sum(x.hi, y.hi);
    dd_t t = two_sum(x.lo, y.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}

static inline dd_t dd_mul_d(dd_t x, double b)
{
    dd_t p = two_prod(x.hi, b);
    p.lo += x.lo * b;
    return fast_two_sum(p.hi, p.lo);
}

static inline dd_t dd_mul(dd_t x, dd_t y)
{
    dd_t p = two_prod(x.hi, y.hi);
    p.lo += x.hi * y.lo + x.lo * y.hi;
    return fast_two_sum(p.hi, p.lo);
}

static inline dd_t dd_neg(dd_t x) { dd_t r = {-x.hi, -x.lo}; return r; }

/* dd dot-product GEMM: out_hi = round(C), out_lo = residual (may be NULL).
 * C is read for beta (not when beta == 0) and NOT written; results go to out_hi. */
void orc_gemm_dd_d(char ta, char tb, int M, int N, int K, double alpha,
                   const double *A, int lda, const double *B, int ldb,
                   double beta, const double *C, int ldc,
                   double *out_hi, double *out_lo, int ldo)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            dd_t s = {0.0, 0.0};
            for (int p = 0; p < K; ++p)
                s = dd_add(s, two_prod(A[OPIDX(ta, i, p, lda)], B[OPIDX(tb, p, j, ldb)]));
            dd_t r = dd_mul_d(s, alpha);
            if (beta != 0.0)
                r = dd_add(r, two_prod(beta, C[i + (int64_t)j * ldc]));
            out_hi[i + (int64_t)j * ldo] = r.hi;
            if (out_lo) out_lo[i + (int64_t)j * ldo] = r.lo;
        }
}

/* complex dd GEMM on interleaved storage; outputs interleaved hi and lo */
void orc_gemm_dd_z(char ta, char tb, int M, int N, int K, const double *alpha,
                   const double *A, int lda, const double *B, int ldb,
                   const double *beta, const double *C, int ldc,
                   double *out_hi, double *out_lo, int ldo)
{
    int beta_zero = (beta[0] == 0.0 && beta[1] == 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            dd_t sr = {0, 0}, si = {0, 0};
            for (int p = 0; p < K; ++p) {
                int64_t ia = OPIDX(ta, i, p, lda), ib = OPIDX(tb, p, j, ldb);
                double xr = A[2 * ia], xi = A[2 * ia + 1];
                double yr = B[2 * ib], yi = B[2 * ib + 1];
                if (ta == 'C') xi = -xi;
                if (tb == 'C') yi = -yi;
                sr = dd_add(sr, two_prod(xr, yr));
                sr = dd_add(sr, two_prod(-xi, yi));
                si = dd_add(si, two_prod(xr, yi));
                si = dd_add(si, two_prod(xi, yr));
            }
            dd_t rr = dd_add(dd_mul_d(sr, alpha[0]), dd_neg(dd_mul_d(si, alpha[1])));
            dd_t ri = dd_add(dd_mul_d(si, alpha[0]), dd_mul_d(sr, alpha[1]));
            if (!beta_zero) {
                const double *c = &C[2 * (i + (int64_t)j * ldc)];
                rr = dd_add(rr, dd_add(two_prod(beta[0], c[0]), two_prod(-beta[1], c[1])));
                ri = dd_add(ri, dd_add(two_prod(beta[0], c[1]), two_prod(beta[1], c[0])));
            }
            int64_t o = 2 * (i + (int64_t)j * ldo);
            out_hi[o] = rr.hi;
            out_hi[o + 1] = ri.hi;
            if (out_lo) {
                out_lo[o] = rr.lo;
                out_lo[o + 1] = ri.lo;
            }
        }
}

/* exposed dd primitives so the tests can pin them against exact rationals */
void orc_dd_two_sum(double a, double b, double *s, double *e)
{
    dd_t r = two_sum(a, b);
    *s = r.hi;
    *e = r.lo;
}

void orc_dd_two_prod(double a, double b, double *p, double *e)
{
    dd_t r = two_prod(a, b);
    *p = r.hi;
    *e = r.lo;
}

void orc_dd_dot(const double *x, const double *y, int n, double *hi, double *lo)
{
    dd_t s = {0, 0};
    for (int p = 0; p < n; ++p) s = dd_add(s, two_prod(x[p], y[p]));
    *hi = s.hi;
    *lo = s.lo;
}

/* ------------------------------------------------------------------ */
/* Tiled GEMM restating the reference's tile decomposition             */
/* (blas/blasx_dgemm.c:28-94, blas/blasx_mem_control.c:24-42):         */
/* C tile (i,k), K-steps j = 0..z in order, each step one tile product */
/* with beta_inner = beta at j==0 else 1.  Used to pin tiling index    */
/* math: on integer inputs it must equal orc_gemm_d bit for bit.       */
/* ------------------------------------------------------------------ */
static void orc_margin(int rows, int cols, int T, int br, int bc, int *nr, int *nc)
{
    /* restates blas/blasx_mem_control.c:24-42 */
    *nr = (br == rows / T) ? rows % T : T;
    *nc = (bc == cols / T) ? cols % T : T;
}

void orc_gemm_tiled_d(char ta, char tb, int M, int N, int K, int T, double alpha,
                      const double *A, int lda, const double *B, int ldb,
                      double beta, double *C, int ldc)
{
    int nrowa = (ta == 'N') ? M : K, ncola = (ta == 'N') ? K : M;
    int nrowb = (tb == 'N') ? K : N, ncolb = (tb == 'N') ? N : K;
    /* tile grid: x,y,z = ceil(dim/T)-1 (blas/blasx_dgemm.c:284-296) */
    int x = (M + T - 1) / T - 1, y = (N + T - 1) / T - 1, z = (K + T - 1) / T - 1;
    int task_num = (x + 1) * (y + 1);
    for (int task = task_num - 1; task >= 0; --task) { /* queue order: high to low */
        int i = task / (y + 1), k = task % (y + 1);
        int mc, nc;
        orc_margin(M, N, T, i, k, &mc, &nc);
        double *Ct = &C[(int64_t)i * T + (int64_t)k * T * ldc];
        for (int j = 0; j <= z; ++j) {
            int nra, nca, nrb, ncb;
            const double *At, *Bt;
            if (ta == 'N') {
                orc_margin(nrowa, ncola, T, i, j, &nra, &nca);
                At = &A[(int64_t)i * T + (int64_t)j * T * lda];
            } else {
                orc_margin(nrowa, ncola, T, j, i, &nra, &nca);
                At = &A[(int64_t)j * T + (int64_t)i * T * lda];
            }
            if (tb == 'N') {
                orc_margin(nrowb, ncolb, T, j, k, &nrb, &ncb);
                Bt = &B[(int64_t)j * T + (int64_t)k * T * ldb];
            } else {
                orc_margin(nrowb, ncolb, T, k, j, &nrb, &ncb);
                Bt = &B[(int64_t)k * T + (int64_t)j * T * ldb];
            }
            int ka = (ta == 'N') ? nca : nra;
            double beta_inner = (j == 0) ? beta : 1.0;
            orc_gemm_d(ta, tb, mc, nc, ka, alpha, At, lda, Bt, ldb, beta_inner, Ct, ldc);
        }
    }
}

/* FLOP formulas, include/flops.h:218-221 (with the complex count the
 * reference's zgemm.c:95 print should have used) */
double orc_flops_gemm(int complex_, double m, double n, double k)
{
    return (complex_ ? 8.0 : 2.0) * m * n * k;
}

/* ------------------------------------------------------------------ */
/* Split-precision DGEMM (Ozaki scheme, INT8 digits) restated on the   */
/* CPU.  No reference counterpart; follows the published method and    */
/* the exact evaluation order of the library:                           */
/*   e_i = frexp exponent of max_k |op(A)_ik| (0 for a zero row),       */
/*   digits d_t = trunc(r*128), r <- r*128 - d_t, t = 1..s,             */
/*   for g = s+1 down to 2, pairs p = 1..s with q = g-p in 1..s, taken  */
/*   in chunks of `segmax` pairs: W (+)= 2^-7g * (integer chunk sum),   */
/*   C = alpha * 2^(e_i+f_j) * W  (+ fma(beta, C, .)).                  */
/* Digits and integer sums are exact, so the result is reproducible   */
/* bit for bit.                                                         */
/* ------------------------------------------------------------------ */
static int orc_exp_of(double amax)
{
    if (amax == 0.0 || !isfinite(amax)) return 0;
    int e;
    frexp(amax, &e);
    return e;
}

void orc_ozaki_digits(double x, int e, int s, int8_t *d)
{
    double r = ldexp(x, -e);
    for (int t = 0; t < s; ++t) {
        r *= 128.0;
        double q = trunc(r);
        d[t] = (int8_t)(int)q;
        r -= q;
    }
}

void orc_dgemm_ozaki(char ta, char tb, int M, int N, int K, double alpha, const double *A,
                     int lda, const double *B, int ldb, double beta, double *C, int ldc, int s,
                     int segmax)
{
    int *ea = (int *)malloc(sizeof(int) * (M > 0 ? M : 1));
    int *eb = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    int8_t *DA = (int8_t *)calloc((size_t)s * M * (K > 0 ? K : 1), 1);
    int8_t *DB = (int8_t *)calloc((size_t)s * N * (K > 0 ? K : 1), 1);
    double *W = (double *)calloc((size_t)M * N + 1, sizeof(double));
    for (int i = 0; i < M; ++i) {
        double m = 0;
        for (int p = 0; p < K; ++p) m = fmax(m, fabs(A[OPIDX(ta, i, p, lda)]));
        ea[i] = orc_exp_of(m);
        for (int p = 0; p < K; ++p) {
            int8_t d[9];
            orc_ozaki_digits(A[OPIDX(ta, i, p, lda)], ea[i], s, d);
            for (int t = 0; t < s; ++t) DA[((size_t)t * M + i) * K + p] = d[t];
        }
    }
    for (int j = 0; j < N; ++j) {
        double m = 0;
        for (int p = 0; p < K; ++p) m = fmax(m, fabs(B[OPIDX(tb, p, j, ldb)]));
        eb[j] = orc_exp_of(m);
        for (int p = 0; p < K; ++p) {
            int8_t d[9];
            orc_ozaki_digits(B[OPIDX(tb, p, j, ldb)], eb[j], s, d);
            for (int t = 0; t < s; ++t) DB[((size_t)t * N + j) * K + p] = d[t];
        }
    }
    int first = 1;
    for (int g = s + 1; g >= 2; --g) {
        int pairs[16][2], np = 0;
        for (int p = 1; p <= s; ++p) {
            int q = g - p;
            if (q >= 1 && q <= s) {
                pairs[np][0] = p;
                pairs[np][1] = q;
                ++np;
            }
        }
        double scale = ldexp(1.0, -7 * g);
        for (int off = 0; off < np; off += segmax) {
            int cnt = np - off < segmax ? np - off : segmax;
#pragma omp parallel for schedule(static)
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < M; ++i) {
                    int64_t acc = 0;
                    for (int c = 0; c < cnt; ++c) {
                        const int8_t *da = DA + ((size_t)(pairs[off + c][0] - 1) * M + i) * K;
                        const int8_t *db = DB + ((size_t)(pairs[off + c][1] - 1) * N + j) * K;
                        for (int p = 0; p < K; ++p) acc += (int64_t)da[p] * db[p];
                    }
                    double add = scale * (double)acc;
                    double *w = &W[i + (size_t)j * M];
                    *w = first ? add : *w + add;
                }
            first = 0;
        }
    }
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            double v = alpha * ldexp(W[i + (size_t)j * M], ea[i] + eb[j]);
            double *c = &C[i + (int64_t)j * ldc];
            *c = (beta == 0.0) ? v : fma(beta, *c, v);
        }
    free(ea);
    free(eb);
    free(DA);
    free(DB);
    free(W);
}

/* complex split GEMM restated: planes (conjugate = negated imaginary plane),
 * Pr = Ar Br - Ai Bi, Pi = Ar Bi + Ai Br by four real split products in
 * this order, then C = alpha P + beta C with the library's FMA order. */
void orc_zgemm_ozaki(char ta, char tb, int M, int N, int K, const double *alpha, const double *A,
                     int lda, const double *B, int ldb, const double *beta, double *C, int ldc,
                     int s, int segmax)
{
    const int ra = ta == 'N' ? M : K, ca = ta == 'N' ? K : M;
    const int rb = tb == 'N' ? K : N, cb = tb == 'N' ? N : K;
    double *Ar = (double *)malloc(sizeof(double) * ((size_t)ra * ca + 1));
    double *Ai = (double *)malloc(sizeof(double) * ((size_t)ra * ca + 1));
    double *Br = (double *)malloc(sizeof(double) * ((size_t)rb * cb + 1));
    double *Bi = (double *)malloc(sizeof(double) * ((size_t)rb * cb + 1));
    double *Pr = (double *)calloc((size_t)M * N + 1, sizeof(double));
    double *Pi = (double *)calloc((size_t)M * N + 1, sizeof(double));
    for (int j = 0; j < ca; ++j)
        for (int i = 0; i < ra; ++i) {
            const double *z = &A[2 * (i + (int64_t)j * lda)];
            Ar[i + (size_t)j * ra] = z[0];
            Ai[i + (size_t)j * ra] = ta == 'C' ? -z[1] : z[1];
        }
    for (int j = 0; j < cb; ++j)
        for (int i = 0; i < rb; ++i) {
            const double *z = &B[2 * (i + (int64_t)j * ldb)];
            Br[i + (size_t)j * rb] = z[0];
            Bi[i + (size_t)j * rb] = tb == 'C' ? -z[1] : z[1];
        }
    const char rta = ta == 'N' ? 'N' : 'T', rtb = tb == 'N' ? 'N' : 'T';
    orc_dgemm_ozaki(rta, rtb, M, N, K, 1.0, Ar, ra, Br, rb, 0.0, Pr, M, s, segmax);
    orc_dgemm_ozaki(rta, rtb, M, N, K, -1.0, Ai, ra, Bi, rb, 1.0, Pr, M, s, segmax);
    orc_dgemm_ozaki(rta, rtb, M, N, K, 1.0, Ar, ra, Bi, rb, 0.0, Pi, M, s, segmax);
    orc_dgemm_ozaki(rta, rtb, M, N, K, 1.0, Ai, ra, Br, rb, 1.0, Pi, M, s, segmax);
    const int bz = beta[0] == 0.0 && beta[1] == 0.0;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            const size_t t = i + (size_t)j * M;
            double *c = &C[2 * (i + (int64_t)j * ldc)];
            const double tr = fma(alpha[0], Pr[t], -(alpha[1] * Pi[t]));
            const double ti = fma(alpha[0], Pi[t], alpha[1] * Pr[t]);
            if (bz) {
                c[0] = tr;
                c[1] = ti;
            } else {
                const double cr = c[0], ci = c[1];
                c[0] = fma(beta[0], cr, fma(-beta[1], ci, tr));
                c[1] = fma(beta[0], ci, fma(beta[1], cr, ti));
            }
        }
    free(Ar);
    free(Ai);
    free(Br);
    free(Bi);
    free(Pr);
    free(Pi);
}

/* ------------------------------------------------------------------ */
/* CRT split DGEMM (Ozaki scheme II: Ozaki, Uchino, Imamura 2025,      */
/* integer matrix products modulo small coprime moduli, recombined by   */
/* the Chinese remainder theorem).  No reference counterpart; restates  */
/* the method in the library's exact evaluation order:                  */
/*   moduli m_1..m_L from ORC_CRT_MODS (pairwise coprime, <= 256),       */
/*   beta = gamma = floor((sum log2 m_l - 1) / 2 - 2^-20),              */
/*   e_i = orc_norm_exp(sum_k op(A)_ik^2): ||op(A)_i||_2 < 2^e_i, and   */
/*   f_j likewise for the columns of op(B) (Cauchy-Schwarz scaling),    */
/*   A'_ik = trunc(ldexp(op(A)_ik, beta - e_i)),  ||A'_i|| < 2^beta,    */
/*   a_l = smod(A', m_l) with smod(v, m) in [-floor(m/2), ceil(m/2)-1], */
/*   R_l = smod(a_l b_l^T, m_l)  (exact integer products),              */
/*   floating CRT: M = prod m_l with bit length E, w_l = (M/m_l) y_l,   */
/*   y_l = (M/m_l)^-1 mod m_l; w_l and M are cut into 40-bit chunks,    */
/*   chunk k holding bits [max(0, E-40k), E-40(k-1)), k = 1..4, so that  */
/*     Q_k = sum_l Wk_l R_l       (exact: |Q_k| < 20 2^47 chunk units)  */
/*     Md = ((M_4 + M_3) + M_2) + M_1,  kq = rint((Q_1 + Q_2) / Md)     */
/*        (computed as (Q_1 + Q_2) * (1 / Md)), |kq| < 2^12             */
/*     D_k = Q_k - kq M_k         (exact: < 2^53 chunk units)           */
/*     S = ((D_1 + D_2) + D_3) + D_4                                    */
/*   and C = alpha * ldexp(S, e_i + f_j - 2 beta)  (+ fma(beta, C, .)). */
/* |A'B'| <= ||A'_i|| ||B'_j|| < 2^(2 beta) < M/2 (worst-case bits     */
/* relative to the row maximum: orc_crt_beta), so Q = P + kq M with    */
/* |P/M| < 1/2 - 2^-21 while (Q_1 + Q_2)/Md is within 2^-38 of Q/M:    */
/* kq                                                                   */
/* is exact.  The partial sum D_1 + .. + D_j equals P minus the lower  */
/* chunks, so it is exact whenever it is small, and S = P (1 + 3u) for */
/* every P, with no cancellation loss.                                  */
/* Garner's mixed-radix digits (orc_crt_digits) give the exact integer  */
/* and check the residues independently.                                */
/* ------------------------------------------------------------------ */
static const int ORC_CRT_MODS[20] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                     223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

int orc_crt_modulus(int l) { return (l >= 0 && l < 20) ? ORC_CRT_MODS[l] : 0; }

/* scaling bits relative to the row 2-norm: 2^(2 gamma + 1) < M (M is
 * never a power of two, and the 2^-20 keeps 2^(2 gamma + 1) below
 * M (1 - 2^-20)) */
int orc_crt_gamma(int L)
{
    double lm = 0.0;
    for (int l = 0; l < L; ++l) lm += log2((double)ORC_CRT_MODS[l]);
    return (int)floor((lm - 1.0) / 2.0 - 0x1p-20);
}

/* smallest e with ss (1 + 2^-30) < 4^e: from the binary exponent k of
 * ss (1 + 2^-30) (below 2^k, at least 2^(k-1)), e = ceil(k / 2) */
int orc_norm_exp(double ss)
{
    const double u = ss * (1.0 + 0x1p-30);
    if (u == 0.0 || !isfinite(u)) return 0;
    int k;
    frexp(u, &k);
    return (k + 1) >> 1;
}

int orc_crt_beta(int K, int L)
{
    double lm = 0.0;
    for (int l = 0; l < L; ++l) lm += log2((double)ORC_CRT_MODS[l]);
    return (int)floor((lm - log2((double)(K > 0 ? K : 1)) - 1.0) / 2.0 - 0x1p-20);
}

int64_t orc_smod(int64_t v, int64_t m)
{
    int64_t r = v % m;
    if (r < 0) r += m;
    if (r >= (m + 1) / 2) r -= m;
    return r;
}

/* residue of an integer-valued double of any magnitude, with int64 math */
int orc_crt_resid(double v, int m)
{
    int t = 0;
    if (fabs(v) >= 0x1p53) {
        t = ilogb(v) - 52;
        v = ldexp(v, -t);
    }
    int64_t r = orc_smod((int64_t)v, m), p = 1 % m;
    for (int k = 0; k < t; ++k) p = (p * 2) % m;
    return (int)orc_smod(r * p, m);
}

static int orc_inv_mod(int a, int m)
{
    a %= m;
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return 0;
}

/* Garner digits of residues r[0..L-1] (symmetric); returns digits v[] */
void orc_crt_digits(const int *r, int L, int *v)
{
    for (int l = 0; l < L; ++l) {
        const int m = ORC_CRT_MODS[l];
        int64_t t = r[l];
        for (int j = 0; j < l; ++j)
            t = orc_smod((t - v[j]) * orc_inv_mod(ORC_CRT_MODS[j], m), m);
        v[l] = (int)orc_smod(t, m);
    }
}

double orc_crt_horner(const int *v, int L)
{
    double S = (double)v[L - 1];
    for (int l = L - 2; l >= 0; --l) S = fma(S, (double)ORC_CRT_MODS[l], (double)v[l]);
    return S;
}

/* little-endian base-2^32 naturals, 6 limbs (M < 2^160) */
typedef struct {
    uint32_t d[6];
} orc_big;

static void big_mul_small(orc_big *x, uint32_t f)
{
    uint64_t carry = 0;
    for (int i = 0; i < 6; ++i) {
        uint64_t t = (uint64_t)x->d[i] * f + carry;
        x->d[i] = (uint32_t)t;
        carry = t >> 32;
    }
}

static void big_div_small(orc_big *x, uint32_t q)
{
    uint64_t rem = 0;
    for (int i = 5; i >= 0; --i) {
        uint64_t cur = (rem << 32) | x->d[i];
        x->d[i] = (uint32_t)(cur / q);
        rem = cur % q;
    }
}

static uint32_t big_mod_small(const orc_big *x, uint32_t q)
{
    uint64_t rem = 0;
    for (int i = 5; i >= 0; --i) rem = ((rem << 32) | x->d[i]) % q;
    return (uint32_t)rem;
}

static int big_bitlen(const orc_big *x)
{
    for (int i = 5; i >= 0; --i)
        if (x->d[i])
            for (int b = 31; b >= 0; --b)
                if (x->d[i] >> b & 1) return 32 * i + b + 1;
    return 0;
}

/* bits [lo, hi) of x as an exact double scaled by 2^lo */
static double big_bits(const orc_big *x, int lo, int hi)
{
    double v = 0.0;
    for (int b = hi - 1; b >= lo; --b) v = v * 2.0 + (double)(x->d[b / 32] >> (b % 32) & 1);
    return ldexp(v, lo);
}

/* floating CRT constants: Wk[k*L + l], Mk[k] (k = 0..3), inverse of Md */
void orc_crt_consts(int L, double *Wk, double *Mk, double *invM)
{
    orc_big M = {{1, 0, 0, 0, 0, 0}};
    for (int l = 0; l < L; ++l) big_mul_small(&M, (uint32_t)ORC_CRT_MODS[l]);
    const int E = big_bitlen(&M);
    for (int k = 0; k < 4; ++k) {
        const int hi = E - 40 * k, lo = E - 40 * (k + 1) > 0 ? E - 40 * (k + 1) : 0;
        Mk[k] = hi > 0 ? big_bits(&M, lo, hi) : 0.0;
    }
    for (int l = 0; l < L; ++l) {
        const uint32_t m = (uint32_t)ORC_CRT_MODS[l];
        orc_big w = M;
        big_div_small(&w, m);                            /* M / m_l */
        const uint32_t y = (uint32_t)orc_inv_mod((int)big_mod_small(&w, m), (int)m);
        big_mul_small(&w, y);                            /* w_l = (M/m_l) y_l < M */
        for (int k = 0; k < 4; ++k) {
            const int hi = E - 40 * k, lo = E - 40 * (k + 1) > 0 ? E - 40 * (k + 1) : 0;
            Wk[k * L + l] = hi > 0 ? big_bits(&w, lo, hi) : 0.0;
        }
    }
    const double Md = ((Mk[3] + Mk[2]) + Mk[1]) + Mk[0];
    *invM = 1.0 / Md;
}

/* floating CRT of residues r[0..L-1] with the constants above */
double orc_crt_float(const int *r, int L, const double *Wk, const double *Mk, double invM)
{
    double Q[4];
    for (int k = 0; k < 4; ++k) {
        double q = 0.0;
        for (int l = 0; l < L; ++l) q = fma(Wk[k * L + l], (double)r[l], q);
        Q[k] = q;
    }
    const double kq = rint((Q[0] + Q[1]) * invM);
    double D[4];
    for (int k = 0; k < 4; ++k) D[k] = fma(-kq, Mk[k], Q[k]);
    return ((D[0] + D[1]) + D[2]) + D[3];
}

void orc_dgemm_crt(char ta, char tb, int M, int N, int K, double alpha, const double *A, int lda,
                   const double *B, int ldb, double beta, double *C, int ldc, int L)
{
    const int bb = orc_crt_gamma(L);
    double Wk[4 * 20], Mk[4], invM;
    orc_crt_consts(L, Wk, Mk, &invM);
    int *ea = (int *)malloc(sizeof(int) * (M > 0 ? M : 1));
    int *eb = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    int8_t *RA = (int8_t *)calloc((size_t)L * M * (K > 0 ? K : 1), 1);
    int8_t *RB = (int8_t *)calloc((size_t)L * N * (K > 0 ? K : 1), 1);
    for (int i = 0; i < M; ++i) {
        double ss = 0;
        for (int p = 0; p < K; ++p) ss = fma(A[OPIDX(ta, i, p, lda)], A[OPIDX(ta, i, p, lda)], ss);
        ea[i] = orc_norm_exp(ss);
        for (int p = 0; p < K; ++p) {
            const double v = trunc(ldexp(A[OPIDX(ta, i, p, lda)], bb - ea[i]));
            for (int l = 0; l < L; ++l)
                RA[((size_t)l * M + i) * K + p] = (int8_t)orc_crt_resid(v, ORC_CRT_MODS[l]);
        }
    }
    for (int j = 0; j < N; ++j) {
        double ss = 0;
        for (int p = 0; p < K; ++p) ss = fma(B[OPIDX(tb, p, j, ldb)], B[OPIDX(tb, p, j, ldb)], ss);
        eb[j] = orc_norm_exp(ss);
        for (int p = 0; p < K; ++p) {
            const double v = trunc(ldexp(B[OPIDX(tb, p, j, ldb)], bb - eb[j]));
            for (int l = 0; l < L; ++l)
                RB[((size_t)l * N + j) * K + p] = (int8_t)orc_crt_resid(v, ORC_CRT_MODS[l]);
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j) {
        int r[20];
        for (int i = 0; i < M; ++i) {
            for (int l = 0; l < L; ++l) {
                const int8_t *ra = RA + ((size_t)l * M + i) * K;
                const int8_t *rb = RB + ((size_t)l * N + j) * K;
                int64_t acc = 0;
                for (int p = 0; p < K; ++p) acc += (int64_t)ra[p] * rb[p];
                r[l] = (int)orc_smod(acc, ORC_CRT_MODS[l]);
            }
            const double S = orc_crt_float(r, L, Wk, Mk, invM);
            const double x = alpha * ldexp(S, ea[i] + eb[j] - 2 * bb);
            double *c = &C[i + (int64_t)j * ldc];
            *c = (beta == 0.0) ? x : fma(beta, *c, x);
        }
    }
    free(ea);
    free(eb);
    free(RA);
    free(RB);
}

/* complex CRT split restated: op(A) rows / op(B) columns share one exponent
 * for both parts (of the 2-norm over the 2K real parts), beta =
 * orc_crt_gamma(L) since |Re P|, |Im P| <= ||a_i|| ||b_j|| < 2^(2 beta)
 * (Cauchy-Schwarz over 2K terms); per modulus three exact products on the
 * residue planes (Karatsuba, exact in modular arithmetic):
 *   R1 = ar br, R2 = ai bi, R3 = (ar + ai)(br + bi)   (each smod m),
 *   Re = smod(R1 - R2), Im = smod(R3 - R1 - R2),
 * floating CRT of Re and Im, scaling by 2^(e_i + f_j - 2 beta), then
 * C = alpha x + beta C with the library's complex FMA order. */
void orc_zgemm_crt(char ta, char tb, int M, int N, int K, const double *alpha, const double *A,
                   int lda, const double *B, int ldb, const double *beta, double *C, int ldc,
                   int L)
{
    const int bb = orc_crt_gamma(L);
    double Wk[4 * 20], Mk[4], invM;
    orc_crt_consts(L, Wk, Mk, &invM);
    const size_t pa = (size_t)M * (K > 0 ? K : 1), pb = (size_t)N * (K > 0 ? K : 1);
    int8_t *RA = (int8_t *)calloc((size_t)3 * L * pa, 1);
    int8_t *RB = (int8_t *)calloc((size_t)3 * L * pb, 1);
    int *ea = (int *)malloc(sizeof(int) * (M > 0 ? M : 1));
    int *eb = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    for (int side = 0; side < 2; ++side) {
        const int rows = side == 0 ? M : N;
        const char t = side == 0 ? ta : tb;
        const double *X = side == 0 ? A : B;
        const int ld = side == 0 ? lda : ldb;
        int8_t *R = side == 0 ? RA : RB;
        int *e = side == 0 ? ea : eb;
        const size_t plane = side == 0 ? pa : pb;
        for (int i = 0; i < rows; ++i) {
            double ss = 0;
            for (int p = 0; p < K; ++p) {
                const int64_t o = side == 0 ? OPIDX(t, i, p, ld) : OPIDX(t, p, i, ld);
                ss = fma(X[2 * o + 1], X[2 * o + 1], fma(X[2 * o], X[2 * o], ss));
            }
            e[i] = orc_norm_exp(ss);
            for (int p = 0; p < K; ++p) {
                const int64_t o = side == 0 ? OPIDX(t, i, p, ld) : OPIDX(t, p, i, ld);
                const double re = X[2 * o], im = t == 'C' ? -X[2 * o + 1] : X[2 * o + 1];
                const double vr = trunc(ldexp(re, bb - e[i])), vi = trunc(ldexp(im, bb - e[i]));
                for (int l = 0; l < L; ++l) {
                    const int m = ORC_CRT_MODS[l];
                    const int r = orc_crt_resid(vr, m), q = orc_crt_resid(vi, m);
                    R[(size_t)(3 * l) * plane + (size_t)i * K + p] = (int8_t)r;
                    R[(size_t)(3 * l + 1) * plane + (size_t)i * K + p] = (int8_t)q;
                    R[(size_t)(3 * l + 2) * plane + (size_t)i * K + p] = (int8_t)orc_smod(r + q, m);
                }
            }
        }
    }
    const int bz = beta[0] == 0.0 && beta[1] == 0.0;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j) {
        int rre[20], rim[20];
        for (int i = 0; i < M; ++i) {
            for (int l = 0; l < L; ++l) {
                const int m = ORC_CRT_MODS[l];
                int64_t P[3];
                for (int c = 0; c < 3; ++c) {
                    const int8_t *ra = RA + (size_t)(3 * l + c) * pa + (size_t)i * K;
                    const int8_t *rb = RB + (size_t)(3 * l + c) * pb + (size_t)j * K;
                    int64_t acc = 0;
                    for (int p = 0; p < K; ++p) acc += (int64_t)ra[p] * rb[p];
                    P[c] = orc_smod(acc, m);
                }
                rre[l] = (int)orc_smod(P[0] - P[1], m);
                rim[l] = (int)orc_smod(P[2] - P[0] - P[1], m);
            }
            const int sh = ea[i] + eb[j] - 2 * bb;
            const double xr = ldexp(orc_crt_float(rre, L, Wk, Mk, invM), sh);
            const double xi = ldexp(orc_crt_float(rim, L, Wk, Mk, invM), sh);
            const double tr = fma(alpha[0], xr, -(alpha[1] * xi));
            const double ti = fma(alpha[0], xi, alpha[1] * xr);
            double *c = &C[2 * (i + (int64_t)j * ldc)];
            if (bz) {
                c[0] = tr;
                c[1] = ti;
            } else {
                const double cr = c[0], ci = c[1];
                c[0] = fma(beta[0], cr, fma(-beta[1], ci, tr));
                c[1] = fma(beta[0], ci, fma(beta[1], cr, ti));
            }
        }
    }
    free(RA);
    free(RB);
    free(ea);
    free(eb);
}

/* ------------------------------------------------------------------ */
/* BF16x3 split DGEMM restated on the CPU (no reference counterpart):  */
/*   e_i, f_j as the digit scheme (frexp exponent of the row / column   */
/*   maximum); x = 2^-e a in (-1,1);                                    */
/*   s1 = rn8(x), s2 = rn8(x - s1), s3 = rn8(x - s1 - s2), rn8 rounding */
/*   to 8 significant bits, nearest, ties to even, and giving 0 below   */
/*   2^-126 (bf16's normal range);                                      */
/*   W = sum over groups g = 4, 3, 2 of sum_{p+q=g} sum_k s_p s_q,      */
/*   here in FP64 (the library sums each group in FP32 on the matrix    */
/*   cores: the two agree within that rounding, and bit for bit when    */
/*   every partial sum is exact);                                       */
/*   C = alpha 2^(e_i+f_j) W (+ fma(beta, C, .)).                       */
/* ------------------------------------------------------------------ */
double orc_bf16_rn8(double x)
{
    if (x == 0.0 || !isfinite(x)) return x;
    int e;
    const double m = frexp(x, &e); /* |m| in [0.5, 1) */
    /* m * 2^8 has 8 integer bits; round half to even */
    const double t = ldexp(m, 8), r = nearbyint(t); /* default mode: nearest-even */
    const double y = ldexp(r, e - 8);
    return fabs(y) < ldexp(1.0, -126) ? 0.0 : y;
}

void orc_bf16x3_pieces(double x, double *out)
{
    const double s1 = orc_bf16_rn8(x), r1 = x - s1;
    const double s2 = orc_bf16_rn8(r1), s3 = orc_bf16_rn8(r1 - s2);
    out[0] = s1;
    out[1] = s2;
    out[2] = s3;
}

void orc_dgemm_bf16x3(char ta, char tb, int M, int N, int K, double alpha, const double *A,
                      int lda, const double *B, int ldb, double beta, double *C, int ldc)
{
    int *ea = (int *)calloc(M + 1, sizeof(int)), *eb = (int *)calloc(N + 1, sizeof(int));
    double *PA = (double *)calloc((size_t)3 * M * (K > 0 ? K : 1), sizeof(double));
    double *PB = (double *)calloc((size_t)3 * N * (K > 0 ? K : 1), sizeof(double));
    for (int i = 0; i < M; ++i) {
        double m = 0;
        for (int p = 0; p < K; ++p) m = fmax(m, fabs(A[OPIDX(ta, i, p, lda)]));
        ea[i] = orc_exp_of(m);
        for (int p = 0; p < K; ++p) {
            double q[3];
            orc_bf16x3_pieces(ldexp(A[OPIDX(ta, i, p, lda)], -ea[i]), q);
            for (int t = 0; t < 3; ++t) PA[((size_t)t * M + i) * K + p] = q[t];
        }
    }
    for (int j = 0; j < N; ++j) {
        double m = 0;
        for (int p = 0; p < K; ++p) m = fmax(m, fabs(B[OPIDX(tb, p, j, ldb)]));
        eb[j] = orc_exp_of(m);
        for (int p = 0; p < K; ++p) {
            double q[3];
            orc_bf16x3_pieces(ldexp(B[OPIDX(tb, p, j, ldb)], -eb[j]), q);
            for (int t = 0; t < 3; ++t) PB[((size_t)t * N + j) * K + p] = q[t];
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            double w = 0;
            for (int g = 4; g >= 2; --g) {
                double sg = 0;
                for (int p = 1; p <= 3; ++p) {
                    const int q = g - p;
                    if (q < 1 || q > 3) continue;
                    const double *pa = PA + ((size_t)(p - 1) * M + i) * K;
                    const double *pb = PB + ((size_t)(q - 1) * N + j) * K;
                    for (int k = 0; k < K; ++k) sg += pa[k] * pb[k];
                }
                w += sg;
            }
            const double v = alpha * ldexp(w, ea[i] + eb[j]);
            double *c = &C[i + (int64_t)j * ldc];
            *c = beta == 0.0 ? v : fma(beta, *c, v);
        }
    free(ea);
    free(eb);
    free(PA);
    free(PB);
}
