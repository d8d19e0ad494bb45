/*
 * verify.c -- host-side helpers for measuring accuracy and generating the
 * benchmark inputs of the split-precision workload:
 *   * blasx_fill_dynrange_host: a_ij = (u - 0.5) * exp(phi * g),
 *     u ~ U(0,1), g ~ N(0,1) by Box-Muller, from the seeded SplitMix64
 *     counters (glibc log/cos/exp on the host, so device math never decides
 *     the inputs);
 *   * blasx_dd_gemm_block_host: a double-double (TwoSum/TwoProd, Ogita,
 *     Rump & Oishi 2005) GEMM of one block of C, the accuracy reference the
 *     split-precision results are compared against.
 * Both are multithreaded over columns.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <unistd.h>

#include "internal.h"

static inline uint64_t sm64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline double u11(uint64_t key) { return ((double)(sm64(key) >> 11) * 0x1.0p-53) * 2.0 - 1.0; }
static inline double u01_open(uint64_t key) { return ((double)(sm64(key) >> 11) + 1.0) * 0x1.0p-53; }

typedef struct {
    double *p;
    int64_t rows, cols, ld, c0, c1;
    uint64_t base;
    double phi;
} dyn_job_t;

static void *dyn_worker(void *arg)
{
    dyn_job_t *j = (dyn_job_t *)arg;
    for (int64_t c = j->c0; c < j->c1; ++c)
        for (int64_t r = 0; r < j->rows; ++r) {
            const uint64_t idx = (uint64_t)(r + c * j->rows);
            const double u = (u11(j->base ^ (3 * idx)) + 1.0) * 0.5;
            const double u1 = u01_open(j->base ^ (3 * idx + 1));
            const double u2 = u01_open(j->base ^ (3 * idx + 2));
            const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
            j->p[r + c * j->ld] = (u - 0.5) * exp(j->phi * g);
        }
    return NULL;
}

static int nthreads_for(int64_t work)
{
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    const char *e = getenv("OMP_NUM_THREADS");
    if (e && atoi(e) > 0 && atoi(e) < n) n = atoi(e);
    if (n < 1) n = 1;
    if (n > 64) n = 64;
    if (work < (1 << 18)) n = 1;
    return (int)n;
}

BX_EXPORT int blasx_fill_dynrange_host(double *h, int64_t rows, int64_t cols, int64_t ld,
                                       uint64_t seed, uint32_t id, double phi)
{
    if (rows < 0 || cols < 0 || ld < rows) return BLASX_EINVAL;
    const int nt = nthreads_for(rows * cols);
    pthread_t th[64];
    dyn_job_t jobs[64];
    for (int t = 0; t < nt; ++t) {
        jobs[t] = (dyn_job_t){h, rows, cols, ld, cols * t / nt, cols * (t + 1) / nt,
                              seed ^ ((uint64_t)id << 56), phi};
        if (nt == 1) dyn_worker(&jobs[t]);
        else pthread_create(&th[t], NULL, dyn_worker, &jobs[t]);
    }
    if (nt > 1)
        for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    return 0;
}

/* ---- double-double GEMM of one block ---- */
typedef struct { double hi, lo; } dd;

static inline dd two_sum(double a, double b)
{
    const double s = a + b, bb = s - a;
    return (dd){s, (a - (s - bb)) + (b - bb)};
}
static inline dd fast2(double a, double b)
{
    const double s = a + b;
    return (dd){s, b - (s - a)};
}
static inline dd two_prod(double a, double b)
{
    const double p = a * b;
    return (dd){p, fma(a, b, -p)};
}
static inline dd dd_add(dd x, dd y)
{
    dd s = two_sum(x.hi, y.hi), t = two_sum(x.lo, y.lo);
    s.lo += t.hi;
    s = fast2(s.hi, s.lo);
    s.lo += t.lo;
    return fast2(s.hi, s.lo);
}
static inline dd dd_mul_d(dd x, double b)
{
    dd p = two_prod(x.hi, b);
    p.lo += x.lo * b;
    return fast2(p.hi, p.lo);
}

typedef struct {
    int ta, tb, m, n0, n1, K;
    double alpha, beta;
    const double *A, *B, *C0;
    int64_t lda, ldb, ldc, ldo;
    double *out;
} dd_job_t;

static void *dd_worker(void *arg)
{
    dd_job_t *j = (dd_job_t *)arg;
    for (int c = j->n0; c < j->n1; ++c)
        for (int r = 0; r < j->m; ++r) {
            dd s = {0, 0};
            for (int p = 0; p < j->K; ++p) {
                const double a = j->ta ? j->A[p + (int64_t)r * j->lda] : j->A[r + (int64_t)p * j->lda];
                const double b = j->tb ? j->B[c + (int64_t)p * j->ldb] : j->B[p + (int64_t)c * j->ldb];
                s = dd_add(s, two_prod(a, b));
            }
            dd v = dd_mul_d(s, j->alpha);
            if (j->beta != 0.0 && j->C0) v = dd_add(v, two_prod(j->beta, j->C0[r + (int64_t)c * j->ldc]));
            j->out[r + (int64_t)c * j->ldo] = v.hi + v.lo;
        }
    return NULL;
}

/* out[m x n] = round(alpha op(A) op(B) + beta C0) evaluated in double-double */
BX_EXPORT int blasx_dd_gemm_block_host(int transa, int transb, int m, int n, int K, double alpha,
                                       const double *A, int64_t lda, const double *B,
                                       int64_t ldb, double beta, const double *C0, int64_t ldc,
                                       double *out, int64_t ldo)
{
    if (m < 0 || n < 0 || K < 0 || ldo < m) return BLASX_EINVAL;
    const int nt = nthreads_for((int64_t)m * n * (K > 0 ? K : 1) / 64);
    pthread_t th[64];
    dd_job_t jobs[64];
    for (int t = 0; t < nt; ++t) {
        jobs[t] = (dd_job_t){transa != 111, transb != 111, m, (int)((int64_t)n * t / nt),
                             (int)((int64_t)n * (t + 1) / nt), K, alpha, beta, A, B, C0, lda,
                             ldb, ldc, ldo, out};
        if (nt == 1) dd_worker(&jobs[t]);
        else pthread_create(&th[t], NULL, dd_worker, &jobs[t]);
    }
    if (nt > 1)
        for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    return 0;
}

/* Entries (ri[a], ci[b]) of the seeded uniform matrix with `rows` rows, the
 * values blasx_fill_uniform_host / _device put there (element (i,j) from
 * counter i + j*rows; complex: counters 2e and 2e+1), column-major into out
 * (nr x nc, ld nr) of the element type.  Lets a benchmark check sampled
 * rows and columns of a full-size product without a second copy of the
 * operands. */
BX_EXPORT int blasx_sample_uniform_host(int dtype, void *out, int64_t rows, int nr,
                                        const int64_t *ri, int nc, const int64_t *ci,
                                        uint64_t seed, uint32_t id)
{
    if (nr < 0 || nc < 0 || (nr * nc > 0 && (!out || !ri || !ci))) return BLASX_EINVAL;
    if (dtype != BLASX_S && dtype != BLASX_D && dtype != BLASX_Z) return BLASX_EINVAL;
    const uint64_t base = seed ^ ((uint64_t)id << 56);
    for (int b = 0; b < nc; ++b)
        for (int a = 0; a < nr; ++a) {
            if (ri[a] < 0 || ri[a] >= rows || ci[b] < 0) return BLASX_EINVAL;
            const uint64_t e = (uint64_t)(ri[a] + ci[b] * rows);
            const int64_t o = a + (int64_t)b * nr;
            if (dtype == BLASX_Z) {
                ((double *)out)[2 * o] = u11(base ^ (2 * e));
                ((double *)out)[2 * o + 1] = u11(base ^ (2 * e + 1));
            } else if (dtype == BLASX_S) {
                ((float *)out)[o] = (float)u11(base ^ e);
            } else {
                ((double *)out)[o] = u11(base ^ e);
            }
        }
    return 0;
}
