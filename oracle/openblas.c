/*
 * openblas.c -- the reference's CPU BLAS path, loaded the way the reference
 * loads it: dlopen(CPUBLAS) then dlsym("cblas_?gemm") through that handle
 * (blas/blasx_internal.c:26-36, blas/blasx_tile_resource.c:30-39,
 * blas/blasx_globalpointers.c:458-464).  Test/benchmark infrastructure only:
 * used by tests/ as a cross-check and by bench.py as the CPU baseline.
 *
 * Accepts an LP64 library exporting cblas_dgemm, or the ILP64 build that
 * numpy bundles (OpenBLAS 0.3.29, symbols scipy_cblas_dgemm64_ with 64-bit
 * integer arguments).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <glob.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef void (*lp64_dgemm_t)(int, int, int, int, int, int, double, const double *, int,
                             const double *, int, double, double *, int);
typedef void (*lp64_sgemm_t)(int, int, int, int, int, int, float, const float *, int,
                             const float *, int, float, float *, int);
typedef void (*lp64_zgemm_t)(int, int, int, int, int, int, const void *, const void *, int,
                             const void *, int, const void *, void *, int);
typedef void (*ilp64_dgemm_t)(int, int, int, int64_t, int64_t, int64_t, double, const double *,
                              int64_t, const double *, int64_t, double, double *, int64_t);
typedef void (*ilp64_sgemm_t)(int, int, int, int64_t, int64_t, int64_t, float, const float *,
                              int64_t, const float *, int64_t, float, float *, int64_t);
typedef void (*ilp64_zgemm_t)(int, int, int, int64_t, int64_t, int64_t, const void *,
                              const void *, int64_t, const void *, int64_t, const void *, void *,
                              int64_t);

static void *g_handle;
static int g_ilp64;
static void *g_dgemm, *g_sgemm, *g_zgemm, *g_set_threads, *g_get_config;
static char g_path[1024];

static int try_open(const char *path)
{
    void *h = dlopen(path, RTLD_LAZY | RTLD_LOCAL);
    if (!h) return 0;
    void *d = dlsym(h, "cblas_dgemm");
    if (d) {
        g_ilp64 = 0;
        g_dgemm = d;
        g_sgemm = dlsym(h, "cblas_sgemm");
        g_zgemm = dlsym(h, "cblas_zgemm");
        g_set_threads = dlsym(h, "openblas_set_num_threads");
        g_get_config = dlsym(h, "openblas_get_config");
    } else if ((d = dlsym(h, "scipy_cblas_dgemm64_"))) {
        g_ilp64 = 1;
        g_dgemm = d;
        g_sgemm = dlsym(h, "scipy_cblas_sgemm64_");
        g_zgemm = dlsym(h, "scipy_cblas_zgemm64_");
        g_set_threads = dlsym(h, "scipy_openblas_set_num_threads64_");
        g_get_config = dlsym(h, "scipy_openblas_get_config64_");
    } else {
        dlclose(h);
        return 0;
    }
    g_handle = h;
    snprintf(g_path, sizeof g_path, "%s", path);
    return 1;
}

/* path may be NULL: then $ORACLE_CPUBLAS, libopenblas.so*, numpy's bundle */
int orc_openblas_load(const char *path)
{
    if (g_handle) return 1;
    if (path && *path) return try_open(path);
    const char *env = getenv("ORACLE_CPUBLAS");
    if (env && *env && try_open(env)) return 1;
    const char *names[] = {"libopenblas.so.0", "libopenblas.so", NULL};
    for (int i = 0; names[i]; ++i)
        if (try_open(names[i])) return 1;
    const char *pats[] = {"/usr/local/lib/python3*/dist-packages/numpy.libs/libscipy_openblas64_*.so",
                          "/usr/lib/python3*/dist-packages/numpy.libs/libscipy_openblas64_*.so",
                          "/usr/local/lib/python3*/site-packages/numpy.libs/libscipy_openblas64_*.so",
                          NULL};
    for (int i = 0; pats[i]; ++i) {
        glob_t g;
        if (glob(pats[i], 0, NULL, &g) == 0) {
            for (size_t k = 0; k < g.gl_pathc; ++k)
                if (try_open(g.gl_pathv[k])) {
                    globfree(&g);
                    return 1;
                }
            globfree(&g);
        }
    }
    return 0;
}

const char *orc_openblas_path(void) { return g_path; }
int orc_openblas_is_ilp64(void) { return g_ilp64; }

const char *orc_openblas_config(void)
{
    if (!g_get_config) return "";
    return ((const char *(*)(void))g_get_config)();
}

void orc_openblas_set_threads(int n)
{
    if (!g_set_threads) return;
    if (g_ilp64)
        ((void (*)(int64_t))g_set_threads)((int64_t)n);
    else
        ((void (*)(int))g_set_threads)(n);
}

/* ColMajor only; trans codes are the CBLAS enum values 111/112/113 */
void orc_openblas_dgemm(int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                        const double *A, int64_t lda, const double *B, int64_t ldb,
                        double beta, double *C, int64_t ldc)
{
    if (g_ilp64)
        ((ilp64_dgemm_t)g_dgemm)(102, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    else
        ((lp64_dgemm_t)g_dgemm)(102, ta, tb, (int)M, (int)N, (int)K, alpha, A, (int)lda, B,
                                (int)ldb, beta, C, (int)ldc);
}

void orc_openblas_sgemm(int ta, int tb, int64_t M, int64_t N, int64_t K, float alpha,
                        const float *A, int64_t lda, const float *B, int64_t ldb, float beta,
                        float *C, int64_t ldc)
{
    if (g_ilp64)
        ((ilp64_sgemm_t)g_sgemm)(102, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    else
        ((lp64_sgemm_t)g_sgemm)(102, ta, tb, (int)M, (int)N, (int)K, alpha, A, (int)lda, B,
                                (int)ldb, beta, C, (int)ldc);
}

void orc_openblas_zgemm(int ta, int tb, int64_t M, int64_t N, int64_t K, const double *alpha,
                        const double *A, int64_t lda, const double *B, int64_t ldb,
                        const double *beta, double *C, int64_t ldc)
{
    if (g_ilp64)
        ((ilp64_zgemm_t)g_zgemm)(102, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    else
        ((lp64_zgemm_t)g_zgemm)(102, ta, tb, (int)M, (int)N, (int)K, alpha, A, (int)lda, B,
                                (int)ldb, beta, C, (int)ldc);
}
