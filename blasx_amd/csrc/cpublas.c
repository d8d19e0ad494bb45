/*
 * cpublas.c -- the CPU BLAS forwarder: small problems (and CGEMM) go to a
 * dlopen'ed OpenBLAS, exactly as the reference's type-1 dispatch does
 * (dlopen(CPUBLAS) + dlsym("cblas_?gemm") through the handle,
 * blas/blasx_internal.c:26-36, blas/blasx_tile_resource.c:30-39,
 * blas/blasx_globalpointers.c:450-480).  Symbols are looked up through the
 * handle, never RTLD_NEXT, so under LD_PRELOAD the library cannot recurse
 * into its own cblas_dgemm.
 *
 * Differences by design: the library path is a run-time choice
 * ($BLASX_CPUBLAS, then libopenblas.so*, then the ILP64 OpenBLAS that numpy
 * bundles) instead of a compile-time -DCPUBLAS; a missing CPU BLAS is an
 * error message, not exit() (blas/blasx_internal.c:31-34).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <glob.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

typedef void (*lp64_real_t)(int, int, int, int, int, int, double, const void *, int,
                            const void *, int, double, void *, int);
typedef void (*lp64_sreal_t)(int, int, int, int, int, int, float, const void *, int,
                             const void *, int, float, void *, int);
typedef void (*lp64_cplx_t)(int, int, int, int, int, int, const void *, const void *, int,
                            const void *, int, const void *, void *, int);
typedef void (*ilp64_real_t)(int, int, int, int64_t, int64_t, int64_t, double, const void *,
                             int64_t, const void *, int64_t, double, void *, int64_t);
typedef void (*ilp64_sreal_t)(int, int, int, int64_t, int64_t, int64_t, float, const void *,
                              int64_t, const void *, int64_t, float, void *, int64_t);
typedef void (*ilp64_cplx_t)(int, int, int, int64_t, int64_t, int64_t, const void *,
                             const void *, int64_t, const void *, int64_t, const void *, void *,
                             int64_t);

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static void *g_handle;
static int g_ilp64;
static void *g_fn[4]; /* indexed by BLASX_S/D/Z/C */
static char g_path[1024];
static int g_tried;

static int try_open(const char *path)
{
    void *h = dlopen(path, RTLD_LAZY | RTLD_LOCAL);
    if (!h) return 0;
    const char *lp[4] = {"cblas_sgemm", "cblas_dgemm", "cblas_zgemm", "cblas_cgemm"};
    const char *ilp[4] = {"scipy_cblas_sgemm64_", "scipy_cblas_dgemm64_", "scipy_cblas_zgemm64_",
                          "scipy_cblas_cgemm64_"};
    void *f = dlsym(h, lp[1]);
    int ilp64 = 0;
    if (!f) {
        f = dlsym(h, ilp[1]);
        ilp64 = 1;
    }
    /* never bind to ourselves (a libblasx on the search path) */
    Dl_info self, other;
    if (f && dladdr((void *)&try_open, &self) && dladdr(f, &other) &&
        self.dli_fbase == other.dli_fbase)
        f = NULL;
    if (!f) {
        dlclose(h);
        return 0;
    }
    for (int i = 0; i < 4; ++i) g_fn[i] = dlsym(h, ilp64 ? ilp[i] : lp[i]);
    g_ilp64 = ilp64;
    g_handle = h;
    snprintf(g_path, sizeof g_path, "%s", path);
    bx_log(1, "CPU BLAS: %s (%s)", path, ilp64 ? "ILP64" : "LP64");
    return 1;
}

int bx_cpublas_load(void)
{
    pthread_mutex_lock(&g_lock);
    if (g_handle || g_tried) {
        int ok = g_handle != NULL;
        pthread_mutex_unlock(&g_lock);
        return ok ? 0 : BLASX_ECPU;
    }
    g_tried = 1;
    int ok = 0;
    const char *env = getenv("BLASX_CPUBLAS");
    if (env && *env) ok = try_open(env);
    const char *names[] = {"libopenblas.so.0", "libopenblas.so", NULL};
    for (int i = 0; !ok && names[i]; ++i) ok = try_open(names[i]);
    const char *pats[] = {"/usr/local/lib/python3*/dist-packages/numpy.libs/libscipy_openblas64_*.so",
                          "/usr/lib/python3*/dist-packages/numpy.libs/libscipy_openblas64_*.so",
                          "/usr/local/lib/python3*/site-packages/numpy.libs/libscipy_openblas64_*.so",
                          NULL};
    for (int i = 0; !ok && pats[i]; ++i) {
        glob_t g;
        if (glob(pats[i], 0, NULL, &g) == 0) {
            for (size_t k = 0; !ok && k < g.gl_pathc; ++k) ok = try_open(g.gl_pathv[k]);
            globfree(&g);
        }
    }
    pthread_mutex_unlock(&g_lock);
    if (!ok) bx_log(0, "no CPU BLAS found (set BLASX_CPUBLAS)");
    return ok ? 0 : BLASX_ECPU;
}

/* Resolve one routine of the CPU BLAS by its LP64 name ("cblas_dgemv",
 * "dgemv_"); the ILP64 build numpy bundles names it scipy_<name>64_.  NULL
 * (with a message) when the CPU BLAS or the routine is missing, or when the
 * lookup would land in this library. */
void *bx_cpublas_sym(const char *name)
{
    if (bx_cpublas_load() != 0) return NULL;
    char buf[160];
    if (g_ilp64)
        snprintf(buf, sizeof buf, "scipy_%s64_", name);
    else
        snprintf(buf, sizeof buf, "%s", name);
    void *f = dlsym(g_handle, buf);
    Dl_info self, other;
    if (f && dladdr((void *)&bx_cpublas_sym, &self) && dladdr(f, &other) &&
        self.dli_fbase == other.dli_fbase)
        f = NULL;
    if (!f) bx_log(0, "CPU BLAS %s has no %s", g_path, buf);
    return f;
}

int bx_cpublas_ilp64(void) { return g_ilp64; }

BX_EXPORT const char *blasx_cpublas_path(void)
{
    bx_cpublas_load();
    return g_path;
}

/* Order is passed through unchanged, as the reference's CPU path does
 * (blas/dgemm.c:74); arguments were validated by the caller. */
int bx_cpu_gemm(int dt, int order, int ta, int tb, int M, int N, int K, const void *alpha,
                const void *A, int lda, const void *B, int ldb, const void *beta, void *C,
                int ldc)
{
    if (bx_cpublas_load() != 0 || !g_fn[dt]) return BLASX_ECPU;
    void *f = g_fn[dt];
    if (dt == BLASX_D) {
        double a = *(const double *)alpha, b = *(const double *)beta;
        if (g_ilp64)
            ((ilp64_real_t)f)(order, ta, tb, M, N, K, a, A, lda, B, ldb, b, C, ldc);
        else
            ((lp64_real_t)f)(order, ta, tb, M, N, K, a, A, lda, B, ldb, b, C, ldc);
    } else if (dt == BLASX_S) {
        float a = *(const float *)alpha, b = *(const float *)beta;
        if (g_ilp64)
            ((ilp64_sreal_t)f)(order, ta, tb, M, N, K, a, A, lda, B, ldb, b, C, ldc);
        else
            ((lp64_sreal_t)f)(order, ta, tb, M, N, K, a, A, lda, B, ldb, b, C, ldc);
    } else {
        if (g_ilp64)
            ((ilp64_cplx_t)f)(order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
        else
            ((lp64_cplx_t)f)(order, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    }
    BX_STAT_ADD(calls_cpu, 1);
    return 0;
}
