/*
 * crt.c -- CRT split DGEMM (Ozaki scheme II), host side.  No counterpart in
 * the reference; the method and its exactness argument are stated in
 * hip/ozaki_kernels.hip.
 *
 * One INT8 GEMM per modulus instead of one per digit pair.  Rows of op(A)
 * and columns of op(B) are scaled by their 2-norms (Cauchy-Schwarz: |sum_k
 * A'_ik B'_kj| <= ||A'_i|| ||B'_j|| < 2^(2 gamma) < M / 2, gamma =
 * floor((log2 M - 1) / 2), M = product of the moduli), so every entry keeps
 * gamma bits relative to its row's 2-norm.  In the worst case (all entries
 * of a row equal in magnitude) that is beta = floor((log2 M - log2 K - 1) /
 * 2) bits relative to the row maximum (blasx_crt_bits), e.g. K = 16384:
 * L = 16 -> 55 bits, L = 18 -> 62 bits, where the digit scheme needs
 * s(s+1)/2 = 36 products for 56 bits; rows whose RMS is below their maximum
 * keep log2(max / RMS) bits more.
 *
 * Workspace (HBM, grown on demand and kept): L residue planes of op(A)
 * (Mp x Kp) and of op(B)^T (Np x Kp), L int8 planes of the residue products
 * (Mp x Np), exponents, the exponent scan scratch and the plane table.
 * One call at a time per process (a mutex guards the workspace).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include "internal.h"

int bxh_oz_block_m(void);
int bxh_oz_block_k(void);
int bxh_oz_exp_chunks(void);
int bxh_oz_exponents(const double *X, int64_t ld, int n, int len, int contiguous, int norm,
                     int *e, double *part, void *stream);
int bxh_crt_split(const double *X, int64_t ld, int n, int len, int contiguous, const int *e,
                  int beta, int L, const int *moduli, int8_t *D, int np, int kp, void *stream);
int bxh_crt_gemm(const void *segs_dev, int L, int per, const int *moduli, int kp, int mp, int np,
                 int8_t *R, void *stream);
int bxh_oz_exponents_z(const double *Z, int64_t ld, int n, int len, int contiguous, int norm,
                       int *e, double *part, void *stream);
int bxh_zcrt_split(const double *Z, int64_t ld, int n, int len, int contiguous, int conj,
                   const int *e, int beta, int L, const int *moduli, int8_t *D, int np, int kp,
                   void *stream);
int bxh_zcrt_recon(int m, int n, const int8_t *R, int64_t ldr, int64_t plane, const void *recon,
                   const int *moduli, const int *ea, const int *eb, int beta_bits,
                   const double *alpha, const double *beta, double *C, int64_t ldc,
                   void *stream);
int bxh_crt_recon(int m, int n, const int8_t *R, int64_t ldr, int64_t plane, const void *recon,
                  const int *ea, const int *eb, int beta_bits, double alpha, double beta,
                  double *C, int64_t ldc, void *stream);

/* pairwise coprime moduli <= 256, largest first (256 = 2^8, 255 = 3*5*17,
 * 253 = 11*23, 247 = 13*19, 217 = 7*31, the rest prime) */
static const int CRT_MODS[20] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                 223, 217, 211, 199, 197, 193, 191, 181, 179, 173};
#define CRT_MAX_L 20
/* int32 accumulation of int8 residues (|r| <= 128) is exact for K * 2^14 < 2^31 */
#define CRT_MAX_K 131071

/* layout shared with the recon kernel's argument */
typedef struct {
    int L;
    double W[4][20];
    double Mk[4];
    double invM;
} crt_recon_t;

/* scaling bits relative to the row 2-norm */
int bx_crt_gamma(int L)
{
    double lm = 0.0;
    for (int l = 0; l < L; ++l) lm += log2((double)CRT_MODS[l]);
    return (int)floor((lm - 1.0) / 2.0 - 0x1p-20);
}

/* worst-case bits relative to the row maximum */
int bx_crt_beta(int K, int L)
{
    double lm = 0.0;
    for (int l = 0; l < L; ++l) lm += log2((double)CRT_MODS[l]);
    return (int)floor((lm - log2((double)(K > 0 ? K : 1)) - 1.0) / 2.0 - 0x1p-20);
}

/* ---- naturals below 2^192 as 12 little-endian 16-bit digits in uint32 --- */
typedef struct {
    uint32_t g[12];
} nat_t;

static void nat_mul(nat_t *x, uint32_t f) /* f < 2^16 */
{
    uint32_t c = 0;
    for (int i = 0; i < 12; ++i) {
        const uint32_t t = x->g[i] * f + c;
        x->g[i] = t & 0xFFFF;
        c = t >> 16;
    }
}

static uint32_t nat_divmod(nat_t *x, uint32_t q) /* x /= q, returns x mod q */
{
    uint32_t r = 0;
    for (int i = 11; i >= 0; --i) {
        const uint32_t cur = (r << 16) | x->g[i];
        x->g[i] = cur / q;
        r = cur % q;
    }
    return r;
}

static int nat_bits(const nat_t *x)
{
    for (int i = 11; i >= 0; --i)
        if (x->g[i]) {
            int b = 15;
            while (!(x->g[i] >> b & 1)) --b;
            return 16 * i + b + 1;
        }
    return 0;
}

/* bits [lo, hi) of x, scaled by 2^lo, as an exact double (hi - lo <= 53) */
static double nat_field(const nat_t *x, int lo, int hi)
{
    uint64_t v = 0;
    for (int b = hi - 1; b >= lo; --b) v = (v << 1) | (x->g[b >> 4] >> (b & 15) & 1);
    return ldexp((double)v, lo);
}

static int inv_mod(int a, int m)
{
    /* extended Euclid on small integers */
    int r0 = m, r1 = a % m, t0 = 0, t1 = 1;
    while (r1) {
        const int q = r0 / r1, r2 = r0 - q * r1, t2 = t0 - q * t1;
        r0 = r1;
        r1 = r2;
        t0 = t1;
        t1 = t2;
    }
    return t0 < 0 ? t0 + m : t0;
}

/* weights w_l = (M/m_l)((M/m_l)^-1 mod m_l) and M, cut into 40-bit chunks
 * from the top (chunk k: bits [max(0, E-40(k+1)), E-40k)) */
void bx_crt_consts(int L, void *out)
{
    crt_recon_t *c = (crt_recon_t *)out;
    memset(c, 0, sizeof *c);
    c->L = L;
    nat_t M = {{1}};
    for (int l = 0; l < L; ++l) nat_mul(&M, (uint32_t)CRT_MODS[l]);
    const int E = nat_bits(&M);
    for (int k = 0; k < 4; ++k) {
        const int hi = E - 40 * k, lo = E - 40 * (k + 1) > 0 ? E - 40 * (k + 1) : 0;
        c->Mk[k] = hi > 0 ? nat_field(&M, lo, hi) : 0.0;
    }
    for (int l = 0; l < L; ++l) {
        nat_t w = M;
        const uint32_t m = (uint32_t)CRT_MODS[l];
        nat_divmod(&w, m);
        nat_t t = w;
        const int y = inv_mod((int)nat_divmod(&t, m), (int)m);
        nat_mul(&w, (uint32_t)y);
        for (int k = 0; k < 4; ++k) {
            const int hi = E - 40 * k, lo = E - 40 * (k + 1) > 0 ? E - 40 * (k + 1) : 0;
            c->W[k][l] = hi > 0 ? nat_field(&w, lo, hi) : 0.0;
        }
    }
    const double Md = ((c->Mk[3] + c->Mk[2]) + c->Mk[1]) + c->Mk[0];
    c->invM = 1.0 / Md;
}

static pthread_mutex_t g_crt_lock = PTHREAD_MUTEX_INITIALIZER;
static struct {
    void *dev;
    size_t dev_cap;
    void *segs_host, *segs_dev;
    void *done;
} g_crt;

static size_t rup(size_t v, size_t a) { return (v + a - 1) / a * a; }

/* smallest L whose beta reaches 56 bits (the digit scheme's 8 slices) */
int bx_crt_default_moduli(int K)
{
    for (int L = 2; L <= CRT_MAX_L; ++L)
        if (bx_crt_beta(K, L) >= 56) return L;
    return CRT_MAX_L;
}

int bx_dgemm_crt(int opa, int opb, int M, int N, int K, double alpha, const double *A,
                 int64_t lda, const double *B, int64_t ldb, double beta, double *C, int64_t ldc,
                 int L, void *stream)
{
    if (M == 0 || N == 0) return 0;
    if (K == 0 || alpha == 0.0) {
        double b2[2] = {beta, 0.0};
        return bxh_scale(BLASX_D, M, N, b2, C, ldc, stream);
    }
    if (K > CRT_MAX_K) return BLASX_ERANGE;
    const int bb = bx_crt_gamma(L);
    const int pad = (M >= 4096 && N >= 4096) ? 2 * bxh_oz_block_m() : bxh_oz_block_m();
    const int mp = (int)rup(M, pad), np = (int)rup(N, pad);
    const int kp = (int)rup(K, 2 * bxh_oz_block_k()); /* an even number of K tiles */

    const size_t a8 = (size_t)L * mp * kp, b8 = (size_t)L * np * kp;
    const size_t r8 = (size_t)L * mp * np;
    const size_t off_b8 = rup(a8, 256), off_r = off_b8 + rup(b8, 256);
    const size_t off_ea = off_r + rup(r8, 256), off_eb = off_ea + rup((size_t)M * 4, 256);
    const size_t off_part = off_eb + rup((size_t)N * 4, 256);
    const size_t total = off_part + (size_t)bxh_oz_exp_chunks() * (M > N ? M : N) * sizeof(double);
    const size_t segs_bytes = (size_t)3 * CRT_MAX_L * 2 * sizeof(void *);

    crt_recon_t rc_tab;
    bx_crt_consts(L, &rc_tab);

    pthread_mutex_lock(&g_crt_lock);
    int rc = 0;
    if (!g_crt.done) rc = bxh_event_create(&g_crt.done, 0);
    if (!rc && !g_crt.segs_host) rc = bxh_host_malloc(&g_crt.segs_host, segs_bytes);
    if (!rc && !g_crt.segs_dev) rc = bxh_malloc(&g_crt.segs_dev, segs_bytes);
    if (!rc && total > g_crt.dev_cap) {
        bxh_event_sync(g_crt.done); /* the old buffer may still be read */
        bxh_free(g_crt.dev);
        g_crt.dev = NULL;
        g_crt.dev_cap = 0;
        rc = bxh_malloc(&g_crt.dev, total);
        if (!rc) g_crt.dev_cap = total;
    }
    if (rc) {
        pthread_mutex_unlock(&g_crt_lock);
        return rc;
    }
    /* the plane table is rewritten below: the previous call must be done with it */
    bxh_event_sync(g_crt.done);

    char *base = (char *)g_crt.dev;
    int8_t *DA = (int8_t *)base, *DB = (int8_t *)(base + off_b8), *R = (int8_t *)(base + off_r);
    int *ea = (int *)(base + off_ea), *eb = (int *)(base + off_eb);
    double *part = (double *)(base + off_part);

    const int a_contig = opa != BX_OP_N; /* A^T: row i of op(A) is column i of A */
    const int b_contig = opb == BX_OP_N; /* B: column j of op(B) is column j of B */
    rc = bxh_oz_exponents(A, lda, M, K, a_contig, 1, ea, part, stream);
    if (!rc) rc = bxh_oz_exponents(B, ldb, N, K, b_contig, 1, eb, part, stream);
    if (!rc) rc = bxh_crt_split(A, lda, M, K, a_contig, ea, bb, L, CRT_MODS, DA, mp, kp, stream);
    if (!rc) rc = bxh_crt_split(B, ldb, N, K, b_contig, eb, bb, L, CRT_MODS, DB, np, kp, stream);
    const int8_t **tab = (const int8_t **)g_crt.segs_host;
    for (int l = 0; l < L; ++l) {
        tab[2 * l] = DA + (size_t)l * mp * kp;
        tab[2 * l + 1] = DB + (size_t)l * np * kp;
    }
    if (!rc) rc = bxh_memcpy_h2d_async(g_crt.segs_dev, g_crt.segs_host,
                                       (size_t)L * 2 * sizeof(void *), stream);
    if (!rc) rc = bxh_crt_gemm(g_crt.segs_dev, L, 1, CRT_MODS, kp, mp, np, R, stream);
    if (!rc)
        rc = bxh_crt_recon(M, N, R, mp, (int64_t)mp * np, &rc_tab, ea, eb, bb, alpha, beta, C,
                           ldc, stream);
    if (!rc) rc = bxh_event_record(g_crt.done, stream);
    pthread_mutex_unlock(&g_crt_lock);
    if (!rc) {
        BX_STAT_ADD(kernel_launches, 8);
        bx_stat_flops(2.0 * M * (double)N * K);
    }
    return rc;
}

/* complex: 3 products per modulus (Karatsuba, exact in residues), see
 * hip/ozaki_kernels.hip; the worst-case bits count 2K real terms */
int bx_crt_default_moduli_z(int K)
{
    for (int L = 2; L <= CRT_MAX_L; ++L)
        if (bx_crt_beta(2 * K, L) >= 56) return L;
    return CRT_MAX_L;
}

int bx_zgemm_crt(int opa, int opb, int M, int N, int K, const double *alpha, const double *A,
                 int64_t lda, const double *B, int64_t ldb, const double *beta, double *C,
                 int64_t ldc, int L, void *stream)
{
    if (M == 0 || N == 0) return 0;
    if (K == 0 || (alpha[0] == 0.0 && alpha[1] == 0.0))
        return bxh_scale(BLASX_Z, M, N, beta, C, ldc, stream);
    if (K > CRT_MAX_K) return BLASX_ERANGE;
    /* |Re P|, |Im P| <= ||a_i|| ||b_j|| over the 2K real parts */
    const int bb = bx_crt_gamma(L);
    const int pad = (M >= 4096 && N >= 4096) ? 2 * bxh_oz_block_m() : bxh_oz_block_m();
    const int mp = (int)rup(M, pad), np = (int)rup(N, pad);
    const int kp = (int)rup(K, 2 * bxh_oz_block_k()); /* an even number of K tiles */
    const int P = 3 * L;

    const size_t a8 = (size_t)P * mp * kp, b8 = (size_t)P * np * kp, r8 = (size_t)P * mp * np;
    const size_t off_b8 = rup(a8, 256), off_r = off_b8 + rup(b8, 256);
    const size_t off_ea = off_r + rup(r8, 256), off_eb = off_ea + rup((size_t)M * 4, 256);
    const size_t off_part = off_eb + rup((size_t)N * 4, 256);
    const size_t total =
        off_part + (size_t)2 * bxh_oz_exp_chunks() * (M > N ? M : N) * sizeof(double);
    const size_t segs_bytes = (size_t)3 * CRT_MAX_L * 2 * sizeof(void *);

    crt_recon_t rc_tab;
    bx_crt_consts(L, &rc_tab);

    pthread_mutex_lock(&g_crt_lock);
    int rc = 0;
    if (!g_crt.done) rc = bxh_event_create(&g_crt.done, 0);
    if (!rc && !g_crt.segs_host) rc = bxh_host_malloc(&g_crt.segs_host, segs_bytes);
    if (!rc && !g_crt.segs_dev) rc = bxh_malloc(&g_crt.segs_dev, segs_bytes);
    if (!rc && total > g_crt.dev_cap) {
        bxh_event_sync(g_crt.done);
        bxh_free(g_crt.dev);
        g_crt.dev = NULL;
        g_crt.dev_cap = 0;
        rc = bxh_malloc(&g_crt.dev, total);
        if (!rc) g_crt.dev_cap = total;
    }
    if (rc) {
        pthread_mutex_unlock(&g_crt_lock);
        return rc;
    }
    bxh_event_sync(g_crt.done);

    char *base = (char *)g_crt.dev;
    int8_t *DA = (int8_t *)base, *DB = (int8_t *)(base + off_b8), *R = (int8_t *)(base + off_r);
    int *ea = (int *)(base + off_ea), *eb = (int *)(base + off_eb);
    double *part = (double *)(base + off_part);

    const int a_contig = opa != BX_OP_N, b_contig = opb == BX_OP_N;
    rc = bxh_oz_exponents_z(A, lda, M, K, a_contig, 1, ea, part, stream);
    if (!rc) rc = bxh_oz_exponents_z(B, ldb, N, K, b_contig, 1, eb, part, stream);
    if (!rc)
        rc = bxh_zcrt_split(A, lda, M, K, a_contig, opa == BX_OP_C, ea, bb, L, CRT_MODS, DA, mp,
                            kp, stream);
    if (!rc)
        rc = bxh_zcrt_split(B, ldb, N, K, b_contig, opb == BX_OP_C, eb, bb, L, CRT_MODS, DB, np,
                            kp, stream);
    const int8_t **tab = (const int8_t **)g_crt.segs_host;
    for (int p = 0; p < P; ++p) {
        tab[2 * p] = DA + (size_t)p * mp * kp;
        tab[2 * p + 1] = DB + (size_t)p * np * kp;
    }
    if (!rc) rc = bxh_memcpy_h2d_async(g_crt.segs_dev, g_crt.segs_host,
                                       (size_t)P * 2 * sizeof(void *), stream);
    if (!rc) rc = bxh_crt_gemm(g_crt.segs_dev, L, 3, CRT_MODS, kp, mp, np, R, stream);
    if (!rc)
        rc = bxh_zcrt_recon(M, N, R, mp, (int64_t)mp * np, &rc_tab, CRT_MODS, ea, eb, bb, alpha,
                            beta, C, ldc, stream);
    if (!rc) rc = bxh_event_record(g_crt.done, stream);
    pthread_mutex_unlock(&g_crt_lock);
    if (!rc) {
        BX_STAT_ADD(kernel_launches, 10);
        bx_stat_flops(8.0 * M * (double)N * K);
    }
    return rc;
}

BX_EXPORT int blasx_zgemm_crt_device(int transa, int transb, int M, int N, int K,
                                     const double *alpha, const double *dA, int64_t lda,
                                     const double *dB, int64_t ldb, const double *beta,
                                     double *dC, int64_t ldc, int moduli, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    const int opa = transa == 111 ? BX_OP_N : transa == 112 ? BX_OP_T : transa == 113 ? BX_OP_C : -1;
    const int opb = transb == 111 ? BX_OP_N : transb == 112 ? BX_OP_T : transb == 113 ? BX_OP_C : -1;
    if (opa < 0 || opb < 0 || M < 0 || N < 0 || K < 0) return BLASX_EINVAL;
    int64_t nrowa = opa == BX_OP_N ? M : K, nrowb = opb == BX_OP_N ? K : N;
    if (lda < (nrowa > 1 ? nrowa : 1) || ldb < (nrowb > 1 ? nrowb : 1) || ldc < (M > 1 ? M : 1))
        return BLASX_EINVAL;
    if (moduli == 0) moduli = bx_crt_default_moduli_z(K);
    if (moduli < 2 || moduli > CRT_MAX_L) return BLASX_EINVAL;
    return bx_zgemm_crt(opa, opb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, moduli,
                        stream ? stream : g_bx.s_comp);
}

BX_EXPORT int blasx_dgemm_crt_device(int transa, int transb, int M, int N, int K, double alpha,
                                     const double *dA, int64_t lda, const double *dB,
                                     int64_t ldb, double beta, double *dC, int64_t ldc,
                                     int moduli, void *stream)
{
    if (bx_ensure_init() != 0) return BLASX_ENOGPU;
    const int opa = transa == 111 ? BX_OP_N : (transa == 112 || transa == 113) ? BX_OP_T : -1;
    const int opb = transb == 111 ? BX_OP_N : (transb == 112 || transb == 113) ? BX_OP_T : -1;
    if (opa < 0 || opb < 0 || M < 0 || N < 0 || K < 0) return BLASX_EINVAL;
    int64_t nrowa = opa == BX_OP_N ? M : K, nrowb = opb == BX_OP_N ? K : N;
    if (lda < (nrowa > 1 ? nrowa : 1) || ldb < (nrowb > 1 ? nrowb : 1) || ldc < (M > 1 ? M : 1))
        return BLASX_EINVAL;
    if (moduli == 0) moduli = bx_crt_default_moduli(K);
    if (moduli < 2 || moduli > CRT_MAX_L) return BLASX_EINVAL;
    return bx_dgemm_crt(opa, opb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, moduli,
                        stream ? stream : g_bx.s_comp);
}

BX_EXPORT int blasx_crt_bits(int K, int moduli)
{
    if (moduli == 0) moduli = bx_crt_default_moduli(K);
    if (moduli < 2 || moduli > CRT_MAX_L || K < 0) return -1;
    return bx_crt_beta(K, moduli);
}
