// gemm_kernels.hip -- hand-written CDNA4 (gfx950) GEMM kernels for the
// per-tile product the reference hands to cuBLAS (cublasDgemm at
// blas/blasx_dgemm.c:87-94, cublasSgemm at blas/blasx_sgemm.c:87,
// cublasZgemm at blas/blasx_zgemm.c:88).
//
// C[m x n] = alpha * sum_s op(A_s) op(B_s) + beta * C, column-major.
//
// One engine serves every precision:
//   * FP64: V_MFMA_F64_16X16X4_F64, block tile 128x128x16
//   * FP32: V_MFMA_F32_16X16X4_F32, block tile 128x128x32
//   * complex128: four FP64 MFMAs per complex MAC ("4M"); conjugation is a
//     sign flip applied when the tile is staged, block tile 128x64x8
// 256 threads = 4 wave64s in a 2x2 grid.  Each wave owns a 64x64 (real) or
// 64x32 (complex) sub-tile held as 16 (8x2) MFMA accumulators.
// Global -> register prefetch of tile k+1 overlaps the MFMAs on tile k,
// which reads from one of two LDS buffers (one barrier per K tile).
// LDS holds both operands K-major-outer ([k][row]) so each lane reads one
// 16-byte vector per MFMA group; rows inside a wave sub-tile are permuted so
// that vector holds the operands of VEC adjacent MFMA blocks.
// A problem's K range may be split into segments with separate base
// pointers: the tile runtime passes one segment per cached K tile, so one
// launch covers a whole row of tile products without re-reading C.
#include <hip/hip_runtime.h>
#include <pthread.h>
#include <stdint.h>

#include "../hiplayer.h"
#include "blasx_ext.h"
#include "gmem.h"

namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

struct cplx {
    double re, im;
};

template <typename E> struct Cfg;
// OUT_ROW(lg, v): the 16x16 MFMA output row held by accumulator element v
// of a lane in lane-group lg = lane / 16 (column = lane % 16).  Measured on
// gfx950: the FP32 16x16x4 MFMA returns rows 4*lg + v, the FP64 16x16x4
// MFMA returns rows lg + 4*v.
template <> struct Cfg<double> {
    static constexpr int BM = 128, BN = 128, BK = 16, VEC = 2, PADA = 2, PADB = 2;
    typedef d4_t acc_t;
    static __device__ __forceinline__ int out_row(int lg, int v) { return lg + 4 * v; }
};
template <> struct Cfg<float> {
    static constexpr int BM = 128, BN = 128, BK = 32, VEC = 4, PADA = 4, PADB = 4;
    typedef f4_t acc_t;
    static __device__ __forceinline__ int out_row(int lg, int v) { return 4 * lg + v; }
};
template <> struct Cfg<cplx> {
    static constexpr int BM = 128, BN = 64, BK = 8, VEC = 1, PADA = 1, PADB = 1;
    typedef d4_t acc_t;
    static __device__ __forceinline__ int out_row(int lg, int v) { return lg + 4 * v; }
};

constexpr int NT = 256;
enum { MAJ_MN = 0, MAJ_K = 1 };

template <typename E, int VEC> struct alignas(16) Vec {
    E e[VEC];
};
// N consecutive elements as one access of N * sizeof(E) bytes
template <typename E, int N> struct alignas(sizeof(E) * N) Lv {
    E e[N];
};

template <typename E> __device__ __forceinline__ E zero_of() { return E(0); }
template <> __device__ __forceinline__ cplx zero_of<cplx>() { return cplx{0.0, 0.0}; }

template <typename E> __device__ __forceinline__ E conj_of(E x) { return x; }
template <> __device__ __forceinline__ cplx conj_of<cplx>(cplx x) { return cplx{x.re, -x.im}; }

// The epilogue's arithmetic, one rounding sequence for every kernel (explicit
// FMAs: no result depends on how the compiler contracts a kernel's code).
// alpha * s, then + beta * c in one FMA; c is not used when beta_zero.
template <typename E>
__device__ __forceinline__ E out_real(E alpha, E s, E beta, E c, bool beta_zero)
{
    const E x = alpha * s;
    return beta_zero ? x : fma(beta, c, x);
}
// complex: alpha (sr + i sim), then + beta c
__device__ __forceinline__ cplx out_cplx(double ar, double ai, double sr, double sim, double br,
                                         double bi, cplx c, bool beta_zero)
{
    double orr = fma(ar, sr, -(ai * sim)), oi = fma(ar, sim, ai * sr);
    if (!beta_zero) {
        orr = fma(br, c.re, fma(-bi, c.im, orr));
        oi = fma(br, c.im, fma(bi, c.re, oi));
    }
    return cplx{orr, oi};
}

__device__ __forceinline__ d4_t mfma(double a, double b, d4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma(float a, float b, f4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Tile loader: ROWS x BK tile of op(X) (rows = the M or N dimension).
//   MAJ_MN: element (r, k) at base[r + k*ld]   (op(A)=A, op(B)=B^T)
//   MAJ_K : element (r, k) at base[k + r*ld]   (op(A)=A^T/A^H, op(B)=B)
// Out-of-range elements are staged as zero.
// ---------------------------------------------------------------------------
template <typename E, int VEC, int ROWS, int BK, int PAD, int MAJ, bool CONJ> struct Loader {
    static constexpr int NL = ROWS * BK / (NT * VEC);
    static constexpr int STR = ROWS + PAD;
    static_assert(NL >= 1, "tile too small for the block");
    Vec<E, VEC> r[NL];

    __device__ __forceinline__ void load(const E *__restrict__ base, int64_t ld, int r0, int rlim,
                                         int64_t k0, int64_t klim, bool vec_ok)
    {
        const int tid = threadIdx.x;
        // Interior tile with aligned operands (block-uniform test): straight
        // vector loads, no per-element bounds checks or branches.
        if (vec_ok && r0 + ROWS <= rlim && k0 + BK <= klim) {
            if constexpr (MAJ == MAJ_MN) {
                constexpr int TPR = ROWS / VEC, RPP = NT / TPR;
                const E *p0 = base + (r0 + (tid % TPR) * VEC) + (k0 + tid / TPR) * ld;
                const int64_t step = (int64_t)RPP * ld;
#pragma unroll
                for (int p = 0; p < NL; ++p) r[p] = gload((const Vec<E, VEC> *)(p0 + p * step));
            } else {
                constexpr int TPR = BK / VEC, RPP = NT / TPR;
                const E *p0 = base + (k0 + (tid % TPR) * VEC) + (int64_t)(r0 + tid / TPR) * ld;
                const int64_t step = (int64_t)RPP * ld;
#pragma unroll
                for (int p = 0; p < NL; ++p) r[p] = gload((const Vec<E, VEC> *)(p0 + p * step));
            }
            if constexpr (CONJ) {
#pragma unroll
                for (int p = 0; p < NL; ++p)
#pragma unroll
                    for (int e = 0; e < VEC; ++e) r[p].e[e] = conj_of(r[p].e[e]);
            }
            return;
        }
        if constexpr (MAJ == MAJ_MN) {
            constexpr int TPR = ROWS / VEC, RPP = NT / TPR;
            const int rr = (tid % TPR) * VEC, kq = tid / TPR;
#pragma unroll
            for (int p = 0; p < NL; ++p) {
                const int64_t gk = k0 + kq + p * RPP;
                const int gr = r0 + rr;
                if (vec_ok && gk < klim && gr + VEC <= rlim) {
                    r[p] = gload((const Vec<E, VEC> *)(base + gr + gk * ld));
                } else {
#pragma unroll
                    for (int e = 0; e < VEC; ++e)
                        r[p].e[e] = (gk < klim && gr + e < rlim) ? gload(base + gr + e + gk * ld)
                                                                 : zero_of<E>();
                }
            }
        } else {
            constexpr int TPR = BK / VEC, RPP = NT / TPR;
            const int kv = (tid % TPR) * VEC, rq = tid / TPR;
#pragma unroll
            for (int p = 0; p < NL; ++p) {
                const int gr = r0 + rq + p * RPP;
                const int64_t gk = k0 + kv;
                if (vec_ok && gr < rlim && gk + VEC <= klim) {
                    r[p] = gload((const Vec<E, VEC> *)(base + gk + (int64_t)gr * ld));
                } else {
#pragma unroll
                    for (int e = 0; e < VEC; ++e)
                        r[p].e[e] = (gr < rlim && gk + e < klim) ? gload(base + gk + e + (int64_t)gr * ld)
                                                                 : zero_of<E>();
                }
            }
        }
        if constexpr (CONJ) {
#pragma unroll
            for (int p = 0; p < NL; ++p)
#pragma unroll
                for (int e = 0; e < VEC; ++e) r[p].e[e] = conj_of(r[p].e[e]);
        }
    }

    // registers -> LDS laid out S[k * STR + row]
    __device__ __forceinline__ void store(E *__restrict__ S) const
    {
        const int tid = threadIdx.x;
        if constexpr (MAJ == MAJ_MN) {
            constexpr int TPR = ROWS / VEC, RPP = NT / TPR;
            const int rr = (tid % TPR) * VEC, kq = tid / TPR;
#pragma unroll
            for (int p = 0; p < NL; ++p)
                *(Vec<E, VEC> *)(S + (kq + p * RPP) * STR + rr) = r[p];
        } else {
            constexpr int TPR = BK / VEC, RPP = NT / TPR;
            const int kv = (tid % TPR) * VEC, rq = tid / TPR;
#pragma unroll
            for (int p = 0; p < NL; ++p)
#pragma unroll
                for (int e = 0; e < VEC; ++e) S[(kv + e) * STR + rq + p * RPP] = r[p].e[e];
        }
    }
};

__host__ __device__ __forceinline__ bool aligned16(const void *p, int64_t ld, int esize)
{
    return ((reinterpret_cast<uintptr_t>(p) & 15) == 0) && (((ld * esize) & 15) == 0);
}

// block tile index -> (tm, tn), walking groups of 8 tile rows so that the
// blocks resident at one time share A row panels and B column panels
__device__ __forceinline__ void tile_coords(int t, int tiles_m, int tiles_n, int &tm, int &tn)
{
    constexpr int G = 8;
    const int per_group = G * tiles_n;
    const int g = t / per_group;
    const int first = g * G;
    const int gsz = min(tiles_m - first, G);
    const int r = t - g * per_group;
    tm = first + r % gsz;
    tn = r / gsz;
}

// ---------------------------------------------------------------------------
// LDS -> MFMA for one staged K tile (real types)
// ---------------------------------------------------------------------------
// Block tile of the register-staged kernels: 2x2 waves, each WTM x WTN
// (real types; 64 x 64 is the large tile, 32 x 32 and 64 x 32 the small ones
// the grouped launches may pick per problem), or Cfg<cplx>'s 128 x 64.  A
// lane reads RV consecutive rows of a stage with one LDS read and so feeds
// RV MFMA blocks: rows inside a wave sub-tile are permuted that way, and
// un-permuted in the epilogue.
template <typename E, int WTM = 64, int WTN = 64> struct Blk {
    static constexpr bool CPLX = sizeof(E) == 16;
    static_assert(!CPLX || (WTM == 64 && WTN == 64), "complex: one tile shape");
    static constexpr int BM = CPLX ? Cfg<E>::BM : 2 * WTM;
    static constexpr int BN = CPLX ? Cfg<E>::BN : 2 * WTN;
    static constexpr int FM = WTM / 16, FN = WTN / 16; // MFMA blocks per wave
    static constexpr int VEC = Cfg<E>::VEC;
    static constexpr int RVA = VEC < FM ? VEC : FM, RVB = VEC < FN ? VEC : FN;
    static constexpr int STRA = BM + Cfg<E>::PADA, STRB = BN + Cfg<E>::PADB;
    static constexpr int A_ELEMS = Cfg<E>::BK * STRA, B_ELEMS = Cfg<E>::BK * STRB;
    static constexpr int ELEMS = 2 * (A_ELEMS + B_ELEMS);
};

// The K value lane group lg (lane / 16) feeds at K step s of a K tile: the
// order the DMA-staged kernels' LDS layouts need (dma64::frag, dmas::frag),
// used by the register-staged block too, so that every kernel and block
// shape sums a C element's products in one order and gives the same bits.
template <typename E> __device__ __forceinline__ int k_of(int s, int lg)
{
    if constexpr (sizeof(E) == 8)
        return 2 * (4 * (s >> 1) + lg) + (s & 1);
    else
        return 4 * (lg + 4 * (s >> 2)) + (s & 3);
}

template <typename E, int WTM, int WTN>
__device__ __forceinline__ void mma_real(const E *__restrict__ As, const E *__restrict__ Bs,
                                         typename Cfg<E>::acc_t (&acc)[WTM / 16][WTN / 16],
                                         int wm, int wn, int lane)
{
    typedef Blk<E, WTM, WTN> T_;
    constexpr int BK = Cfg<E>::BK, FM = T_::FM, FN = T_::FN, RVA = T_::RVA, RVB = T_::RVB;
    const int l16 = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int s = 0; s < BK / 4; ++s) {
        const int k = k_of<E>(s, lg);
        E a[FM], b[FN];
#pragma unroll
        for (int h = 0; h < FM / RVA; ++h) {
            Lv<E, RVA> va = *(const Lv<E, RVA> *)(As + k * T_::STRA + wm * WTM + h * 16 * RVA +
                                                   RVA * l16);
#pragma unroll
            for (int e = 0; e < RVA; ++e) a[h * RVA + e] = va.e[e];
        }
#pragma unroll
        for (int h = 0; h < FN / RVB; ++h) {
            Lv<E, RVB> vb = *(const Lv<E, RVB> *)(Bs + k * T_::STRB + wn * WTN + h * 16 * RVB +
                                                   RVB * l16);
#pragma unroll
            for (int e = 0; e < RVB; ++e) b[h * RVB + e] = vb.e[e];
        }
#pragma unroll
        for (int mb = 0; mb < FM; ++mb)
#pragma unroll
            for (int nb = 0; nb < FN; ++nb) acc[mb][nb] = mfma(a[mb], b[nb], acc[mb][nb]);
    }
}

// complex: acc[mb][nb][0] = real part, [1] = imaginary part
__device__ __forceinline__ void mma_cplx(const cplx *__restrict__ As, const cplx *__restrict__ Bs,
                                         d4_t (&acc)[4][2][2], int wm, int wn, int lane)
{
    constexpr int BK = Cfg<cplx>::BK;
    constexpr int STRA = Cfg<cplx>::BM + Cfg<cplx>::PADA, STRB = Cfg<cplx>::BN + Cfg<cplx>::PADB;
    const int l16 = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int s = 0; s < BK / 4; ++s) {
        const int k = s * 4 + lg;
        cplx a[4], b[2];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) a[mb] = As[k * STRA + wm * 64 + mb * 16 + l16];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) b[nb] = Bs[k * STRB + wn * 32 + nb * 16 + l16];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const double nai = -a[mb].im;
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                acc[mb][nb][0] = mfma(a[mb].re, b[nb].re, acc[mb][nb][0]);
                acc[mb][nb][1] = mfma(a[mb].re, b[nb].im, acc[mb][nb][1]);
                acc[mb][nb][0] = mfma(nai, b[nb].im, acc[mb][nb][0]);
                acc[mb][nb][1] = mfma(a[mb].im, b[nb].re, acc[mb][nb][1]);
            }
        }
    }
}

template <typename E> struct Smem {
    static constexpr int A_ELEMS = Blk<E>::A_ELEMS, B_ELEMS = Blk<E>::B_ELEMS;
    static constexpr int ELEMS = Blk<E>::ELEMS;
};

// segment i of the table, or the by-value segment when there is no table;
// fields are read through the global address space one by one (selecting
// between a table pointer and a local would make every access FLAT)
__device__ __forceinline__ bxh_seg_t seg_at(const bxh_seg_t *segs, const bxh_seg_t &seg0, int i)
{
    if (!segs) return seg0;
    const __attribute__((address_space(1))) bxh_seg_t *g =
        (const __attribute__((address_space(1))) bxh_seg_t *)(segs + i);
    bxh_seg_t s;
    s.A = g->A;
    s.B = g->B;
    s.lda = g->lda;
    s.ldb = g->ldb;
    s.k = g->k;
    return s;
}

// ---------------------------------------------------------------------------
// One block tile of one problem.
// ---------------------------------------------------------------------------
// whether a rank-k update writes element (row, col) of its diagonal block
__device__ __forceinline__ bool tri_keep(int tri, int row, int col)
{
    return (!(tri & BX_TRI_LOWER) || row >= col) && (!(tri & BX_TRI_UPPER) || row <= col);
}

template <typename E, int MAJA, int MAJB, bool CJA, bool CJB, int WTM = 64, int WTN = 64>
__device__ __forceinline__ void gemm_block(const bxh_problem_t &P, const bxh_seg_t *segs,
                                           const bxh_seg_t &seg0, int tm, int tn, E *smem)
{
    typedef Cfg<E> C_;
    typedef Blk<E, WTM, WTN> T_;
    constexpr int BM = T_::BM, BN = T_::BN, BK = C_::BK, VEC = C_::VEC;
    constexpr int FM = T_::FM, FN = T_::FN, RVA = T_::RVA, RVB = T_::RVB;
    constexpr bool CPLX = sizeof(E) == 16;
    E *As0 = smem;
    E *Bs0 = smem + 2 * T_::A_ELEMS;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = tm * BM, n0 = tn * BN;
    const int nseg = P.nseg;

    Loader<E, VEC, BM, BK, C_::PADA, MAJA, CJA> la;
    Loader<E, VEC, BN, BK, C_::PADB, MAJB, CJB> lb;

    typedef typename C_::acc_t acc_t;
    acc_t accr[FM][FN];
    d4_t accz[4][2][2];
    if constexpr (CPLX) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) accz[i][j][0] = accz[i][j][1] = d4_t{0, 0, 0, 0};
    } else {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) accr[i][j] = acc_t{0, 0, 0, 0};
    }

    auto load_both = [&](int si, int64_t k0) {
        const bxh_seg_t S = seg_at(segs, seg0, si);
        const E *A = reinterpret_cast<const E *>(S.A);
        const E *B = reinterpret_cast<const E *>(S.B);
        la.load(A, S.lda, m0, P.m, k0, S.k, aligned16(A, S.lda, sizeof(E)));
        lb.load(B, S.ldb, n0, P.n, k0, S.k, aligned16(B, S.ldb, sizeof(E)));
    };

    // first non-empty segment
    int si = 0;
    int64_t k0 = 0;
    while (si < nseg && seg_at(segs, seg0, si).k <= 0) ++si;
    bool have = si < nseg;
    if (have) {
        load_both(si, k0);
        la.store(As0);
        lb.store(Bs0);
    }
    __syncthreads();
    int buf = 0;
    while (have) {
        int nsi = si;
        int64_t nk0 = k0 + BK;
        while (nsi < nseg && nk0 >= seg_at(segs, seg0, nsi).k) {
            ++nsi;
            nk0 = 0;
        }
        const bool nhave = nsi < nseg;
        if (nhave) load_both(nsi, nk0);
        const E *As = As0 + buf * T_::A_ELEMS;
        const E *Bs = Bs0 + buf * T_::B_ELEMS;
        if constexpr (CPLX)
            mma_cplx(reinterpret_cast<const cplx *>(As), reinterpret_cast<const cplx *>(Bs), accz,
                     wm, wn, lane);
        else
            mma_real<E, WTM, WTN>(As, Bs, accr, wm, wn, lane);
        if (nhave) {
            la.store(As0 + (buf ^ 1) * T_::A_ELEMS);
            lb.store(Bs0 + (buf ^ 1) * T_::B_ELEMS);
        }
        __syncthreads();
        si = nsi;
        k0 = nk0;
        have = nhave;
        buf ^= 1;
    }

    // ---- epilogue: C = alpha * acc + beta * C (on the stored triangle of a
    // rank-k update's diagonal block only) ----
    E *Cp = reinterpret_cast<E *>(P.C);
    const int64_t ldc = P.ldc;
    const int l16 = lane & 15, lg = lane >> 4;
    const int tri = P.tri, toff = P.tri_off;
    if constexpr (CPLX) {
        const double ar = P.alpha_re, ai = P.alpha_im, br = P.beta_re, bi = P.beta_im;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const int col = n0 + wn * 32 + nb * 16 + l16;
                if (col >= P.n) continue;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = m0 + wm * 64 + mb * 16 + Cfg<cplx>::out_row(lg, v);
                    if (row >= P.m || !tri_keep(tri, row + toff, col)) continue;
                    cplx *c = Cp + row + (int64_t)col * ldc;
                    const cplx c0 = P.beta_zero ? cplx{0.0, 0.0} : gload(c);
                    cplx o = out_cplx(ar, ai, accz[mb][nb][0][v], accz[mb][nb][1][v], br, bi, c0,
                                      P.beta_zero);
                    if ((tri & BX_TRI_HERM) && row + toff == col) o.im = 0.0;
                    gstore(c, o);
                }
            }
    } else {
        const E alpha = (E)P.alpha_re, beta = (E)P.beta_re;
        const bool cvec = aligned16(Cp, ldc, sizeof(E));
        (void)VEC;
#pragma unroll
        for (int h = 0; h < FM / RVA; ++h)
#pragma unroll
            for (int nb = 0; nb < FN; ++nb) {
                const int col = n0 + wn * WTN + (nb / RVB) * 16 * RVB + RVB * l16 + nb % RVB;
                if (col >= P.n) continue;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rbase = m0 + wm * WTM + h * 16 * RVA + RVA * C_::out_row(lg, v);
                    E *c = Cp + rbase + (int64_t)col * ldc;
                    if (cvec && rbase + RVA <= P.m && tri == 0) {
                        Lv<E, RVA> out;
                        if (P.beta_zero) {
#pragma unroll
                            for (int e = 0; e < RVA; ++e)
                                out.e[e] = out_real(alpha, accr[h * RVA + e][nb][v], beta, E(0), true);
                        } else {
                            const Lv<E, RVA> old = gload((const Lv<E, RVA> *)c);
#pragma unroll
                            for (int e = 0; e < RVA; ++e)
                                out.e[e] = out_real(alpha, accr[h * RVA + e][nb][v], beta, old.e[e],
                                                    false);
                        }
                        gstore((Lv<E, RVA> *)c, out);
                    } else {
#pragma unroll
                        for (int e = 0; e < RVA; ++e) {
                            if (rbase + e >= P.m || !tri_keep(tri, rbase + e + toff, col)) continue;
                            const E old = P.beta_zero ? E(0) : gload(c + e);
                            gstore(c + e, out_real(alpha, accr[h * RVA + e][nb][v], beta, old,
                                                   (bool)P.beta_zero));
                        }
                    }
                }
            }
    }
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
template <typename E, int MAJA, int MAJB, bool CJA, bool CJB>
__global__ __launch_bounds__(NT, 2) void k_gemm_plain(bxh_problem_t P, bxh_seg_t S)
{
    __shared__ __attribute__((aligned(16))) E smem[Smem<E>::ELEMS];
    int b = blockIdx.x;
    const int nb = gridDim.x;
    // The dispatcher places block b on XCD b % 8.  Renumber so each XCD
    // works on a contiguous range of tiles (shared panels stay in its L2).
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
    int tm, tn;
    tile_coords(b, P.tiles_m, P.tiles_n, tm, tn);
    gemm_block<E, MAJA, MAJB, CJA, CJB>(P, nullptr, S, tm, tn, smem);
}

// ---------------------------------------------------------------------------
// FP64 GEMM with global -> LDS DMA staging, for problems without edge tiles
// (M, N multiples of 128, K a multiple of 16, 16-byte aligned operands).
//
// Why: measured on the 32768^3 NN problem, the register-staged kernel above
// ran at 67.3 TFLOP/s, 75.5 with its LDS stores removed and 75.9 with MFMAs
// and LDS reads only: staging each K tile through VGPRs and DS_WRITEs costs
// 11%.  Here every operand byte goes from global memory straight to LDS
// (GLOBAL_LOAD_LDS_DWORDX4: 16 bytes per lane at M0 + 16 lane), which
// constrains the LDS layouts to what one instruction writes, 1 KB runs:
//  * an operand whose rows are contiguous in memory (MAJ_MN: op(A) = A,
//    op(B) = B^T) is staged [k][row], rows of 128 doubles padded to 136 (a
//    K row is one 1 KB run); the MFMA reads one double per lane;
//  * an operand with K contiguous (MAJ_K) is staged [row][k], 16 doubles per
//    row as eight 16-byte chunks, chunk c at position c ^ ((row >> 1) & 7)
//    (the 16 lanes of a fragment read then cover all 64 banks); one 16-byte
//    read gives a lane the operands of two consecutive K steps.
// K values are assigned to the MFMA's lane groups so that both layouts
// serve them: K step s, lane group lg uses k = 2 (4 (s / 2) + lg) + s % 2.
// Two LDS stages (35 KB each at most), two blocks per CU; step t starts the
// DMA of tile t + 1 into the other stage (read last in step t - 1, before
// the last barrier), runs the 64 MFMAs of tile t, waits for its own DMAs
// and ends with one barrier.  The DMA is inline assembly, and the waits are
// explicit (`s_waitcnt vmcnt(0)`, bare `s_barrier`): with the builtin the
// compiler would also wait for the in-flight DMA before every LDS read.
// ---------------------------------------------------------------------------
namespace dma64 {
constexpr int BM = 128, BK = 16, STR_MN = 136;
template <int MAJ> constexpr int stage_doubles() { return MAJ == MAJ_MN ? BK * STR_MN : BM * BK; }

// element (r, k) of a MAJ_K stage, in doubles
__device__ __forceinline__ int pos_k(int r, int k)
{
    return r * BK + ((((k >> 1) ^ ((r >> 1) & 7))) << 1) + (k & 1);
}

// 16 bytes per lane from global memory to LDS (M0 + 16 lane).  M0 is a
// reserved register the compiler does not allocate; nothing else in this
// kernel uses M0 (checked in its gfx950 assembly).
__device__ __forceinline__ void dma16(const double *g, const double *lds_wave_base)
{
    const uint32_t l =
        (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const double *)lds_wave_base;
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off"
                 :
                 : "s"(__builtin_amdgcn_readfirstlane(l)), "v"(g)
                 : "memory");
}

__device__ __forceinline__ void wait_dma()
{
    // vmcnt(0), the other counters at their maximum
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xF << 8));
}

// this wave's four DMA instructions for one operand's K tile k0 (rows r0..)
// EDGE: the operand has rlim rows; a lane whose rows lie beyond reads the
// last ones instead (rows r0 + i >= rlim of the stage then hold copies that
// feed only output rows or columns the epilogue does not write).  A MAJ_MN
// lane reads rows 2j, 2j + 1: clamped to the even j with 2j <= rlim - 1, its
// 16 bytes stay inside the 16-byte granule of element rlim - 1 (even ld,
// aligned base), so inside the allocation.
template <int MAJ, bool EDGE = false>
__device__ __forceinline__ void dma_tile(const double *X, int64_t ld, int r0, int64_t k0,
                                         double *stage, int wave, int lane, int rlim = 0)
{
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int q = wave * 4 + i; // 1 KB run of the stage
        if constexpr (MAJ == MAJ_MN) {
            // K row q: doubles 2 lane, 2 lane + 1 of rows r0 .. r0 + 127
            int row = r0 + 2 * lane;
            if constexpr (EDGE) row = min(row, (rlim - 1) & ~1);
            dma16(X + row + (k0 + q) * ld, stage + q * STR_MN);
        } else {
            // rows 8q .. 8q + 7, lane -> row 8q + lane / 8, position lane % 8
            const int r = 8 * q + (lane >> 3);
            const int c = (lane & 7) ^ ((r >> 1) & 7);
            int row = r0 + r;
            if constexpr (EDGE) row = min(row, rlim - 1);
            dma16(X + (k0 + 2 * c) + (int64_t)row * ld, stage + q * 128);
        }
    }
}

// The last kr < 16 columns of K (from k0) of one operand, staged by
// ordinary bounds-checked loads in the layout dma_tile writes, columns kr..15
// zero: a K tile the 16-byte DMA cannot take.  256 threads, 8 elements each,
// consecutive threads on consecutive addresses.
// EDGE: rows r0 + row >= rlim are staged as zero too.
template <int MAJ, bool EDGE = false>
__device__ __forceinline__ void stage_rem(const double *X, int64_t ld, int r0, int64_t k0, int kr,
                                          double *stage, int tid, int rlim = 0)
{
    // two loads in flight per lane: the 128 accumulators stay live across this
#pragma unroll 2
    for (int i = 0; i < 8; ++i) {
        const int e = tid + 256 * i;
        if constexpr (MAJ == MAJ_MN) {
            const int row = e & 127, k = e >> 7;
            const bool in = k < kr && (!EDGE || r0 + row < rlim);
            stage[k * STR_MN + row] = in ? gload(X + (r0 + row) + (k0 + k) * ld) : 0.0;
        } else {
            const int k = e & 15, row = e >> 4;
            const bool in = k < kr && (!EDGE || r0 + row < rlim);
            stage[pos_k(row, k)] = in ? gload(X + (k0 + k) + (int64_t)(r0 + row) * ld) : 0.0;
        }
    }
}

// the operand value lane (l16, lg) feeds block b of K step s
template <int MAJ>
__device__ __forceinline__ double frag(const double *stage, int row, int s, int lg)
{
    const int k = 2 * (4 * (s >> 1) + lg) + (s & 1);
    if constexpr (MAJ == MAJ_MN)
        return stage[k * STR_MN + row];
    else
        return stage[pos_k(row, k)];
}
} // namespace dma64

namespace dma64 {
// LDS the DMA block needs at most (doubles): two stages of the larger layout
constexpr int SMEM_DOUBLES = 2 * 2 * (BK * STR_MN);
} // namespace dma64

// One 128x128 block tile (tm, tn) of P over all its segments (a plain GEMM
// has one, the tile runtime one per cached K tile), operands DMA-staged; the
// K tiles of consecutive segments run as one pipeline.
// MASK: P may be a rank-k diagonal block (P.tri), whose epilogue writes only
// the stored triangle; the plain kernels instantiate it without.
// KTAIL: segments may have any K; each one's last K % 16 columns take one
// more step staged by stage_rem (the plain kernels for ragged K).
// EDGE (with KTAIL): the tile may cross row P.m or column P.n: operand rows
// beyond are clamped to the last ones (dma_tile) and the epilogue writes
// only the elements inside C.
template <int MAJA, int MAJB, bool MASK = false, bool NESTED = true, bool KTAIL = false,
          bool EDGE = false>
__device__ __forceinline__ void dma64_block(const bxh_problem_t &P, const bxh_seg_t *segs,
                                            const bxh_seg_t &seg0, int tm, int tn, double *smem)
{
    using namespace dma64;
    constexpr int SA = stage_doubles<MAJA>(), SB = stage_doubles<MAJB>();
    const int m0 = tm * BM, n0 = tn * BM;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lg = lane >> 4;

    d4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = d4_t{0, 0, 0, 0};

    auto stageA = [&](int buf) { return smem + buf * SA; };
    auto stageB = [&](int buf) { return smem + 2 * SA + buf * SB; };
    if constexpr (KTAIL) {
        // Per segment: its full K tiles in the DMA pipeline below, then, if
        // K % 16 != 0, one step on its last columns staged by stage_rem (the
        // prefetch of the next segment's first tile moves to that step).  No
        // DMA is in flight while the tail loads are issued or waited for.
        auto issue = [&](const bxh_seg_t &sg, int64_t k, int b) {
            dma_tile<MAJA, EDGE>((const double *)sg.A, sg.lda, m0, k, stageA(b), wave, lane, P.m);
            dma_tile<MAJB, EDGE>((const double *)sg.B, sg.ldb, n0, k, stageB(b), wave, lane, P.n);
        };
        auto multiply = [&](int b) {
            const double *As = stageA(b), *Bs = stageB(b);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                double a[4], bb[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a[i] = frag<MAJA>(As, wm * 64 + i * 16 + l16, s, lg);
                    bb[i] = frag<MAJB>(Bs, wn * 64 + i * 16 + l16, s, lg);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = mfma(a[i], bb[j], acc[i][j]);
            }
        };
        auto find = [&](int i, bxh_seg_t &out) {
            for (; i < P.nseg; ++i) {
                out = seg_at(segs, seg0, i);
                if (out.k > 0) break;
            }
            return i;
        };
        bxh_seg_t cur = seg0;
        int si = find(0, cur);
        // stage 0 holds the current segment's first full tile when it has one
        if (si < P.nseg && cur.k >= BK) {
            issue(cur, 0, 0);
            wait_dma();
            __builtin_amdgcn_s_barrier();
        }
        int buf = 0;
        while (si < P.nseg) {
            const int64_t kfull = cur.k & ~(int64_t)(BK - 1);
            const int kr = (int)(cur.k - kfull);
            bxh_seg_t nxt = cur;
            const int sn = find(si + 1, nxt);
            const bool pre = sn < P.nseg && nxt.k >= BK;
            if (kfull > 0) {
                for (int64_t k0 = 0; k0 + BK < kfull; k0 += BK) {
                    issue(cur, k0 + BK, buf ^ 1);
                    multiply(buf);
                    wait_dma();
                    __builtin_amdgcn_s_barrier();
                    buf ^= 1;
                }
                if (!kr && pre) issue(nxt, 0, buf ^ 1);
                multiply(buf);
                wait_dma();
                __builtin_amdgcn_s_barrier();
                buf ^= 1;
            }
            if (kr) {
                // stage `buf` was last read two steps ago, before the last barrier
                stage_rem<MAJA, EDGE>((const double *)cur.A, cur.lda, m0, kfull, kr, stageA(buf),
                                      tid, P.m);
                stage_rem<MAJB, EDGE>((const double *)cur.B, cur.ldb, n0, kfull, kr, stageB(buf),
                                      tid, P.n);
                __syncthreads();
                if (pre) issue(nxt, 0, buf ^ 1);
                multiply(buf);
                wait_dma();
                __builtin_amdgcn_s_barrier();
                buf ^= 1;
            }
            si = sn;
            cur = nxt;
        }
    } else if constexpr (NESTED) {
        // NESTED picks the K loop: nested per segment, or one loop walking the
        // segments (kept where it measured faster: see the kernels below)
        // The segments run as one pipeline of K tiles: stage `buf` holds the
        // tile being multiplied while the next one is DMA'd into the other.  An
        // inner loop takes every tile of one segment but its last, with that
        // segment's fields fixed (so addresses advance by a constant step); the
        // last tile prefetches the first of the next non-empty segment.  A
        // segment's table entry is read once.
        auto issue = [&](const bxh_seg_t &sg, int64_t k, int b) {
            dma_tile<MAJA>((const double *)sg.A, sg.lda, m0, k, stageA(b), wave, lane);
            dma_tile<MAJB>((const double *)sg.B, sg.ldb, n0, k, stageB(b), wave, lane);
        };
        auto multiply = [&](int b) {
            const double *As = stageA(b), *Bs = stageB(b);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                double a[4], bb[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a[i] = frag<MAJA>(As, wm * 64 + i * 16 + l16, s, lg);
                    bb[i] = frag<MAJB>(Bs, wn * 64 + i * 16 + l16, s, lg);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = mfma(a[i], bb[j], acc[i][j]);
            }
        };
        // the first non-empty segment at index i or later (P.nseg if none),
        // its fields in `out`
        auto find = [&](int i, bxh_seg_t &out) {
            for (; i < P.nseg; ++i) {
                out = seg_at(segs, seg0, i);
                if (out.k > 0) break;
            }
            return i;
        };
        bxh_seg_t cur = seg0;
        int si = find(0, cur);
        if (si < P.nseg) {
            issue(cur, 0, 0);
            wait_dma();
            __builtin_amdgcn_s_barrier();
        }
        int buf = 0;
        while (si < P.nseg) {
            int64_t k0 = 0;
            for (; k0 + BK < cur.k; k0 += BK) {
                issue(cur, k0 + BK, buf ^ 1);
                multiply(buf);
                wait_dma();
                __builtin_amdgcn_s_barrier();
                buf ^= 1;
            }
            bxh_seg_t nxt = cur;
            const int sn = find(si + 1, nxt);
            if (sn < P.nseg) issue(nxt, 0, buf ^ 1);
            multiply(buf);
            wait_dma();
            __builtin_amdgcn_s_barrier();
            buf ^= 1;
            si = sn;
            cur = nxt;
        }
    } else {
        // The K tile being DMA'd is tile k0 / 16 of segment si, whose fields are
        // kept in `cur` (the table is read once per segment).  advance() moves to
        // the next 16 of K, else to the first non-empty later segment; si reaches
        // P.nseg when there is none.
        int si = 0;
        int64_t k0 = 0;
        bxh_seg_t cur = seg_at(segs, seg0, 0);
        auto next_nonempty = [&]() {
            while (si < P.nseg) {
                cur = seg_at(segs, seg0, si);
                if (cur.k > 0) return;
                ++si;
            }
        };
        auto advance = [&]() {
            k0 += BK;
            if (k0 < cur.k) return;
            k0 = 0;
            ++si;
            next_nonempty();
        };
        auto issue = [&](int buf) {
            dma_tile<MAJA>((const double *)cur.A, cur.lda, m0, k0, stageA(buf), wave, lane);
            dma_tile<MAJB>((const double *)cur.B, cur.ldb, n0, k0, stageB(buf), wave, lane);
        };
        next_nonempty();
        const bool any = si < P.nseg;
        if (any) {
            issue(0);
            wait_dma();
            __builtin_amdgcn_s_barrier();
        }
        int buf = 0;
        bool more = any; // the tile in stage `buf` is valid
        while (more) {
            advance();
            more = si < P.nseg;
            if (more) issue(buf ^ 1);
            const double *As = stageA(buf), *Bs = stageB(buf);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                double a[4], bb[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a[i] = frag<MAJA>(As, wm * 64 + i * 16 + l16, s, lg);
                    bb[i] = frag<MAJB>(Bs, wn * 64 + i * 16 + l16, s, lg);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = mfma(a[i], bb[j], acc[i][j]);
            }
            wait_dma();
            __builtin_amdgcn_s_barrier();
            buf ^= 1;
        }
    }

    double *Cp = (double *)P.C;
    const double alpha = P.alpha_re, beta = P.beta_re;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + wn * 64 + j * 16 + l16;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm * 64 + i * 16 + Cfg<double>::out_row(lg, v);
                if constexpr (MASK)
                    if (!tri_keep(P.tri, row + P.tri_off, col)) continue;
                if constexpr (EDGE)
                    if (row >= P.m || col >= P.n) continue;
                double *c = Cp + row + (int64_t)col * P.ldc;
                const double old = P.beta_zero ? 0.0 : gload(c);
                gstore(c, out_real(alpha, acc[i][j][v], beta, old, (bool)P.beta_zero));
            }
        }
}

template <int MAJA, int MAJB>
__global__ __launch_bounds__(NT, 2) void k_dgemm_dma(bxh_problem_t P, bxh_seg_t S)
{
    __shared__ __attribute__((aligned(16))) double smem[dma64::SMEM_DOUBLES];
    int b = blockIdx.x;
    const int nb = gridDim.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
    int tm, tn;
    tile_coords(b, P.tiles_m, P.tiles_n, tm, tn);
    // the nested K loop for NN only (1% faster at 32768^3); NT and TN ran 7-11%
    // slower with it and TT no faster, so they keep the segment-walking loop
    dma64_block<MAJA, MAJB, false, MAJA == MAJ_MN && MAJB == MAJ_K>(P, nullptr, S, tm, tn, smem);
}

// A ragged DGEMM: any M, N and K, aligned operands (edge tiles clamp their
// operand rows and mask their stores, K tails take one register-staged step).
template <int MAJA, int MAJB>
__global__ __launch_bounds__(NT, 2) void k_dgemm_dma_kt(bxh_problem_t P, bxh_seg_t S)
{
    __shared__ __attribute__((aligned(16))) double smem[dma64::SMEM_DOUBLES];
    // blocks go to the 8 XCDs round-robin (block b on XCD b % 8): give XCD x
    // the contiguous tile range [x q + min(x, r), ...) of its q or q + 1
    // blocks (nb = 8 q + r), so its blocks share A and B panels in its L2
    // for any grid size
    const int nb = gridDim.x, x = blockIdx.x & 7, q = nb >> 3, r = nb & 7;
    const int b = x * q + (x < r ? x : r) + (blockIdx.x >> 3);
    int tm, tn;
    tile_coords(b, P.tiles_m, P.tiles_n, tm, tn);
    dma64_block<MAJA, MAJB, false, true, true, true>(P, nullptr, S, tm, tn, smem);
}

// ---------------------------------------------------------------------------
// ZGEMM with global -> LDS DMA staging (M a multiple of 128, N of 64, K of 8;
// complex128 operands are always 16-byte aligned).  Same scheme as the FP64
// kernel; a complex element is exactly one DMA lane's 16 bytes and one
// 16-byte fragment read, so K steps map to lane groups directly (step s,
// group lg: k = 4 s + lg).  A row-contiguous operand is staged [k][row]
// unpadded (a fragment read is 16 lanes x 16 bytes = one 256-byte LDS pass);
// a K-contiguous one [row][k], 8 elements per row, element k of row r at
// position k ^ ((r >> 1) & 7).  Conjugation is a sign flip of the imaginary
// part as the fragment is read.  4M: four FP64 MFMAs per complex MAC.
// ---------------------------------------------------------------------------
namespace dmaz {
constexpr int BM = 128, BN = 64, BK = 8;
template <int MAJ, int ROWS> constexpr int stage_elems() { return BK * ROWS; }

template <int MAJ, int ROWS>
__device__ __forceinline__ void dma_tile(const cplx *X, int64_t ld, int r0, int64_t k0, cplx *stage,
                                         int wave, int lane)
{
    constexpr int RUNS = ROWS * BK / 64; // 1 KB runs per stage
#pragma unroll
    for (int i = 0; i < RUNS / 4; ++i) {
        const int q = wave * (RUNS / 4) + i;
        if constexpr (MAJ == MAJ_MN) {
            // K row q / (ROWS / 64), elements 64 (q % (ROWS / 64)) + lane
            constexpr int PER = ROWS / 64;
            const int kk = q / PER, rr = (q % PER) * 64 + lane;
            dma64::dma16((const double *)(X + (r0 + rr) + (k0 + kk) * ld),
                         (const double *)(stage + kk * ROWS + (q % PER) * 64));
        } else {
            // rows 8q .. 8q + 7, lane -> row 8q + lane / 8, position lane % 8
            const int r = 8 * q + (lane >> 3);
            const int k = (lane & 7) ^ ((r >> 1) & 7);
            dma64::dma16((const double *)(X + (k0 + k) + (int64_t)(r0 + r) * ld),
                         (const double *)(stage + q * 64));
        }
    }
}

template <int MAJ, int ROWS, bool CJ>
__device__ __forceinline__ cplx frag(const cplx *stage, int row, int s, int lg)
{
    const int k = 4 * s + lg;
    const cplx v = MAJ == MAJ_MN ? stage[k * ROWS + row]
                                 : stage[row * BK + (k ^ ((row >> 1) & 7))];
    return CJ ? cplx{v.re, -v.im} : v;
}
} // namespace dmaz

template <int MAJA, int MAJB, bool CJA, bool CJB, bool MASK = false, bool NESTED = true>
__device__ __forceinline__ void dmaz_block(const bxh_problem_t &P, const bxh_seg_t *segs,
                                           const bxh_seg_t &seg0, int tm, int tn, cplx *smem);

template <int MAJA, int MAJB, bool CJA, bool CJB>
__global__ __launch_bounds__(NT, 2) void k_zgemm_dma(bxh_problem_t P, bxh_seg_t S)
{
    __shared__ __attribute__((aligned(16))) cplx smem[2 * (dmaz::BM + dmaz::BN) * dmaz::BK];
    int b = blockIdx.x;
    const int nb = gridDim.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
    int tm, tn;
    tile_coords(b, P.tiles_m, P.tiles_n, tm, tn);
    // the segment-walking K loop: the nested one ran 6% slower on every
    // transpose at 32768^3
    dmaz_block<MAJA, MAJB, CJA, CJB, false, false>(P, nullptr, S, tm, tn, smem);
}

// ---------------------------------------------------------------------------
// SGEMM with global -> LDS DMA staging (M, N multiples of 128, K of 32,
// 16-byte aligned operands).  Same scheme as the FP64 kernel; a DMA lane
// carries four floats:
//  * a row-contiguous operand is staged [k][row], 128 floats per K row, its
//    16-byte chunks (four rows) at position chunk ^ (4 ((k >> 2) & 3)) so
//    that the four lane groups of a fragment read (K rows 4 apart) fall on
//    different banks; one DMA instruction fills two K rows;
//  * a K-contiguous operand is staged [row][k], 32 floats per row as eight
//    16-byte chunks (four K values), chunk c at position c ^ ((row >> 1) & 7);
//    one 16-byte read gives a lane four K steps.
// K step s, lane group lg: k = 4 (lg + 4 (s / 4)) + s % 4.
// ---------------------------------------------------------------------------
namespace dmas {
constexpr int BM = 128, BK = 32;
constexpr int STAGE = BM * BK; // floats per operand per stage (both layouts)

__device__ __forceinline__ int pos_mn(int row, int k)
{
    return k * BM + ((((row >> 2) ^ (((k >> 2) & 3) << 2))) << 2) + (row & 3);
}
__device__ __forceinline__ int pos_k(int row, int k)
{
    return row * BK + ((((k >> 2) ^ ((row >> 1) & 7))) << 2) + (k & 3);
}

template <int MAJ>
__device__ __forceinline__ void dma_tile(const float *X, int64_t ld, int r0, int64_t k0,
                                         float *stage, int wave, int lane)
{
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int q = wave * 4 + i; // 1 KB run of the stage
        if constexpr (MAJ == MAJ_MN) {
            // K rows 2q, 2q + 1; lane -> K row 2q + lane / 32, position lane % 32
            const int k = 2 * q + (lane >> 5), p = lane & 31;
            const int chunk = p ^ (((k >> 2) & 3) << 2);
            dma64::dma16((const double *)(X + (r0 + 4 * chunk) + (k0 + k) * ld),
                         (const double *)(stage + q * 256));
        } else {
            // rows 8q .. 8q + 7, lane -> row 8q + lane / 8, position lane % 8
            const int r = 8 * q + (lane >> 3);
            const int c = (lane & 7) ^ ((r >> 1) & 7);
            dma64::dma16((const double *)(X + (k0 + 4 * c) + (int64_t)(r0 + r) * ld),
                         (const double *)(stage + q * 256));
        }
    }
}

template <int MAJ> __device__ __forceinline__ float frag(const float *stage, int row, int s, int lg)
{
    const int k = 4 * (lg + 4 * (s >> 2)) + (s & 3);
    return MAJ == MAJ_MN ? stage[pos_mn(row, k)] : stage[pos_k(row, k)];
}
} // namespace dmas

// the segment walk shared by the DMA-staged blocks: the K tile being DMA'd
// is tile k0 / BK of segment si, whose fields are kept in `cur` (the table
// is read once per segment); advance() moves to the next BK of K, else to
// the first non-empty later segment, and si reaches nseg when there is none
template <int BK> struct SegWalk {
    const bxh_seg_t *segs;
    const bxh_seg_t &seg0;
    int nseg, si;
    int64_t k0;
    bxh_seg_t cur;
    __device__ __forceinline__ SegWalk(const bxh_seg_t *s, const bxh_seg_t &s0, int n)
        : segs(s), seg0(s0), nseg(n), si(0), k0(0)
    {
        next_nonempty();
    }
    __device__ __forceinline__ void next_nonempty()
    {
        while (si < nseg) {
            cur = seg_at(segs, seg0, si);
            if (cur.k > 0) return;
            ++si;
        }
    }
    __device__ __forceinline__ bool valid() const { return si < nseg; }
    __device__ __forceinline__ void advance()
    {
        k0 += BK;
        if (k0 < cur.k) return;
        k0 = 0;
        ++si;
        next_nonempty();
    }
};

// every segment of P has K a multiple of bk (or is empty) and 16-byte
// aligned operands, and P has no edge tiles (bm x bn blocks)
// Whether block tile (tm, tn) of P can run on the DMA-staged block: a full
// bm x bn tile (the DMA blocks write C without bounds checks) and every
// segment 16-byte aligned with K a multiple of bk, or any K when the block
// stages K tails itself (ktail: the FP64 block).  A problem with edge tiles
// or a ragged K then runs its interior on the DMA block and only its edge
// tiles on the register-staged block.
__device__ __forceinline__ bool dma_tile_ok(const bxh_problem_t &P, const bxh_seg_t *segs,
                                            int tm, int tn, int bm, int bn, int bk, int esize,
                                            bool ktail)
{
    if ((tm + 1) * bm > P.m || (tn + 1) * bn > P.n || P.nseg < 1) return false;
    const bxh_seg_t unused = {nullptr, nullptr, 0, 0, 0};
    for (int i = 0; i < P.nseg; ++i) {
        const bxh_seg_t S = seg_at(segs, unused, i);
        if (S.k <= 0) continue;
        if ((!ktail && S.k % bk) || !aligned16(S.A, S.lda, esize) ||
            !aligned16(S.B, S.ldb, esize))
            return false;
    }
    return true;
}

__device__ __forceinline__ bool dma_eligible(const bxh_problem_t &P, const bxh_seg_t *segs,
                                             const bxh_seg_t &seg0, int bm, int bn, int bk,
                                             int esize)
{
    if (P.m % bm || P.n % bn || P.nseg < 1) return false;
    for (int i = 0; i < P.nseg; ++i) {
        const bxh_seg_t S = seg_at(segs, seg0, i);
        if (S.k <= 0) continue;
        if (S.k % bk || !aligned16(S.A, S.lda, esize) || !aligned16(S.B, S.ldb, esize))
            return false;
    }
    return true;
}

template <int MAJA, int MAJB, bool MASK = false, bool NESTED = true>
__device__ __forceinline__ void dmas_block(const bxh_problem_t &P, const bxh_seg_t *segs,
                                           const bxh_seg_t &seg0, int tm, int tn, float *smem)
{
    using namespace dmas;
    const int m0 = tm * BM, n0 = tn * BM;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lg = lane >> 4;

    f4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{0, 0, 0, 0};

    auto stageA = [&](int buf) { return smem + buf * STAGE; };
    auto stageB = [&](int buf) { return smem + 2 * STAGE + buf * STAGE; };
    if constexpr (NESTED) {
        // K tiles of all segments as one two-stage pipeline; the inner loop
        // takes one segment's tiles but its last with the segment's fields
        // fixed, the last prefetches the next non-empty segment's first tile
        auto issue = [&](const bxh_seg_t &sg, int64_t k, int b) {
            dma_tile<MAJA>((const float *)sg.A, sg.lda, m0, k, stageA(b), wave, lane);
            dma_tile<MAJB>((const float *)sg.B, sg.ldb, n0, k, stageB(b), wave, lane);
        };
        auto multiply = [&](int b) {
            const float *As = stageA(b), *Bs = stageB(b);
#pragma unroll
            for (int s = 0; s < BK / 4; ++s) {
                float a[4], bb[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a[i] = frag<MAJA>(As, wm * 64 + i * 16 + l16, s, lg);
                    bb[i] = frag<MAJB>(Bs, wn * 64 + i * 16 + l16, s, lg);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = mfma(a[i], bb[j], acc[i][j]);
            }
        };
        auto find = [&](int i, bxh_seg_t &out) {
            for (; i < P.nseg; ++i) {
                out = seg_at(segs, seg0, i);
                if (out.k > 0) break;
            }
            return i;
        };
        bxh_seg_t cur = seg0;
        int si = find(0, cur);
        if (si < P.nseg) {
            issue(cur, 0, 0);
            dma64::wait_dma();
            __builtin_amdgcn_s_barrier();
        }
        int buf = 0;
        while (si < P.nseg) {
            int64_t k0 = 0;
            for (; k0 + BK < cur.k; k0 += BK) {
                issue(cur, k0 + BK, buf ^ 1);
                multiply(buf);
                dma64::wait_dma();
                __builtin_amdgcn_s_barrier();
                buf ^= 1;
            }
            bxh_seg_t nxt = cur;
            const int sn = find(si + 1, nxt);
            if (sn < P.nseg) issue(nxt, 0, buf ^ 1);
            multiply(buf);
            dma64::wait_dma();
            __builtin_amdgcn_s_barrier();
            buf ^= 1;
            si = sn;
            cur = nxt;
        }
    } else {
        SegWalk<BK> w(segs, seg0, P.nseg);
        auto issue = [&](int buf) {
            dma_tile<MAJA>((const float *)w.cur.A, w.cur.lda, m0, w.k0, stageA(buf), wave, lane);
            dma_tile<MAJB>((const float *)w.cur.B, w.cur.ldb, n0, w.k0, stageB(buf), wave, lane);
        };
        bool more = w.valid();
        if (more) {
            issue(0);
            dma64::wait_dma();
            __builtin_amdgcn_s_barrier();
        }
        int buf = 0;
        while (more) {
            w.advance();
            more = w.valid();
            if (more) issue(buf ^ 1);
            const float *As = stageA(buf), *Bs = stageB(buf);
#pragma unroll
            for (int s = 0; s < BK / 4; ++s) {
                float a[4], bb[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a[i] = frag<MAJA>(As, wm * 64 + i * 16 + l16, s, lg);
                    bb[i] = frag<MAJB>(Bs, wn * 64 + i * 16 + l16, s, lg);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = mfma(a[i], bb[j], acc[i][j]);
            }
            dma64::wait_dma();
            __builtin_amdgcn_s_barrier();
            buf ^= 1;
        }
    }

    float *Cp = (float *)P.C;
    const float alpha = (float)P.alpha_re, beta = (float)P.beta_re;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + wn * 64 + j * 16 + l16;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm * 64 + i * 16 + Cfg<float>::out_row(lg, v);
                if constexpr (MASK)
                    if (!tri_keep(P.tri, row + P.tri_off, col)) continue;
                float *c = Cp + row + (int64_t)col * P.ldc;
                const float old = P.beta_zero ? 0.f : gload(c);
                gstore(c, out_real(alpha, acc[i][j][v], beta, old, (bool)P.beta_zero));
            }
        }
}

template <int MAJA, int MAJB>
__global__ __launch_bounds__(NT, 2) void k_sgemm_dma(bxh_problem_t P, bxh_seg_t S)
{
    __shared__ __attribute__((aligned(16))) float smem[2 * 2 * dmas::STAGE];
    int b = blockIdx.x;
    const int nb = gridDim.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
    int tm, tn;
    tile_coords(b, P.tiles_m, P.tiles_n, tm, tn);
    // the nested K loop: 3-6% faster than the segment-walking one on every
    // transpose at 32768^3
    dmas_block<MAJA, MAJB>(P, nullptr, S, tm, tn, smem);
}

template <int MAJA, int MAJB, bool CJA, bool CJB, bool MASK, bool NESTED>
__device__ __forceinline__ void dmaz_block(const bxh_problem_t &P, const bxh_seg_t *segs,
                                           const bxh_seg_t &seg0, int tm, int tn, cplx *smem)
{
    using namespace dmaz;
    constexpr int SA = stage_elems<MAJA, BM>(), SB = stage_elems<MAJB, BN>();
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lg = lane >> 4;

    d4_t acc[4][2][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j][0] = acc[i][j][1] = d4_t{0, 0, 0, 0};

    auto stageA = [&](int buf) { return smem + buf * SA; };
    auto stageB = [&](int buf) { return smem + 2 * SA + buf * SB; };
    if constexpr (NESTED) {
        // K tiles of all segments as one two-stage pipeline; the inner loop
        // takes one segment's tiles but its last with the segment's fields
        // fixed, the last prefetches the next non-empty segment's first tile
        auto issue = [&](const bxh_seg_t &sg, int64_t k, int b) {
            dma_tile<MAJA, BM>((const cplx *)sg.A, sg.lda, m0, k, stageA(b), wave, lane);
            dma_tile<MAJB, BN>((const cplx *)sg.B, sg.ldb, n0, k, stageB(b), wave, lane);
        };
        auto multiply = [&](int b) {
            const cplx *As = stageA(b), *Bs = stageB(b);
#pragma unroll
            for (int s = 0; s < BK / 4; ++s) {
                cplx a[4], bb[2];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = frag<MAJA, BM, CJA>(As, wm * 64 + i * 16 + l16, s, lg);
#pragma unroll
                for (int j = 0; j < 2; ++j) bb[j] = frag<MAJB, BN, CJB>(Bs, wn * 32 + j * 16 + l16, s, lg);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const double nai = -a[i].im;
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        acc[i][j][0] = mfma(a[i].re, bb[j].re, acc[i][j][0]);
                        acc[i][j][1] = mfma(a[i].re, bb[j].im, acc[i][j][1]);
                        acc[i][j][0] = mfma(nai, bb[j].im, acc[i][j][0]);
                        acc[i][j][1] = mfma(a[i].im, bb[j].re, acc[i][j][1]);
                    }
                }
            }
        };
        auto find = [&](int i, bxh_seg_t &out) {
            for (; i < P.nseg; ++i) {
                out = seg_at(segs, seg0, i);
                if (out.k > 0) break;
            }
            return i;
        };
        bxh_seg_t cur = seg0;
        int si = find(0, cur);
        if (si < P.nseg) {
            issue(cur, 0, 0);
            dma64::wait_dma();
            __builtin_amdgcn_s_barrier();
        }
        int buf = 0;
        while (si < P.nseg) {
            int64_t k0 = 0;
            for (; k0 + BK < cur.k; k0 += BK) {
                issue(cur, k0 + BK, buf ^ 1);
                multiply(buf);
                dma64::wait_dma();
                __builtin_amdgcn_s_barrier();
                buf ^= 1;
            }
            bxh_seg_t nxt = cur;
            const int sn = find(si + 1, nxt);
            if (sn < P.nseg) issue(nxt, 0, buf ^ 1);
            multiply(buf);
            dma64::wait_dma();
            __builtin_amdgcn_s_barrier();
            buf ^= 1;
            si = sn;
            cur = nxt;
        }
    } else {
        SegWalk<BK> w(segs, seg0, P.nseg);
        auto issue = [&](int buf) {
            dma_tile<MAJA, BM>((const cplx *)w.cur.A, w.cur.lda, m0, w.k0, stageA(buf), wave, lane);
            dma_tile<MAJB, BN>((const cplx *)w.cur.B, w.cur.ldb, n0, w.k0, stageB(buf), wave, lane);
        };
        bool more = w.valid();
        if (more) {
            issue(0);
            dma64::wait_dma();
            __builtin_amdgcn_s_barrier();
        }
        int buf = 0;
        while (more) {
            w.advance();
            more = w.valid();
            if (more) issue(buf ^ 1);
            const cplx *As = stageA(buf), *Bs = stageB(buf);
#pragma unroll
            for (int s = 0; s < BK / 4; ++s) {
                cplx a[4], bb[2];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = frag<MAJA, BM, CJA>(As, wm * 64 + i * 16 + l16, s, lg);
#pragma unroll
                for (int j = 0; j < 2; ++j) bb[j] = frag<MAJB, BN, CJB>(Bs, wn * 32 + j * 16 + l16, s, lg);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const double nai = -a[i].im;
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        acc[i][j][0] = mfma(a[i].re, bb[j].re, acc[i][j][0]);
                        acc[i][j][1] = mfma(a[i].re, bb[j].im, acc[i][j][1]);
                        acc[i][j][0] = mfma(nai, bb[j].im, acc[i][j][0]);
                        acc[i][j][1] = mfma(a[i].im, bb[j].re, acc[i][j][1]);
                    }
                }
            }
            dma64::wait_dma();
            __builtin_amdgcn_s_barrier();
            buf ^= 1;
        }
    }

    cplx *Cp = (cplx *)P.C;
    const double ar = P.alpha_re, ai = P.alpha_im, br = P.beta_re, bi = P.beta_im;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn * 32 + j * 16 + l16;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm * 64 + i * 16 + Cfg<cplx>::out_row(lg, v);
                if constexpr (MASK)
                    if (!tri_keep(P.tri, row + P.tri_off, col)) continue;
                cplx *c = Cp + row + (int64_t)col * P.ldc;
                const cplx c0 = P.beta_zero ? cplx{0.0, 0.0} : gload(c);
                cplx o = out_cplx(ar, ai, acc[i][j][0][v], acc[i][j][1][v], br, bi, c0,
                                  P.beta_zero);
                if constexpr (MASK)
                    if ((P.tri & BX_TRI_HERM) && row + P.tri_off == col) o.im = 0.0;
                gstore(c, o);
            }
        }
}

template <typename E, int MAJA, int MAJB, bool CJA, bool CJB>
__device__ __forceinline__ void run_tile(const bxh_problem_t &P, const bxh_seg_t *segs, int local,
                                         E *smem)
{
    int tm, tn;
    tile_coords(local, P.tiles_m, P.tiles_n, tm, tn);
    // segments come from the table; the by-value segment is never read
    // when a table is given (a problem may have zero segments)
    const bxh_seg_t unused = {nullptr, nullptr, 0, 0, 0};
    if constexpr (sizeof(E) != 16) {
        // the smaller block tiles a grouped launch may pick per problem
        if (P.shape == BXH_SHAPE_128x64) {
            gemm_block<E, MAJA, MAJB, CJA, CJB, 64, 32>(P, segs + P.seg_off, unused, tm, tn, smem);
            return;
        }
        if (P.shape == BXH_SHAPE_64x64) {
            gemm_block<E, MAJA, MAJB, CJA, CJB, 32, 32>(P, segs + P.seg_off, unused, tm, tn, smem);
            return;
        }
    }
    gemm_block<E, MAJA, MAJB, CJA, CJB>(P, segs + P.seg_off, unused, tm, tn, smem);
}

// MODE 1 runs only the problems the DMA-staged block of the type can take
// (no edge tiles, aligned segments) and MODE 2 only the others: two
// launches, so that neither kernel carries both blocks' registers (one
// kernel with both spilled).
template <typename E, int MODE>
__device__ __forceinline__ void grouped_body(const bxh_problem_t *__restrict__ probs, int nprob,
                                             const bxh_seg_t *__restrict__ segs, int total,
                                             E *smem);

// Grid-stride over the tiles of many problems (grouped GEMM / tile runtime).
// LDS of a grouped kernel: MODE 1 runs the DMA-staged block of its type,
// the other modes the register-staged block
template <typename E, int MODE> constexpr int grouped_smem_elems()
{
    if (MODE == 2) return Smem<E>::ELEMS;
    if (sizeof(E) == 8) return dma64::SMEM_DOUBLES;
    if (sizeof(E) == 4) return 2 * 2 * dmas::STAGE;
    return 2 * (dmaz::BM + dmaz::BN) * dmaz::BK;
}

template <typename E, int MODE>
__global__ __launch_bounds__(NT, 2) void k_gemm_grouped(const bxh_problem_t *__restrict__ probs,
                                                        int nprob,
                                                        const bxh_seg_t *__restrict__ segs,
                                                        int total)
{
    __shared__ __attribute__((aligned(16))) E smem[grouped_smem_elems<E, MODE>()];
    grouped_body<E, MODE>(probs, nprob, segs, total, smem);
}

// Persistent variant for graph replay: the problem count and tile total are
// read from device memory, so a captured launch is replayed with whatever
// table the host uploaded before the replay.  counts[0] = nprob,
// counts[1] = total tiles.
template <typename E, int MODE>
__global__ __launch_bounds__(NT, 2) void k_gemm_grouped_dyn(const bxh_problem_t *__restrict__ probs,
                                                            const bxh_seg_t *__restrict__ segs,
                                                            const int *__restrict__ counts)
{
    __shared__ __attribute__((aligned(16))) E smem[grouped_smem_elems<E, MODE>()];
    grouped_body<E, MODE>(probs, counts[0], segs, counts[1], smem);
}

template <typename E, int MODE>
__device__ __forceinline__ void grouped_body(const bxh_problem_t *__restrict__ probs, int nprob,
                                             const bxh_seg_t *__restrict__ segs, int total,
                                             E *smem)
{
    constexpr bool CPLX = sizeof(E) == 16;
    static_assert(MODE == 1 || MODE == 2 || (MODE == 3 && sizeof(E) == 8), "grouped mode");
    const bxh_seg_t unused = {nullptr, nullptr, 0, 0, 0};
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (probs[mid].tile_start <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const bxh_problem_t P = probs[lo];
        const int local = t - P.tile_start;
        const bxh_seg_t *ps = segs + P.seg_off;
        int ttm, ttn;
        tile_coords(local, P.tiles_m, P.tiles_n, ttm, ttn);
        // problems on the smaller block tiles run on the register-staged block
        const bool big = sizeof(E) == 16 || P.shape == BXH_SHAPE_128x128;
        const bool dma = !big                ? false
                         : sizeof(E) == 8    ? dma_tile_ok(P, ps, ttm, ttn, 128, 128, 16, 8, false)
                         : sizeof(E) == 4 ? dma_tile_ok(P, ps, ttm, ttn, 128, 128, 32, 4, false)
                                          : dma_tile_ok(P, ps, ttm, ttn, 128, 64, 8, 16, false);
        // MODE 3: FP64 tiles the DMA block takes once it stages their K tails
        // (the tile runtime's last K tile, a ragged K; one segment needs K >=
        // 16, as the plain path's ragged split does); a third launch, without
        // the triangle mask: inlined beside MODE 1's blocks, or with the
        // mask, its MFMA loops spilled registers
        const bool tail = sizeof(E) == 8 && big && !dma && !P.tri &&
                          (P.nseg > 1 || (P.nseg == 1 && seg_at(ps, unused, 0).k >= 16)) &&
                          dma_tile_ok(P, ps, ttm, ttn, 128, 128, 16, 8, true);
        if (MODE == 1 && !dma) continue; // another launch's tile
        if (MODE == 2 && (dma || tail)) continue;
        if (MODE == 3 && !tail) continue;
        if (P.tri) {
            // a block tile of a rank-k diagonal block wholly off the stored
            // triangle writes nothing: skip it before any load
            int bm = Blk<E>::BM, bn = Blk<E>::BN;
            if constexpr (sizeof(E) != 16) {
                if (P.shape == BXH_SHAPE_128x64) {
                    bm = Blk<E, 64, 32>::BM;
                    bn = Blk<E, 64, 32>::BN;
                } else if (P.shape == BXH_SHAPE_64x64) {
                    bm = Blk<E, 32, 32>::BM;
                    bn = Blk<E, 32, 32>::BN;
                }
            }
            const int r0 = ttm * bm + P.tri_off, c0 = ttn * bn;
            if ((P.tri & BX_TRI_LOWER) && r0 + bm - 1 < c0) continue;
            if ((P.tri & BX_TRI_UPPER) && r0 > c0 + bn - 1) continue;
        }
        if constexpr (MODE == 3) {
            const int ta = P.opa != 0, tb = P.opb != 0;
            double *sd = reinterpret_cast<double *>(smem);
            switch (ta * 2 + tb) {
            case 0: dma64_block<MAJ_MN, MAJ_K, false, true, true>(P, ps, unused, ttm, ttn, sd); break;
            case 1: dma64_block<MAJ_MN, MAJ_MN, false, true, true>(P, ps, unused, ttm, ttn, sd); break;
            case 2: dma64_block<MAJ_K, MAJ_K, false, true, true>(P, ps, unused, ttm, ttn, sd); break;
            default: dma64_block<MAJ_K, MAJ_MN, false, true, true>(P, ps, unused, ttm, ttn, sd); break;
            }
        } else if constexpr (MODE == 1) {
            // block tiles of the DMA-staged blocks equal the register-staged ones
            int tm, tn;
            tile_coords(local, P.tiles_m, P.tiles_n, tm, tn);
            if constexpr (CPLX) {
                cplx *sz = reinterpret_cast<cplx *>(smem);
                switch (P.opa * 3 + P.opb) {
                case 0: dmaz_block<MAJ_MN, MAJ_K, false, false, true>(P, ps, unused, tm, tn, sz); break;
                case 1: dmaz_block<MAJ_MN, MAJ_MN, false, false, true>(P, ps, unused, tm, tn, sz); break;
                case 2: dmaz_block<MAJ_MN, MAJ_MN, false, true, true>(P, ps, unused, tm, tn, sz); break;
                case 3: dmaz_block<MAJ_K, MAJ_K, false, false, true>(P, ps, unused, tm, tn, sz); break;
                case 4: dmaz_block<MAJ_K, MAJ_MN, false, false, true>(P, ps, unused, tm, tn, sz); break;
                case 5: dmaz_block<MAJ_K, MAJ_MN, false, true, true>(P, ps, unused, tm, tn, sz); break;
                case 6: dmaz_block<MAJ_K, MAJ_K, true, false, true>(P, ps, unused, tm, tn, sz); break;
                case 7: dmaz_block<MAJ_K, MAJ_MN, true, false, true>(P, ps, unused, tm, tn, sz); break;
                default: dmaz_block<MAJ_K, MAJ_MN, true, true, true>(P, ps, unused, tm, tn, sz); break;
                }
            } else {
                const int ta = P.opa != 0, tb = P.opb != 0;
                if constexpr (sizeof(E) == 8) {
                    double *sd = reinterpret_cast<double *>(smem);
                    if (P.nseg == 1) {
                        // one K segment (rank-k, TRMM panels): passed by value
                        // as the plain kernels pass theirs, which keeps the
                        // segment walk out of the K loop (half its scalar and
                        // vector instructions)
                        const bxh_seg_t s1 = seg_at(ps, unused, 0);
                        switch (ta * 2 + tb) {
                        case 0: dma64_block<MAJ_MN, MAJ_K, true>(P, nullptr, s1, tm, tn, sd); break;
                        case 1: dma64_block<MAJ_MN, MAJ_MN, true>(P, nullptr, s1, tm, tn, sd); break;
                        case 2: dma64_block<MAJ_K, MAJ_K, true>(P, nullptr, s1, tm, tn, sd); break;
                        default: dma64_block<MAJ_K, MAJ_MN, true>(P, nullptr, s1, tm, tn, sd); break;
                        }
                        continue;
                    }
                    switch (ta * 2 + tb) {
                    case 0: dma64_block<MAJ_MN, MAJ_K, true>(P, ps, unused, tm, tn, sd); break;
                    case 1: dma64_block<MAJ_MN, MAJ_MN, true>(P, ps, unused, tm, tn, sd); break;
                    case 2: dma64_block<MAJ_K, MAJ_K, true>(P, ps, unused, tm, tn, sd); break;
                    default: dma64_block<MAJ_K, MAJ_MN, true>(P, ps, unused, tm, tn, sd); break;
                    }
                } else {
                    float *sf = reinterpret_cast<float *>(smem);
                    switch (ta * 2 + tb) {
                    case 0: dmas_block<MAJ_MN, MAJ_K, true>(P, ps, unused, tm, tn, sf); break;
                    case 1: dmas_block<MAJ_MN, MAJ_MN, true>(P, ps, unused, tm, tn, sf); break;
                    case 2: dmas_block<MAJ_K, MAJ_K, true>(P, ps, unused, tm, tn, sf); break;
                    default: dmas_block<MAJ_K, MAJ_MN, true>(P, ps, unused, tm, tn, sf); break;
                    }
                }
            }
        } else if constexpr (CPLX) {
            switch (P.opa * 3 + P.opb) {
            case 0: run_tile<E, MAJ_MN, MAJ_K, false, false>(P, segs, local, smem); break;
            case 1: run_tile<E, MAJ_MN, MAJ_MN, false, false>(P, segs, local, smem); break;
            case 2: run_tile<E, MAJ_MN, MAJ_MN, false, true>(P, segs, local, smem); break;
            case 3: run_tile<E, MAJ_K, MAJ_K, false, false>(P, segs, local, smem); break;
            case 4: run_tile<E, MAJ_K, MAJ_MN, false, false>(P, segs, local, smem); break;
            case 5: run_tile<E, MAJ_K, MAJ_MN, false, true>(P, segs, local, smem); break;
            case 6: run_tile<E, MAJ_K, MAJ_K, true, false>(P, segs, local, smem); break;
            case 7: run_tile<E, MAJ_K, MAJ_MN, true, false>(P, segs, local, smem); break;
            default: run_tile<E, MAJ_K, MAJ_MN, true, true>(P, segs, local, smem); break;
            }
        } else {
            const int ta = P.opa != 0, tb = P.opb != 0;
            switch (ta * 2 + tb) {
            case 0: run_tile<E, MAJ_MN, MAJ_K, false, false>(P, segs, local, smem); break;
            case 1: run_tile<E, MAJ_MN, MAJ_MN, false, false>(P, segs, local, smem); break;
            case 2: run_tile<E, MAJ_K, MAJ_K, false, false>(P, segs, local, smem); break;
            default: run_tile<E, MAJ_K, MAJ_MN, false, false>(P, segs, local, smem); break;
            }
        }
    }
}

template <typename E>
__global__ void k_scale(int m, int n, E beta, int beta_zero, int tri, E *C, int64_t ldc)
{
    const int64_t total = (int64_t)m * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i % m, c = i / m;
        if (!tri_keep(tri, (int)r, (int)c)) continue;
        E *p = C + r + c * ldc;
        *p = beta_zero ? E(0) : beta * *p;
    }
}

__global__ void k_scale_z(int m, int n, double br, double bi, int beta_zero, int tri, cplx *C,
                          int64_t ldc)
{
    const int64_t total = (int64_t)m * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i % m, c = i / m;
        if (!tri_keep(tri, (int)r, (int)c)) continue;
        cplx *p = C + r + c * ldc;
        if (beta_zero) {
            *p = cplx{0.0, 0.0};
        } else {
            const cplx v = *p;
            *p = cplx{br * v.re - bi * v.im, br * v.im + bi * v.re};
        }
        if ((tri & BX_TRI_HERM) && r == c) p->im = 0.0;
    }
}

// SplitMix64 (Steele, Lea, Flood 2014); the host generator in fill.c is the same function
__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double u11(uint64_t key)
{
    return ((double)(splitmix64(key) >> 11) * 0x1.0p-53) * 2.0 - 1.0;
}

__global__ void k_fill(int dtype, void *p, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                       uint32_t id)
{
    const int64_t total = rows * cols;
    const uint64_t base = seed ^ ((uint64_t)id << 56);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i % rows, c = i / rows;
        const int64_t o = r + c * ld;
        const uint64_t e = (uint64_t)i;
        if (dtype == BLASX_D) {
            ((double *)p)[o] = u11(base ^ e);
        } else if (dtype == BLASX_S) {
            ((float *)p)[o] = (float)u11(base ^ e);
        } else {
            ((double *)p)[2 * o] = u11(base ^ (2 * e));
            ((double *)p)[2 * o + 1] = u11(base ^ (2 * e + 1));
        }
    }
}

// plain streaming copy (16-byte vectors); also the reference kernel for
// calibrating HBM byte counters
__global__ void k_copy16(const int4 *__restrict__ src, int4 *__restrict__ dst, int64_t n16)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

thread_local char g_err[256];

int fail(const char *what, hipError_t e)
{
    snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
    return BLASX_EHIP;
}

template <typename E, int MAJA, int MAJB, bool CJA, bool CJB>
int launch_plain(const bxh_problem_t &P, const bxh_seg_t &S, hipStream_t st)
{
    const int grid = P.tiles_m * P.tiles_n;
    hipLaunchKernelGGL((k_gemm_plain<E, MAJA, MAJB, CJA, CJB>), dim3(grid), dim3(NT), 0, st, P, S);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail("gemm launch", e);
}

template <int MAJA, int MAJB>
int launch_dgemm_dma(const bxh_problem_t &P, const bxh_seg_t &S, hipStream_t st)
{
    hipLaunchKernelGGL((k_dgemm_dma<MAJA, MAJB>), dim3(P.tiles_m * P.tiles_n), dim3(NT), 0, st, P,
                       S);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail("dgemm dma launch", e);
}

// the DMA-staged FP64 kernel's conditions: no edge tiles, 16-byte aligned
// operands and leading dimensions
bool dgemm_dma_ok(const bxh_problem_t &P, const bxh_seg_t &S)
{
    return P.m % 128 == 0 && P.n % 128 == 0 && S.k >= 16 && S.k % 16 == 0 &&
           aligned16(S.A, S.lda, 8) && aligned16(S.B, S.ldb, 8);
}

// ---- ragged DGEMM on the DMA-staged kernel ----
// A problem with edge tiles (M or N not a multiple of 128), any K, or
// operands whose base or leading dimension is not 16-byte aligned runs as one
// launch of k_dgemm_dma_kt over all its tiles (edge tiles clamp their operand
// rows and mask their stores).  A misaligned operand is first copied into an
// aligned workspace (one HBM pass, leading dimension rounded up to 8 doubles)
// when the product is large enough to pay for it.

__global__ void k_repack(const double *__restrict__ src, int64_t lds, double *__restrict__ dst,
                         int64_t ldd, int64_t rows, int64_t cols)
{
    for (int64_t c = blockIdx.y; c < cols; c += gridDim.y)
        for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows;
             r += (int64_t)gridDim.x * blockDim.x)
            dst[r + c * ldd] = src[r + c * lds];
}

void launch_repack(const double *src, int64_t lds, double *dst, int64_t ldd, int64_t rows,
                   int64_t cols, hipStream_t st)
{
    const int gx = (int)((rows + 255) / 256 < 16 ? (rows + 255) / 256 : 16);
    const int gy = (int)(cols < 8192 ? cols : 8192);
    hipLaunchKernelGGL(k_repack, dim3(gx, gy), dim3(256), 0, st, src, lds, dst, ldd, rows, cols);
}

// A device workspace reused across calls: a call's copies wait for the
// previous call's kernels (event on the device), growth waits on the host.
struct Workspace {
    void *p = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
};
Workspace g_rws;
pthread_mutex_t g_rws_lock = PTHREAD_MUTEX_INITIALIZER;

template <int MAJA, int MAJB>
int launch_dgemm_kt(const bxh_problem_t &P, const bxh_seg_t &S, hipStream_t st)
{
    hipLaunchKernelGGL((k_dgemm_dma_kt<MAJA, MAJB>), dim3(P.tiles_m * P.tiles_n), dim3(NT), 0, st,
                       P, S);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail("dgemm dma (ragged) launch", e);
}

// rows x cols of the stored operand behind op(A) (m x k) / op(B) (k x n)
inline void stored_dims(int op, int64_t r, int64_t c, int64_t &rows, int64_t &cols)
{
    rows = op ? c : r;
    cols = op ? r : c;
}

// 1: ran; 0: not taken (the caller runs the plain kernel); <0: error
int dgemm_ragged(int ta, int tb, const bxh_problem_t &P, const bxh_seg_t &S, hipStream_t st)
{
    if (P.m < 128 || P.n < 128 || S.k < 16 || P.tri) return 0;
    const bool okA = aligned16(S.A, S.lda, 8), okB = aligned16(S.B, S.ldb, 8);
    // a copy of A costs ~16 MK bytes against 2 MNK flops: worth it from N ~ 4096
    if ((!okA && P.n < 4096) || (!okB && P.m < 4096)) return 0;
    int64_t ra, ca, rb, cb;
    stored_dims(ta, P.m, S.k, ra, ca);
    stored_dims(tb, S.k, P.n, rb, cb);
    const int64_t lda2 = (ra + 7) & ~(int64_t)7, ldb2 = (rb + 7) & ~(int64_t)7;
    const size_t bytesA = okA ? 0 : (size_t)lda2 * ca * 8, bytesB = okB ? 0 : (size_t)ldb2 * cb * 8;
    bxh_seg_t S2 = S;
    const bool ws = bytesA + bytesB > 0;
    if (ws) {
        pthread_mutex_lock(&g_rws_lock);
        hipError_t e = hipSuccess;
        if (!g_rws.done) e = hipEventCreateWithFlags(&g_rws.done, hipEventDisableTiming);
        if (e == hipSuccess && bytesA + bytesB > g_rws.cap) {
            e = hipEventSynchronize(g_rws.done);
            if (e == hipSuccess && g_rws.p) e = hipFree(g_rws.p);
            g_rws.p = nullptr;
            g_rws.cap = 0;
            if (e == hipSuccess) e = hipMalloc(&g_rws.p, bytesA + bytesB);
            if (e == hipSuccess) g_rws.cap = bytesA + bytesB;
        }
        if (e == hipSuccess) e = hipStreamWaitEvent(st, g_rws.done, 0);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            pthread_mutex_unlock(&g_rws_lock);
            return fail("ragged dgemm workspace", e);
        }
        double *wa = (double *)g_rws.p, *wb = (double *)((char *)g_rws.p + bytesA);
        if (!okA) {
            launch_repack((const double *)S.A, S.lda, wa, lda2, ra, ca, st);
            S2.A = wa;
            S2.lda = lda2;
        }
        if (!okB) {
            launch_repack((const double *)S.B, S.ldb, wb, ldb2, rb, cb, st);
            S2.B = wb;
            S2.ldb = ldb2;
        }
    }
    int rc;
    switch (ta * 2 + tb) {
    case 0: rc = launch_dgemm_kt<MAJ_MN, MAJ_K>(P, S2, st); break;
    case 1: rc = launch_dgemm_kt<MAJ_MN, MAJ_MN>(P, S2, st); break;
    case 2: rc = launch_dgemm_kt<MAJ_K, MAJ_K>(P, S2, st); break;
    default: rc = launch_dgemm_kt<MAJ_K, MAJ_MN>(P, S2, st); break;
    }
    if (ws) {
        hipError_t e = hipEventRecord(g_rws.done, st);
        pthread_mutex_unlock(&g_rws_lock);
        if (!rc && e != hipSuccess) rc = fail("ragged dgemm event", e);
    }
    return rc ? rc : 1;
}

template <int MAJA, int MAJB>
int launch_sgemm_dma(const bxh_problem_t &P, const bxh_seg_t &S, hipStream_t st)
{
    hipLaunchKernelGGL((k_sgemm_dma<MAJA, MAJB>), dim3(P.tiles_m * P.tiles_n), dim3(NT), 0, st, P,
                       S);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail("sgemm dma launch", e);
}

template <typename E> int dispatch_real(int opa, int opb, const bxh_problem_t &P, const bxh_seg_t &S,
                                        hipStream_t st)
{
    const int ta = opa != 0, tb = opb != 0;
    if constexpr (sizeof(E) == 4) {
        if (P.m % dmas::BM == 0 && P.n % dmas::BM == 0 && S.k >= dmas::BK &&
            S.k % dmas::BK == 0 && aligned16(S.A, S.lda, 4) && aligned16(S.B, S.ldb, 4)) {
            switch (ta * 2 + tb) {
            case 0: return launch_sgemm_dma<MAJ_MN, MAJ_K>(P, S, st);
            case 1: return launch_sgemm_dma<MAJ_MN, MAJ_MN>(P, S, st);
            case 2: return launch_sgemm_dma<MAJ_K, MAJ_K>(P, S, st);
            default: return launch_sgemm_dma<MAJ_K, MAJ_MN>(P, S, st);
            }
        }
    }
    if constexpr (sizeof(E) == 8) {
        if (dgemm_dma_ok(P, S)) {
            switch (ta * 2 + tb) {
            case 0: return launch_dgemm_dma<MAJ_MN, MAJ_K>(P, S, st);
            case 1: return launch_dgemm_dma<MAJ_MN, MAJ_MN>(P, S, st);
            case 2: return launch_dgemm_dma<MAJ_K, MAJ_K>(P, S, st);
            default: return launch_dgemm_dma<MAJ_K, MAJ_MN>(P, S, st);
            }
        }
        const int r = dgemm_ragged(ta, tb, P, S, st);
        if (r != 0) return r < 0 ? r : 0;
    }
    switch (ta * 2 + tb) {
    case 0: return launch_plain<E, MAJ_MN, MAJ_K, false, false>(P, S, st);
    case 1: return launch_plain<E, MAJ_MN, MAJ_MN, false, false>(P, S, st);
    case 2: return launch_plain<E, MAJ_K, MAJ_K, false, false>(P, S, st);
    default: return launch_plain<E, MAJ_K, MAJ_MN, false, false>(P, S, st);
    }
}

template <int MAJA, int MAJB, bool CJA, bool CJB>
int launch_zgemm_dma(const bxh_problem_t &P, const bxh_seg_t &S, hipStream_t st)
{
    hipLaunchKernelGGL((k_zgemm_dma<MAJA, MAJB, CJA, CJB>), dim3(P.tiles_m * P.tiles_n), dim3(NT),
                       0, st, P, S);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail("zgemm dma launch", e);
}

int dispatch_cplx(int opa, int opb, const bxh_problem_t &P, const bxh_seg_t &S, hipStream_t st)
{
    // the DMA-staged kernel: no edge tiles, 16-byte aligned operands
    if (P.m % dmaz::BM == 0 && P.n % dmaz::BN == 0 && S.k >= dmaz::BK && S.k % dmaz::BK == 0 &&
        aligned16(S.A, S.lda, 16) && aligned16(S.B, S.ldb, 16)) {
        switch (opa * 3 + opb) {
        case 0: return launch_zgemm_dma<MAJ_MN, MAJ_K, false, false>(P, S, st);
        case 1: return launch_zgemm_dma<MAJ_MN, MAJ_MN, false, false>(P, S, st);
        case 2: return launch_zgemm_dma<MAJ_MN, MAJ_MN, false, true>(P, S, st);
        case 3: return launch_zgemm_dma<MAJ_K, MAJ_K, false, false>(P, S, st);
        case 4: return launch_zgemm_dma<MAJ_K, MAJ_MN, false, false>(P, S, st);
        case 5: return launch_zgemm_dma<MAJ_K, MAJ_MN, false, true>(P, S, st);
        case 6: return launch_zgemm_dma<MAJ_K, MAJ_K, true, false>(P, S, st);
        case 7: return launch_zgemm_dma<MAJ_K, MAJ_MN, true, false>(P, S, st);
        default: return launch_zgemm_dma<MAJ_K, MAJ_MN, true, true>(P, S, st);
        }
    }
    switch (opa * 3 + opb) {
    case 0: return launch_plain<cplx, MAJ_MN, MAJ_K, false, false>(P, S, st);
    case 1: return launch_plain<cplx, MAJ_MN, MAJ_MN, false, false>(P, S, st);
    case 2: return launch_plain<cplx, MAJ_MN, MAJ_MN, false, true>(P, S, st);
    case 3: return launch_plain<cplx, MAJ_K, MAJ_K, false, false>(P, S, st);
    case 4: return launch_plain<cplx, MAJ_K, MAJ_MN, false, false>(P, S, st);
    case 5: return launch_plain<cplx, MAJ_K, MAJ_MN, false, true>(P, S, st);
    case 6: return launch_plain<cplx, MAJ_K, MAJ_K, true, false>(P, S, st);
    case 7: return launch_plain<cplx, MAJ_K, MAJ_MN, true, false>(P, S, st);
    default: return launch_plain<cplx, MAJ_K, MAJ_MN, true, true>(P, S, st);
    }
}

} // namespace

extern "C" {

const char *bxh_kernel_error(void) { return g_err; }

void bxh_tile_dims(int dtype, int shape, int *bm, int *bn)
{
    if (dtype == BLASX_Z || shape == BXH_SHAPE_128x128 || shape < 0 || shape >= BXH_NSHAPES) {
        *bm = bxh_tile_m(dtype);
        *bn = bxh_tile_n(dtype);
        return;
    }
    *bm = shape == BXH_SHAPE_64x64 ? Blk<double, 32, 32>::BM : Blk<double, 64, 32>::BM;
    *bn = Blk<double, 64, 32>::BN; // the same for both types
}

int bxh_tile_m(int dtype)
{
    return dtype == BLASX_S ? Cfg<float>::BM : dtype == BLASX_D ? Cfg<double>::BM : Cfg<cplx>::BM;
}

int bxh_tile_n(int dtype)
{
    return dtype == BLASX_S ? Cfg<float>::BN : dtype == BLASX_D ? Cfg<double>::BN : Cfg<cplx>::BN;
}

int bxh_gemm(int dtype, int opa, int opb, int m, int n, int k, const double *alpha,
             const void *A, int64_t lda, const void *B, int64_t ldb, const double *beta, void *C,
             int64_t ldc, void *stream)
{
    if (m <= 0 || n <= 0) return 0;
    if (opa < 0 || opa > 2 || opb < 0 || opb > 2 || k < 0) return BLASX_EINVAL;
    bxh_problem_t P;
    P.C = C;
    P.ldc = ldc;
    P.m = m;
    P.n = n;
    P.nseg = 1;
    P.seg_off = 0;
    P.opa = opa;
    P.opb = opb;
    P.tiles_m = (m + bxh_tile_m(dtype) - 1) / bxh_tile_m(dtype);
    P.tiles_n = (n + bxh_tile_n(dtype) - 1) / bxh_tile_n(dtype);
    P.tile_start = 0;
    P.alpha_re = alpha[0];
    P.alpha_im = dtype == BLASX_Z ? alpha[1] : 0.0;
    P.beta_re = beta[0];
    P.beta_im = dtype == BLASX_Z ? beta[1] : 0.0;
    P.beta_zero = (P.beta_re == 0.0 && P.beta_im == 0.0);
    P.tri = 0;
    P.tri_off = 0;
    P.shape = BXH_SHAPE_128x128;
    P.pad0 = 0;
    bxh_seg_t S;
    S.A = A;
    S.B = B;
    S.lda = lda;
    S.ldb = ldb;
    S.k = k;
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
    case BLASX_D: return dispatch_real<double>(opa, opb, P, S, st);
    case BLASX_S: return dispatch_real<float>(opa, opb, P, S, st);
    case BLASX_Z: return dispatch_cplx(opa, opb, P, S, st);
    default: return BLASX_EINVAL;
    }
}

int bxh_gemm_grouped(int dtype, const bxh_problem_t *probs_dev, int nprob,
                     const bxh_seg_t *segs_dev, int total_tiles, int grid, void *stream)
{
    if (nprob <= 0 || total_tiles <= 0) return 0;
    if (grid <= 0 || grid > total_tiles) grid = total_tiles;
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
    case BLASX_D:
        hipLaunchKernelGGL((k_gemm_grouped<double, 1>), dim3(grid), dim3(NT), 0, st, probs_dev,
                           nprob, segs_dev, total_tiles);
        hipLaunchKernelGGL((k_gemm_grouped<double, 3>), dim3(grid), dim3(NT), 0, st, probs_dev,
                           nprob, segs_dev, total_tiles);
        hipLaunchKernelGGL((k_gemm_grouped<double, 2>), dim3(grid), dim3(NT), 0, st, probs_dev,
                           nprob, segs_dev, total_tiles);
        break;
    case BLASX_S:
        hipLaunchKernelGGL((k_gemm_grouped<float, 1>), dim3(grid), dim3(NT), 0, st, probs_dev,
                           nprob, segs_dev, total_tiles);
        hipLaunchKernelGGL((k_gemm_grouped<float, 2>), dim3(grid), dim3(NT), 0, st, probs_dev,
                           nprob, segs_dev, total_tiles);
        break;
    case BLASX_Z:
        hipLaunchKernelGGL((k_gemm_grouped<cplx, 1>), dim3(grid), dim3(NT), 0, st, probs_dev,
                           nprob, segs_dev, total_tiles);
        hipLaunchKernelGGL((k_gemm_grouped<cplx, 2>), dim3(grid), dim3(NT), 0, st, probs_dev,
                           nprob, segs_dev, total_tiles);
        break;
    default: return BLASX_EINVAL;
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail("grouped gemm launch", e);
}

int bxh_gemm_grouped_dyn(int dtype, const bxh_problem_t *probs_dev, const bxh_seg_t *segs_dev,
                         const int *counts_dev, int grid, void *stream)
{
    if (grid <= 0) return BLASX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
    case BLASX_D:
        hipLaunchKernelGGL((k_gemm_grouped_dyn<double, 1>), dim3(grid), dim3(NT), 0, st,
                           probs_dev, segs_dev, counts_dev);
        hipLaunchKernelGGL((k_gemm_grouped_dyn<double, 3>), dim3(grid), dim3(NT), 0, st,
                           probs_dev, segs_dev, counts_dev);
        hipLaunchKernelGGL((k_gemm_grouped_dyn<double, 2>), dim3(grid), dim3(NT), 0, st,
                           probs_dev, segs_dev, counts_dev);
        break;
    case BLASX_S:
        hipLaunchKernelGGL((k_gemm_grouped_dyn<float, 1>), dim3(grid), dim3(NT), 0, st,
                           probs_dev, segs_dev, counts_dev);
        hipLaunchKernelGGL((k_gemm_grouped_dyn<float, 2>), dim3(grid), dim3(NT), 0, st,
                           probs_dev, segs_dev, counts_dev);
        break;
    case BLASX_Z:
        hipLaunchKernelGGL((k_gemm_grouped_dyn<cplx, 1>), dim3(grid), dim3(NT), 0, st,
                           probs_dev, segs_dev, counts_dev);
        hipLaunchKernelGGL((k_gemm_grouped_dyn<cplx, 2>), dim3(grid), dim3(NT), 0, st,
                           probs_dev, segs_dev, counts_dev);
        break;
    default: return BLASX_EINVAL;
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail("grouped gemm launch", e);
}

int bxh_scale(int dtype, int m, int n, const double *beta, void *C, int64_t ldc, void *stream)
{
    return bxh_scale_tri(dtype, m, n, 0, beta, C, ldc, stream);
}

int bxh_scale_tri(int dtype, int m, int n, int tri, const double *beta, void *C, int64_t ldc,
                  void *stream)
{
    if (m <= 0 || n <= 0) return 0;
    const int64_t total = (int64_t)m * n;
    int grid = (int)((total + 255) / 256);
    if (grid > 65536) grid = 65536;
    hipStream_t st = (hipStream_t)stream;
    const int bz = (beta[0] == 0.0 && (dtype != BLASX_Z || beta[1] == 0.0));
    switch (dtype) {
    case BLASX_D:
        hipLaunchKernelGGL(k_scale<double>, dim3(grid), dim3(256), 0, st, m, n, beta[0], bz, tri,
                           (double *)C, ldc);
        break;
    case BLASX_S:
        hipLaunchKernelGGL(k_scale<float>, dim3(grid), dim3(256), 0, st, m, n, (float)beta[0], bz,
                           tri, (float *)C, ldc);
        break;
    case BLASX_Z:
        hipLaunchKernelGGL(k_scale_z, dim3(grid), dim3(256), 0, st, m, n, beta[0], beta[1], bz,
                           tri, (cplx *)C, ldc);
        break;
    default: return BLASX_EINVAL;
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail("scale launch", e);
}

int bxh_copy_kernel(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return 0;
    if ((bytes & 15) || (((uintptr_t)dst | (uintptr_t)src) & 15)) return BLASX_EINVAL;
    const int64_t n16 = (int64_t)(bytes / 16);
    int64_t grid = (n16 + 255) / 256;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream,
                       (const int4 *)src, (int4 *)dst, n16);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail("copy launch", e);
}

int bxh_fill_uniform(int dtype, void *p, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                     uint32_t id, void *stream)
{
    if (rows <= 0 || cols <= 0) return 0;
    if (dtype != BLASX_S && dtype != BLASX_D && dtype != BLASX_Z) return BLASX_EINVAL;
    const int64_t total = rows * cols;
    int grid = (int)((total + 255) / 256);
    if (grid > 262144) grid = 262144;
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, (hipStream_t)stream, dtype, p, rows, cols,
                       ld, seed, id);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail("fill launch", e);
}

} // extern "C"
