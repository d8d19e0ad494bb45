// ozaki_kernels.hip -- split-precision DGEMM on the INT8 matrix cores of
// gfx950 (Ozaki scheme: Ozaki, Ogita, Oishi & Rump 2012, error-free
// splitting of matrix products; INT8 realisation after Ootomo, Ozaki &
// Yokota 2024).  No counterpart in the reference.
//
// For C = alpha*op(A)*op(B) + beta*C in FP64:
//  1. scale: row i of op(A) by 2^-e_i, column j of op(B) by 2^-f_j, where
//     e_i, f_j make every scaled entry lie in (-1, 1)  (k_rowcol_exponent);
//  2. split: r <- 128 r; d_p = trunc(r); r <- r - d_p, p = 1..s.  Every
//     step is exact in FP64, |d_p| <= 127, and
//        a = 2^e (sum_p d_p 2^-7p + r_s 2^-7s),  |r_s| < 1   (k_split)
//     digits go to int8 slice matrices, K contiguous, zero padded;
//  3. products: for each digit group g = p + q <= s + 1, the exact integer
//        S_g = sum_{p+q=g} D^A_p (D^B_q)^T
//     is one INT8-MFMA launch whose K range runs over the g-1 slice pairs
//     (int32 accumulation is exact while (g-1)*K*127^2 < 2^31; the host
//     splits longer ranges), folded into an FP64 workspace as
//     W += 2^-7g * S_g, smallest groups first  (k_ozaki_group);
//  4. C = alpha * 2^(e_i + f_j) * W + beta * C  (k_ozaki_finish).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "../hiplayer.h"
#include "blasx_ext.h"
#include "gmem.h"

namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// The matrix-core instruction of a tile.  INT8: exact int32 sums of
// V_MFMA_I32_16X16X64_I8 products (the digit and CRT schemes).  BF16: FP32
// sums of V_MFMA_F32_16X16X32_BF16 products (the BF16x3 split).  Both take 16
// bytes per lane per operand fragment -- 16 int8 or 8 bf16 consecutive in K,
// K block lane / 16 -- and give lane (l16, lg) rows 4 lg .. 4 lg + 3 of
// column l16, so one byte layout, one pipeline and one epilogue serve both.
struct MacI8 {
    typedef i32x4 acc_t;
    __device__ static acc_t zero() { return i32x4{0, 0, 0, 0}; }
    __device__ static acc_t mac(i32x4 a, i32x4 b, acc_t c)
    {
        return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
    }
};
struct MacBF16 {
    typedef f32x4 acc_t;
    __device__ static acc_t zero() { return f32x4{0.f, 0.f, 0.f, 0.f}; }
    __device__ static acc_t mac(i32x4 a, i32x4 b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    }
};

constexpr int OZ_BM = 128, OZ_BN = 128, OZ_BK = 64;
// LDS layout of a staged K tile: row r (of the M or N extent), 16-byte
// chunk c (k = 16c .. 16c+15) at byte lds_at(r, c): unpadded 64-byte rows
// with the chunk index XORed by bits 1-2 of the row.  Measured on gfx950
// with the LDS bank-conflict counter, this is conflict-free both for the
// MFMA fragment reads (ds_read_b128, lane -> row lane%16, chunk lane/16) and
// for the staging writes (lane -> row lane/4, chunk lane%4); a padded
// 80-byte row stride conflicts 2-way on the reads and 3-way on the writes.
constexpr int OZ_STR = OZ_BK;
__device__ __forceinline__ int lds_at(int r, int c) { return r * OZ_STR + ((c ^ ((r >> 1) & 3)) << 4); }
constexpr int OZ_NT = 256;

// Slice / residue planes are stored as panels: for each block of OZ_PR = 128
// rows and each K tile of OZ_BK = 64 bytes, the 128 x 64 piece is one
// contiguous 8 KB chunk (row-major inside), chunks ordered by row block, then
// K tile.  A block's K-tile load is then a contiguous, fully used stretch
// of cache lines instead of 128 half-used lines kp bytes apart.
constexpr int OZ_PR = 128;
constexpr int OZ_CHUNK = OZ_PR * OZ_BK; // 8 KB
__host__ __device__ __forceinline__ int64_t panel_off(int i, int k, int kp)
{
    return ((int64_t)(i / OZ_PR) * (kp / OZ_BK) + k / OZ_BK) * OZ_CHUNK + (i % OZ_PR) * OZ_BK +
           k % OZ_BK;
}

// ---------------------------------------------------------------------------
// 1. exponents, two kinds (NORM = 0 / 1):
//    max:  e[i] = smallest integer with max_k |x(i,k)| * 2^-e < 1
//          (the digit scheme: every scaled entry in (-1, 1));
//    norm: e[i] = smallest integer with ss_i (1 + 2^-30) < 4^e, ss_i =
//          sum_k x(i,k)^2 as computed (the CRT scheme's Cauchy-Schwarz
//          scaling: ||x_i||_2 < 2^e, so every entry too; the factor covers
//          the rounding of the computed sum for K < 2^22).  From the
//          exponent k of ss_i (1 + 2^-30) (frexp: below 2^k, at least
//          2^(k-1)), e = ceil(k / 2), with no square root.
//    mode 0: x(i,k) = X[i + k*ld]  (reduce along the strided dimension:
//            one thread per i, coalesced across threads)
//    mode 1: x(i,k) = X[k + i*ld]  (reduce along the contiguous dimension:
//            one block per i)
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ int exp_of(double amax)
{
    if (amax == 0.0 || !__builtin_isfinite(amax)) return 0;
    int e;
    frexp(amax, &e); // amax = m * 2^e, m in [0.5, 1)
    return e;
}

__host__ __device__ __forceinline__ int norm_exp_of(double ss)
{
    const int k = exp_of(ss * (1.0 + 0x1p-30));
    return (k + 1) >> 1; // ceil(k / 2), k of either sign
}

template <int NORM> __device__ __forceinline__ double red_op(double acc, double x)
{
    return NORM ? fma(x, x, acc) : fmax(acc, fabs(x));
}
template <int NORM> __device__ __forceinline__ double red_join(double a, double b)
{
    return NORM ? a + b : fmax(a, b);
}

// mode 0 runs in two passes so that the whole GPU streams X: pass 1 gives
// block (bx, by) rows i in [64 bx, 64 bx + 64) and the by-th chunk of k,
// 64 rows x 4 k per step (coalesced 512-byte row segments), and writes the
// chunk's max |x| (sum of squares) to part[by * n + i]; pass 2 combines the
// chunks.  fmax over non-negative values is associative, so the max equals
// the single sequential max (NaN entries are skipped by fmax in both).
constexpr int EXP_CHUNKS = 32;

template <int NORM>
__global__ void k_amax_strided(const double *__restrict__ X, int64_t ld, int n, int len,
                               int chunk, double *__restrict__ part)
{
    __shared__ double red[4][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + tx;
    const int k0 = blockIdx.y * chunk, k1 = min(len, k0 + chunk);
    double m = 0.0;
    if (i < n)
        for (int k = k0 + ty; k < k1; k += 4) m = red_op<NORM>(m, X[i + (int64_t)k * ld]);
    red[ty][tx] = m;
    __syncthreads();
    if (ty == 0 && i < n)
        part[(int64_t)blockIdx.y * n + i] =
            red_join<NORM>(red_join<NORM>(red[0][tx], red[1][tx]),
                           red_join<NORM>(red[2][tx], red[3][tx]));
}

// per = 2: rows 2i and 2i+1 of the scan (real and imaginary parts of a
// complex row) give one exponent
template <int NORM>
__global__ void k_exp_parts(const double *__restrict__ part, int chunks, int n, int per,
                            int *__restrict__ e)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t rows = (int64_t)n * per;
    double m = 0.0;
    for (int c = 0; c < chunks; ++c)
        for (int r = 0; r < per; ++r) m = red_join<NORM>(m, part[c * rows + (int64_t)i * per + r]);
    e[i] = NORM ? norm_exp_of(m) : exp_of(m);
}

template <int NORM>
__global__ void k_exp_contig(const double *__restrict__ X, int64_t ld, int n, int len,
                             int *__restrict__ e)
{
    __shared__ double red[256];
    const int i = blockIdx.x;
    double m = 0.0;
    const double *row = X + (int64_t)i * ld;
    for (int k = threadIdx.x; k < len; k += blockDim.x) m = red_op<NORM>(m, row[k]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            red[threadIdx.x] = red_join<NORM>(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) e[i] = NORM ? norm_exp_of(red[0]) : exp_of(red[0]);
}

// ---------------------------------------------------------------------------
// 2. split into s int8 slices, layout D[p][i][k] (k contiguous, row stride
//    kp, slice stride np*kp), zero for i >= n or k >= len.
//    strided (mode 0): x(i,k) = X[i + k*ld]; staged through LDS so that both
//    the FP64 reads (along i) and the int8 writes (along k) are coalesced.
// ---------------------------------------------------------------------------
constexpr int SP_T = 64; // 64 x 64 element tile per block

__global__ void k_split_strided(const double *__restrict__ X, int64_t ld, int n, int len,
                                const int *__restrict__ e, int s, int8_t *__restrict__ D,
                                int np, int kp)
{
    __shared__ int8_t tile[9][SP_T][SP_T + 4];
    const int i0 = blockIdx.x * SP_T, k0 = blockIdx.y * SP_T;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6; // 64 x 4
    for (int kk = ty; kk < SP_T; kk += 4) {
        const int i = i0 + tx, k = k0 + kk;
        // |r| < 1; r*128 is exact, trunc is exact, the fractional remainder
        // is exactly representable: every digit step is error free
        double r = (i < n && k < len) ? ldexp(X[i + (int64_t)k * ld], -e[i]) : 0.0;
        for (int p = 0; p < s; ++p) {
            r *= 128.0;
            const double t = trunc(r);
            tile[p][tx][kk] = (int8_t)(int)t;
            r -= t;
        }
    }
    __syncthreads();
    // write: each thread writes 16 consecutive k bytes of one row
    const int row = threadIdx.x >> 2, kc = (threadIdx.x & 3) * 16;
    const int i = i0 + row;
    if (i >= np) return;
    for (int p = 0; p < s; ++p) {
        int8_t *dst = D + (int64_t)p * np * kp + panel_off(i, k0 + kc, kp);
        uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < 16; ++b)
            w[b >> 2] |= ((uint32_t)(uint8_t)tile[p][row][kc + b]) << (8 * (b & 3));
        *(i32x4 *)dst = i32x4{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
    }
}

// contiguous (mode 1): x(i,k) = X[k + i*ld]; one thread per 16 consecutive k
__global__ void k_split_contig(const double *__restrict__ X, int64_t ld, int n, int len,
                               const int *__restrict__ e, int s, int8_t *__restrict__ D, int np,
                               int kp)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int chunks = kp / 16;
    const int64_t total = (int64_t)np * chunks;
    if (t >= total) return;
    const int i = (int)(t / chunks), k0 = (int)(t % chunks) * 16;
    double r[16];
    const int ei = i < n ? e[i] : 0;
#pragma unroll
    for (int b = 0; b < 16; ++b)
        r[b] = (i < n && k0 + b < len) ? ldexp(X[k0 + b + (int64_t)i * ld], -ei) : 0.0;
    for (int p = 0; p < s; ++p) {
        uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            r[b] *= 128.0;
            const double tq = trunc(r[b]);
            r[b] -= tq;
            w[b >> 2] |= ((uint32_t)(uint8_t)(int8_t)(int)tq) << (8 * (b & 3));
        }
        *(i32x4 *)(D + (int64_t)p * np * kp + panel_off(i, k0, kp)) =
            i32x4{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
    }
}

// ---------------------------------------------------------------------------
// 3. one digit group: S = sum over segments of DA_seg * DB_seg^T (int32),
//    W = (first ? 0 : W) + scale * S.  DA/DB rows are K-contiguous and padded
//    to whole block tiles, so loads need no bounds checks.
//    256 threads = 2x2 waves, wave tile 64x64 = 4x4 MFMA_I32_16X16X64_I8.
// ---------------------------------------------------------------------------
// Block tile index -> (tm, tn), walking groups of 8 tile rows column by
// column.  The blocks of one XCD run a contiguous index range at a time
// (see the XCD renumbering in the kernels), so those resident together form
// an 8 x (blocks / 8) patch: per K step they read 8 A tiles and blocks/8 B
// tiles through their shared L2 instead of 1 A tile and one B tile each.
__host__ __device__ inline void group_tile(int t, int tiles_m, int tiles_n, int &tm, int &tn)
{
    constexpr int G = 8;
    const int per_group = G * tiles_n;
    const int g = t / per_group;
    const int first = g * G;
    const int gsz = tiles_m - first < G ? tiles_m - first : G;
    const int r = t - g * per_group;
    tm = first + r % gsz;
    tn = r / gsz;
}

struct OzSeg {
    const int8_t *A;
    const int8_t *B;
};

// block b of nb -> (m0, n0) of its BT x BT tile: XCD renumbering, then
// group_tile
template <int BT> __host__ __device__ inline void block_origin_of(int b, int nb, int n, int &m0, int &n0)
{
    const int tiles_n = (n + BT - 1) / BT;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3); // contiguous tiles per XCD
    int tm, tn;
    group_tile(b, nb / tiles_n, tiles_n, tm, tn);
    m0 = tm * BT;
    n0 = tn * BT;
}
template <int BT> __device__ __forceinline__ void block_origin(int n, int &m0, int &n0)
{
    block_origin_of<BT>((int)blockIdx.x, (int)gridDim.x, n, m0, n0);
}

// 128x128 tile: acc[mb][nb] of wave (wm, wn) holds rows
// m0 + 64 wm + 16 mb + 4 (lane/16) + v, column n0 + 64 wn + 16 nb + lane%16
template <typename MAC = MacI8>
__device__ __forceinline__ void tile128(const OzSeg *__restrict__ segs, int nseg, int kp, int m0,
                                        int n0, typename MAC::acc_t (&acc)[4][4])
{
    __shared__ __attribute__((aligned(16))) int8_t sA[2][OZ_BM * OZ_STR];
    __shared__ __attribute__((aligned(16))) int8_t sB[2][OZ_BN * OZ_STR];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int l16 = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = MAC::zero();

    // loader: 128 rows x 64 bytes per operand = 512 x 16 B, two per thread
    const int lrow = tid >> 2, lk = (tid & 3) * 16;
    i32x4 ra[2], rb[2];
    // one contiguous 8 KB panel chunk per operand: rows lrow and lrow + 64
    auto fetch = [&](int seg, int k0) {
        const int8_t *pa = segs[seg].A + panel_off(m0, k0, kp) + tid * 16;
        const int8_t *pb = segs[seg].B + panel_off(n0, k0, kp) + tid * 16;
        ra[0] = gload((const i32x4 *)pa);
        ra[1] = gload((const i32x4 *)(pa + OZ_CHUNK / 2));
        rb[0] = gload((const i32x4 *)pb);
        rb[1] = gload((const i32x4 *)(pb + OZ_CHUNK / 2));
    };
    auto lstore = [&](int buf) {
        *(i32x4 *)&sA[buf][lds_at(lrow, lk >> 4)] = ra[0];
        *(i32x4 *)&sA[buf][lds_at(lrow + 64, lk >> 4)] = ra[1];
        *(i32x4 *)&sB[buf][lds_at(lrow, lk >> 4)] = rb[0];
        *(i32x4 *)&sB[buf][lds_at(lrow + 64, lk >> 4)] = rb[1];
    };

    const int ktiles = kp / OZ_BK;
    const int total = nseg * ktiles;
    fetch(0, 0);
    lstore(0);
    __syncthreads();
    for (int it = 0; it < total; ++it) {
        const int buf = it & 1;
        const bool more = it + 1 < total;
        if (more) fetch((it + 1) / ktiles, ((it + 1) % ktiles) * OZ_BK);
        i32x4 fa[4], fb[4];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
            fa[mb] = *(const i32x4 *)&sA[buf][lds_at(wm * 64 + mb * 16 + l16, lg)];
#pragma unroll
        for (int nb2 = 0; nb2 < 4; ++nb2)
            fb[nb2] = *(const i32x4 *)&sB[buf][lds_at(wn * 64 + nb2 * 16 + l16, lg)];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int nb2 = 0; nb2 < 4; ++nb2)
                acc[mb][nb2] = MAC::mac(fa[mb], fb[nb2], acc[mb][nb2]);
        if (more) lstore(buf ^ 1);
        __syncthreads();
    }

}

// 256x256 tile, 8 waves (two per SIMD) of 128x64: 32 MFMA_I32_16X16X64_I8
// accumulators (128 VGPRs) per wave.  acc[mb][nb] of wave (wm, wn) =
// (wave & 1, wave >> 1) holds rows m0 + 128 wm + 16 mb + 4 (lane/16) + v,
// column n0 + 64 wn + 16 nb + lane%16.
//
// LDS traffic, not the matrix cores, bounds an INT8 tile on gfx950: per
// 64-byte K step the tile stores 32 KB into LDS and its waves read 96 KB.
// Measured on this tile (16384^3, one modulus): removing only the LDS stores
// of a register-staged version lifted it from 3.0 to 4.7 POP/s.  So the
// operands go from global memory straight into LDS (GLOBAL_LOAD_LDS_DWORDX4:
// no VGPR round trip, no DS_WRITE), and every LDS read is issued a step or
// 24 MFMAs before the MFMA that needs it.
//
// Pipeline (one segment's K range; K tile t = 64 bytes of K):
//  * four LDS stages of 256 rows x 64 bytes per operand (128 KB), tile t in
//    stage t % 4;
//  * step t: start the DMA of tile t + 3 into stage (t + 3) % 4 (the stage of
//    tile t - 1, whose reads all completed before the last barrier), read the
//    last two A fragments of tile t and issue the 4 MFMAs of row block 0;
//    then wait until those reads have returned and tile t + 1's DMA has
//    landed (tile t + 2's and t + 3's may still run) and meet at the step's
//    one barrier while the MFMAs run; then the other 28 MFMAs of tile t while
//    reading tile t + 1's fragments: the B fragments into the second B
//    register set first, A fragment mb < 6 after the MFMAs of row block mb.
//  * a DMA instruction writes 64 lanes x 16 bytes to LDS at M0 + 16 lane;
//    wave w's instruction i fills bytes 1024 (2w + i) .. of a stage (rows
//    16 (2w + i) ..), so each lane fetches the global 16 bytes that the LDS
//    layout puts at its position.
// The DMA is issued as inline assembly because the compiler, seeing the
// builtin's LDS write, waits for every outstanding DMA before each LDS read
// (it cannot tell the stages apart), which would expose the full global
// latency every step; the waits are explicit here instead (`vm_wait`), and
// the barrier is a bare S_BARRIER for the same reason (__syncthreads would
// also wait for every DMA).  Measured against the register-staged version
// on random data: bit-identical, 2.61 vs 2.46 POP/s.
constexpr int OZ_BIG = 256;
constexpr int OZ_NT_BIG = 512;
constexpr int OZ_STAGE = OZ_BIG * OZ_STR; // bytes per operand per stage
constexpr int OZ_NSTAGE = 4;

template <int V> struct Const {
    static constexpr int value = V;
};

// S_WAITCNT with only the vector-memory counter limited to N
template <int N> __device__ __forceinline__ void vm_wait()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// S_WAITCNT with only the LDS/scalar-memory counter limited to 0
__device__ __forceinline__ void lgkm_wait0()
{
    __builtin_amdgcn_s_waitcnt(0xF | (0x3 << 14) | (0x7 << 4));
}

// 16 bytes per lane from global memory to LDS (M0 + 16 lane); `lds` is the
// wave's destination, a wave-uniform LDS address held in a scalar register.
// M0 is a reserved register the compiler does not allocate (a clobber of it
// is ignored); nothing else in the kernels that call this uses M0 (checked
// in their gfx950 assembly: every M0 write is this statement's).
__device__ __forceinline__ void dma16_at(const int8_t *g, uint32_t lds)
{
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off"
                 :
                 : "s"(lds), "v"(g)
                 : "memory");
}

template <typename MAC = MacI8>
__device__ __forceinline__ void tile256_seg(const int8_t *__restrict__ gA,
                                            const int8_t *__restrict__ gB, int kp, int m0, int n0,
                                            typename MAC::acc_t (&acc)[8][4], int8_t *sA,
                                            int8_t *sB)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int l16 = lane & 15, lg = lane >> 4;
    const int ktiles = kp / OZ_BK; // even and >= 2 (the host pads K to 128)
    const int last = ktiles - 1;
    const int64_t next_rows = (int64_t)ktiles * OZ_CHUNK;
    // DMA: instruction i of this wave fills tile rows r = 16 (2 wave + i) +
    // lane / 4 at LDS chunk position lane % 4, which holds logical 16-byte
    // chunk c = position ^ ((r >> 1) & 3) (the lds_at swizzle)
    // per DMA instruction i: this lane's global source at K tile 0 and the
    // wave's LDS destination in stage 0 (scalar, M0 takes it as is)
    const int8_t *srcA[2], *srcB[2];
    uint32_t dstA[2], dstB[2];
    const uint32_t lA = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) int8_t *)sA;
    const uint32_t lB = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) int8_t *)sB;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int r = (wave * 2 + i) * 16 + (lane >> 2);
        const int c = (lane & 3) ^ ((r >> 1) & 3);
        const int64_t goff = (int64_t)(r >> 7) * next_rows + (r & 127) * OZ_BK + c * 16;
        srcA[i] = gA + panel_off(m0, 0, kp) + goff;
        srcB[i] = gB + panel_off(n0, 0, kp) + goff;
        dstA[i] = __builtin_amdgcn_readfirstlane(lA + (wave * 2 + i) * 1024);
        dstB[i] = __builtin_amdgcn_readfirstlane(lB + (wave * 2 + i) * 1024);
    }
    auto dma = [&](int f, int stage) {
        const int64_t off = (int64_t)f * OZ_CHUNK;
        const uint32_t so = (uint32_t)stage * OZ_STAGE;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            dma16_at(srcA[i] + off, dstA[i] + so);
            dma16_at(srcB[i] + off, dstB[i] + so);
        }
    };
    const int a_row = wm * 128 + l16, b_row = wn * 64 + l16;
    auto rdA = [&](int stage, int mb) {
        return *(const i32x4 *)&sA[stage * OZ_STAGE + lds_at(a_row + mb * 16, lg)];
    };
    auto rdB = [&](int stage, int nb) {
        return *(const i32x4 *)&sB[stage * OZ_STAGE + lds_at(b_row + nb * 16, lg)];
    };

    i32x4 fa[8], fb[2][4];
    dma(0, 0);
    dma(1, 1);
    dma(min(2, last), 2);
    vm_wait<8>(); // tile 0 (the first step waits for tile 1)
    __builtin_amdgcn_s_barrier();
#pragma unroll
    for (int mb = 0; mb < 6; ++mb) fa[mb] = rdA(0, mb);
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) fb[0][nb] = rdB(0, nb);
    // enter the loop with no LDS or scalar load outstanding (a scalar load
    // hoisted above the loop would otherwise make every step's first wait a
    // full drain)
    lgkm_wait0();

    int cur = 0; // stage of tile t
    auto step = [&](int t, auto par_c) {
        constexpr int P = decltype(par_c)::value;
        const int nxt = (cur + 1) & 3, wr = (cur + 3) & 3;
        dma(min(t + 3, last), wr); // tile t + 3
        fa[6] = rdA(cur, 6);
        fa[7] = rdA(cur, 7);
        // row 0 first: its MFMAs keep the matrix core busy across the barrier
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) acc[0][nb] = MAC::mac(fa[0], fb[P][nb], acc[0][nb]);
        __builtin_amdgcn_sched_group_barrier(0x100, 2, 0); // last A fragments of tile t
        __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
        __builtin_amdgcn_sched_barrier(0);
        // this wave's reads of tile t are done and its DMAs of tile t + 1
        // have landed (tile t + 2's and t + 3's may run): after the barrier
        // every wave may read tile t + 1, and the next step's DMA may
        // overwrite tile t's stage
        lgkm_wait0();
        vm_wait<8>();
        __builtin_amdgcn_s_barrier();
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) fb[P ^ 1][nb] = rdB(nxt, nb);
        fa[0] = rdA(nxt, 0);
#pragma unroll
        for (int mb = 1; mb < 8; ++mb) {
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
                acc[mb][nb] = MAC::mac(fa[mb], fb[P][nb], acc[mb][nb]);
            if (mb < 6) fa[mb] = rdA(nxt, mb);
        }
        // pin that order (the scheduler would otherwise hoist the reads of
        // tile t + 1 above the MFMAs that still need the registers)
        __builtin_amdgcn_sched_group_barrier(0x100, 5, 0);
#pragma unroll
        for (int mb = 1; mb < 6; ++mb) {
            __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);
        __builtin_amdgcn_sched_barrier(0);
        cur = nxt;
    };
    for (int t = 0; t < ktiles; t += 2) {
        step(t, Const<0>{});
        step(t + 1, Const<1>{});
    }
    // drain the DMAs past the end (clamped re-loads of the last tile) before
    // the stages are reused or the workgroup's LDS is released
    vm_wait<0>();
    __builtin_amdgcn_s_barrier();
}

template <typename MAC = MacI8>
__device__ __forceinline__ void tile256(const OzSeg *__restrict__ segs, int nseg, int kp, int m0,
                                        int n0, typename MAC::acc_t (&acc)[8][4])
{
    __shared__ __attribute__((aligned(16))) int8_t sA[OZ_NSTAGE * OZ_STAGE];
    __shared__ __attribute__((aligned(16))) int8_t sB[OZ_NSTAGE * OZ_STAGE];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = MAC::zero();
    // each segment restarts the pipeline (its final barrier has retired every
    // read and DMA of the stages the next prologue overwrites)
    for (int s = 0; s < nseg; ++s)
        tile256_seg<MAC>(segs[s].A, segs[s].B, kp, m0, n0, acc, sA, sB);
}

// W = (first ? 0 : W) + scale * (tile sum); rows >= m / columns >= n skipped.
// Wave (wm, wn) = (wave & 1, wave >> 1) owns rows m0 + WTM wm + 16 mb + ...,
// columns n0 + WTN wn + 16 nb + lane % 16.
template <int WTM, int WTN, int NBM, int NBN, typename ACC = i32x4>
__device__ __forceinline__ void fold_w(const ACC (&acc)[NBM][NBN], int m0, int n0, int m, int n,
                                       double *__restrict__ W, int64_t ldw, double scale,
                                       int first)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int mb = 0; mb < NBM; ++mb)
#pragma unroll
        for (int nb2 = 0; nb2 < NBN; ++nb2) {
            const int col = n0 + wn * WTN + nb2 * 16 + l16;
            if (col >= n) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm * WTM + mb * 16 + 4 * lg + v;
                if (row >= m) continue;
                double *w = W + row + (int64_t)col * ldw;
                const double add = scale * (double)acc[mb][nb2][v];
                *w = first ? add : *w + add;
            }
        }
}

__global__ __launch_bounds__(OZ_NT, 2) void k_ozaki_group(const OzSeg *__restrict__ segs,
                                                          int nseg, int kp, int m, int n,
                                                          double *__restrict__ W, int64_t ldw,
                                                          double scale, int first)
{
    int m0, n0;
    block_origin<OZ_BM>(n, m0, n0);
    i32x4 acc[4][4];
    tile128(segs, nseg, kp, m0, n0, acc);
    fold_w<64, 64, 4, 4>(acc, m0, n0, m, n, W, ldw, scale, first);
}

__global__ __launch_bounds__(OZ_NT_BIG, 1) void k_ozaki_group_big(const OzSeg *__restrict__ segs,
                                                                  int nseg, int kp, int m, int n,
                                                                  double *__restrict__ W,
                                                                  int64_t ldw, double scale,
                                                                  int first)
{
    int m0, n0;
    block_origin<OZ_BIG>(n, m0, n0);
    i32x4 acc[8][4];
    tile256(segs, nseg, kp, m0, n0, acc);
    fold_w<128, 64, 8, 4>(acc, m0, n0, m, n, W, ldw, scale, first);
}

// ---------------------------------------------------------------------------
// BF16x3 split (the BF16 variant beside the INT8 digit and CRT schemes): the
// scaled entry x = 2^-e_i a_ik in (-1, 1) is cut into three bf16 pieces
//   s1 = rn8(x), s2 = rn8(x - s1), s3 = rn8(x - s1 - s2),
// rn8 rounding to 8 significant bits (nearest, ties to even) on the double's
// bit pattern; each subtraction is exact in FP64, so x = s1 + s2 + s3 + r
// with |r| <= 2^-25 |x| (pieces below bf16's normal range are 0).  Products
// s_p s_q with p + q <= 4 run on the BF16 matrix cores with FP32 sums; the
// FP64 workspace adds the three groups (p + q = 4, 3, 2).  Planes [p][i][k]
// of bf16 in the panel layout, K counted in bytes (kp2 = 2 Kp).
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ double bf16_rn8(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    u += 0x00000FFFFFFFFFFFull + ((u >> 45) & 1);
    u &= ~0x00001FFFFFFFFFFFull;
    double y;
    memcpy(&y, &u, 8);
    return fabs(y) < 0x1p-126 ? 0.0 : y;
}

// bf16 bits of a value rn8 produced (exact: 8 significant bits, normal range)
__host__ __device__ __forceinline__ uint16_t bf16_bits(double y)
{
    const float f = (float)y;
    uint32_t b;
    memcpy(&b, &f, 4);
    return (uint16_t)(b >> 16);
}

__host__ __device__ __forceinline__ void bf16x3_pieces(double x, uint16_t *out)
{
    const double s1 = bf16_rn8(x), r1 = x - s1;
    const double s2 = bf16_rn8(r1), s3 = bf16_rn8(r1 - s2);
    out[0] = bf16_bits(s1);
    out[1] = bf16_bits(s2);
    out[2] = bf16_bits(s3);
}

// contiguous rows (x(i,k) = X[k + i*ld]): one thread per 8 consecutive k
__global__ void k_bf16_split_contig(const double *__restrict__ X, int64_t ld, int n, int len,
                                    const int *__restrict__ e, int8_t *__restrict__ D, int np,
                                    int kp2)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int chunks = kp2 / 16;
    if (t >= (int64_t)np * chunks) return;
    const int i = (int)(t / chunks), k0 = (int)(t % chunks) * 8;
    const int ei = i < n ? e[i] : 0;
    uint16_t pc[3][8];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const double x = (i < n && k0 + b < len) ? ldexp(X[k0 + b + (int64_t)i * ld], -ei) : 0.0;
        uint16_t q[3];
        bf16x3_pieces(x, q);
        pc[0][b] = q[0];
        pc[1][b] = q[1];
        pc[2][b] = q[2];
    }
    for (int p = 0; p < 3; ++p) {
        i32x4 w;
        for (int h = 0; h < 4; ++h) w[h] = (int)((uint32_t)pc[p][2 * h] | ((uint32_t)pc[p][2 * h + 1] << 16));
        *(i32x4 *)(D + (int64_t)p * np * kp2 + panel_off(i, 2 * k0, kp2)) = w;
    }
}

// strided rows (x(i,k) = X[i + k*ld]): a 64 x 32 tile through LDS so that
// the FP64 reads (along i) and the bf16 writes (along k) are both coalesced
__global__ void k_bf16_split_strided(const double *__restrict__ X, int64_t ld, int n, int len,
                                     const int *__restrict__ e, int8_t *__restrict__ D, int np,
                                     int kp2)
{
    __shared__ uint16_t tile[3][64][34];
    const int i0 = blockIdx.x * 64, k0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6; // 64 x 4
    for (int kk = ty; kk < 32; kk += 4) {
        const int i = i0 + tx, k = k0 + kk;
        const double x = (i < n && k < len) ? ldexp(X[i + (int64_t)k * ld], -e[i]) : 0.0;
        uint16_t q[3];
        bf16x3_pieces(x, q);
        tile[0][tx][kk] = q[0];
        tile[1][tx][kk] = q[1];
        tile[2][tx][kk] = q[2];
    }
    __syncthreads();
    // write: each thread 8 consecutive k (16 bytes) of one row
    const int row = threadIdx.x >> 2, kc = (threadIdx.x & 3) * 8;
    const int i = i0 + row;
    if (i >= np) return;
    for (int p = 0; p < 3; ++p) {
        i32x4 w;
        for (int h = 0; h < 4; ++h)
            w[h] = (int)((uint32_t)tile[p][row][kc + 2 * h] |
                         ((uint32_t)tile[p][row][kc + 2 * h + 1] << 16));
        *(i32x4 *)(D + (int64_t)p * np * kp2 + panel_off(i, 2 * (k0 + kc), kp2)) = w;
    }
}

// W = (first ? 0 : W) + sum over segments of S_A S_B^T on the BF16 matrix
// cores (FP32 sums per tile)
__global__ __launch_bounds__(OZ_NT, 2) void k_bf16_group(const OzSeg *__restrict__ segs, int nseg,
                                                         int kp2, int m, int n,
                                                         double *__restrict__ W, int64_t ldw,
                                                         int first)
{
    int m0, n0;
    block_origin<OZ_BM>(n, m0, n0);
    f32x4 acc[4][4];
    tile128<MacBF16>(segs, nseg, kp2, m0, n0, acc);
    fold_w<64, 64, 4, 4, f32x4>(acc, m0, n0, m, n, W, ldw, 1.0, first);
}

__global__ __launch_bounds__(OZ_NT_BIG, 1) void k_bf16_group_big(const OzSeg *__restrict__ segs,
                                                                 int nseg, int kp2, int m, int n,
                                                                 double *__restrict__ W,
                                                                 int64_t ldw, int first)
{
    int m0, n0;
    block_origin<OZ_BIG>(n, m0, n0);
    f32x4 acc[8][4];
    tile256<MacBF16>(segs, nseg, kp2, m0, n0, acc);
    fold_w<128, 64, 8, 4, f32x4>(acc, m0, n0, m, n, W, ldw, 1.0, first);
}

// 4. C = alpha * 2^(e_i + f_j) * W + beta * C
__global__ void k_ozaki_finish(int m, int n, const double *__restrict__ W, int64_t ldw,
                               const int *__restrict__ ea, const int *__restrict__ eb,
                               double alpha, double beta, int beta_zero, double *__restrict__ C,
                               int64_t ldc)
{
    const int64_t total = (int64_t)m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % m), j = (int)(t / m);
        const double v = alpha * ldexp(W[i + (int64_t)j * ldw], ea[i] + eb[j]);
        double *c = C + i + (int64_t)j * ldc;
        *c = beta_zero ? v : fma(beta, *c, v);
    }
}

// complex support: split a complex matrix (rows x cols, ld) into dense real
// and imaginary planes (ld = rows); conj negates the imaginary plane
__global__ void k_deinterleave(const double *__restrict__ Z, int64_t ld, int rows, int cols,
                               int conj, double *__restrict__ re, double *__restrict__ im)
{
    const int64_t total = (int64_t)rows * cols;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % rows), j = (int)(t / rows);
        const double *z = Z + 2 * (i + (int64_t)j * ld);
        re[t] = z[0];
        im[t] = conj ? -z[1] : z[1];
    }
}

// C = alpha * (Pr + i Pi) + beta * C with a fixed FMA order
__global__ void k_zcombine(int m, int n, const double *__restrict__ pr, const double *__restrict__ pi,
                           double ar, double ai, double br, double bi, int beta_zero,
                           double *__restrict__ C, int64_t ldc)
{
    const int64_t total = (int64_t)m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % m), j = (int)(t / m);
        double *c = C + 2 * (i + (int64_t)j * ldc);
        const double tr = fma(ar, pr[t], -(ai * pi[t]));
        const double ti = fma(ar, pi[t], ai * pr[t]);
        if (beta_zero) {
            c[0] = tr;
            c[1] = ti;
        } else {
            const double cr = c[0], cim = c[1];
            c[0] = fma(br, cr, fma(-bi, cim, tr));
            c[1] = fma(br, cim, fma(bi, cr, ti));
        }
    }
}

// ---------------------------------------------------------------------------
// CRT split (Ozaki scheme II: Ozaki, Uchino, Imamura 2025).  Entries are
// scaled to integers A'_ik = trunc(a_ik 2^(beta - e_i)) with ||A'_i||_2 <
// 2^beta (e_i from the row's 2-norm); for each modulus m_l (pairwise
// coprime, <= 256) the symmetric residues a_l = smod(A', m_l) in
// [-floor(m/2), ceil(m/2) - 1] fit int8, one exact INT8 GEMM per modulus
// gives R_l = smod(a_l b_l^T, m_l), and the Chinese remainder theorem
// recovers the integer product (|A'B'| < 2^(2 beta) < M/2 by Cauchy-Schwarz,
// M = prod m_l, beta = floor((log2 M - 1) / 2)) in floating point:
//   Q_k = sum_l Wk_l R_l      (w_l = (M/m_l)((M/m_l)^-1 mod m_l) cut into
//                              40-bit chunks k = 1..4: exact sums)
//   kq  = rint((Q_1 + Q_2) / M)  (within 2^-38 of Q/M = kq + P/M, |P/M| <
//         1/2 - 2^-21: exact),  D_k = Q_k - kq M_k  (exact)
//   S   = ((D_1 + D_2) + D_3) + D_4 = A'B' (1 + 3u)
// and C = alpha 2^(e_i + f_j - 2 beta) S + beta C.
// ---------------------------------------------------------------------------
struct CrtMods {
    int L;
    int m[20];
    float rinvf[20];     // 1 / m_l, rounded
    int8_t c16[20];      // 2^16 smod m_l
    uint32_t c7[3][20];  // bytes 2^(7 j) smod m_l, j < 12 (byte j % 4 of word j / 4)
};

// r in [lo - m, hi + m] -> [lo, hi], lo = -floor(m/2), hi = ceil(m/2) - 1
__host__ __device__ __forceinline__ int smod_fix(int r, int m)
{
    const int lo = -(m >> 1), hi = ((m + 1) >> 1) - 1;
    r = r > hi ? r - m : r;
    r = r < lo ? r + m : r;
    return r;
}

// |x| < 2^22: the float quotient is within 1 of round(x / m)
__host__ __device__ __forceinline__ int smod_small(int x, int m, float rinvf)
{
    const int q = (int)rintf((float)x * rinvf);
    return smod_fix(x - q * m, m);
}

// The symmetric residue of s modulo m (odd m, or 256) in the low byte of the
// result, for |s| < 2^22 + 2^17.  t = fl(s fl(1/m) + 1.5 2^23), one rounding
// in the FMA, has 2^22 + q in its low 24 bits with q = round(s / m) exactly:
// for odd m, s / m is at least 1/(2m) from every half-integer, and the
// product's error |s| 2^-24 / m is below 0.26/m.  s - (2^22 + q) m = (s - q
// m) - 2^22 m, and 2^22 m = 0 mod 256, so the low byte is that of the
// symmetric residue s - q m.  m = 256: the low byte is s's own, whatever q.
//
// The product (2^22 + q) m is a full-rate 24-bit multiply of the low 24 bits
// of t by m (m wave-uniform).  Written out in assembly: only the product's
// low byte is used, so the compiler drops the operands' range and emits the
// quarter-rate 32-bit multiply.
__host__ __device__ __forceinline__ uint32_t mul_u24(uint32_t a, uint32_t m)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "s"(m), "v"(a));
    return r;
#else
    return (a & 0xFFFFFFu) * (m & 0xFFFFFFu);
#endif
}
__host__ __device__ __forceinline__ uint32_t resid_byte(int s, int m, float rinvf)
{
    const float t = fmaf((float)s, rinvf, 0x1.8p23f);
    return (uint32_t)s - mul_u24(__builtin_bit_cast(uint32_t, t), (uint32_t)m);
}

// any int32 x: x = 2^16 h + l (h = x >> 16, l = x & 0xFFFF) and s = h (2^16
// smod m) + l = x (mod m) with |s| <= 2^15 127 + 2^16 - 1 < 2^22 + 2^17
__host__ __device__ __forceinline__ uint32_t resid_byte_i32(int x, int m, int c16, float rinvf)
{
    return resid_byte((x >> 16) * c16 + (x & 0xFFFF), m, rinvf);
}

// four low bytes -> one word (byte v from b_v): three byte permutes
__host__ __device__ __forceinline__ uint32_t pack4(uint32_t b0, uint32_t b1, uint32_t b2,
                                                   uint32_t b3)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t lo = __builtin_amdgcn_perm(b1, b0, 0x0c0c0400u); // b0.0 b1.0 0 0
    const uint32_t hi = __builtin_amdgcn_perm(b3, b2, 0x0c0c0400u);
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
#else
    return (b0 & 0xFF) | (b1 & 0xFF) << 8 | (b2 & 0xFF) << 16 | b3 << 24;
#endif
}

// sum over the four byte lanes of a * b (signed bytes) plus c: V_DOT4_I32_I8
__host__ __device__ __forceinline__ int dot4_i8(uint32_t a, uint32_t b, int c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sdot4((int)a, (int)b, c, false);
#else
    for (int v = 0; v < 4; ++v) c += (int8_t)(a >> (8 * v)) * (int8_t)(b >> (8 * v));
    return c;
#endif
}

// the same with c = 0, b wave-uniform.  The builtin with a zero accumulator
// becomes a register move and the accumulating form; the three-operand form
// takes the zero as an inline constant (register operations only).
__host__ __device__ __forceinline__ int dot4_i8_first(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    int r;
    asm("v_dot4_i32_i8 %0, %1, %2, 0" : "=v"(r) : "v"(a), "s"(b));
    return r;
#else
    return dot4_i8(a, b, 0);
#endif
}

// An integer-valued double v, |v| < 2^(28 ND), as 4 ND signed 7-bit digits:
// |v| = sum_j d_j 2^(7 j) with d_j in [0, 128), byte j % 4 of word j / 4
// holding sign(v) d_j in two's complement.  Non-finite values and |v| >=
// 2^(28 ND) are taken as zero.  Every step is exact: |v| is cut into 28-bit
// integers (trunc of exact power-of-two multiples, FMA remainders below 2^56
// and 2^28), each spread to four bytes of 7 bits; the bytewise negation
// (0x80 - d) ^ 0x80 = -d (mod 256) never borrows across bytes (d < 0x80).
template <int ND> struct Dig7 {
    uint32_t w[ND];
};
__host__ __device__ __forceinline__ uint32_t spread7(uint32_t x) // x < 2^28
{
    x = (x & 0x3FFFu) | ((x << 2) & 0x3FFF0000u);        // 14-bit halves at bits 0, 16
    return (x & 0x007F007Fu) | ((x << 1) & 0x7F007F00u); // 7-bit digits at bytes
}
template <int ND> __host__ __device__ __forceinline__ Dig7<ND> digits7(double v)
{
    static_assert(ND == 2 || ND == 3, "digit words");
    double a = fabs(v);
    if (!(a < (ND == 2 ? 0x1p56 : 0x1p84))) a = 0.0;
    uint32_t c[3];
    if (ND == 3) {
        const double top = trunc(a * 0x1p-56);
        a = fma(-top, 0x1p56, a);
        c[2] = (uint32_t)top;
    }
    const double hd = trunc(a * 0x1p-28);
    c[1] = (uint32_t)hd;
    c[0] = (uint32_t)fma(-hd, 0x1p28, a);
    Dig7<ND> r;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        const uint32_t w = spread7(c[j]);
        r.w[j] = v < 0 ? (0x80808080u - w) ^ 0x80808080u : w;
    }
    return r;
}

// Symmetric residue of v modulo m (odd, or 256) from its digits, in the low
// byte: s = sum_j digit_j (2^(7 j) smod m) = v (mod m) on the dot-product
// unit, exact (|s| <= 12 * 127 * 128 < 2^18), then resid_byte.  The
// constants of one modulus are read once (`ModConst`) and reused for every
// element.
struct ModConst {
    uint32_t c[3];
    float rinv;
    int m;
};
__host__ __device__ __forceinline__ ModConst mod_const(const CrtMods &md, int l)
{
    ModConst k;
#pragma unroll
    for (int j = 0; j < 3; ++j) k.c[j] = md.c7[j][l];
    k.rinv = md.rinvf[l];
    k.m = md.m[l];
    return k;
}
template <int ND>
__host__ __device__ __forceinline__ int digit_sum(const Dig7<ND> &x, const ModConst &k)
{
    int s = dot4_i8_first(x.w[0], k.c[0]);
#pragma unroll
    for (int j = 1; j < ND; ++j) s = dot4_i8(x.w[j], k.c[j], s);
    return s;
}
template <int ND>
__host__ __device__ __forceinline__ int digit_sum_add(const Dig7<ND> &x, const ModConst &k, int s)
{
#pragma unroll
    for (int j = 0; j < ND; ++j) s = dot4_i8(x.w[j], k.c[j], s);
    return s;
}
template <int ND>
__host__ __device__ __forceinline__ int digit_resid(const Dig7<ND> &x, const ModConst &k)
{
    return (int8_t)resid_byte(digit_sum<ND>(x, k), k.m, k.rinv);
}

// one plane's residues of 16 consecutive-k elements, packed for a 16-byte
// store
template <int ND>
__device__ __forceinline__ i32x4 pack_resid16(const Dig7<ND> (&v)[16], const ModConst &k)
{
    uint32_t b[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) b[e] = resid_byte(digit_sum<ND>(v[e], k), k.m, k.rinv);
    return i32x4{(int)pack4(b[0], b[1], b[2], b[3]), (int)pack4(b[4], b[5], b[6], b[7]),
                 (int)pack4(b[8], b[9], b[10], b[11]), (int)pack4(b[12], b[13], b[14], b[15])};
}

// contiguous (mode 1): x(i,k) = X[k + i*ld]; one thread per 16 consecutive
// k of one row; all L residue planes (plane stride np*kp).  Threads walk the
// planes' panel layout: the 4 chunks of a row's 64-byte K tile, then the
// 128 rows of its row block, then K tiles, then row blocks, so a wave's
// stores are 1 KB contiguous and 4 threads read a row's 512 bytes
template <int ND>
__global__ __launch_bounds__(256) void k_crt_split_contig(const double *__restrict__ X, int64_t ld, int n, int len,
                                   const int *__restrict__ e, int beta, CrtMods md,
                                   int8_t *__restrict__ D, int np, int kp)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int ktiles = kp / OZ_BK;
    const int64_t rest = t >> 9; // (row block, K tile)
    const int rb = (int)(rest / ktiles), kt = (int)(rest % ktiles);
    const int i = rb * OZ_PR + (int)((t >> 2) & (OZ_PR - 1)), k0 = kt * OZ_BK + (int)(t & 3) * 16;
    if (i >= np) return;
    Dig7<ND> v[16];
    const int sh = i < n ? beta - e[i] : 0;
    const double *x = X + k0 + (int64_t)i * ld;
    if (i < n && k0 + 16 <= len && ((((uintptr_t)x) & 15) == 0)) {
        // 16-byte loads, all in flight together
        double2 w[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) w[b] = *(const double2 *)(x + 2 * b);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            v[2 * b] = digits7<ND>(trunc(ldexp(w[b].x, sh)));
            v[2 * b + 1] = digits7<ND>(trunc(ldexp(w[b].y, sh)));
        }
    } else {
#pragma unroll
        for (int b = 0; b < 16; ++b)
            v[b] = digits7<ND>((i < n && k0 + b < len) ? trunc(ldexp(x[b], sh)) : 0.0);
    }
    for (int l = 0; l < md.L; ++l)
        *(i32x4 *)(D + (int64_t)l * np * kp + panel_off(i, k0, kp)) =
            pack_resid16<ND>(v, mod_const(md, l));
}

// strided (mode 0): x(i,k) = X[i + k*ld]; block = 64 rows x 64 k, thread =
// one row x 16 consecutive k, lanes in the planes' panel order (4 K chunks
// of a row, then 16 rows): a wave's FP64 reads are four 128-byte runs along
// i per k, and its store to each plane is 1 KB contiguous
template <int ND>
__global__ __launch_bounds__(256) void k_crt_split_strided(const double *__restrict__ X, int64_t ld,
                                                           int n, int len,
                                                           const int *__restrict__ e, int beta,
                                                           CrtMods md, int8_t *__restrict__ D,
                                                           int np, int kp)
{
    const int i = blockIdx.x * 64 + (threadIdx.x >> 2);
    const int k0 = blockIdx.y * 64 + (threadIdx.x & 3) * 16;
    if (i >= np) return;
    Dig7<ND> v[16];
    const int sh = i < n ? beta - e[i] : 0;
    const double *x = X + i + (int64_t)k0 * ld;
    if (i < n && k0 + 16 <= len) {
        // all 16 loads in flight together
        double xs[16];
#pragma unroll
        for (int b = 0; b < 16; ++b) xs[b] = x[b * ld];
#pragma unroll
        for (int b = 0; b < 16; ++b) v[b] = digits7<ND>(trunc(ldexp(xs[b], sh)));
    } else {
#pragma unroll
        for (int b = 0; b < 16; ++b)
            v[b] = digits7<ND>((i < n && k0 + b < len) ? trunc(ldexp(x[b * ld], sh)) : 0.0);
    }
    for (int l = 0; l < md.L; ++l)
        *(i32x4 *)(D + (int64_t)l * np * kp + panel_off(i, k0, kp)) =
            pack_resid16<ND>(v, mod_const(md, l));
}

// complex operands: planes 3l (real part), 3l+1 (imaginary part), 3l+2
// (their sum) of modulus l, from one shared exponent per row; conj negates
// the imaginary part.  Plane p: c = p % 3 of modulus p / 3.
template <int ND>
__device__ __forceinline__ i32x4 zcrt_pack16(const Dig7<ND> (&vr)[16], const Dig7<ND> (&vi)[16],
                                             int p, const CrtMods &md)
{
    const int l = p / 3, c = p - 3 * l;
    const ModConst k = mod_const(md, l);
    if (c == 0) return pack_resid16<ND>(vr, k);
    if (c == 1) return pack_resid16<ND>(vi, k);
    // the sum's digit sums add exactly (|s| < 2^19)
    uint32_t b[16];
#pragma unroll
    for (int e = 0; e < 16; ++e)
        b[e] = resid_byte(digit_sum_add<ND>(vi[e], k, digit_sum<ND>(vr[e], k)), k.m, k.rinv);
    return i32x4{(int)pack4(b[0], b[1], b[2], b[3]), (int)pack4(b[4], b[5], b[6], b[7]),
                 (int)pack4(b[8], b[9], b[10], b[11]), (int)pack4(b[12], b[13], b[14], b[15])};
}

// contiguous (mode 1): z(i,k) = Z[k + i*ld]; one thread per 16 k of one row
template <int ND>
__global__ __launch_bounds__(256) void k_zcrt_split_contig(const double *__restrict__ Z,
                                                           int64_t ld, int n, int len, int conj,
                                                           const int *__restrict__ e, int beta,
                                                           CrtMods md, int8_t *__restrict__ D,
                                                           int np, int kp)
{
    // threads in the planes' panel order, as k_crt_split_contig
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int ktiles = kp / OZ_BK;
    const int64_t rest = t >> 9;
    const int rb = (int)(rest / ktiles), kt = (int)(rest % ktiles);
    const int i = rb * OZ_PR + (int)((t >> 2) & (OZ_PR - 1)), k0 = kt * OZ_BK + (int)(t & 3) * 16;
    if (i >= np) return;
    Dig7<ND> vr[16], vi[16];
    const int sh = i < n ? beta - e[i] : 0;
    const double *z0 = Z + 2 * (k0 + (int64_t)i * ld);
    if (i < n && k0 + 16 <= len && ((((uintptr_t)z0) & 15) == 0)) {
        // 16-byte loads, all in flight together
        double2 zs[16];
#pragma unroll
        for (int b = 0; b < 16; ++b) zs[b] = *(const double2 *)(z0 + 2 * b);
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            vr[b] = digits7<ND>(trunc(ldexp(zs[b].x, sh)));
            vi[b] = digits7<ND>(trunc(ldexp(conj ? -zs[b].y : zs[b].y, sh)));
        }
    } else {
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            const bool in = i < n && k0 + b < len;
            const double *z = z0 + 2 * b;
            vr[b] = digits7<ND>(in ? trunc(ldexp(z[0], sh)) : 0.0);
            vi[b] = digits7<ND>(in ? trunc(ldexp(conj ? -z[1] : z[1], sh)) : 0.0);
        }
    }
    for (int p = 0; p < 3 * md.L; ++p)
        *(i32x4 *)(D + (int64_t)p * np * kp + panel_off(i, k0, kp)) = zcrt_pack16<ND>(vr, vi, p, md);
}

// strided (mode 0): z(i,k) = Z[i + k*ld]; threads as k_crt_split_strided
template <int ND>
__global__ __launch_bounds__(256) void k_zcrt_split_strided(const double *__restrict__ Z,
                                                            int64_t ld, int n, int len, int conj,
                                                            const int *__restrict__ e, int beta,
                                                            CrtMods md, int8_t *__restrict__ D,
                                                            int np, int kp)
{
    const int i = blockIdx.x * 64 + (threadIdx.x >> 2);
    const int k0 = blockIdx.y * 64 + (threadIdx.x & 3) * 16;
    if (i >= np) return;
    Dig7<ND> vr[16], vi[16];
    const int sh = i < n ? beta - e[i] : 0;
    const double *z0 = Z + 2 * (i + (int64_t)k0 * ld);
    if (i < n && k0 + 16 <= len && ((((uintptr_t)Z) & 15) == 0)) {
        // 16-byte loads (every element 16-byte aligned), all in flight together
        double2 zs[16];
#pragma unroll
        for (int b = 0; b < 16; ++b) zs[b] = *(const double2 *)(z0 + 2 * b * ld);
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            vr[b] = digits7<ND>(trunc(ldexp(zs[b].x, sh)));
            vi[b] = digits7<ND>(trunc(ldexp(conj ? -zs[b].y : zs[b].y, sh)));
        }
    } else {
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            const bool in = i < n && k0 + b < len;
            const double *z = z0 + 2 * b * ld;
            vr[b] = digits7<ND>(in ? trunc(ldexp(z[0], sh)) : 0.0);
            vi[b] = digits7<ND>(in ? trunc(ldexp(conj ? -z[1] : z[1], sh)) : 0.0);
        }
    }
    for (int p = 0; p < 3 * md.L; ++p)
        *(i32x4 *)(D + (int64_t)p * np * kp + panel_off(i, k0, kp)) = zcrt_pack16<ND>(vr, vi, p, md);
}

// residues of one tile sum -> int8 plane (column-major, ld ldr); every lane
// owns 4 consecutive rows of one column: one 32-bit store each
template <int WTM, int WTN, int NBM, int NBN>
__device__ __forceinline__ void store_resid(const i32x4 (&acc)[NBM][NBN], int m0, int n0,
                                            const CrtMods &md, int l, int8_t *__restrict__ R,
                                            int64_t ldr)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lg = lane >> 4;
    const int mod = md.m[l], c16 = md.c16[l];
    const float rinvf = md.rinvf[l];
#pragma unroll
    for (int mb = 0; mb < NBM; ++mb)
#pragma unroll
        for (int nb2 = 0; nb2 < NBN; ++nb2) {
            const int col = n0 + wn * WTN + nb2 * 16 + l16;
            const int row = m0 + wm * WTM + mb * 16 + 4 * lg;
            uint32_t b[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) b[v] = resid_byte_i32(acc[mb][nb2][v], mod, c16, rinvf);
            *(uint32_t *)(R + row + (int64_t)col * ldr) = pack4(b[0], b[1], b[2], b[3]);
        }
}

// one INT8 GEMM per modulus (blockIdx.y) over padded operands; R planes are
// mp x np (ld mp), so no bounds checks
// product p = blockIdx.y uses modulus p / per (per = 3 for complex planes)
__global__ __launch_bounds__(OZ_NT, 2) void k_crt_gemm(const OzSeg *__restrict__ segs, int kp,
                                                       int np, int per, CrtMods md,
                                                       int8_t *__restrict__ R, int64_t ldr,
                                                       int64_t plane)
{
    int m0, n0;
    block_origin<OZ_BM>(np, m0, n0);
    const int p = blockIdx.y, l = p / per;
    i32x4 acc[4][4];
    tile128(segs + p, 1, kp, m0, n0, acc);
    store_resid<64, 64, 4, 4>(acc, m0, n0, md, l, R + p * plane, ldr);
}

// Tile of this block for the CRT products: the operand panels of the tiles
// that all XCDs run at one time should stay in the Infinity Cache while
// they are reused.  Each XCD runs 32 tiles at a time (one per CU) as a
// footprint of 4 tile rows x 8 tile columns (12 panels in its L2); the 8
// footprints of a time step cover 32x8 tiles, two steps a super-tile of 32x16
// (32 + 16 panels of 4 MB at K = 16384), and the super-tiles go row by row,
// so the 32 row panels are read once per row of super-tiles.  Measured at
// 16384^3, 18 moduli: 3.4% faster than XCD ranges of 8 full tile rows (16x16
// super-tiles: 2.8%); the digit-split and BF16x3 kernels, whose segments
// run K several times over and whose epilogue adds into an FP64 W, ran
// 2-3% slower with it and keep block_origin.  Shapes that do not divide
// fall back to 16x16 super-tiles, then to block_origin.
__host__ __device__ inline void crt_tile_origin_of(int b, int nb, int np, int &m0, int &n0)
{
    const int tiles_n = np / OZ_BIG, tiles_m = nb / tiles_n;
    const int x = b & 7, q = b >> 3, c = q & 31;
    if ((nb & 7) == 0 && tiles_m % 32 == 0 && tiles_n % 16 == 0) {
        // XCD x: rows 4 x .. 4 x + 3 of the super-tile, half h of its columns
        const int sup = q >> 6, h = (q >> 5) & 1, sn = tiles_n / 16;
        m0 = (32 * (sup / sn) + 4 * x + (c & 3)) * OZ_BIG;
        n0 = (16 * (sup % sn) + 8 * h + (c >> 2)) * OZ_BIG;
    } else if ((nb & 7) == 0 && tiles_m % 16 == 0 && tiles_n % 16 == 0) {
        // 16x16 super-tiles, one per time step: XCD x rows 4 (x & 3) ..,
        // columns 8 (x >> 2) ..
        const int sup = q >> 5, sn = tiles_n / 16;
        m0 = (16 * (sup / sn) + 4 * (x & 3) + (c & 3)) * OZ_BIG;
        n0 = (16 * (sup % sn) + 8 * (x >> 2) + (c >> 2)) * OZ_BIG;
    } else {
        block_origin_of<OZ_BIG>(b, nb, np, m0, n0);
    }
}
__device__ __forceinline__ void crt_tile_origin(int np, int &m0, int &n0)
{
    crt_tile_origin_of((int)blockIdx.x, (int)gridDim.x, np, m0, n0);
}

// product p0 + blockIdx.y.  The host launches one product at a time when
// one fills the GPU: measured at 16384^3, a single launch over all products
// ran the same blocks 5% slower.
__global__ __launch_bounds__(OZ_NT_BIG, 1) void k_crt_gemm_big(const OzSeg *__restrict__ segs,
                                                               int kp, int np, int per, CrtMods md,
                                                               int8_t *__restrict__ R, int64_t ldr,
                                                               int64_t plane, int p0)
{
    int m0, n0;
    crt_tile_origin(np, m0, n0);
    const int p = p0 + (int)blockIdx.y, l = p / per;
    i32x4 acc[8][4];
    tile256(segs + p, 1, kp, m0, n0, acc);
    store_resid<128, 64, 8, 4>(acc, m0, n0, md, l, R + p * plane, ldr);
}

struct CrtRecon {
    int L;
    double W[4][20]; // 40-bit chunks of the CRT weights
    double Mk[4];    // 40-bit chunks of M
    double invM;     // 1 / (((M_4 + M_3) + M_2) + M_1)
};

// one thread per 4 consecutive rows of one column (one 32-bit load per
// plane; the planes are mp x np with mp a multiple of 128, so the 4 bytes
// are always inside the plane)
__device__ __forceinline__ double crt_float(const double (&q)[4], const CrtRecon &c)
{
    const double kq = rint((q[0] + q[1]) * c.invM);
    return ((fma(-kq, c.Mk[0], q[0]) + fma(-kq, c.Mk[1], q[1])) + fma(-kq, c.Mk[2], q[2])) +
           fma(-kq, c.Mk[3], q[3]);
}

__global__ void k_crt_recon(int m, int n, const int8_t *__restrict__ R, int64_t ldr,
                            int64_t plane, CrtRecon c, const int *__restrict__ ea,
                            const int *__restrict__ eb, int two_beta, double alpha, double beta,
                            int beta_zero, double *__restrict__ C, int64_t ldc)
{
    const int m4 = (m + 3) / 4;
    const int64_t total = (int64_t)m4 * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i0 = 4 * (int)(t % m4), j = (int)(t / m4);
        const int8_t *r = R + i0 + (int64_t)j * ldr;
        double q[4][4] = {};
        // six planes per iteration: their loads are issued together
#pragma unroll 6
        for (int l = 0; l < c.L; ++l) {
            const uint32_t w = gload((const uint32_t *)(r + l * plane));
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double rl = (double)(int8_t)(w >> (8 * v));
#pragma unroll
                for (int k = 0; k < 4; ++k) q[v][k] = fma(c.W[k][l], rl, q[v][k]);
            }
        }
        const int fj = eb[j] - two_beta;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + v;
            if (i >= m) break;
            const double x = alpha * ldexp(crt_float(q[v], c), ea[i] + fj);
            double *cp = C + i + (int64_t)j * ldc;
            *cp = beta_zero ? x : fma(beta, *cp, x);
        }
    }
}

// complex: per modulus Re = smod(R1 - R2), Im = smod(R3 - R1 - R2) from the
// three product planes 3l, 3l+1, 3l+2; CRT of both; C = alpha x + beta C
// with the complex FMA order of k_zcombine
__global__ void k_zcrt_recon(int m, int n, const int8_t *__restrict__ R, int64_t ldr,
                             int64_t plane, CrtRecon c, CrtMods md, const int *__restrict__ ea,
                             const int *__restrict__ eb, int two_beta, double ar, double ai,
                             double br, double bi, int beta_zero, double *__restrict__ C,
                             int64_t ldc)
{
    const int m4 = (m + 3) / 4;
    const int64_t total = (int64_t)m4 * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i0 = 4 * (int)(t % m4), j = (int)(t / m4);
        const int8_t *r = R + i0 + (int64_t)j * ldr;
        double q[4][2][4] = {};
        for (int l = 0; l < c.L; ++l) {
            const int mod = md.m[l];
            const uint32_t w1 = gload((const uint32_t *)(r + (3 * l) * plane));
            const uint32_t w2 = gload((const uint32_t *)(r + (3 * l + 1) * plane));
            const uint32_t w3 = gload((const uint32_t *)(r + (3 * l + 2) * plane));
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r1 = (int8_t)(w1 >> (8 * v)), r2 = (int8_t)(w2 >> (8 * v)),
                          r3 = (int8_t)(w3 >> (8 * v));
                const double vre = (double)smod_small(r1 - r2, mod, md.rinvf[l]);
                const double vim = (double)smod_small(r3 - r1 - r2, mod, md.rinvf[l]);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    q[v][0][k] = fma(c.W[k][l], vre, q[v][0][k]);
                    q[v][1][k] = fma(c.W[k][l], vim, q[v][1][k]);
                }
            }
        }
        const int fj = eb[j] - two_beta;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + v;
            if (i >= m) break;
            const int sh = ea[i] + fj;
            const double x0 = ldexp(crt_float(q[v][0], c), sh), x1 = ldexp(crt_float(q[v][1], c), sh);
            const double tr = fma(ar, x0, -(ai * x1));
            const double ti = fma(ar, x1, ai * x0);
            double *cp = C + 2 * (i + (int64_t)j * ldc);
            if (beta_zero) {
                cp[0] = tr;
                cp[1] = ti;
            } else {
                const double cr = cp[0], cim = cp[1];
                cp[0] = fma(br, cr, fma(-bi, cim, tr));
                cp[1] = fma(br, cim, fma(bi, cr, ti));
            }
        }
    }
}

thread_local char g_oz_err[256];
int oz_fail(const char *what, hipError_t e)
{
    snprintf(g_oz_err, sizeof g_oz_err, "%s: %s", what, hipGetErrorString(e));
    return BLASX_EHIP;
}

} // namespace

namespace {
template <int NORM>
void exp_launch(const double *X, int64_t ld, int n, int len, int contiguous, int per,
                       int *e, double *part, hipStream_t st)
{
    if (contiguous) {
        hipLaunchKernelGGL(k_exp_contig<NORM>, dim3(n), dim3(256), 0, st, X, ld, n, len, e);
        return;
    }
    int chunk = (len + EXP_CHUNKS - 1) / EXP_CHUNKS;
    chunk = chunk < 4 ? 4 : (chunk + 3) / 4 * 4;
    const int chunks = len > 0 ? (len + chunk - 1) / chunk : 1;
    // per = 2: 2n real rows (re of row i at 2i, im at 2i+1)
    hipLaunchKernelGGL(k_amax_strided<NORM>, dim3((per * n + 63) / 64, chunks), dim3(256), 0, st,
                       X, ld, per * n, len, chunk, part);
    hipLaunchKernelGGL(k_exp_parts<NORM>, dim3((n + 255) / 256), dim3(256), 0, st, part, chunks, n,
                       per, e);
}

} // namespace

extern "C" {

const char *bxh_ozaki_error(void) { return g_oz_err; }

int bxh_oz_block_m(void) { return OZ_BM; }
int bxh_oz_block_k(void) { return OZ_BK; }

// exponents of the rows of op(X) (n rows, len = reduced extent):
// contiguous != 0: x(i,k) = X[k + i*ld], else x(i,k) = X[i + k*ld]
// part: EXP_CHUNKS * n doubles of scratch (contiguous == 0 only)
// norm = 0: max exponents (digit scheme); 1: 2-norm exponents (CRT scheme)
int bxh_oz_exponents(const double *X, int64_t ld, int n, int len, int contiguous, int norm,
                     int *e, double *part, void *stream)
{
    if (n <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (norm)
        exp_launch<1>(X, ld, n, len, contiguous, 1, e, part, st);
    else
        exp_launch<0>(X, ld, n, len, contiguous, 1, e, part, st);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("exponent launch", r);
}

// complex rows (interleaved re/im, ld in complex elements): one exponent per
// row from max(|re|, |im|) (norm = 0) or sum(re^2 + im^2) (norm = 1); part:
// EXP_CHUNKS * 2n doubles (contiguous == 0)
int bxh_oz_exponents_z(const double *Z, int64_t ld, int n, int len, int contiguous, int norm,
                       int *e, double *part, void *stream)
{
    if (n <= 0) return 0;
    if (contiguous) // row i: 2*len consecutive doubles
        return bxh_oz_exponents(Z, 2 * ld, n, 2 * len, 1, norm, e, part, stream);
    hipStream_t st = (hipStream_t)stream;
    if (norm)
        exp_launch<1>(Z, 2 * ld, n, len, 0, 2, e, part, st);
    else
        exp_launch<0>(Z, 2 * ld, n, len, 0, 2, e, part, st);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("exponent launch", r);
}

int bxh_oz_exp_chunks(void) { return EXP_CHUNKS; }

int bxh_oz_split(const double *X, int64_t ld, int n, int len, int contiguous, const int *e,
                 int s, int8_t *D, int np, int kp, void *stream)
{
    if (s < 1 || s > 9 || np % 64 || kp % 64) return BLASX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (contiguous) {
        const int64_t total = (int64_t)np * (kp / 16);
        hipLaunchKernelGGL(k_split_contig, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                           X, ld, n, len, e, s, D, np, kp);
    } else {
        hipLaunchKernelGGL(k_split_strided, dim3(np / SP_T, kp / SP_T), dim3(256), 0, st, X, ld,
                           n, len, e, s, D, np, kp);
    }
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("split launch", r);
}

// segs_dev: nseg (A slice, B slice) pointers, each slice np x kp / mp x kp
int bxh_oz_group(const void *segs_dev, int nseg, int kp, int m, int n, int mp, int np, double *W,
                 int64_t ldw, double scale, int first, void *stream)
{
    if (nseg <= 0) return 0;
    if (kp <= 0 || mp % OZ_BM || np % OZ_BN || kp % OZ_BK) return BLASX_EINVAL;
    // the 256x256 engine runs K tiles in pairs
    const int big_tiles = (mp % OZ_BIG == 0 && np % OZ_BIG == 0 && (kp / OZ_BK) % 2 == 0)
                              ? (mp / OZ_BIG) * (np / OZ_BIG)
                              : 0;
    if (big_tiles >= 256) { // enough 256x256 tiles to give every CU one
        hipLaunchKernelGGL(k_ozaki_group_big, dim3(big_tiles), dim3(OZ_NT_BIG), 0,
                           (hipStream_t)stream, (const OzSeg *)segs_dev, nseg, kp, m, n, W, ldw,
                           scale, first);
    } else {
        const int grid = (mp / OZ_BM) * (np / OZ_BN);
        hipLaunchKernelGGL(k_ozaki_group, dim3(grid), dim3(OZ_NT), 0, (hipStream_t)stream,
                           (const OzSeg *)segs_dev, nseg, kp, m, n, W, ldw, scale, first);
    }
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("ozaki group launch", r);
}

// ---- BF16x3 split ------------------------------------------------------------
int bxh_bf16_split(const double *X, int64_t ld, int n, int len, int contiguous, const int *e,
                   int8_t *D, int np, int kp2, void *stream)
{
    if (np % 64 || kp2 % 64) return BLASX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (contiguous) {
        const int64_t total = (int64_t)np * (kp2 / 16);
        hipLaunchKernelGGL(k_bf16_split_contig, dim3((unsigned)((total + 255) / 256)), dim3(256),
                           0, st, X, ld, n, len, e, D, np, kp2);
    } else {
        hipLaunchKernelGGL(k_bf16_split_strided, dim3(np / 64, kp2 / 64), dim3(256), 0, st, X, ld,
                           n, len, e, D, np, kp2);
    }
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("bf16 split launch", r);
}

int bxh_bf16_group(const void *segs_dev, int nseg, int kp2, int m, int n, int mp, int np,
                   double *W, int64_t ldw, int first, void *stream)
{
    if (nseg <= 0) return 0;
    if (kp2 <= 0 || mp % OZ_BM || np % OZ_BN || kp2 % OZ_BK) return BLASX_EINVAL;
    const int big_tiles = (mp % OZ_BIG == 0 && np % OZ_BIG == 0 && (kp2 / OZ_BK) % 2 == 0)
                              ? (mp / OZ_BIG) * (np / OZ_BIG)
                              : 0;
    if (big_tiles >= 256) {
        hipLaunchKernelGGL(k_bf16_group_big, dim3(big_tiles), dim3(OZ_NT_BIG), 0,
                           (hipStream_t)stream, (const OzSeg *)segs_dev, nseg, kp2, m, n, W, ldw,
                           first);
    } else {
        const int grid = (mp / OZ_BM) * (np / OZ_BN);
        hipLaunchKernelGGL(k_bf16_group, dim3(grid), dim3(OZ_NT), 0, (hipStream_t)stream,
                           (const OzSeg *)segs_dev, nseg, kp2, m, n, W, ldw, first);
    }
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("bf16 group launch", r);
}

// the host build of the piece split, for the tests
void bxh_bf16x3_pieces_host(double x, uint16_t *out) { bf16x3_pieces(x, out); }

// ---- CRT split -------------------------------------------------------------
// host view of the reconstruction constants (crt.c fills the same layout)
struct bxh_crt_recon_host {
    int L;
    double W[4][20];
    double Mk[4];
    double invM;
};
static_assert(sizeof(bxh_crt_recon_host) == sizeof(CrtRecon), "layout");

static CrtMods crt_mods(int L, const int *m)
{
    CrtMods md;
    memset(&md, 0, sizeof md);
    md.L = L;
    for (int l = 0; l < L; ++l) {
        md.m[l] = m[l];
        md.rinvf[l] = 1.0f / (float)m[l];
        md.c16[l] = smod_fix((int)((1u << 16) % (unsigned)m[l]), m[l]);
        uint64_t p = 1; // 2^(7 j) mod m
        for (int j = 0; j < 12; ++j) {
            const uint32_t c = (uint8_t)(int8_t)smod_fix((int)p, m[l]);
            md.c7[j / 4][l] |= c << (8 * (j % 4));
            p = (p << 7) % (uint64_t)m[l];
        }
    }
    return md;
}

// residue planes of op(X) rows (n x len, padded np x kp), modes as bxh_oz_split
int bxh_crt_split(const double *X, int64_t ld, int n, int len, int contiguous, const int *e,
                  int beta, int L, const int *moduli, int8_t *D, int np, int kp, void *stream)
{
    if (L < 1 || L > 20 || np % 64 || kp % 64) return BLASX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const CrtMods md = crt_mods(L, moduli);
    // |A'| < 2^beta: the fewest 28-bit digit words that cover beta bits
    const int ND = beta <= 56 ? 2 : 3;
    if (contiguous) {
        // whole 128-row blocks (the planes' panels), 4 threads per 64-byte K tile
        const int64_t total = (int64_t)((np + OZ_PR - 1) / OZ_PR) * OZ_PR * (kp / 16);
        const dim3 grid((unsigned)((total + 255) / 256));
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, X, ld, n, len, e, beta, md, D, np, kp);
        };
        if (ND == 2) go(k_crt_split_contig<2>);
        else go(k_crt_split_contig<3>);
    } else {
        const dim3 grid(np / 64, kp / 64);
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, X, ld, n, len, e, beta, md, D, np, kp);
        };
        if (ND == 2) go(k_crt_split_strided<2>);
        else go(k_crt_split_strided<3>);
    }
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("crt split launch", r);
}

// segs_dev: per*L (A plane, B plane) pointers; R: per*L planes of mp x np
// int8 (ld mp); product p uses modulus p / per
int bxh_crt_gemm(const void *segs_dev, int L, int per, const int *moduli, int kp, int mp, int np,
                 int8_t *R, void *stream)
{
    if (L < 1 || L > 20 || per < 1 || per > 3 || kp <= 0 || mp % OZ_BM || np % OZ_BN ||
        kp % OZ_BK)
        return BLASX_EINVAL;
    const CrtMods md = crt_mods(L, moduli);
    const int64_t plane = (int64_t)mp * np;
    const int nprod = per * L;
    // the 256x256 engine runs K tiles in pairs
    const int big_tiles = (mp % OZ_BIG == 0 && np % OZ_BIG == 0 && (kp / OZ_BK) % 2 == 0)
                              ? (mp / OZ_BIG) * (np / OZ_BIG)
                              : 0;
    if (big_tiles * nprod >= 256) {
        // one launch per product when one product fills the GPU, else all
        // products in one launch
        const int per_launch = big_tiles >= 256 ? 1 : nprod;
        for (int p = 0; p < nprod; p += per_launch)
            hipLaunchKernelGGL(k_crt_gemm_big, dim3(big_tiles, per_launch), dim3(OZ_NT_BIG), 0,
                               (hipStream_t)stream, (const OzSeg *)segs_dev, kp, np, per, md, R,
                               (int64_t)mp, plane, p);
    } else {
        hipLaunchKernelGGL(k_crt_gemm, dim3((mp / OZ_BM) * (np / OZ_BN), nprod), dim3(OZ_NT), 0,
                           (hipStream_t)stream, (const OzSeg *)segs_dev, kp, np, per, md, R,
                           (int64_t)mp, plane);
    }
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("crt gemm launch", r);
}

int bxh_crt_recon(int m, int n, const int8_t *R, int64_t ldr, int64_t plane, const void *recon,
                  const int *ea, const int *eb, int beta_bits, double alpha, double beta,
                  double *C, int64_t ldc, void *stream)
{
    if (m <= 0 || n <= 0) return 0;
    CrtRecon c;
    memcpy(&c, recon, sizeof c);
    const int64_t total = (int64_t)((m + 3) / 4) * n; // 4 rows per thread
    int grid = (int)((total + 255) / 256);
    if (grid > 262144) grid = 262144;
    hipLaunchKernelGGL(k_crt_recon, dim3(grid), dim3(256), 0, (hipStream_t)stream, m, n, R, ldr,
                       plane, c, ea, eb, 2 * beta_bits, alpha, beta, beta == 0.0 ? 1 : 0, C, ldc);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("crt recon launch", r);
}

// complex operand rows: 3L planes (re, im, re + im per modulus)
int bxh_zcrt_split(const double *Z, int64_t ld, int n, int len, int contiguous, int conj,
                   const int *e, int beta, int L, const int *moduli, int8_t *D, int np, int kp,
                   void *stream)
{
    if (L < 1 || L > 20 || np % 64 || kp % 64) return BLASX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const CrtMods md = crt_mods(L, moduli);
    // |A'| < 2^beta: the fewest 28-bit digit words that cover beta bits
    const int ND = beta <= 56 ? 2 : 3;
    if (contiguous) {
        // whole 128-row blocks (the planes' panels), 4 threads per 64-byte K tile
        const int64_t total = (int64_t)((np + OZ_PR - 1) / OZ_PR) * OZ_PR * (kp / 16);
        const dim3 grid((unsigned)((total + 255) / 256));
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, Z, ld, n, len, conj, e, beta, md, D,
                               np, kp);
        };
        if (ND == 2) go(k_zcrt_split_contig<2>);
        else go(k_zcrt_split_contig<3>);
    } else {
        const dim3 grid(np / 64, kp / 64);
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, Z, ld, n, len, conj, e, beta, md, D,
                               np, kp);
        };
        if (ND == 2) go(k_zcrt_split_strided<2>);
        else go(k_zcrt_split_strided<3>);
    }
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("zcrt split launch", r);
}

int bxh_zcrt_recon(int m, int n, const int8_t *R, int64_t ldr, int64_t plane, const void *recon,
                   const int *moduli, const int *ea, const int *eb, int beta_bits,
                   const double *alpha, const double *beta, double *C, int64_t ldc,
                   void *stream)
{
    if (m <= 0 || n <= 0) return 0;
    CrtRecon c;
    memcpy(&c, recon, sizeof c);
    const CrtMods md = crt_mods(c.L, moduli);
    const int64_t total = (int64_t)((m + 3) / 4) * n; // 4 rows per thread
    int grid = (int)((total + 255) / 256);
    if (grid > 262144) grid = 262144;
    const int bz = beta[0] == 0.0 && beta[1] == 0.0;
    hipLaunchKernelGGL(k_zcrt_recon, dim3(grid), dim3(256), 0, (hipStream_t)stream, m, n, R, ldr,
                       plane, c, md, ea, eb, 2 * beta_bits, alpha[0], alpha[1], beta[0], beta[1],
                       bz, C, ldc);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("zcrt recon launch", r);
}

// host-side check of the residue helpers the kernels use (same code, CPU)
int bxh_crt_resid_host(double v, int m)
{
    const CrtMods md = crt_mods(1, &m);
    return digit_resid<3>(digits7<3>(v), mod_const(md, 0));
}
int bxh_crt_resid2_host(double v, int m)
{
    const CrtMods md = crt_mods(1, &m);
    return digit_resid<2>(digits7<2>(v), mod_const(md, 0));
}
int bxh_crt_norm_exp_host(double ss) { return norm_exp_of(ss); }
// the INT8 CRT GEMM's block -> tile origin (same code, CPU): nb blocks over
// np columns of BIG-square tiles
void bxh_crt_tile_origin_host(int b, int nb, int np, int *m0, int *n0)
{
    crt_tile_origin_of(b, nb, np, *m0, *n0);
}
int bxh_crt_smod_i32_host(int x, int m)
{
    const CrtMods md = crt_mods(1, &m);
    return (int8_t)resid_byte_i32(x, md.m[0], md.c16[0], md.rinvf[0]);
}

int bxh_oz_deinterleave(const double *Z, int64_t ld, int rows, int cols, int conj, double *re,
                        double *im, void *stream)
{
    const int64_t total = (int64_t)rows * cols;
    if (total <= 0) return 0;
    int grid = (int)((total + 255) / 256);
    if (grid > 262144) grid = 262144;
    hipLaunchKernelGGL(k_deinterleave, dim3(grid), dim3(256), 0, (hipStream_t)stream, Z, ld,
                       rows, cols, conj, re, im);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("deinterleave launch", r);
}

int bxh_oz_zcombine(int m, int n, const double *pr, const double *pi, const double *alpha,
                    const double *beta, double *C, int64_t ldc, void *stream)
{
    const int64_t total = (int64_t)m * n;
    if (total <= 0) return 0;
    int grid = (int)((total + 255) / 256);
    if (grid > 262144) grid = 262144;
    const int bz = beta[0] == 0.0 && beta[1] == 0.0;
    hipLaunchKernelGGL(k_zcombine, dim3(grid), dim3(256), 0, (hipStream_t)stream, m, n, pr, pi,
                       alpha[0], alpha[1], beta[0], beta[1], bz, C, ldc);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("zcombine launch", r);
}

int bxh_oz_finish(int m, int n, const double *W, int64_t ldw, const int *ea, const int *eb,
                  double alpha, double beta, double *C, int64_t ldc, void *stream)
{
    if (m <= 0 || n <= 0) return 0;
    const int64_t total = (int64_t)m * n;
    int grid = (int)((total + 255) / 256);
    if (grid > 262144) grid = 262144;
    hipLaunchKernelGGL(k_ozaki_finish, dim3(grid), dim3(256), 0, (hipStream_t)stream, m, n, W,
                       ldw, ea, eb, alpha, beta, beta == 0.0 ? 1 : 0, C, ldc);
    hipError_t r = hipGetLastError();
    return r == hipSuccess ? 0 : oz_fail("ozaki finish launch", r);
}

} // extern "C"
