// gemm.hip -- the k-sum contraction G = A B on the matrix cores (gfx950).
//
// A [M x K] row-major: rows (4 iw + {Re explosion, Im explosion, Re force,
// Im force}) of the weighted k-kernels; B [K x N] row-major: J_n(k_i r_j);
// C [M x N].  Split-K over grid.z: every slice writes its partial product to
// P[z][M][N] and k_gemm_reduce sums the slices in FP64 in a fixed order.
// kend (optional, one entry per BM-row tile) limits the K range of a tile;
// nk_w (optional) masks A per row: row r is taken as zero at and beyond
// column nk_w[r / 4] (four operand rows per frequency), so entries past a
// frequency's wavenumber cut are never read.
//
//   k_gemm_f64    v_mfma_f64_16x16x4f64, 64x64 block tile, 4 waves of 32x32
//   k_gemm_f32    v_mfma_f32_32x32x2f32, 128x128 block tile, 4 waves of 64x64
//   k_gemm_bf16x3 v_mfma_f32_32x32x16_bf16 on split operands
//                 a = a_hi + a_lo, b = b_hi + b_lo (bf16 each):
//                 a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi, FP32 accumulation
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "kernels.h"

namespace grtm {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ FP64
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K, int kslice, const double *__restrict__ A,
                                                  long lda, const double *__restrict__ B, long ldb,
                                                  double *__restrict__ P, const int *__restrict__ kend,
                                                  const int *__restrict__ nk_w)
{
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kt_end = min(K, kend ? kend[blockIdx.y] : K);
    const int kb = blockIdx.z * kslice, ke = min(kt_end, kb + kslice);
    v4d acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = v4d{0, 0, 0, 0};
    for (int k0 = kb; k0 < ke; k0 += BK) {
        for (int e = tid; e < BM * BK; e += 256) {
            const int mm = e / BK, kk = e % BK, gm = m0 + mm, gk = k0 + kk;
            const int kr = (gm < M && nk_w) ? min(ke, nk_w[gm >> 2]) : ke;
            As[kk][mm] = (gm < M && gk < kr) ? A[(size_t)gm * lda + gk] : 0.0;
        }
        for (int e = tid; e < BN * BK; e += 256) {
            const int kk = e / BN, nn = e % BN, gk = k0 + kk, gn = n0 + nn;
            Bs[kk][nn] = (gk < ke && gn < N) ? B[(size_t)gk * ldb + gn] : 0.0;
        }
        __syncthreads();
        const int i = lane & 15, kq = lane >> 4;
        for (int kk = 0; kk < BK; kk += 4) {
            double a[2], b[2];
            for (int t = 0; t < 2; t++) a[t] = As[kk + kq][wm * 32 + t * 16 + i];
            for (int t = 0; t < 2; t++) b[t] = Bs[kk + kq][wn * 32 + t * 16 + i];
            for (int tm = 0; tm < 2; tm++)
                for (int tn = 0; tn < 2; tn++)
                    acc[tm][tn] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[tm], b[tn], acc[tm][tn], 0, 0, 0);
        }
        __syncthreads();
    }
    double *Pz = P + (size_t)blockIdx.z * M * N;
    for (int tm = 0; tm < 2; tm++)
        for (int tn = 0; tn < 2; tn++)
            for (int r = 0; r < 4; r++) {
                // accumulator r of lane l holds row (l / 16) + 4 r, column l % 16
                const int row = m0 + wm * 32 + tm * 16 + (lane >> 4) + 4 * r;
                const int col = n0 + wn * 32 + tn * 16 + (lane & 15);
                if (row < M && col < N) Pz[(size_t)row * N + col] = acc[tm][tn][r];
            }
}

// ------------------------------------------------------------------ FP32
__global__ __launch_bounds__(256) void k_gemm_f32(int M, int N, int K, int kslice, const float *__restrict__ A,
                                                  long lda, const float *__restrict__ B, long ldb,
                                                  float *__restrict__ P, const int *__restrict__ kend,
                                                  const int *__restrict__ nk_w)
{
    constexpr int BM = 128, BN = 128, BK = 16;
    __shared__ float As[BK][BM + 4];
    __shared__ float Bs[BK][BN + 4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kt_end = min(K, kend ? kend[blockIdx.y] : K);
    const int kb = blockIdx.z * kslice, ke = min(kt_end, kb + kslice);
    v16f acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int r = 0; r < 16; r++) acc[a][b][r] = 0.f;
    for (int k0 = kb; k0 < ke; k0 += BK) {
        {  // A tile: 128 rows x 16 k, 8 per thread
            const int mm = tid >> 1, kk0 = (tid & 1) * 8, gm = m0 + mm;
            const int kr = (gm < M && nk_w) ? min(ke, nk_w[gm >> 2]) : ke;
            for (int q = 0; q < 8; q++) {
                const int gk = k0 + kk0 + q;
                As[kk0 + q][mm] = (gm < M && gk < kr) ? A[(size_t)gm * lda + gk] : 0.f;
            }
        }
        {  // B tile: 16 k x 128 cols, 8 per thread
            const int kk = tid >> 4, nn0 = (tid & 15) * 8, gk = k0 + kk;
            for (int q = 0; q < 8; q++) {
                const int gn = n0 + nn0 + q;
                Bs[kk][nn0 + q] = (gk < ke && gn < N) ? B[(size_t)gk * ldb + gn] : 0.f;
            }
        }
        __syncthreads();
        const int i = lane & 31, kq = lane >> 5;
        for (int kk = 0; kk < BK; kk += 2) {
            float a[2], b[2];
            for (int t = 0; t < 2; t++) a[t] = As[kk + kq][wm * 64 + t * 32 + i];
            for (int t = 0; t < 2; t++) b[t] = Bs[kk + kq][wn * 64 + t * 32 + i];
            for (int tm = 0; tm < 2; tm++)
                for (int tn = 0; tn < 2; tn++)
                    acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[tm], b[tn], acc[tm][tn], 0, 0, 0);
        }
        __syncthreads();
    }
    float *Pz = P + (size_t)blockIdx.z * M * N;
    for (int tm = 0; tm < 2; tm++)
        for (int tn = 0; tn < 2; tn++)
            for (int r = 0; r < 16; r++) {
                const int row = m0 + wm * 64 + tm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int col = n0 + wn * 64 + tn * 32 + (lane & 31);
                if (row < M && col < N) Pz[(size_t)row * N + col] = acc[tm][tn][r];
            }
}

// ------------------------------------------------------------------ FP32, streamed
// Up to five contractions C_p = A_p B_p with a shared N, K and strides and
// their own row counts, in one launch: grid z = p S + slice (contraction
// outermost, then slice), grid y = row tiles of the largest M (blocks past a
// smaller contraction's rows return at once).
constexpr int kGemmSetMax = 5;
template <typename T>
struct GemmSetT {
    const T *A[kGemmSetMax];
    const T *B[kGemmSetMax];
    T *P[kGemmSetMax];             // split-K partials of contraction p: [S][M_p][N]
    T *C[kGemmSetMax];
    const int *kend[kGemmSetMax];  // K extent per kKendRows operand rows (null: K)
    int M[kGemmSetMax];            // rows of contraction p
    int rpw[kGemmSetMax];          // rows per frequency row of nk_w
    // resident batch layout of every contraction's A (null: rows of lda):
    // row r at foff[r / rpw] + (r % rpw) fext[r / rpw]
    const long long *foff;
    const int *fext;
    // FP16 split contraction: per frequency row (operand row r / rpw) the
    // row's largest |A| as float bits, which sets the row's scale
    const unsigned *rpk;
};
using GemmSet = GemmSetT<float>;

// Rows covered by one kend entry (k_kend_tiles over 32 frequency rows of four
// operand rows, k_kend_sets over 128 rows).
constexpr int kKendRows = 128;

// K extent of the operand rows [m0, m0 + bm) of a contraction with M rows:
// the largest extent of the kend entries they touch; K without kend.
__device__ __forceinline__ int tile_extent(const int *__restrict__ kend, int M, int K, int m0, int bm)
{
    if (!kend) return K;
    const int j1 = (min(m0 + bm, M) - 1) / kKendRows;
    int kt = 0;
    for (int j = m0 / kKendRows; j <= j1; j++) kt = max(kt, kend[j]);
    return min(K, kt);
}

// The first `lim` values of a 4-vector of A, the rest zero (a row's values
// at or past its cut are zeros of the product).
__device__ __forceinline__ float4 cut_tail(float4 v, int lim)
{
    return make_float4(lim > 0 ? v.x : 0.f, lim > 1 ? v.y : 0.f, lim > 2 ? v.z : 0.f, lim > 3 ? v.w : 0.f);
}
// Operand loads of the contraction tiles go to registers unconditionally, from
// the operand's own address where it is valid and from the buffer's first
// element elsewhere; the masking happens when the registers are stored to LDS.
// A load under a branch would join the zero it replaces at the branch's end,
// and the compiler waits for the load right there: each slab's global loads
// would then stall the wave instead of staying in flight through the current
// slab's products.

// Block tile BM x BN of 8 waves, wave tile WM x WN (WM/32 x WN/32 MFMA
// 32x32x2 tiles), K slab 32.  The next slab's 16-byte global loads are in
// flight while the matrix cores work on the current one (two LDS buffers),
// and within a slab the operands of the next four k steps are read from LDS
// while the current four are multiplied.  Lanes 0-31 of every MFMA take k in
// [0, 16) of the slab and lanes 32-63 k in [16, 32), so a lane's A operands
// for four steps are one 16-byte LDS read.  Split-K: slice z covers
// [z kc, (z+1) kc); a slice at or past the block's K extent returns without
// writing (k_gemm_reduce_set sums only the live slices).  The slice index is
// the outermost grid dimension, so blocks running together read the same K
// range of B.  A row r is read up to its frequency row's cut nk_w[r / rpw]
// (16-byte loads, the entries past the cut replaced by zeros).  Requires lda
// and ldb multiples of 4, N <= ldb, 16-byte aligned A and B.
template <int BM, int BN, int WM, int WN>
struct F32Tile {
    static constexpr int NT = (BM / WM) * (BN / WN) * 64, BK = 32, PA = BK + 4, PB = BN + 4;
    static constexpr int TM = WM / 32, TN = WN / 32, AE = BM * BK / 4 / NT, BE = BK * BN / 4 / NT, BNV = BN / 4;
    static_assert(AE * NT * 4 == BM * BK && BE * NT * 4 == BK * BN, "staging must cover the slab");
    struct Smem {
        float As[2][BM][PA];
        float Bs[2][BK][PB];
    };

    // acc = A[m0 : m0 + BM, kb : ke] B[kb : ke, n0 : n0 + BN], A row r read up
    // to nk_w[r / rpw] (all of [kb, ke) without nk_w), rows at or past M zero.
    // Ends with a barrier, so the shared buffers may be refilled at once.
    static __device__ __forceinline__ void mma(const float *__restrict__ A, const float *__restrict__ B, long lda,
                                               long ldb, int M, int N, int m0, int n0, int kb, int ke, int rpw,
                                               const int *__restrict__ nk_w, Smem &sm, v16f (&acc)[TM][TN],
                                               const long long *__restrict__ foff = nullptr,
                                               const int *__restrict__ fext = nullptr)
    {
        const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
        const int wm = wv / (BN / WN), wn = wv % (BN / WN);
        // staging: A item idx = tid + NT e -> row idx >> 3, k offset (idx & 7) * 4;
        // B item -> k row idx / BNV, column offset (idx % BNV) * 4
        int a_lim[AE];
        size_t a_row[AE];
#pragma unroll
        for (int e = 0; e < AE; e++) {
            const int gm = m0 + ((tid + NT * e) >> 3);
            a_lim[e] = gm < M ? (nk_w ? min(ke, nk_w[gm / rpw]) : ke) : 0;
            // resident layout: a row's stored length bounds nk_w (and is a multiple of 4)
            a_row[e] = foff ? (gm < M ? size_t(foff[gm / rpw]) + size_t(gm % rpw) * size_t(fext[gm / rpw]) : 0)
                            : size_t(gm) * size_t(lda);
        }
        float4 ra[AE], rb[BE];
        int rlim[AE];   // values of ra[e] at or past the row's cut are zeroed as they are stored
        bool bval[BE];  // rb[e] is a value of B (else it is stored as zeros)
        auto load = [&](int k0) {
#pragma unroll
            for (int e = 0; e < AE; e++) {
                const int idx = tid + NT * e, gk = k0 + (idx & 7) * 4;
                rlim[e] = a_lim[e] - gk;
                // gk + 3 < lda (or the row's stored length) where rlim > 0: both
                // are multiples of 4 and gk is below them
                ra[e] = *reinterpret_cast<const float4 *>(A + (rlim[e] > 0 ? a_row[e] + gk : 0));
            }
#pragma unroll
            for (int e = 0; e < BE; e++) {
                const int idx = tid + NT * e, gk = k0 + idx / BNV, gn = n0 + (idx % BNV) * 4;
                // gn + 3 < ldb: gn and ldb are multiples of 4 and gn < N <= ldb;
                // columns at or past N only reach accumulators that are not written
                bval[e] = gk < ke && gn < N;
                rb[e] = *reinterpret_cast<const float4 *>(B + (bval[e] ? (size_t)gk * ldb + gn : 0));
            }
        };
        auto store = [&](int buf) {
#pragma unroll
            for (int e = 0; e < AE; e++) {
                const int idx = tid + NT * e;
                *reinterpret_cast<float4 *>(&sm.As[buf][idx >> 3][(idx & 7) * 4]) = cut_tail(ra[e], rlim[e]);
            }
#pragma unroll
            for (int e = 0; e < BE; e++) {
                const int idx = tid + NT * e;
                *reinterpret_cast<float4 *>(&sm.Bs[buf][idx / BNV][(idx % BNV) * 4]) =
                    bval[e] ? rb[e] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        };
#pragma unroll
        for (int a = 0; a < TM; a++)
#pragma unroll
            for (int b = 0; b < TN; b++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[a][b][r] = 0.f;
        load(kb);
        store(0);
        __syncthreads();
        const int i = lane & 31, kq = lane >> 5;
        int buf = 0;
        for (int k0 = kb; k0 < ke; k0 += BK) {
            const bool more = k0 + BK < ke;
            if (more) load(k0 + BK);
            float4 fa[2][TM];
            float fb[2][4][TN];
            auto frag = [&](int q, int s) {
#pragma unroll
                for (int t = 0; t < TM; t++)
                    fa[s][t] = *reinterpret_cast<const float4 *>(&sm.As[buf][wm * WM + t * 32 + i][16 * kq + 4 * q]);
#pragma unroll
                for (int j = 0; j < 4; j++)
#pragma unroll
                    for (int t = 0; t < TN; t++) fb[s][j][t] = sm.Bs[buf][16 * kq + 4 * q + j][wn * WN + t * 32 + i];
            };
            frag(0, 0);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int s = q & 1;
                if (q < 3) frag(q + 1, s ^ 1);
#pragma unroll
                for (int j = 0; j < 4; j++)
#pragma unroll
                    for (int tm = 0; tm < TM; tm++) {
                        const float a = j == 0 ? fa[s][tm].x : j == 1 ? fa[s][tm].y : j == 2 ? fa[s][tm].z : fa[s][tm].w;
#pragma unroll
                        for (int tn = 0; tn < TN; tn++)
                            acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, fb[s][j][tn], acc[tm][tn], 0, 0, 0);
                    }
            }
            if (more) store(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }

    // dst[r][c] = acc for the tile's rows r < nrow and columns c < ncol
    // (tile-local indices), row stride ld.
    static __device__ __forceinline__ void put(const v16f (&acc)[TM][TN], float *__restrict__ dst, long ld, int nrow,
                                               int ncol)
    {
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const int wm = wv / (BN / WN), wn = wv % (BN / WN);
#pragma unroll
        for (int tm = 0; tm < TM; tm++)
#pragma unroll
            for (int tn = 0; tn < TN; tn++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int row = wm * WM + tm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    const int col = wn * WN + tn * 32 + (lane & 31);
                    if (row < nrow && col < ncol) dst[(size_t)row * ld + col] = acc[tm][tn][r];
                }
    }
};

// Block tile BM x BN of 8 waves, wave tile WM x WN (WM/32 x WN/32 MFMA
// 32x32x2 tiles), K slab 32.  The next slab's 16-byte global loads are in
// flight while the matrix cores work on the current one (two LDS buffers),
// and within a slab the operands of the next four k steps are read from LDS
// while the current four are multiplied.  Lanes 0-31 of every MFMA take k in
// [0, 16) of the slab and lanes 32-63 k in [16, 32), so a lane's A operands
// for four steps are one 16-byte LDS read.  Split-K: slice z covers
// [z kc, (z+1) kc); a slice at or past the block's K extent returns without
// writing (k_gemm_reduce_set sums only the live slices).  The slice index is
// the outermost grid dimension, so blocks running together read the same K
// range of B.  A row r is read up to its frequency row's cut nk_w[r / rpw]
// (16-byte loads, the entries past the cut replaced by zeros).  Requires lda
// and ldb multiples of 4, N <= ldb, 16-byte aligned A and B.
template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void k_gemm_f32t(int N, int K, int kc, int S, GemmSet gp,
                                                                       long lda, long ldb,
                                                                       const int *__restrict__ nk_w)
{
    using Tl = F32Tile<BM, BN, WM, WN>;
    __shared__ __attribute__((aligned(16))) typename Tl::Smem sm;
    const int pi = blockIdx.z / S, z = blockIdx.z - pi * S;
    const int M = gp.M[pi];
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    if (m0 >= M) return;  // uniform: a contraction with fewer row tiles
    const int kt = tile_extent(gp.kend[pi], M, K, m0, BM);
    const int kb = z * kc;
    if (kb >= kt) return;  // uniform: this slice is past the block's extent
    v16f acc[Tl::TM][Tl::TN];
    Tl::mma(gp.A[pi], gp.B[pi], lda, ldb, M, N, m0, n0, kb, min(kt, kb + kc), gp.rpw[pi], nk_w, sm, acc);
    Tl::put(acc, gp.P[pi] + (size_t)z * M * N + (size_t)m0 * N + n0, N, M - m0, N - n0);
}

// ------------------------------------------------------------------ FP32, stream-K
// The ragged (learned-grid) contraction set with the K work balanced over a
// fixed number of persistent blocks.  The block tiles of all contractions,
// in order (contraction, row tile, column tile), have n_t = ceil(kt / 32)
// K slabs each (kt the tile's extent); off[t] is their exclusive prefix sum,
// G = off[T] the total.  Block b of nb' = min(nb, G) takes the slabs
// [b G / nb', (b + 1) G / nb') (at least one each): a tile it covers whole is
// written to C directly, a tile it covers in part leaves a partial in its
// slot 0 (its first tile) or slot 1 (its last tile), which k_streamk_reduce
// sums in block order.
struct StreamK {
    int cum[kGemmSetMax + 1];  // first tile of each contraction; cum[n] = T
    int ncol;                  // column tiles per row tile
    int nb;                    // persistent blocks
};

__device__ __forceinline__ int sk_block_of(long g, long G, int nb) { return int(((g + 1) * nb - 1) / G); }

// off[0 .. T]: exclusive prefix sum of the tiles' slab counts; one block.
template <int BM>
__global__ __launch_bounds__(1024) void k_streamk_plan(int n, int K, GemmSet gp, StreamK sk, int *__restrict__ off)
{
    __shared__ int part[1024];
    const int T = sk.cum[n], tid = threadIdx.x;
    const int per = (T + 1023) / 1024, t0 = tid * per;
    int sum = 0;
    for (int t = t0; t < min(T, t0 + per); t++) {
        int p = 0;
        while (p + 1 < n && t >= sk.cum[p + 1]) p++;
        const int m0 = (t - sk.cum[p]) / sk.ncol * BM;
        sum += (tile_extent(gp.kend[p], gp.M[p], K, m0, BM) + 31) / 32;
    }
    part[tid] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d *= 2) {
        const int v = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = tid ? part[tid - 1] : 0;
    for (int t = t0; t < min(T, t0 + per); t++) {
        off[t] = run;
        int p = 0;
        while (p + 1 < n && t >= sk.cum[p + 1]) p++;
        const int m0 = (t - sk.cum[p]) / sk.ncol * BM;
        run += (tile_extent(gp.kend[p], gp.M[p], K, m0, BM) + 31) / 32;
    }
    if (tid == 1023) off[T] = part[1023];
}

template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void k_gemm_f32sk(int n, int N, int K, GemmSet gp, StreamK sk,
                                                                        long lda, long ldb,
                                                                        const int *__restrict__ nk_w,
                                                                        const int *__restrict__ off,
                                                                        float *__restrict__ part)
{
    using Tl = F32Tile<BM, BN, WM, WN>;
    __shared__ __attribute__((aligned(16))) typename Tl::Smem sm;
    const int T = sk.cum[n];
    const long G = off[T];
    const int b = blockIdx.x, nbe = int(min((long)sk.nb, G));
    if (b >= nbe) return;  // uniform
    long lo = G * b / nbe;
    const long hi = G * (b + 1) / nbe, lo_b = lo;
    // the tile holding slab lo: the last t with off[t] <= lo
    int t = 0;
    for (int step = 1 << 30; step; step >>= 1)
        if (t + step <= T && off[t + step] <= lo) t += step;
    while (lo < hi) {
        while (off[t + 1] <= lo) t++;  // tiles without slabs
        const long t0 = off[t], t1 = off[t + 1];
        const long seg = min(hi, t1);
        int p = 0;
        while (p + 1 < n && t >= sk.cum[p + 1]) p++;
        const int M = gp.M[p];
        const int m0 = (t - sk.cum[p]) / sk.ncol * BM, n0 = (t - sk.cum[p]) % sk.ncol * BN;
        const int kt = tile_extent(gp.kend[p], M, K, m0, BM);
        v16f acc[Tl::TM][Tl::TN];
        Tl::mma(gp.A[p], gp.B[p], lda, ldb, M, N, m0, n0, int(lo - t0) * 32, min(kt, int(seg - t0) * 32), gp.rpw[p],
                nk_w, sm, acc, gp.foff, gp.fext);
        const bool whole = lo == t0 && seg == t1;
        float *dst = whole ? gp.C[p] + (size_t)m0 * N + n0 : part + ((size_t)b * 2 + (lo == lo_b ? 0 : 1)) * BM * BN;
        Tl::put(acc, dst, whole ? N : BN, whole ? M - m0 : BM, whole ? N - n0 : BN);
        lo = seg;
        t++;
    }
}

// C of every tile split between blocks: the sum, in FP64 and block order, of
// its partials; C of a tile without slabs: zero.  Thread e covers element e
// of the concatenated outputs; tiles written whole are left alone.
// row slices per tile of the reduce (enough workgroups for a few hundred tiles)
constexpr int kReduceSlices = 8;
template <int BM, int BN>
__global__ __launch_bounds__(256) void k_streamk_reduce(int n, int N, GemmSet gs, StreamK sk,
                                                        const int *__restrict__ off, const float *__restrict__ part)
{
    // workgroups (tile, slice of BM / gridDim.y rows): blockIdx.x runs over
    // all contractions' tiles; the tile's bounds and blocks are found once
    // per workgroup, not per element
    const int t = blockIdx.x;
    const int i0 = blockIdx.y * (BM / gridDim.y) * BN, i1 = i0 + (BM / gridDim.y) * BN;
    int p = 0;
    while (p + 1 < n && t >= sk.cum[p + 1]) p++;
    const int tl = t - sk.cum[p];
    const int m0 = tl / sk.ncol * BM, n0 = tl % sk.ncol * BN;
    const int M = gs.M[p];
    float *__restrict__ C = gs.C[p];
    const long G = off[sk.cum[n]], t0 = off[t], t1 = off[t + 1];
    if (t1 == t0) {  // empty K extent: the contraction kernel wrote nothing
        for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            const int r = m0 + i / BN, c = n0 + i % BN;
            if (r < M && c < N) C[(size_t)r * N + c] = 0.f;
        }
        return;
    }
    const int nbe = int(min((long)sk.nb, G));
    const int b0 = sk_block_of(t0, G, nbe), b1 = sk_block_of(t1 - 1, G, nbe);
    if (b0 == b1) return;  // written whole by block b0
    // block b0 left its part in slot 1 when the tile is not its first
    const int slot0 = (G * b0 / nbe < t0) ? 1 : 0;
    // 16-byte path: each thread's NV groups of 4 columns, all loads of one
    // block's partials issued before they are added (the sums keep the
    // scalar path's order, so both give the same bits)
    constexpr int NV = BM * BN / 4 / 256 / kReduceSlices;
    static_assert(NV >= 1 && NV * 4 * 256 * kReduceSlices == BM * BN, "reduce slice");
    if (gridDim.y == kReduceSlices && blockDim.x == 256 && (N & 3) == 0 &&
        ((reinterpret_cast<uintptr_t>(C) | reinterpret_cast<uintptr_t>(part)) & 15) == 0) {
        int e[NV];
        bool ok[NV];
        double s[NV][4];
#pragma unroll
        for (int j = 0; j < NV; j++) {
            e[j] = i0 + (int(threadIdx.x) + 256 * j) * 4;
            ok[j] = m0 + e[j] / BN < M && n0 + e[j] % BN < N;  // N % 4 == 0: all four columns in range
        }
        float4 v[NV];
#pragma unroll
        for (int j = 0; j < NV; j++)
            v[j] = ok[j] ? *reinterpret_cast<const float4 *>(part + ((size_t)b0 * 2 + slot0) * BM * BN + e[j])
                         : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int j = 0; j < NV; j++) s[j][0] = v[j].x, s[j][1] = v[j].y, s[j][2] = v[j].z, s[j][3] = v[j].w;
        for (int b = b0 + 1; b <= b1; b++) {
#pragma unroll
            for (int j = 0; j < NV; j++)
                v[j] = ok[j] ? *reinterpret_cast<const float4 *>(part + (size_t)b * 2 * BM * BN + e[j])
                             : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int j = 0; j < NV; j++) s[j][0] += v[j].x, s[j][1] += v[j].y, s[j][2] += v[j].z, s[j][3] += v[j].w;
        }
#pragma unroll
        for (int j = 0; j < NV; j++)
            if (ok[j])
                *reinterpret_cast<float4 *>(C + (size_t)(m0 + e[j] / BN) * N + n0 + e[j] % BN) =
                    make_float4(float(s[j][0]), float(s[j][1]), float(s[j][2]), float(s[j][3]));
        return;
    }
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const int r = m0 + i / BN, c = n0 + i % BN;
        if (r >= M || c >= N) continue;
        double s = double(part[((size_t)b0 * 2 + slot0) * BM * BN + i]);
        for (int b = b0 + 1; b <= b1; b++) s += double(part[(size_t)b * 2 * BM * BN + i]);
        C[(size_t)r * N + c] = float(s);
    }
}

// FP32 block tile for N columns: 128 rows by 128 columns (waves of 32 x 64),
// or by 64 columns (waves of 32 x 32) for N <= 64; two resident blocks of 8
// waves per compute unit.  On the 256-receiver contraction pair the 128 x 128
// tile and a 256 x 256 tile run the dense grid within 2% of each other, and
// the 128 x 128 tile runs the learned (ragged) grid 1.4x faster: its finer
// row tiles follow the per-tile K extents more closely.
constexpr int kTBM = 128, kTBK = 32;
static int f32_bn(int N) { return N > 64 ? 128 : 64; }

// Sum of the live split-K slices of every contraction of a set, in FP64, in
// slice order: C_p[row] = sum over z < ceil(kt / kc) of P_p[z][row], kt the K
// extent of the row's bm-row block tile.  Thread t covers element t of the
// concatenated outputs.
template <typename T>
__global__ void k_gemm_reduce_set(int n, int N, int K, int kc, int bm, GemmSetT<T> gs, long total)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int pi = 0;
    long e = t;
    while (pi + 1 < n && e >= (long)gs.M[pi] * N) {
        e -= (long)gs.M[pi] * N;
        pi++;
    }
    const long MN = (long)gs.M[pi] * N;
    const int row = int(e / N);
    const int kt = tile_extent(gs.kend[pi], gs.M[pi], K, row / bm * bm, bm);
    const int nz = (kt + kc - 1) / kc;
    const T *P = gs.P[pi];
    double s = 0;
    for (int zz = 0; zz < nz; zz++) s += double(P[(size_t)zz * MN + e]);
    gs.C[pi][e] = T(s);
}

// ------------------------------------------------------------------ FP64, streamed
// Two contractions in FP64 on v_mfma_f64_16x16x4f64, as k_gemm_f32t: block
// tile 128 x 128 (A is read N / 128 times instead of N / 64), 4 waves of
// 64 x 64 (4 x 4 MFMA tiles), K slab 16 with the next slab's 16-byte loads in
// flight (two LDS buffers).  Lane l of an MFMA takes A[m0 + l % 16][k + l / 16]
// and B[k + l / 16][n0 + l % 16]; accumulator r of lane l is row
// (l >> 4) + 4 r, column l % 16.  Requires lda, ldb even and 16-byte aligned
// A, B.
constexpr int kDBM = 128, kDBN = 128, kDBK = 16, kDPA = kDBK + 2, kDPB = kDBN + 2;

using GemmSet64 = GemmSetT<double>;

__global__ __launch_bounds__(256) void k_gemm_f64s(int N, int K, int kc, int S, GemmSet64 gp, long lda, long ldb,
                                                   const int *__restrict__ nk_w)
{
    __shared__ __attribute__((aligned(16))) double As[2][kDBM][kDPA];
    __shared__ __attribute__((aligned(16))) double Bs[2][kDBK][kDPB];
    const int pi = blockIdx.z / S, z = blockIdx.z - pi * S;
    const int M = gp.M[pi];
    const int m0 = blockIdx.y * kDBM, n0 = blockIdx.x * kDBN;
    if (m0 >= M) return;  // uniform: a contraction with fewer row tiles
    const double *__restrict__ A = gp.A[pi];
    const double *__restrict__ B = gp.B[pi];
    const int *__restrict__ kend = gp.kend[pi];
    const int rpw = gp.rpw[pi];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int kt = min(K, kend ? kend[blockIdx.y] : K);
    const int kb = z * kc;
    if (kb >= kt) return;  // uniform: this slice is past the tile's extent
    const int ke = min(kt, kb + kc);

    // staging: A items idx = tid + 256 e (e < 4) -> row idx >> 3, k offset (idx & 7) * 2;
    // B items -> k row idx >> 6, column offset (idx & 63) * 2
    int a_lim[4];
    for (int e = 0; e < 4; e++) {
        const int gm = m0 + ((tid + 256 * e) >> 3);
        a_lim[e] = gm < M ? (nk_w ? min(ke, nk_w[gm / rpw]) : ke) : 0;
    }
    double2 ra[4], rb[4];
    auto load = [&](int k0) {
        for (int e = 0; e < 4; e++) {
            const int idx = tid + 256 * e, gm = m0 + (idx >> 3), gk = k0 + (idx & 7) * 2;
            double2 v = make_double2(0.0, 0.0);
            if (gk + 1 < a_lim[e]) v = *reinterpret_cast<const double2 *>(A + (size_t)gm * lda + gk);
            else if (gk < a_lim[e]) v.x = A[(size_t)gm * lda + gk];
            ra[e] = v;
        }
        for (int e = 0; e < 4; e++) {
            const int idx = tid + 256 * e, gk = k0 + (idx >> 6), gn = n0 + (idx & 63) * 2;
            double2 v = make_double2(0.0, 0.0);
            if (gk < ke) {
                const double *r = B + (size_t)gk * ldb;
                if (gn + 1 < N) v = *reinterpret_cast<const double2 *>(r + gn);
                else if (gn < N) v.x = r[gn];
            }
            rb[e] = v;
        }
    };
    auto store = [&](int buf) {
        for (int e = 0; e < 4; e++) {
            const int idx = tid + 256 * e;
            *reinterpret_cast<double2 *>(&As[buf][idx >> 3][(idx & 7) * 2]) = ra[e];
            *reinterpret_cast<double2 *>(&Bs[buf][idx >> 6][(idx & 63) * 2]) = rb[e];
        }
    };

    v4d acc[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) acc[a][b] = v4d{0, 0, 0, 0};
    load(kb);
    store(0);
    __syncthreads();
    const int i = lane & 15, kq = lane >> 4;
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += kDBK) {
        const bool more = k0 + kDBK < ke;
        if (more) load(k0 + kDBK);
        for (int kk = 0; kk < kDBK; kk += 4) {
            double a[4], b[4];
            for (int t = 0; t < 4; t++) a[t] = As[buf][wm * 64 + t * 16 + i][kk + kq];
            for (int t = 0; t < 4; t++) b[t] = Bs[buf][kk + kq][wn * 64 + t * 16 + i];
            for (int tm = 0; tm < 4; tm++)
                for (int tn = 0; tn < 4; tn++)
                    acc[tm][tn] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[tm], b[tn], acc[tm][tn], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    double *Pz = gp.P[pi] + (size_t)z * M * N;
    for (int tm = 0; tm < 4; tm++)
        for (int tn = 0; tn < 4; tn++)
            for (int r = 0; r < 4; r++) {
                const int row = m0 + wm * 64 + tm * 16 + (lane >> 4) + 4 * r;
                const int col = n0 + wn * 64 + tn * 16 + (lane & 15);
                if (row < M && col < N) Pz[(size_t)row * N + col] = acc[tm][tn][r];
            }
}

// ------------------------------------------------------------ split BF16
// FP32 operands split into NPL bfloat16 terms, x = x1 + x2 (+ x3), each the
// bf16 rounding of what the previous terms leave; the products of the pair
// orders that matter are summed by v_mfma_f32_32x32x16_bf16 in FP32:
// NP = 3: x1 y1 + x1 y2 + x2 y1 (NPL = 2, relative error about 2^-16 per
// product), NP = 6: also x1 y3 + x2 y2 + x3 y1 (NPL = 3, about 2^-24: the FP32
// contraction's accuracy at 6 / 16 of its matrix-core time).  The B operand
// (a Bessel table, fixed per plan) is split once into NPL planes of N ldk
// values, zero past K, laid out slab-major: [ldk / 32][N][32], so the B of
// one 32-wavenumber slab of a tile's columns is one contiguous run (with the
// k rows of each column 32 KB apart, every slab touched a separate cache
// line per column and plane); A is split while it is staged to LDS.
__host__ __device__ __forceinline__ size_t split_off(int N, int n, int k)
{
    return size_t(k >> 5) * (size_t(N) << 5) + (size_t(n) << 5) + size_t(k & 31);
}
template <int NPL, typename E>
__device__ __forceinline__ void split_terms(float x, E (&t)[NPL])
{
    float r = x;
#pragma unroll
    for (int p = 0; p < NPL; p++) {
        t[p] = E(r);
        r = r - float(t[p]);
    }
}

// FP16 split (E = _Float16, two terms, three products): fp16 has 11 mantissa
// bits against bf16's 8, so two terms carry 22 bits and the dropped product
// x2 y2 is about 2^-22 of x y (two bf16 terms: 2^-16, which misses the 1e-4
// bar on the 256-receiver workload; three: 2^-24 at six products).  Its
// exponent range is what costs: each frequency row of A is scaled by a power
// of two that takes the row's peak |A| to [2^9, 2^10) (values below 2^-14 of
// the scaled peak go subnormal, an absolute error under 2^-34 of the peak),
// and B by 2^13.  Powers of two are exact, so the products come
// back unscaled exactly as they are written.  B must stay below 8 in
// magnitude (the Bessel tables hold |J_n| <= 1).
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
constexpr float kF16BScale = 8192.f;
__device__ __forceinline__ float f16_row_scale(const unsigned *__restrict__ rpk, int f)
{
    const float pk = __uint_as_float(rpk[f]);
    if (!(pk > 0.f) || !(pk < 3.0e38f)) return 1.f;
    int e;
    (void)frexpf(pk, &e);  // pk in [2^(e-1), 2^e)
    return ldexpf(1.f, max(-100, min(100, 10 - e)));
}

template <int BM, int BN, int WM, int WN, int NPL, int NP, int BK = 32, typename E = __bf16>
struct BSTile {
    static constexpr bool F16 = std::is_same<E, _Float16>::value;
    using V8 = typename std::conditional<F16, v8h, v8bf>::type;
    static constexpr int NT = (BM / WM) * (BN / WN) * 64, KP = BK + 8;
    static constexpr int AR = BK / 4, BR = BK / 8;  // 16-byte A (FP32) and B (bf16) chunks per slab row
    static constexpr int TM = WM / 32, TN = WN / 32, AE = BM * BK / 4 / NT;
    static constexpr int BC = BN * BK / 8;  // 16-byte B chunks per plane
    static_assert(AE * NT * 4 == BM * BK, "A staging must cover the slab");
    static_assert((NPL == 2 && NP == 3) || (NPL == 3 && NP == 6), "split");
    static_assert(BK % 16 == 0, "one MFMA k step is 16");
    struct Smem {
        E As[2][NPL][BM][KP];
        E Bs[2][NPL][BN][KP];
    };

    // acc = A[m0 : m0 + BM, kb : ke] B[kb : ke, n0 : n0 + BN] as F32Tile::mma;
    // S: the split B planes (split_off layout), plane stride N ldk.
    static __device__ __forceinline__ void mma(const float *__restrict__ A, const E *__restrict__ S, long lda,
                                               long ldk, int M, int N, int m0, int n0, int kb, int ke, int rpw,
                                               const int *__restrict__ nk_w, Smem &sm, v16f (&acc)[TM][TN],
                                               const long long *__restrict__ foff, const int *__restrict__ fext,
                                               const unsigned *__restrict__ rpk)
    {
        const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
        const int wm = wv / (BN / WN), wn = wv % (BN / WN);
        int a_lim[AE];
        size_t a_row[AE];
        float a_sc[AE];  // FP16 split: the rows' scales
#pragma unroll
        for (int e = 0; e < AE; e++) {
            const int gm = m0 + (tid + NT * e) / AR;
            a_lim[e] = gm < M ? (nk_w ? min(ke, nk_w[gm / rpw]) : ke) : 0;
            a_row[e] = foff ? (gm < M ? size_t(foff[gm / rpw]) + size_t(gm % rpw) * size_t(fext[gm / rpw]) : 0)
                            : size_t(gm) * size_t(lda);
            if constexpr (F16) a_sc[e] = gm < M ? f16_row_scale(rpk, gm / rpw) : 1.f;
        }
        const size_t pstride = size_t(N) * size_t(ldk);
        // one K slab of the block's A (FP32) and split B (bf16 planes) in registers
        struct Regs {
            float4 ra[AE];
            int rlim[AE];  // values at or past the row's cut are zeroed as they are stored
            uint4 rb[NPL][(BC + NT - 1) / NT];
            bool bval[(BC + NT - 1) / NT];  // rb[.][c] holds B values (else stored as zeros)
        };
        auto load = [&](int k0, Regs &R) {
#pragma unroll
            for (int e = 0; e < AE; e++) {
                const int idx = tid + NT * e, gk = k0 + (idx % AR) * 4;
                R.rlim[e] = a_lim[e] - gk;
                // 16-byte aligned and inside the row where rlim > 0 (gk and the
                // row's length are multiples of 4)
                R.ra[e] = *reinterpret_cast<const float4 *>(A + (R.rlim[e] > 0 ? a_row[e] + gk : 0));
            }
#pragma unroll
            for (int c = 0; c < (BC + NT - 1) / NT; c++) {
                const int q = tid + NT * c, n = q / BR, kq = (q % BR) * 8;
                // k0 + kq + 7 < ldk (a multiple of 32 at least ke) for k0 < ke,
                // within one 32-value run of the slab-major planes; columns past
                // N and slabs past ke are stored as zeros
                R.bval[c] = q < BC && n0 + n < N && k0 < ke;
                const size_t o = R.bval[c] ? split_off(N, n0 + n, k0 + kq) : 0;
#pragma unroll
                for (int p = 0; p < NPL; p++) R.rb[p][c] = *reinterpret_cast<const uint4 *>(S + p * pstride + o);
            }
        };
        auto store = [&](int buf, const Regs &R) {
#pragma unroll
            for (int e = 0; e < AE; e++) {
                const int idx = tid + NT * e, r = idx / AR, k = (idx % AR) * 4;
                const float4 a = cut_tail(R.ra[e], R.rlim[e]);
                float x[4] = {a.x, a.y, a.z, a.w};
                if constexpr (F16)
#pragma unroll
                    for (int j = 0; j < 4; j++) x[j] *= a_sc[e];
                E t[4][NPL];
#pragma unroll
                for (int j = 0; j < 4; j++) split_terms<NPL>(x[j], t[j]);
#pragma unroll
                for (int p = 0; p < NPL; p++) {
                    typedef E v4e __attribute__((ext_vector_type(4)));
                    v4e w;
                    w[0] = t[0][p], w[1] = t[1][p], w[2] = t[2][p], w[3] = t[3][p];
                    *reinterpret_cast<v4e *>(&sm.As[buf][p][r][k]) = w;
                }
            }
#pragma unroll
            for (int c = 0; c < (BC + NT - 1) / NT; c++) {
                const int q = tid + NT * c, n = q / BR, kq = (q % BR) * 8;
                if (q < BC)
#pragma unroll
                    for (int p = 0; p < NPL; p++)
                        *reinterpret_cast<uint4 *>(&sm.Bs[buf][p][n][kq]) =
                            R.bval[c] ? R.rb[p][c] : make_uint4(0u, 0u, 0u, 0u);
            }
        };
        const int i = lane & 31, kq = (lane >> 5) * 8;
        // one slab: its fragments read from LDS buffer buf, the next slab (in
        // R, loaded an iteration earlier) split and written to buf ^ 1, the
        // loads of the slab after it into R, then the products.  The split's
        // VALU work and the LDS writes sit between
        // the fragment reads and the MFMAs of the same basic block, so they
        // issue while the matrix cores run (with the split and the writes
        // after the products, every wave of the block spent each slab first
        // on the products and then on the split, all at once).  The write to
        // buf ^ 1 is unconditional: past the last slab it fills a buffer that
        // is not read again before the next barrier.
        auto step = [&](int buf, Regs &R, int knext) {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 16) {
                V8 fa[NPL][TM], fb[NPL][TN];
#pragma unroll
                for (int p = 0; p < NPL; p++) {
#pragma unroll
                    for (int t = 0; t < TM; t++)
                        fa[p][t] = *reinterpret_cast<const V8 *>(&sm.As[buf][p][wm * WM + t * 32 + i][kk + kq]);
#pragma unroll
                    for (int t = 0; t < TN; t++)
                        fb[p][t] = *reinterpret_cast<const V8 *>(&sm.Bs[buf][p][wn * WN + t * 32 + i][kk + kq]);
                }
                if (kk == 0) {
                    store(buf ^ 1, R);
                    load(knext, R);
                }
                // products (small ones first) outer, tiles inner: consecutive
                // MFMAs of a wave with several tiles go to different accumulators
                constexpr int PA[6] = {2, 1, 0, 1, 0, 0}, PB[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
                for (int q = 6 - NP; q < 6; q++)
#pragma unroll
                    for (int tm = 0; tm < TM; tm++)
#pragma unroll
                        for (int tn = 0; tn < TN; tn++) {
                            if constexpr (F16)
                                acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[PA[q]][tm], fb[PB[q]][tn],
                                                                                     acc[tm][tn], 0, 0, 0);
                            else
                                acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[PA[q]][tm], fb[PB[q]][tn],
                                                                                      acc[tm][tn], 0, 0, 0);
                        }
            }
        };
#pragma unroll
        for (int a = 0; a < TM; a++)
#pragma unroll
            for (int b = 0; b < TN; b++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[a][b][r] = 0.f;
        // Slab k + 1 waits in R (loaded during slab k - 1's products) while
        // slab k is read from LDS; it is split and stored, then slab k + 2's
        // loads go out, then slab k's products.  The loads are issued
        // unconditionally (a slab at or past ke loads clamped addresses and is
        // never read back): a load on a conditional path would make the
        // compiler wait for it where the paths join.
        Regs R;
        load(kb, R);
        store(0, R);
        load(kb + BK, R);
        __syncthreads();
        int buf = 0;
        for (int k0 = kb; k0 < ke; k0 += BK) {
            step(buf, R, k0 + 2 * BK);
            __syncthreads();
            buf ^= 1;
        }
    }

    // dst[r][c] = acc for tile-local rows r < nrow, columns c < ncol, row stride
    // ld; the FP16 split divides out the scales of row m0 + r (rows per
    // frequency rpw) and of B
    static __device__ __forceinline__ void put(const v16f (&acc)[TM][TN], float *__restrict__ dst, long ld, int nrow,
                                               int ncol, const unsigned *__restrict__ rpk, int m0, int rpw)
    {
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const int wm = wv / (BN / WN), wn = wv % (BN / WN);
#pragma unroll
        for (int tm = 0; tm < TM; tm++)
#pragma unroll
            for (int tn = 0; tn < TN; tn++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int row = wm * WM + tm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    const int col = wn * WN + tn * 32 + (lane & 31);
                    if (row < nrow && col < ncol) {
                        float v = acc[tm][tn][r];
                        if constexpr (F16) v *= 1.f / (f16_row_scale(rpk, (m0 + row) / rpw) * kF16BScale);
                        dst[(size_t)row * ld + col] = v;
                    }
                }
    }
};

// Stream-K over the tiles' K slabs as k_gemm_f32sk, with the split-BF16 tile;
// S[p]: the split B planes of contraction p.
struct SplitB {
    const void *S[kGemmSetMax];
};
// WPE: waves per SIMD the registers must leave room for (the launch bound);
// E: __bf16, or _Float16 for the scaled FP16 split
template <int BM, int BN, int WM, int WN, int NPL, int NP, int BK, int WPE, typename E>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, WPE) void k_gemm_bssk(int n, int N, int K, GemmSet gp, SplitB sb,
                                                                       StreamK sk, long lda, long ldk,
                                                                       const int *__restrict__ nk_w,
                                                                       const int *__restrict__ off,
                                                                       float *__restrict__ part)
{
    using Tl = BSTile<BM, BN, WM, WN, NPL, NP, BK, E>;
    __shared__ __attribute__((aligned(16))) typename Tl::Smem sm;
    const int T = sk.cum[n];
    const long G = off[T];
    const int b = blockIdx.x, nbe = int(min((long)sk.nb, G));
    if (b >= nbe) return;  // uniform
    long lo = G * b / nbe;
    const long hi = G * (b + 1) / nbe, lo_b = lo;
    int t = 0;
    for (int step = 1 << 30; step; step >>= 1)
        if (t + step <= T && off[t + step] <= lo) t += step;
    while (lo < hi) {
        while (off[t + 1] <= lo) t++;
        const long t0 = off[t], t1 = off[t + 1];
        const long seg = min(hi, t1);
        int p = 0;
        while (p + 1 < n && t >= sk.cum[p + 1]) p++;
        const int M = gp.M[p];
        const int m0 = (t - sk.cum[p]) / sk.ncol * BM, n0 = (t - sk.cum[p]) % sk.ncol * BN;
        const int kt = tile_extent(gp.kend[p], M, K, m0, BM);
        v16f acc[Tl::TM][Tl::TN];
        Tl::mma(gp.A[p], reinterpret_cast<const E *>(sb.S[p]), lda, ldk, M, N, m0, n0, int(lo - t0) * 32,
                min(kt, int(seg - t0) * 32), gp.rpw[p], nk_w, sm, acc, gp.foff, gp.fext, gp.rpk);
        const bool whole = lo == t0 && seg == t1;
        float *dst = whole ? gp.C[p] + (size_t)m0 * N + n0 : part + ((size_t)b * 2 + (lo == lo_b ? 0 : 1)) * BM * BN;
        Tl::put(acc, dst, whole ? N : BN, whole ? M - m0 : BM, whole ? N - n0 : BN, gp.rpk, m0, gp.rpw[p]);
        lo = seg;
        t++;
    }
}

// B [K][ldb] (FP32, N columns) into NPL bf16 planes of N ldk values (split_off
// layout), zero for K <= k < ldk.
template <int NPL, typename E>
__global__ void k_split_table(int K, int N, const float *__restrict__ B, long ldb, E *__restrict__ S, long ldk)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * ldk) return;
    // t = split_off(N, n, k)
    const size_t slab = t / (size_t(N) << 5), r = t - slab * (size_t(N) << 5);
    const int n = int(r >> 5), k = int(slab * 32 + (r & 31));
    E v[NPL];
    const float x = k < K ? B[(size_t)k * ldb + n] : 0.f;
    split_terms<NPL>(std::is_same<E, _Float16>::value ? x * kF16BScale : x, v);
#pragma unroll
    for (int p = 0; p < NPL; p++) S[(size_t)p * N * ldk + t] = v[p];
}

// ------------------------------------------------------------------ BF16x3
__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo)
{
    hi = __bf16(x);
    lo = __bf16(x - float(hi));
}

__global__ __launch_bounds__(256) void k_gemm_bf16x3(int M, int N, int K, int kslice, const float *__restrict__ A,
                                                     long lda, const float *__restrict__ B, long ldb,
                                                     float *__restrict__ P, const int *__restrict__ kend,
                                                  const int *__restrict__ nk_w)
{
    constexpr int BM = 128, BN = 128, BK = 32, PADK = BK + 8;
    // k-contiguous rows so one lane's 8 consecutive k are one 16-byte read
    __shared__ __attribute__((aligned(16))) __bf16 Ah[BM][PADK], Al[BM][PADK], Bh[BN][PADK], Bl[BN][PADK];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kt_end = min(K, kend ? kend[blockIdx.y] : K);
    const int kb = blockIdx.z * kslice, ke = min(kt_end, kb + kslice);
    v16f acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int r = 0; r < 16; r++) acc[a][b][r] = 0.f;
    for (int k0 = kb; k0 < ke; k0 += BK) {
        {  // A tile 128 x 32: 16 per thread
            const int mm = tid >> 1, kk0 = (tid & 1) * 16, gm = m0 + mm;
            const int kr = (gm < M && nk_w) ? min(ke, nk_w[gm >> 2]) : ke;
            for (int q = 0; q < 16; q++) {
                const int gk = k0 + kk0 + q;
                const float x = (gm < M && gk < kr) ? A[(size_t)gm * lda + gk] : 0.f;
                split_bf16(x, Ah[mm][kk0 + q], Al[mm][kk0 + q]);
            }
        }
        {  // B tile 32 x 128 -> stored [n][k]: 16 per thread
            const int kk = tid >> 3, nn0 = (tid & 7) * 16, gk = k0 + kk;
            for (int q = 0; q < 16; q++) {
                const int gn = n0 + nn0 + q;
                const float x = (gk < ke && gn < N) ? B[(size_t)gk * ldb + gn] : 0.f;
                split_bf16(x, Bh[nn0 + q][kk], Bl[nn0 + q][kk]);
            }
        }
        __syncthreads();
        const int i = lane & 31, kq = (lane >> 5) * 8;
        for (int kk = 0; kk < BK; kk += 16) {
            v8bf ah[2], al[2], bh[2], bl[2];
            for (int t = 0; t < 2; t++) {
                const int r = wm * 64 + t * 32 + i, c = wn * 64 + t * 32 + i;
                ah[t] = *reinterpret_cast<const v8bf *>(&Ah[r][kk + kq]);
                al[t] = *reinterpret_cast<const v8bf *>(&Al[r][kk + kq]);
                bh[t] = *reinterpret_cast<const v8bf *>(&Bh[c][kk + kq]);
                bl[t] = *reinterpret_cast<const v8bf *>(&Bl[c][kk + kq]);
            }
            for (int tm = 0; tm < 2; tm++)
                for (int tn = 0; tn < 2; tn++) {
                    acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[tm], bh[tn], acc[tm][tn], 0, 0, 0);
                    acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[tm], bl[tn], acc[tm][tn], 0, 0, 0);
                    acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[tm], bh[tn], acc[tm][tn], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    float *Pz = P + (size_t)blockIdx.z * M * N;
    for (int tm = 0; tm < 2; tm++)
        for (int tn = 0; tn < 2; tn++)
            for (int r = 0; r < 16; r++) {
                const int row = m0 + wm * 64 + tm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int col = n0 + wn * 64 + tn * 32 + (lane & 31);
                if (row < M && col < N) Pz[(size_t)row * N + col] = acc[tm][tn][r];
            }
}

// ------------------------------------------------------------------ layout probe
// One wave: lane l0 holds a = 1 (all other lanes 0), lane l holds b = 1000 + l;
// out[l*4 + r] = accumulator r of lane l after one v_mfma_f64_16x16x4f64.
__global__ void k_probe_f64(int l0, double *out)
{
    const int l = threadIdx.x;
    const double a = (l == l0) ? 1.0 : 0.0, b = 1000.0 + l;
    v4d acc = v4d{0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    for (int r = 0; r < 4; r++) out[l * 4 + r] = acc[r];
}

hipError_t launch_probe_f64(int l0, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_probe_f64, dim3(1), dim3(64), 0, st, l0, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ split-K sum
template <typename T>
__global__ void k_gemm_reduce(int MN, int S, const T *__restrict__ P, T *__restrict__ Cm)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= MN) return;
    double s = 0;
    for (int z = 0; z < S; z++) s += double(P[(size_t)z * MN + t]);
    Cm[t] = T(s);
}

// Split-K launch plan: slice length a multiple of the K step, enough slices
// to give every compute unit a few blocks.
static void split_plan(int M, int N, int K, int BM, int BN, int BK, int &S, int &kslice)
{
    const int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
    S = 1;
    while (tiles * S < 1024 && S < 64 && (K / (S * 2)) >= 4 * BK) S *= 2;
    kslice = ((K + S - 1) / S + BK - 1) / BK * BK;
}

// Block tiles of n contractions of M[q] rows for the streamed kernels: mode 1
// the FP32 tile (kTBM x f32_bn(N)), mode 0 the FP64 tile (128 x 128); two
// resident blocks per compute unit either way.
static long set_tiles(int mode, int n, const int *M, int N)
{
    const int bm = mode == 1 ? kTBM : kDBM, bn = mode == 1 ? f32_bn(N) : kDBN;
    long tiles = 0;
    for (int q = 0; q < n; q++) tiles += (long)((M[q] + bm - 1) / bm) * ((N + bn - 1) / bn);
    return tiles;
}
static long set_slots(int mode) { return 2L * 256; }

// Slice length of a dense grid: K / S for the power of two S <= smax (slices of
// at least 256) whose blocks fill their last dispatch round best; ties go to
// the smaller S.  min_rounds > 1 first restricts S to splits of at least that
// many dispatch rounds where one exists.
static int dense_slice(long tiles, long slots, int K, int smax, int min_rounds = 1)
{
    int best = K;
    double best_eff = -1;
    for (int pass = 0; pass < 2 && best_eff < 0; pass++)
        for (int S = 1; S <= smax; S *= 2) {
            const int kc = ((K + S - 1) / S + kTBK - 1) / kTBK * kTBK;
            if (S > 1 && kc < 256) break;
            const long blocks = tiles * ((K + kc - 1) / kc);
            if (pass == 0 && blocks < min_rounds * slots) continue;
            const double eff = double(blocks) / double((blocks + slots - 1) / slots * slots);
            if (eff > best_eff + 1e-9) best_eff = eff, best = kc;
        }
    return best;
}

// Slice length of a pair (two contractions of M rows) from the mean K extent
// of the row tiles (k_avg; K for a dense grid).  FP32: the dense grid fills
// its dispatch rounds (dense_slice); a ragged grid takes the power of two
// (at least 256) nearest to kRaggedRounds rounds of equal slices.  FP64: about
// three blocks per compute unit live, a power of two of at least 256.
constexpr double kRaggedRounds = 4.0;
int gemm_pair_slice(int mode, int M, int N, int K, double k_avg)
{
    const int Ms[2] = {M, M};
    const long tiles = set_tiles(mode, 2, Ms, N);
    if (mode == 1) {
        if (k_avg >= K) return dense_slice(tiles, set_slots(mode), K, 16, 2);
        const double want = double(tiles) * std::max(1.0, k_avg) / (kRaggedRounds * set_slots(mode));
        int kc = 256;
        while (kc * 2 <= want * 1.4142 && kc < K) kc *= 2;
        return kc;
    }
    const double want = double(tiles) * std::max(1.0, k_avg) / 768.0;
    int kc = 256;
    while (kc * 2 <= want && kc < K) kc *= 2;
    return kc;
}

size_t gemm_pair_workspace_bytes(int M, int N, int K, int kc)
{
    const size_t S = (size_t)(K + kc - 1) / kc;
    return 2 * S * (size_t)M * N * sizeof(float);
}

// Stream-K blocks: two resident per compute unit, at most one per K slab.
constexpr int kStreamKBlocks = 2 * 256;
static int sk_blocks(long tiles, int K) { return (int)std::min<long>(kStreamKBlocks, tiles * ((K + kTBK - 1) / kTBK)); }

size_t gemm_streamk_workspace_bytes(int n, const int *M, int N, int K)
{
    const long tiles = set_tiles(1, n, M, N);
    return (size_t)sk_blocks(tiles, K) * 2 * kTBM * f32_bn(N) * sizeof(float) + (size_t)(tiles + 1) * sizeof(int);
}

// FP32 set with per-tile K extents: stream-K (k_streamk_plan, k_gemm_f32sk,
// k_streamk_reduce) in a workspace of gemm_streamk_workspace_bytes.
static hipError_t launch_streamk(int n, const GemmSet &gs, int N, int K, long lda, long ldb, float *P,
                                 const int *nk_w, hipStream_t st)
{
    const int bn = f32_bn(N);
    StreamK sk{};
    int t = 0;
    for (int q = 0; q < kGemmSetMax; q++) {
        sk.cum[q] = t;
        if (q < n) t += (gs.M[q] + kTBM - 1) / kTBM * ((N + bn - 1) / bn);
    }
    sk.cum[kGemmSetMax] = t;
    sk.ncol = (N + bn - 1) / bn;
    sk.nb = sk_blocks(t, K);
    float *part = P;
    int *off = reinterpret_cast<int *>(P + (size_t)sk.nb * 2 * kTBM * bn);
    hipLaunchKernelGGL(k_streamk_plan<kTBM>, dim3(1), dim3(1024), 0, st, n, K, gs, sk, off);
    if (bn == 128) {
        hipLaunchKernelGGL((k_gemm_f32sk<kTBM, 128, 32, 64>), dim3(sk.nb), dim3(512), 0, st, n, N, K, gs, sk, lda,
                           ldb, nk_w, off, part);
        hipLaunchKernelGGL((k_streamk_reduce<kTBM, 128>), dim3((unsigned)t, kReduceSlices), dim3(256), 0, st, n, N, gs,
                           sk, off, part);
    } else {
        hipLaunchKernelGGL((k_gemm_f32sk<kTBM, 64, 32, 32>), dim3(sk.nb), dim3(512), 0, st, n, N, K, gs, sk, lda, ldb,
                           nk_w, off, part);
        hipLaunchKernelGGL((k_streamk_reduce<kTBM, 64>), dim3((unsigned)t, kReduceSlices), dim3(256), 0, st, n, N, gs,
                           sk, off, part);
    }
    return hipGetLastError();
}

// Split-BF16 pair (Z and R contractions) by stream-K, one resident block of
// 8 waves per compute unit (the double-buffered split tiles take 82-123 KB of
// LDS).  Workspace: gemm_bsplit_workspace_bytes.
// More than 128 receivers: 128 x 256 block tiles of 64 x 64 wave tiles and
// K slabs of 16 (111 KB of LDS).  A wave then reads NPL (TM + TN) = 12
// fragments from LDS for NP TM TN = 24 products where the 32 x 64 wave tile
// read 9 for 12: at 32 x 64 the LDS reads alone (0.75 KB per 32-cycle
// product per SIMD, against 1 KB of LDS bandwidth) left the matrix cores
// under half busy.  A 256-wide tile also stages and splits each A slab once
// for 256 receivers instead of once per 128.
constexpr int kSplitBlocks = 256;
static int bs_blocks(long tiles, int K) { return (int)std::min<long>(kSplitBlocks, tiles * ((K + kTBK - 1) / kTBK)); }
static int bs_bn(int N) { return N > 128 ? 256 : f32_bn(N); }

size_t gemm_bsplit_workspace_bytes(int M, int N, int K)
{
    const int bn = bs_bn(N);
    const long tiles = 2L * ((M + kTBM - 1) / kTBM) * ((N + bn - 1) / bn);
    return (size_t)bs_blocks(tiles, K) * 2 * kTBM * bn * sizeof(float) + (size_t)(tiles + 1) * sizeof(int);
}

hipError_t launch_split_table(int nprod, int K, int N, const float *B, long ldb, void *S, long ldk, hipStream_t st)
{
    if ((nprod != kSplitBf16x3 && nprod != kSplitBf16x6 && nprod != kSplitF16x3) || ldk < K || ldk % 32)
        return hipErrorInvalidValue;
    const size_t n = (size_t)N * ldk;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (nprod == kSplitBf16x3)
        hipLaunchKernelGGL((k_split_table<2, __bf16>), grid, dim3(256), 0, st, K, N, B, ldb, (__bf16 *)S, ldk);
    else if (nprod == kSplitBf16x6)
        hipLaunchKernelGGL((k_split_table<3, __bf16>), grid, dim3(256), 0, st, K, N, B, ldb, (__bf16 *)S, ldk);
    else
        hipLaunchKernelGGL((k_split_table<2, _Float16>), grid, dim3(256), 0, st, K, N, B, ldb, (_Float16 *)S, ldk);
    return hipGetLastError();
}

hipError_t launch_gemm_pair_bsplit(int nprod, int M, int N, int K, const float *A0, const float *A1, long lda,
                                   const void *S0, const void *S1, long ldk, float *C0, float *C1, float *P,
                                   const int *kend, const int *nk_w, hipStream_t st, const long long *foff,
                                   const int *fext, const unsigned *rpk)
{
    if ((nprod != kSplitBf16x3 && nprod != kSplitBf16x6 && nprod != kSplitF16x3) || ldk < K || ldk % 32 || lda % 4 ||
        (foff && (!fext || !nk_w || !kend)) || (nprod == kSplitF16x3 && !rpk))
        return hipErrorInvalidValue;
    GemmSet gs{};
    gs.A[0] = A0, gs.A[1] = A1, gs.C[0] = C0, gs.C[1] = C1;
    gs.kend[0] = gs.kend[1] = kend;
    gs.M[0] = gs.M[1] = M;
    gs.rpw[0] = gs.rpw[1] = 4;
    gs.foff = foff;
    gs.fext = fext;
    gs.rpk = rpk;
    SplitB sb{};
    sb.S[0] = S0;
    sb.S[1] = S1;
    const int bn = bs_bn(N);
    StreamK sk{};
    int t = 0;
    for (int q = 0; q < kGemmSetMax; q++) {
        sk.cum[q] = t;
        if (q < 2) t += (M + kTBM - 1) / kTBM * ((N + bn - 1) / bn);
    }
    sk.cum[kGemmSetMax] = t;
    sk.ncol = (N + bn - 1) / bn;
    sk.nb = bs_blocks(t, K);
    float *part = P;
    int *off = reinterpret_cast<int *>(P + (size_t)sk.nb * 2 * kTBM * bn);
    hipLaunchKernelGGL(k_streamk_plan<kTBM>, dim3(1), dim3(1024), 0, st, 2, K, gs, sk, off);
#define GRTM_BS(BNV, WMV, WNV, NPL, NP, BKV, WPEV, E)                                                          \
    hipLaunchKernelGGL((k_gemm_bssk<kTBM, BNV, WMV, WNV, NPL, NP, BKV, WPEV, E>), dim3(sk.nb), dim3(512), 0, st, 2, N, \
                       K, gs, sb, sk, lda, ldk, nk_w, off, part)
    if (bn == 256) {
        if (nprod == kSplitBf16x6) GRTM_BS(256, 64, 64, 3, 6, 16, 2, __bf16);
        else if (nprod == kSplitF16x3) GRTM_BS(256, 64, 64, 2, 3, 16, 2, _Float16);
        else GRTM_BS(256, 64, 64, 2, 3, 16, 2, __bf16);
        hipLaunchKernelGGL((k_streamk_reduce<kTBM, 256>), dim3((unsigned)t, kReduceSlices), dim3(256), 0, st, 2, N, gs,
                           sk, off, part);
    } else if (bn == 128) {
        if (nprod == kSplitBf16x6) GRTM_BS(128, 32, 64, 3, 6, 32, 2, __bf16);
        else if (nprod == kSplitF16x3) GRTM_BS(128, 32, 64, 2, 3, 32, 2, _Float16);
        else GRTM_BS(128, 32, 64, 2, 3, 32, 2, __bf16);
        hipLaunchKernelGGL((k_streamk_reduce<kTBM, 128>), dim3((unsigned)t, kReduceSlices), dim3(256), 0, st, 2, N, gs,
                           sk, off, part);
    } else {
        if (nprod == kSplitBf16x6) GRTM_BS(64, 32, 32, 3, 6, 32, 2, __bf16);
        else if (nprod == kSplitF16x3) GRTM_BS(64, 32, 32, 2, 3, 32, 2, _Float16);
        else GRTM_BS(64, 32, 32, 2, 3, 32, 2, __bf16);
        hipLaunchKernelGGL((k_streamk_reduce<kTBM, 64>), dim3((unsigned)t, kReduceSlices), dim3(256), 0, st, 2, N, gs,
                           sk, off, part);
    }
#undef GRTM_BS
    return hipGetLastError();
}

// Fills the partial pointers of a set from one workspace and launches the
// streamed kernel and the reduce; an FP32 set with per-tile K extents runs
// stream-K instead (kc unused).
template <typename T>
static hipError_t launch_gemm_set_t(int n, GemmSetT<T> gs, int N, int K, int kc, long lda, long ldb, T *P,
                                    const int *nk_w, hipStream_t st)
{
    if (n < 1 || n > kGemmSetMax) return hipErrorInvalidValue;
    if constexpr (sizeof(T) == 4) {
        bool ragged = false;
        for (int q = 0; q < n; q++) ragged = ragged || gs.kend[q];
        if (ragged) {
            for (int q = 0; q < n; q++)
                if (gs.M[q] < 1 || gs.rpw[q] < 1) return hipErrorInvalidValue;
            return launch_streamk(n, gs, N, K, lda, ldb, P, nk_w, st);
        }
    }
    const int bm = sizeof(T) == 4 ? kTBM : kDBM, bn = sizeof(T) == 4 ? f32_bn(N) : kDBN;
    const int S = (K + kc - 1) / kc;
    int mt = 0;
    long total = 0;
    for (int q = 0; q < n; q++) {
        if (gs.M[q] < 1 || gs.rpw[q] < 1) return hipErrorInvalidValue;
        gs.P[q] = P + (size_t)S * total;
        total += (long)gs.M[q] * N;
        mt = std::max(mt, (gs.M[q] + bm - 1) / bm);
    }
    if (mt > 65535 || (long)n * S > 65535) return hipErrorInvalidValue;
    dim3 grid((N + bn - 1) / bn, mt, n * S);
    if constexpr (sizeof(T) == 4) {
        if (bn == 128)
            hipLaunchKernelGGL((k_gemm_f32t<kTBM, 128, 32, 64>), grid, dim3(512), 0, st, N, K, kc, S, gs, lda, ldb,
                               nk_w);
        else
            hipLaunchKernelGGL((k_gemm_f32t<kTBM, 64, 32, 32>), grid, dim3(512), 0, st, N, K, kc, S, gs, lda, ldb,
                               nk_w);
    } else {
        hipLaunchKernelGGL(k_gemm_f64s, grid, dim3(256), 0, st, N, K, kc, S, gs, lda, ldb, nk_w);
    }
    hipLaunchKernelGGL(k_gemm_reduce_set<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, n, N, K, kc,
                       bm, gs, total);
    return hipGetLastError();
}

template <typename T>
static GemmSetT<T> pair_set(int M, const T *A0, const T *A1, const T *B0, const T *B1, T *C0, T *C1,
                            const int *kend)
{
    GemmSetT<T> gs{};
    gs.A[0] = A0, gs.A[1] = A1, gs.B[0] = B0, gs.B[1] = B1, gs.C[0] = C0, gs.C[1] = C1;
    gs.kend[0] = gs.kend[1] = kend;
    gs.M[0] = gs.M[1] = M;
    gs.rpw[0] = gs.rpw[1] = 4;
    return gs;
}

// Slice length of a dense set (dense_slice over the set's block tiles).
int gemm_set_slice(int mode, int n, const int *M, int N, int K)
{
    return dense_slice(set_tiles(mode, n, M, N), set_slots(mode), K, mode == 1 ? 16 : 8, mode == 1 ? 2 : 1);
}

size_t gemm_set_workspace_bytes(int n, const int *M, int N, int K, int kc, size_t es)
{
    size_t rows = 0;
    for (int q = 0; q < n; q++) rows += (size_t)M[q];
    return (size_t)((K + kc - 1) / kc) * rows * N * es;
}

hipError_t launch_gemm_pair_f64(int M, int N, int K, int kc, const double *A0, const double *A1, long lda,
                                const double *B0, const double *B1, long ldb, double *C0, double *C1, double *P,
                                const int *kend, const int *nk_w, hipStream_t st)
{
    if (kc < kDBK || kc % 32 || lda % 2 || ldb % 2) return hipErrorInvalidValue;
    return launch_gemm_set_t<double>(2, pair_set(M, A0, A1, B0, B1, C0, C1, kend), N, K, kc, lda, ldb, P, nk_w,
                                     st);
}

bool gemm_pair_ok(long lda, long ldb, const void *A0, const void *A1, const void *B0, const void *B1)
{
    auto al = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
    return lda % 4 == 0 && ldb % 4 == 0 && al(A0) && al(A1) && al(B0) && al(B1);
}

hipError_t launch_gemm_pair_f32(int M, int N, int K, int kc, const float *A0, const float *A1, long lda,
                                const float *B0, const float *B1, long ldb, float *C0, float *C1, float *P,
                                const int *kend, const int *nk_w, hipStream_t st, const long long *foff,
                                const int *fext)
{
    if (kc < kTBK || kc % kTBK) return hipErrorInvalidValue;
    // the resident layout is read by the stream-K kernel only (per-row extents)
    if (foff && (!fext || !nk_w || !kend)) return hipErrorInvalidValue;
    GemmSet gs = pair_set(M, A0, A1, B0, B1, C0, C1, kend);
    gs.foff = foff;
    gs.fext = fext;
    return launch_gemm_set_t<float>(2, gs, N, K, kc, lda, ldb, P, nk_w, st);
}

hipError_t launch_gemm_set(int prec, int n, const int *M, const int *rpw, int N, int K, int kc, const void *const *A,
                           long lda, const void *const *B, long ldb, void *const *C, void *P,
                           const int *const *kend, const int *nk_w, hipStream_t st)
{
    if (n < 1 || n > kGemmSetMax) return hipErrorInvalidValue;
    if (prec == 0) {
        if (kc < kDBK || kc % 32 || lda % 2 || ldb % 2) return hipErrorInvalidValue;
        GemmSet64 gs{};
        for (int q = 0; q < n; q++) {
            gs.A[q] = (const double *)A[q], gs.B[q] = (const double *)B[q], gs.C[q] = (double *)C[q];
            gs.kend[q] = kend ? kend[q] : nullptr, gs.M[q] = M[q], gs.rpw[q] = rpw[q];
        }
        return launch_gemm_set_t<double>(n, gs, N, K, kc, lda, ldb, (double *)P, nk_w, st);
    }
    if (kc < kTBK || kc % kTBK) return hipErrorInvalidValue;
    GemmSet gs{};
    for (int q = 0; q < n; q++) {
        gs.A[q] = (const float *)A[q], gs.B[q] = (const float *)B[q], gs.C[q] = (float *)C[q];
        gs.kend[q] = kend ? kend[q] : nullptr, gs.M[q] = M[q], gs.rpw[q] = rpw[q];
    }
    return launch_gemm_set_t<float>(n, gs, N, K, kc, lda, ldb, (float *)P, nk_w, st);
}

size_t gemm_workspace_bytes(int mode, int M, int N, int K)
{
    int S, ks;
    if (mode == 0) split_plan(M, N, K, 64, 64, 16, S, ks);
    else split_plan(M, N, K, 128, 128, mode == 1 ? 16 : 32, S, ks);
    return (size_t)S * M * N * (mode == 0 ? 8 : 4);
}

// mode 0: FP64 MFMA, 1: FP32 MFMA, 2: BF16x3 MFMA.  P: workspace of
// gemm_workspace_bytes(mode, M, N, K).
hipError_t launch_gemm_mfma(int mode, int M, int N, int K, const void *A, long lda, const void *B, long ldb,
                            void *Cm, void *P, const int *kend, const int *nk_w, hipStream_t st)
{
    int S, ks;
    if (mode == 0) {
        split_plan(M, N, K, 64, 64, 16, S, ks);
        dim3 grid((N + 63) / 64, (M + 63) / 64, S);
        hipLaunchKernelGGL(k_gemm_f64, grid, dim3(256), 0, st, M, N, K, ks, (const double *)A, lda,
                           (const double *)B, ldb, (double *)P, kend, nk_w);
        hipLaunchKernelGGL(k_gemm_reduce<double>, dim3((M * N + 255) / 256), dim3(256), 0, st, M * N, S,
                           (const double *)P, (double *)Cm);
    } else {
        split_plan(M, N, K, 128, 128, mode == 1 ? 16 : 32, S, ks);
        dim3 grid((N + 127) / 128, (M + 127) / 128, S);
        if (mode == 1)
            hipLaunchKernelGGL(k_gemm_f32, grid, dim3(256), 0, st, M, N, K, ks, (const float *)A, lda,
                               (const float *)B, ldb, (float *)P, kend, nk_w);
        else
            hipLaunchKernelGGL(k_gemm_bf16x3, grid, dim3(256), 0, st, M, N, K, ks, (const float *)A, lda,
                               (const float *)B, ldb, (float *)P, kend, nk_w);
        hipLaunchKernelGGL(k_gemm_reduce<float>, dim3((M * N + 255) / 256), dim3(256), 0, st, M * N, S,
                           (const float *)P, (float *)Cm);
    }
    return hipGetLastError();
}

}  // namespace grtm
