// kernels.hip -- HIP kernels of the Green's-function path for gfx950.
//
//   k_cells_zr     per-(omega,k) recursion -> Z/R k-kernels of the explosion
//                  and vertical-force sources, weighted by the k-quadrature
//                  and written as the A operands of the k-sum contraction
//   k_green_finish ordered sum of the per-block partials of the drop-in
//                  path's cell kernel (green.hip)
//   k_fallback_select, k_fallback_cells  row selection and dense completion
//                  of the learned adaptive grid (the residual check runs as
//                  flagged items of the recursion launches)
//   k_bessel_tab   J_n(k r) tables (B operands of the contraction)
//   k_gemm         C = A B on VALU (reference-precision path)
//   k_post_ifft    phase / taper / scale + Hermitian C2R inverse FFT
//                  (Stockham radix 4, in LDS) + exp(omega_I t) scaling; k_post_ifft_cols
//                  the same for column traces assembled from several outputs
// The batch recursion is in batch.hip, the column recursion in columns.hip,
// the cell helpers they share in cellops.h.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cell.h"
#include "cellops.h"
#include "integrand.h"
#include "kernels.h"

namespace grtm {

// Recursion over (frequency row, k tile) work items, one block per item.
// Items come from `tiles` (row << kTileBits | k tile, ordered by row) when
// given, otherwise they are all tiles of rows [w0, w0 + gridDim.x / tpr).  A
// truncated grid launches only the tiles below each row's cut, so its cost
// follows the cells it solves (launching empty blocks for the dropped tiles
// makes workgroup dispatch the bound).
// nk_w (optional): per-row number of wavenumbers solved (adaptive grid);
// row_peak (optional): per-row max |A| as float bits.
// BR: the plan's receiver branch, fixed at compile time.
// dev_count (persistent form): the item list was built on the device; the
// grid is sized to the resident blocks, and each block takes items
// [lo, hi) = [dev_count[which], dev_count[which + 1]) one at a time from the
// counter *ctr until the list is drained.
template <typename T, typename TO, bool COMPAT, int BR>
__global__ __launch_bounds__(kCellThreads) void k_cells_zr(const LayerW<T> *__restrict__ lw_all, int L,
                                                           Geo g, int nk, double dk, double df,
                                                           double omega_i, int w0, int tpr, int k_base,
                                                           const unsigned *__restrict__ tiles,
                                                           TO *__restrict__ A0, TO *__restrict__ A1, long lda,
                                                           unsigned *__restrict__ flag_count,
                                                           unsigned *__restrict__ flag_list,
                                                           const int *__restrict__ nk_w,
                                                           unsigned *__restrict__ row_peak,
                                                           const int *__restrict__ nchk, int stride,
                                                           unsigned *__restrict__ tail_max,
                                                           const unsigned *__restrict__ dev_count, int which,
                                                           unsigned *__restrict__ ctr)
{
    __shared__ LayerW<T> lw[kMaxLayers];
    __shared__ unsigned item_sh;
    const unsigned lo = dev_count ? dev_count[which] : 0u, hi = dev_count ? dev_count[which + 1] : 0u;
    for (int pass = 0;; pass++) {
        unsigned idx;
        if (dev_count) {
            __syncthreads();  // the previous item's reads of lw and item_sh are done
            if (threadIdx.x == 0) item_sh = lo + atomicAdd(ctr, 1u);
            __syncthreads();
            idx = item_sh;
            if (idx >= hi) return;  // uniform over the block: the list is drained
        } else {
            if (pass) return;
            idx = blockIdx.x;
        }
        int iw, kb;
        bool chk = false;
        if (tiles) {
            const unsigned e = tiles[idx];
            chk = (e & kCheckItem) != 0;
            iw = int((e & ~kCheckItem) >> kTileBits);
            kb = int(e & ((1u << kTileBits) - 1));
        } else {
            const int q = int(idx) / tpr;
            iw = w0 + q;
            kb = int(idx) - q * tpr;
        }
        const int nkr = nk_w ? nk_w[iw] : nk;
        const int lim = chk ? nchk[iw] : nkr;      // cells (check cells) of the row
        if (kb * kCellThreads >= lim) continue;  // uniform over the block
        for (int j = threadIdx.x; j < L; j += blockDim.x) lw[j] = lw_all[(size_t)iw * L + j];
        __syncthreads();
        const int j = kb * kCellThreads + threadIdx.x;
        const int ik = chk ? nkr + stride * j + stride / 2 : k_base + j;
        unsigned mag = 0;
        if (j < lim && (!chk || ik < nk)) {
            TO v[8];
            const bool ok = cell_zr<T, TO, COMPAT, BR>(lw, g, iw, ik, dk, df, omega_i, v);
            if (chk) {
                float m = 0.f;
                for (int i = 0; i < 8; i++) m = fmaxf(m, fabsf(float(v[i])));
                mag = __float_as_uint(isfinite(m) ? m : 3.0e38f);  // a non-finite check value fails the row
            } else {
                store_zr(v, iw, j, A0, A1, lda);
                mag = cell_mag_bits(v);
                if (!ok && flag_count) {
                    const unsigned slot = atomicAdd(flag_count, 1u);
                    flag_list[slot] = unsigned(iw) * unsigned(nk) + unsigned(ik);
                }
            }
        }
        unsigned *dst = chk ? tail_max : row_peak;
        if (dst) {
            mag = wave_max_u32(mag);
            if ((threadIdx.x & 63) == 0 && mag) atomicMax(dst + iw, mag);
        }
    }
}

// Recursion items of one model's learned grid, in two steps.  k_grid_scan
// (one block of 1024 threads, rows [8 t, 8 t + 8) per thread, nw <= 8192):
// each row's first tile item and first check item in the list order of the
// host (tiles of rows [0, r64) in row order, then their residual-check items,
// then the tiles of rows [r64, nw), then theirs): off[iw], off[nw + iw]; and
// counts = {0, items of rows [0, r64), all items}.  k_grid_fill
// (one wave per row) writes the items.  The two also start the execution's
// per-row and scalar state (GridInit), which took five launches of their own.
__global__ __launch_bounds__(1024) void k_grid_scan(int nw, int r64, const int *__restrict__ nk_pred,
                                                    const int *__restrict__ nchk, unsigned *__restrict__ off,
                                                    unsigned *__restrict__ counts, GridInit gi, int w_lo)
{
    if (threadIdx.x < kGridZero && gi.zero[threadIdx.x]) *gi.zero[threadIdx.x] = 0u;
    constexpr int R = 8;
    __shared__ unsigned wt[16], wc[16];
    __shared__ unsigned t64_sh, c64_sh;
    const int t = threadIdx.x;
    unsigned nt[R], nc[R], ti = 0, ci = 0;
    for (int q = 0; q < R; q++) {
        const int iw = t * R + q;  // rows below w_lo have no items here
        nt[q] = iw < nw && iw >= w_lo ? unsigned((nk_pred[iw] + kCellThreads - 1) / kCellThreads) : 0u;
        nc[q] = iw < nw && iw >= w_lo ? unsigned((nchk[iw] + kCellThreads - 1) / kCellThreads) : 0u;
        ti += nt[q];
        ci += nc[q];
    }
    // inclusive scans over the threads: within each wave by shuffles, then
    // the 16 wave totals
    const int lane = t & 63, wv = t >> 6;
    unsigned si = ti, sci = ci;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned a = unsigned(__shfl_up(int(si), o, 64)), b = unsigned(__shfl_up(int(sci), o, 64));
        if (lane >= o) si += a, sci += b;
    }
    if (lane == 63) wt[wv] = si, wc[wv] = sci;
    __syncthreads();
    unsigned T = 0, Cn = 0;
    for (int w = 0; w < 16; w++) {
        if (w == wv) si += T, sci += Cn;  // the earlier waves' totals
        T += wt[w];
        Cn += wc[w];
    }
    unsigned pt = si - ti, pc = sci - ci;  // exclusive prefixes of this thread's first row
    for (int q = 0; q < R; q++) {  // the prefixes at row r64
        if (t * R + q == r64) t64_sh = pt, c64_sh = pc;
        pt += nt[q];
        pc += nc[q];
    }
    if (t == 0 && r64 >= nw) t64_sh = T, c64_sh = Cn;
    __syncthreads();
    const unsigned T64 = t64_sh, C64 = c64_sh;
    pt = si - ti, pc = sci - ci;
    for (int q = 0; q < R; q++) {
        const int iw = t * R + q;
        if (iw >= nw) break;
        off[iw] = iw < r64 ? pt : C64 + pt;
        off[nw + iw] = iw < r64 ? T64 + pc : T + pc;
        pt += nt[q];
        pc += nc[q];
    }
    if (t == 0) {
        counts[0] = 0u;
        counts[1] = T64 + C64;
        counts[2] = T + Cn;
    }
}

__global__ __launch_bounds__(256) void k_grid_fill(int nw, const int *__restrict__ nk_pred,
                                                   const int *__restrict__ nchk, const unsigned *__restrict__ off,
                                                   unsigned *__restrict__ items, GridInit gi, int w_lo)
{
    const int iw = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (iw >= nw || iw < w_lo) return;
    if (lane == 0) {
        if (gi.nk_w) gi.nk_w[iw] = nk_pred[iw];
        if (gi.row_peak) gi.row_peak[iw] = 0u;
        if (gi.tail_max) gi.tail_max[iw] = 0u;
    }
    const unsigned nt = unsigned((nk_pred[iw] + kCellThreads - 1) / kCellThreads);
    const unsigned nc = unsigned((nchk[iw] + kCellThreads - 1) / kCellThreads);
    const unsigned tb = off[iw], cb = off[nw + iw];
    for (unsigned k = lane; k < nt; k += 64) items[tb + k] = (unsigned(iw) << kTileBits) | k;
    for (unsigned k = lane; k < nc; k += 64) items[cb + k] = ((unsigned(iw) << kTileBits) | k) | kCheckItem;
}

// Rows whose checked tail exceeds tol * row peak fall back to the dense grid;
// returns the row's final count of wavenumbers.
__device__ __forceinline__ int fallback_pick(int iw, int nw, int nk, double tol, const unsigned *__restrict__ row_peak,
                                             const unsigned *__restrict__ tail_max, int *__restrict__ nk_w,
                                             unsigned *__restrict__ fb_count, unsigned *__restrict__ fb_list)
{
    const float pk = __uint_as_float(row_peak[iw]), tm = __uint_as_float(tail_max[iw]);
    const int n = nk_w[iw];
    if (n < nk && !(tm <= float(tol) * pk)) {
        const unsigned slot = atomicAdd(fb_count, 1u);
        fb_list[slot] = unsigned(iw);
        fb_list[nw + slot] = unsigned(n);  // first wavenumber still to solve
        nk_w[iw] = nk;
        return nk;
    }
    return n;
}
__global__ void k_fallback_select(int nw, int nk, double tol, const unsigned *__restrict__ row_peak,
                                  const unsigned *__restrict__ tail_max, int *__restrict__ nk_w,
                                  unsigned *__restrict__ fb_count, unsigned *__restrict__ fb_list)
{
    const int iw = blockIdx.x * blockDim.x + threadIdx.x;
    if (iw >= nw) return;
    (void)fallback_pick(iw, nw, nk, tol, row_peak, tail_max, nk_w, fb_count, fb_list);
}
// The selection and the contraction tiles' K extents (k_kend_tiles) in one
// launch: wave t takes the rpt <= 64 rows of tile t, selects its fallback
// rows, and forms the tile's extent from their final counts.
__global__ __launch_bounds__(64) void k_fallback_select_kend(int nw, int nk, double tol, int rpt,
                                                             const unsigned *__restrict__ row_peak,
                                                             const unsigned *__restrict__ tail_max,
                                                             int *__restrict__ nk_w, unsigned *__restrict__ fb_count,
                                                             unsigned *__restrict__ fb_list, int *__restrict__ kend)
{
    const int t = blockIdx.x, iw = t * rpt + int(threadIdx.x);
    int n = 0;
    if (int(threadIdx.x) < rpt && iw < nw) n = fallback_pick(iw, nw, nk, tol, row_peak, tail_max, nk_w, fb_count, fb_list);
    const int m = int(wave_max_u32(unsigned(n)));
    if (threadIdx.x == 0) kend[t] = min(nk, (m + 31) / 32 * 32);
}

// Dense completion of fallback rows: wavenumbers [start, nk) of each listed
// row whose index lies in [r_lo, r_hi), in the arithmetic T of that row's
// recursion.  Grid (k blocks, row slots); slots stride over the list.  FP32
// cells that are not finite go to the flag list as (row_base + iw) nk + ik.
// MT: column rows (1: displacement columns, 2: with the strain rows), stored
// through o (A0, A1 unused).
template <typename T, typename TO, bool COMPAT, int MT>
__global__ __launch_bounds__(kCellThreads) void k_fallback_cells(
    const LayerW<T> *__restrict__ lw_all, int L, Geo g, int nk, int nw, double dk, double df, double omega_i,
    int r_lo, int r_hi, TO *__restrict__ A0, TO *__restrict__ A1, long lda, const unsigned *__restrict__ fb_count,
    const unsigned *__restrict__ fb_list, unsigned *__restrict__ flag_count, unsigned *__restrict__ flag_list,
    int row_base, MtOps o, unsigned *__restrict__ row_peak)
{
    __shared__ LayerW<T> lw[kMaxLayers];
    const unsigned n = *fb_count;
    for (unsigned slot = blockIdx.y; slot < n; slot += gridDim.y) {
        const int iw = int(fb_list[slot]);
        if (iw < r_lo || iw >= r_hi) continue;  // uniform over the block
        const int start = int(fb_list[nw + slot]);
        __syncthreads();
        for (int q = threadIdx.x; q < L; q += blockDim.x) lw[q] = lw_all[(size_t)iw * L + q];
        __syncthreads();
        for (int ik = start + blockIdx.x * blockDim.x + threadIdx.x; ik < nk; ik += gridDim.x * blockDim.x) {
            bool ok;
            if constexpr (MT != 0) {
                unsigned mag;
                ok = cell_mt<T, TO, COMPAT, MT == 2>(lw, g, iw, ik, dk, df, omega_i, o, lda, mag);
            } else {
                TO v[8];
                ok = cell_zr<T, TO, COMPAT>(lw, g, iw, ik, dk, df, omega_i, v);
                store_zr(v, iw, ik, A0, A1, lda);
                // the row peak covers the completed tail (it sets the FP16 split's row scale)
                if (row_peak) {
                    const unsigned m = cell_mag_bits(v);
                    if (m) atomicMax(row_peak + iw, m);
                }
            }
            if (!ok && flag_count) {
                const unsigned slot2 = atomicAdd(flag_count, 1u);
                flag_list[slot2] = unsigned(row_base + iw) * unsigned(nk) + unsigned(ik);
            }
        }
    }
}

// K extent of every contraction tile (rows_per_tile frequency rows each),
// rounded up to a multiple of 32.
__global__ void k_kend_tiles(int nw, int rows_per_tile, int nk, const int *__restrict__ nk_w, int *__restrict__ kend)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ntile = (nw + rows_per_tile - 1) / rows_per_tile;
    if (t >= ntile) return;
    int m = 0;
    for (int r = t * rows_per_tile; r < min(nw, (t + 1) * rows_per_tile); r++) m = max(m, nk_w[r]);
    kend[t] = min(nk, (m + 31) / 32 * 32);
}

// K extents of the row tiles of up to five contractions laid end to end in
// kend: contraction q has M_q = rpw_q nw rows (rpw_q per frequency row) and
// ceil(M_q / bm) tiles; each extent is rounded up to a multiple of 32.
struct KendSets {
    int n, bm;
    int rpw[kMtSetsMax];
    int tile0[kMtSetsMax + 1];  // first tile of each contraction; tile0[n] = all tiles
};
__global__ void k_kend_sets(int nw, int nk, KendSets ks, const int *__restrict__ nk_w, int *__restrict__ kend)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ks.tile0[ks.n]) return;
    int q = 0;
    while (q + 1 < ks.n && t >= ks.tile0[q + 1]) q++;
    const int tl = t - ks.tile0[q], rpw = ks.rpw[q];
    const int r0 = tl * ks.bm / rpw, r1 = min(nw - 1, ((tl + 1) * ks.bm - 1) / rpw);
    int m = 0;
    for (int r = r0; r <= r1; r++) m = max(m, nk_w[r]);
    kend[t] = min(nk, (m + 31) / 32 * 32);
}

// Training labels of one frequency row from its solved k-kernels (dense grid):
// m(k) = max |A| over the 8 operand entries of wavenumber k;
// out[0] = peak of m, out[1] = 1 + last k with m(k) > tau peak,
// out[2 + 3i .. 4 + 3i] = (index, half width at half maximum, height / peak)
// of the i-th largest local maximum above 1e-3 peak (index -1 when absent).
template <typename T>
__global__ __launch_bounds__(256) void k_row_labels(int nk, long lda, const T *__restrict__ A0,
                                                    const T *__restrict__ A1, double tau, float *__restrict__ out)
{
    extern __shared__ float m[];
    __shared__ float redv[256];
    __shared__ int redi[256];
    const int iw = blockIdx.x, tid = threadIdx.x;
    for (int k = tid; k < nk; k += 256) {
        float v = 0.f;
        for (int q = 0; q < 4; q++) {
            v = fmaxf(v, fabsf(float(A0[(size_t)(4 * iw + q) * lda + k])));
            v = fmaxf(v, fabsf(float(A1[(size_t)(4 * iw + q) * lda + k])));
        }
        m[k] = isfinite(v) ? v : 0.f;
    }
    __syncthreads();
    // peak
    float pk = 0.f;
    for (int k = tid; k < nk; k += 256) pk = fmaxf(pk, m[k]);
    redv[tid] = pk;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) redv[tid] = fmaxf(redv[tid], redv[tid + o]);
        __syncthreads();
    }
    pk = redv[0];
    __syncthreads();
    // truncation index
    int last = 0;
    for (int k = tid; k < nk; k += 256)
        if (m[k] > float(tau) * pk) last = max(last, k + 1);
    redi[tid] = last;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) redi[tid] = max(redi[tid], redi[tid + o]);
        __syncthreads();
    }
    float *o = out + (size_t)iw * 26;
    if (tid == 0) {
        o[0] = pk;
        o[1] = float(redi[0]);
    }
    __syncthreads();
    // top-8 local maxima; a taken peak suppresses +-2 half widths around it
    for (int round = 0; round < 8; round++) {
        float best = 0.f;
        int bi = -1;
        for (int k = tid + 1; k < nk - 1; k += 256) {
            const float v = m[k];
            if (v > m[k - 1] && v >= m[k + 1] && v > 1e-3f * pk && v > best) {
                best = v;
                bi = k;
            }
        }
        redv[tid] = best;
        redi[tid] = bi;
        __syncthreads();
        for (int s2 = 128; s2 > 0; s2 >>= 1) {
            if (tid < s2 && redv[tid + s2] > redv[tid]) {
                redv[tid] = redv[tid + s2];
                redi[tid] = redi[tid + s2];
            }
            __syncthreads();
        }
        const int c = redi[0];
        const float h = redv[0];
        __syncthreads();
        if (c < 0) {
            if (tid == 0) o[2 + 3 * round] = -1.f, o[3 + 3 * round] = 0.f, o[4 + 3 * round] = 0.f;
            continue;
        }
        // half width at half maximum (scan both sides, bounded)
        int hw = 0;
        if (tid == 0) {
            int l = c, r = c;
            while (l > 0 && c - l < 2048 && m[l] > 0.5f * h) l--;
            while (r < nk - 1 && r - c < 2048 && m[r] > 0.5f * h) r++;
            hw = max(1, (r - l) / 2);
            o[2 + 3 * round] = float(c);
            o[3 + 3 * round] = float(hw);
            o[4 + 3 * round] = h / pk;
            redi[0] = hw;
        }
        __syncthreads();
        hw = redi[0];
        __syncthreads();
        for (int k = max(0, c - 2 * hw) + tid; k <= min(nk - 1, c + 2 * hw); k += 256) m[k] = 0.f;
        __syncthreads();
    }
}

hipError_t launch_row_labels(int prec, int nw, int nk, long lda, const void *A0, const void *A1, double tau,
                             float *out, hipStream_t st)
{
    const size_t lds = sizeof(float) * (size_t)nk;
    if (lds > 65536) {
        hipError_t e = prec == 0 ? hipFuncSetAttribute((const void *)k_row_labels<double>,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)
                                 : hipFuncSetAttribute((const void *)k_row_labels<float>,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (prec == 0)
        hipLaunchKernelGGL(k_row_labels<double>, dim3(nw), dim3(256), lds, st, nk, lda, (const double *)A0,
                           (const double *)A1, tau, out);
    else
        hipLaunchKernelGGL(k_row_labels<float>, dim3(nw), dim3(256), lds, st, nk, lda, (const float *)A0,
                           (const float *)A1, tau, out);
    return hipGetLastError();
}

__global__ void k_copy_i32(const int *__restrict__ a, int *__restrict__ b, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

// FP64 re-solve of the flagged cells (grid-stride over the list); BR: the
// plan's receiver branch, fixed at compile time as in k_cells_zr.
template <typename TO, bool COMPAT, int BR>
__global__ __launch_bounds__(kCellThreads) void k_cells_fix(const LayerW<double> *__restrict__ lw_all, int L,
                                                            Geo g, int nk, double dk, double df, double omega_i,
                                                            TO *__restrict__ A0, TO *__restrict__ A1, long lda,
                                                            const unsigned *__restrict__ flag_count,
                                                            const unsigned *__restrict__ flag_list,
                                                            unsigned *__restrict__ row_peak)
{
    const unsigned n = *flag_count;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const unsigned cell = flag_list[t];
        const int iw = int(cell / unsigned(nk)), ik = int(cell % unsigned(nk));
        TO v[8];
        cell_zr<double, TO, COMPAT, BR>(lw_all + (size_t)iw * L, g, iw, ik, dk, df, omega_i, v);
        store_zr(v, iw, ik, A0, A1, lda);
        // the row peak covers the re-solved values (it sets the FP16 split's row scale)
        if (row_peak) {
            const unsigned m = cell_mag_bits(v);
            if (m) atomicMax(row_peak + iw, m);
        }
    }
}

// Test hook: the listed cells (row * nk + k) get non-finite operands and join
// the flag list, as a FP32 cell that overflowed does in the recursion kernels.
__global__ void k_plant_nonfinite(const unsigned *__restrict__ cells, int n, int nk, float *__restrict__ A0,
                                  float *__restrict__ A1, long lda, unsigned *__restrict__ flag_count,
                                  unsigned *__restrict__ flag_list)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const unsigned cell = cells[t];
    float v[8];
    for (int i = 0; i < 8; i++) v[i] = __builtin_nanf("");
    store_zr(v, int(cell / unsigned(nk)), int(cell % unsigned(nk)), A0, A1, lda);
    flag_list[atomicAdd(flag_count, 1u)] = cell;
}

// Cells of the FP32 rows re-solved in FP64 because they lie next to a branch
// point or a surface-wave pole (grtm_plan_set_flag64).  One block per row
// iw = r64 + blockIdx.x.  Intervals [lo, hi) of wavenumber cells, all below
// the row's cut nk_w[iw]:
//  - every layer's P and S branch points: the vertical wavenumbers
//    sqrt(k^2 - (w/V)^2) (green.cu:184-194) branch where k^2 = (w/V)^2, i.e.
//    nearest the grid at k = Re(w/V); cells within hb of it (hb >= 0);
//  - each pole cluster the network predicted for the row (centre c and half
//    width hw in km^-1, poles[iw][2i], [2i + 1]): cells within pole_widths
//    half widths of c, at most kFlagPoleCells on either side.
// Each cell joins the flag list once (a cell also covered by an earlier
// interval is skipped): a count pass, one atomic reservation per block, a
// write pass.  The order of the row's entries is arbitrary.
constexpr int kFlagPoleCells = 64;
__global__ __launch_bounds__(256) void k_flag_cells(const LayerW<double> *__restrict__ lw_all, int L, int r64, int nk,
                                                    double dk, int hb, const float *__restrict__ poles,
                                                    float pole_widths, const int *__restrict__ nk_w,
                                                    unsigned *__restrict__ flag_count, unsigned *__restrict__ flag_list)
{
    constexpr int kMaxIv = 2 * kMaxLayers + 8;
    __shared__ int lo[kMaxIv], hi[kMaxIv];
    __shared__ unsigned cnt, pos, base;
    const int iw = r64 + blockIdx.x, tid = threadIdx.x;
    const int nkr = nk_w ? nk_w[iw] : nk;
    const int nb = hb >= 0 ? 2 * L : 0, np = poles ? 8 : 0, ni = nb + np;
    if (tid < nb) {
        const LayerW<double> &l = lw_all[(size_t)iw * L + tid / 2];
        const double kb = fabs((tid & 1) ? l.sb.x : l.sa.x);
        const long c = lrint(kb / dk);
        lo[tid] = int(max(0L, c - hb));
        hi[tid] = int(min((long)nkr, c + hb + 1));
    } else if (tid < ni) {
        const int i = tid - nb;
        const float c = poles[(size_t)iw * 16 + 2 * i], w = pole_widths * poles[(size_t)iw * 16 + 2 * i + 1];
        int a = 0, b = 0;
        if (c > 0.f && w > 0.f) {
            const double cc = double(c) / dk, ww = fmin(double(w) / dk, double(kFlagPoleCells));
            a = int(fmin(fmax(0.0, floor(cc - ww)), double(nkr)));
            b = int(fmin(double(nkr), fmax(0.0, ceil(cc + ww) + 1.0)));
        }
        lo[tid] = a;
        hi[tid] = max(a, b);
    }
    if (tid == 0) cnt = 0u, pos = 0u;
    __syncthreads();
    for (int pass = 0; pass < 2; pass++) {
        for (int i = 0; i < ni; i++) {
            for (int k = lo[i] + tid; k < hi[i]; k += blockDim.x) {
                bool fresh = true;
                for (int j = 0; j < i && fresh; j++) fresh = k < lo[j] || k >= hi[j];
                if (!fresh) continue;
                if (pass == 0)
                    atomicAdd(&cnt, 1u);
                else
                    flag_list[base + atomicAdd(&pos, 1u)] = unsigned(iw) * unsigned(nk) + unsigned(k);
            }
        }
        __syncthreads();
        if (pass == 0) {
            if (tid == 0) base = cnt ? atomicAdd(flag_count, cnt) : 0u;
            __syncthreads();
        }
    }
}

__global__ void k_zero_u32(unsigned *p, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0u;
}

__global__ void k_green_finish(int nw, int nblk, const double *__restrict__ partial, double *__restrict__ raw)
{
    // raw[(col * nw + iw) * 2 + {re, im}]
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nw * 60) return;
    const int iw = t / 60, q = t % 60;
    double s = 0;
    for (int b = 0; b < nblk; b++) s += partial[((size_t)iw * nblk + b) * 60 + q];
    const int col = q >> 1, ri = q & 1;
    raw[((size_t)col * nw + iw) * 2 + ri] = s;
}

// ---------------------------------------------------------------- Bessel tables
// J_n0(k r) into B0 and J_n0+1(k r) into B1 (if given), [nk][ldb]; mode 0: the
// reference's forms (A&S polynomials, Miller / upward recurrence), mode 1:
// the math library.
template <typename T>
__global__ void k_bessel_tab(int mode, int n0, int nk, int k0, double dk, const double *__restrict__ dists, int nr,
                             T *__restrict__ B0, T *__restrict__ B1, long ldb)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nk * nr) return;
    const int ik = int(t / nr), ir = int(t % nr);
    const double x = dk * (k0 + ik) * dists[ir];  // phase formed in FP64
    double j0v, j1v;
    if (mode == 0) {
        j0v = n0 == 0 ? as_j0(x) : as_jn(x, n0);
        j1v = n0 == 0 ? as_j1(x) : as_jn(x, n0 + 1);
    } else if (n0 == 0) {
        j0v = j0(x);
        j1v = j1(x);
    } else {
        j0v = lib_jn(n0, x);
        j1v = lib_jn(n0 + 1, x);
    }
    B0[(size_t)ik * ldb + ir] = T(j0v);
    if (B1) B1[(size_t)ik * ldb + ir] = T(j1v);
}

// ---------------------------------------------------------------- VALU GEMM
// C[M x N] = A[M x K] B[K x N], row-major, 64x64 tiles, 256 threads x 4x4.
template <typename T>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const T *__restrict__ A, long lda,
                                              const T *__restrict__ B, long ldb, T *__restrict__ Cm,
                                              long ldc)
{
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ T As[BK][BM + 1];
    __shared__ T Bs[BK][BN + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    T acc[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) acc[i][j] = T(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
        for (int e = threadIdx.x; e < BM * BK; e += 256) {
            int mm = e / BK, kk = e % BK;
            int gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < M && gk < K) ? A[(size_t)gm * lda + gk] : T(0);
        }
        for (int e = threadIdx.x; e < BN * BK; e += 256) {
            int kk = e / BN, nn = e % BN;
            int gk = k0 + kk, gn = n0 + nn;
            Bs[kk][nn] = (gk < K && gn < N) ? B[(size_t)gk * ldb + gn] : T(0);
        }
        __syncthreads();
        for (int kk = 0; kk < BK; kk++) {
            T a[4], b[4];
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
            for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx * 4 + j];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            int gm = m0 + ty * 4 + i, gn = n0 + tx * 4 + j;
            if (gm < M && gn < N) Cm[(size_t)gm * ldc + gn] = acc[i][j];
        }
}

// ---------------------------------------------------------------- post + IFFT
template <typename T>
__device__ __forceinline__ cx<T> post_val(cx<T> v, int iw, const PostParams &pp)
{
    const T *pv = static_cast<const T *>(pp.pv);
    return v * C<T>(pv[2 * iw], pv[2 * iw + 1]);
}

// Phase / taper / scale, Hermitian C2R inverse FFT (Stockham in LDS: radix-4
// stages, one radix-2 stage when log2 of the half length is odd) and time
// scaling of R traces t0 .. t0 + R - 1, whose raw spectrum rows iw are
// load(r, iw).  The R traces are adjacent receivers, which share every
// 128-byte line of the row-major contraction outputs: R consecutive lanes
// load R consecutive values, and all R spectra are staged in LDS before they
// are transformed one after the other.
// kIfftThreads threads (16 waves per compute unit at two workgroups: 30.0
// us for the 1024 traces of the 256-receiver workload, 31.6 at 256 threads;
// a copy of the twiddle table in LDS measured slower, 33.6); LDS
// post_ifft_lds_bytes.
constexpr int kIfftThreads = 512;
template <typename T, int R, typename Load>
__device__ __forceinline__ void post_ifft_group(const PostParams &pp, Load &&load, int t0, T *__restrict__ traces,
                                                T *__restrict__ spec_post)
{
    extern __shared__ unsigned char smem_raw[];
    const int M = pp.nw, N = 2 * pp.nw, Mp = M;
    cx<T> *stage = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *bufA = stage + R * Mp;
    const T *tw = static_cast<const T *>(pp.tw), *ts = static_cast<const T *>(pp.ts);
    // unrolled so that several loads are in flight
#pragma unroll 4
    for (int i = threadIdx.x; i < R * M; i += kIfftThreads) {
        const int r = R == 1 ? 0 : i % R, m = R == 1 ? i : i / R;
        const cx<T> Xm = post_val(load(r, m), m, pp);
        if (spec_post) {
            spec_post[((size_t)(t0 + r) * M + m) * 2] = Xm.x;
            spec_post[((size_t)(t0 + r) * M + m) * 2 + 1] = Xm.y;
        }
        stage[r * Mp + m] = Xm;
    }
    __syncthreads();
    for (int r = 0; r < R; r++) {
        cx<T> *sr = stage + r * Mp;
        // Hermitian fold to a half-length complex sequence
#pragma unroll 4
        for (int m = threadIdx.x; m < M; m += kIfftThreads) {
            cx<T> Xm = sr[m];
            cx<T> Xc = C<T>(0, 0);  // conj(X[M-m]); X[M] (Nyquist) = 0
            if (m > 0) Xc = conj(sr[M - m]);
            if (m == 0) Xm.y = T(0);  // Im X[0] only reaches Im of the output
            const cx<T> Xe = Xm + Xc;
            const cx<T> Xo = (Xm - Xc) * C<T>(tw[2 * m], tw[2 * m + 1]);  // e^{+2 pi i m / N}
            bufA[m] = C<T>(Xe.x - Xo.y, Xe.y + Xo.x);                   // Xe + i Xo
        }
        __syncthreads();
        // Stockham decimation in frequency, inverse sign: a length-n
        // sub-transform at stride s takes x[q + s (p + l n/4)], l = 0..3, to
        // y[q + s (4 p + u)], u = 0..3, twiddled by w^u, w = e^{2 pi i p / n}
        // = tw[p N / n]; the staged spectrum's row is the second buffer
        cx<T> *x = bufA, *y = sr;
        int s = 1, ls = 0, n = M;
        for (; n >= 4; n >>= 2) {
            const int m4 = n >> 2, st = N / n;
#pragma unroll 2
            for (int idx = threadIdx.x; idx < M / 4; idx += kIfftThreads) {
                const int p = idx >> ls, q = idx & (s - 1);
                const cx<T> a0 = x[q + s * p], a1 = x[q + s * (p + m4)];
                const cx<T> a2 = x[q + s * (p + 2 * m4)], a3 = x[q + s * (p + 3 * m4)];
                const cx<T> w1 = C<T>(tw[2 * p * st], tw[2 * p * st + 1]);
                const cx<T> w2 = C<T>(tw[4 * p * st], tw[4 * p * st + 1]);  // 2 p N / n < N / 2
                const cx<T> w3 = w1 * w2;
                const cx<T> b0 = a0 + a2, b1 = a0 - a2, b2 = a1 + a3, d = a1 - a3;
                const cx<T> b3 = C<T>(-d.y, d.x);  // i (a1 - a3)
                cx<T> *o = y + q + s * (4 * p);
                o[0] = b0 + b2;
                o[s] = (b1 + b3) * w1;
                o[2 * s] = (b0 - b2) * w2;
                o[3 * s] = (b1 - b3) * w3;
            }
            __syncthreads();
            cx<T> *tmp = x;
            x = y;
            y = tmp;
            s <<= 2;
            ls += 2;
        }
        if (n == 2) {  // last radix-2 stage: p = 0, unit twiddle
#pragma unroll 4
            for (int q = threadIdx.x; q < s; q += kIfftThreads) {
                const cx<T> a = x[q], b = x[q + s];
                y[q] = a + b;
                y[q + s] = a - b;
            }
            __syncthreads();
            x = y;
        }
        T *out = traces + (size_t)(t0 + r) * N;
#pragma unroll 4
        for (int nn = threadIdx.x; nn < M; nn += kIfftThreads) {
            const cx<T> z = x[nn];
            out[2 * nn] = z.x * ts[2 * nn];
            out[2 * nn + 1] = z.y * ts[2 * nn + 1];
        }
        if (r + 1 < R) __syncthreads();  // the next fold rewrites bufA
    }
}

// Workgroups are dispatched to the 8 XCDs in turn, and each XCD has its own
// L2.  Traces t and t + 1 read the same 128-byte lines of the row-major
// spectra (adjacent receivers), so block b takes trace group
// (b % 8) ceil(ng / 8) + b / 8: each XCD works through a contiguous range
// of traces and fetches each line once.  The grid has 8 ceil(ng / 8) blocks.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_trace(int ng)
{
    const int per = (ng + kXcds - 1) / kXcds;
    return int(blockIdx.x % kXcds) * per + int(blockIdx.x / kXcds);
}
static unsigned xcd_grid(int ng) { return (unsigned)((ng + kXcds - 1) / kXcds * kXcds); }

// One workgroup per group of R trace components (R divides nr).  layout 0:
// raw[(t*nw + iw)*2 + {re,im}] (R = 1); layout 1: rows of the contraction
// outputs: component c = (t / nr) % 2 reads C_c[(4 iw + 2 q + {0,1}) * ldc + ir]
// with q = t / (2 nr), ir = t % nr.  Several models (layout 1): trace t
// belongs to model t / (4 nr), whose contraction outputs start cms elements
// further per model.
template <typename T, int R>
__global__ __launch_bounds__(kIfftThreads) void k_post_ifft(PostParams pp, int layout, const T *__restrict__ raw,
                                                   const T *__restrict__ C0, const T *__restrict__ C1,
                                                   long ldc, int nr, long cms, int ntr, T *__restrict__ traces,
                                                   T *__restrict__ spec_post)
{
    const int g = xcd_trace(ntr / R);
    if (g >= ntr / R) return;
    const int t0 = g * R, M = pp.nw;
    const T *Cc = nullptr;
    if (layout == 1) {
        const int mj = t0 / (4 * nr), tl = t0 - mj * 4 * nr;
        const int q = tl / (2 * nr), c = (tl / nr) % 2, ir0 = tl % nr;
        Cc = (c ? C1 : C0) + (size_t)mj * cms + ir0 + (size_t)(2 * q) * ldc;
    }
    auto load = [&](int r, int iw) -> cx<T> {
        if (layout == 0) {
            const T *p = raw + ((size_t)(t0 + r) * M + iw) * 2;
            return C<T>(p[0], p[1]);
        }
        const T *p = Cc + (size_t)(4 * iw) * ldc + r;
        return C<T>(p[0], p[ldc]);
    };
    post_ifft_group<T, R>(pp, load, t0, traces, spec_post);
}

// Column traces: trace t = column t / nr at receiver t % nr; its raw
// spectrum is the sum of the column's terms scale r^-rpow C_set[row] over
// the contraction outputs (rows 2 (rows_set iw + row) + {0 re, 1 im}).
// Groups of R receivers of one column (R divides nr).
template <typename T, int R>
__global__ __launch_bounds__(kIfftThreads) void k_post_ifft_cols(PostParams pp, const MtCols *__restrict__ mcp, long ldc,
                                                        int nr, int ntr, const double *__restrict__ dists,
                                                        T *__restrict__ traces)
{
    __shared__ double fac[kMtTermsMax][R];  // scale r^-rpow per term and receiver
    const MtCols &mc = *mcp;
    const int g = xcd_trace(ntr / R);
    if (g >= ntr / R) return;
    const int t0 = g * R;
    const int col = t0 / nr, ir0 = t0 - col * nr;
    const int nt = mc.nterm[col];
    if ((int)threadIdx.x < nt * R) {
        const int u = threadIdx.x / R, r = threadIdx.x % R;
        const MtTerm &tm = mc.term[col][u];
        const double rinv = 1.0 / dists[ir0 + r];
        double f = tm.scale;
        for (int e = 0; e < tm.rpow; e++) f *= rinv;
        fac[u][r] = f;
    }
    __syncthreads();
    auto load = [&](int r, int iw) -> cx<T> {
        double re = 0, im = 0;
        for (int u = 0; u < nt; u++) {
            const MtTerm &tm = mc.term[col][u];
            const T *Cs = reinterpret_cast<const T *>(mc.C[tm.set]) + ir0 + r;
            const size_t row = 2 * ((size_t)mc.rows[tm.set] * iw + tm.row);
            const double f = fac[u][r];
            re += f * double(Cs[row * ldc]);
            im += f * double(Cs[(row + 1) * ldc]);
        }
        return C<T>(T(re), T(im));
    };
    post_ifft_group<T, R>(pp, load, t0, traces, (T *)nullptr);
}

// ---------------------------------------------------------------- launchers
// Recursion over frequency rows [w0, w0 + nwr).  prec: 0 FP64 cells with FP64
// A operands, 1 FP32 cells with FP32 A operands (flagged cells appended to the
// list), 2 FP64 cells with FP32 A operands.  lw points at row 0.
int cells_tiles_per_row(int nk) { return (nk + kCellThreads - 1) / kCellThreads; }

unsigned cells_tile(int iw, int kb) { return (unsigned(iw) << kTileBits) | unsigned(kb); }

int cells_tile_limit(int nw, int nk)
{
    return (nw < (1 << (32 - kTileBits)) && cells_tiles_per_row(nk) <= (1 << kTileBits)) ? 1 : 0;
}

// Blocks of `kernel` (kCellThreads threads) resident on the whole device at
// once, capped at ub: the grid of a persistent launch.
static int resident_blocks(const void *kernel, int ub)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kCellThreads, 0) != hipSuccess || per < 1)
        per = 1, cus = cus > 0 ? cus : 256;
    return std::max(1, std::min(ub, per * cus));
}

hipError_t launch_grid_items(int nw, int r64, const int *nk_pred, const int *nchk, unsigned *items, unsigned *off,
                             unsigned *counts, hipStream_t st, const GridInit &gi, bool low_apart)
{
    if (nw < 1 || nw > 8192 || r64 < 0) return hipErrorInvalidValue;
    const int w_lo = low_apart ? std::min(r64, nw) : 0;
    hipLaunchKernelGGL(k_grid_scan, dim3(1), dim3(1024), 0, st, nw, r64, nk_pred, nchk, off, counts, gi, w_lo);
    hipLaunchKernelGGL(k_grid_fill, dim3((nw + 3) / 4), dim3(256), 0, st, nw, nk_pred, nchk, off, items, gi, w_lo);
    return hipGetLastError();
}

__global__ void k_rows_init(int n, const int *__restrict__ nk_src, int *__restrict__ nk_w,
                            unsigned *__restrict__ row_peak, unsigned *__restrict__ tail_max)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    nk_w[r] = nk_src[r];
    row_peak[r] = 0u;
    tail_max[r] = 0u;
}

hipError_t launch_rows_init(int n, const int *nk_src, int *nk_w, unsigned *row_peak, unsigned *tail_max,
                            hipStream_t st)
{
    if (n < 1) return hipSuccess;
    hipLaunchKernelGGL(k_rows_init, dim3((n + 255) / 256), dim3(256), 0, st, n, nk_src, nk_w, row_peak, tail_max);
    return hipGetLastError();
}

hipError_t launch_cells_zr(int prec, int compat, const void *lw, int L, const Geo &g, int nk, double dk,
                           double df, double omega_i, int w0, int nwr, const unsigned *tiles, int ntiles, void *A0,
                           void *A1, long lda, unsigned *flag_count, unsigned *flag_list, const int *nk_w,
                           unsigned *row_peak, hipStream_t st, const int *nchk, int stride, unsigned *tail_max,
                           const unsigned *dev_count, int which, unsigned *ctr, int k_base, bool ctr_zeroed)
{
    const int tpr = cells_tiles_per_row(nk);
    if (!tiles) ntiles = nwr * tpr;
    if (ntiles <= 0) return hipSuccess;
    if (dev_count) {
        if (!tiles || !ctr || which < 0) return hipErrorInvalidValue;
        if (!ctr_zeroed) {
            hipError_t e = launch_zero_u32(ctr, 1, st);
            if (e != hipSuccess) return e;
        }
    }
#define GRTM_ZR_BR(T, TO, CP, BR, FC, FL)                                                                       \
    hipLaunchKernelGGL((k_cells_zr<T, TO, CP, BR>),                                                              \
                       dim3(dev_count ? resident_blocks((const void *)k_cells_zr<T, TO, CP, BR>, ntiles) : ntiles), \
                       dim3(kCellThreads), 0, st, (const LayerW<T> *)lw, L, g, nk, dk, df, omega_i, w0, tpr, k_base, \
                       tiles,                                                                                        \
                       (TO *)A0, (TO *)A1, lda, FC, FL, nk_w, row_peak, nchk, stride, tail_max, dev_count, which, ctr)
#define GRTM_ZR(T, TO, CP, FC, FL)                                                 \
    do {                                                                           \
        switch (g.branch) {                                                        \
        case 1: GRTM_ZR_BR(T, TO, CP, 1, FC, FL); break;                           \
        case 2: GRTM_ZR_BR(T, TO, CP, 2, FC, FL); break;                           \
        case 3: GRTM_ZR_BR(T, TO, CP, 3, FC, FL); break;                           \
        default: GRTM_ZR_BR(T, TO, CP, 4, FC, FL); break;                          \
        }                                                                          \
    } while (0)
    if (prec == 0) {
        if (compat) GRTM_ZR(double, double, true, nullptr, nullptr);
        else GRTM_ZR(double, double, false, nullptr, nullptr);
    } else if (prec == 1) {
        if (compat) GRTM_ZR(float, float, true, flag_count, flag_list);
        else GRTM_ZR(float, float, false, flag_count, flag_list);
    } else {
        if (compat) GRTM_ZR(double, float, true, nullptr, nullptr);
        else GRTM_ZR(double, float, false, nullptr, nullptr);
    }
#undef GRTM_ZR
#undef GRTM_ZR_BR
    return hipGetLastError();
}

hipError_t launch_fallback(int out_f64, int compat, const void *lw64, const void *lw32, int r64, int L, const Geo &g,
                           int nk, int nw, double dk, double df, double omega_i, double tol,
                           unsigned *row_peak, const unsigned *tail_max, int *nk_w, unsigned *fb_count,
                           unsigned *fb_list, void *A0, void *A1, long lda, unsigned *flag_count,
                           unsigned *flag_list, int row_base, hipStream_t st, const MtOps *mto, int deriv,
                           int *kend, int rows_per_tile)
{
    if (deriv && (compat || !mto)) return hipErrorInvalidValue;
    if (kend) {
        if (rows_per_tile < 1 || rows_per_tile > 64) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_fallback_select_kend, dim3((nw + rows_per_tile - 1) / rows_per_tile), dim3(64), 0, st, nw,
                           nk, tol, rows_per_tile, row_peak, tail_max, nk_w, fb_count, fb_list, kend);
    } else {
        hipLaunchKernelGGL(k_fallback_select, dim3((nw + 255) / 256), dim3(256), 0, st, nw, nk, tol, row_peak,
                           tail_max, nk_w, fb_count, fb_list);
    }
    dim3 grid((nk + kCellThreads - 1) / kCellThreads, 16);
    const MtOps o = mto ? *mto : MtOps{};
#define GRTM_FB(T, TO, CP, MT, LWP, LO, HI, FC, FL)                                                             \
    hipLaunchKernelGGL((k_fallback_cells<T, TO, CP, MT>), grid, dim3(kCellThreads), 0, st, (const LayerW<T> *)LWP, \
                       L, g, nk, nw, dk, df, omega_i, LO, HI, (TO *)A0, (TO *)A1, lda, fb_count, fb_list, FC, FL,  \
                       row_base, o, MT ? nullptr : row_peak)
#define GRTM_FB_ALL(MT)                                                                              \
    if (out_f64) {                                                                                   \
        if (compat) GRTM_FB(double, double, true, MT, lw64, 0, nw, nullptr, nullptr);                \
        else GRTM_FB(double, double, false, MT, lw64, 0, nw, nullptr, nullptr);                      \
    } else if (compat) {                                                                             \
        GRTM_FB(double, float, true, MT, lw64, 0, r64, nullptr, nullptr);                            \
        GRTM_FB(float, float, true, MT, lw32, r64, nw, flag_count, flag_list);                       \
    } else {                                                                                         \
        GRTM_FB(double, float, false, MT, lw64, 0, r64, nullptr, nullptr);                           \
        GRTM_FB(float, float, false, MT, lw32, r64, nw, flag_count, flag_list);                      \
    }
    // rows below r64 complete in FP64, the others in FP32 like the rest of their row
    if (mto && deriv) {
        if (out_f64) {
            GRTM_FB(double, double, false, 2, lw64, 0, nw, nullptr, nullptr);
        } else {
            GRTM_FB(double, float, false, 2, lw64, 0, r64, nullptr, nullptr);
            GRTM_FB(float, float, false, 2, lw32, r64, nw, flag_count, flag_list);
        }
    } else if (mto) {
        GRTM_FB_ALL(1)
    } else {
        GRTM_FB_ALL(0)
    }
#undef GRTM_FB_ALL
#undef GRTM_FB
    return hipGetLastError();
}

hipError_t launch_kend_sets(int nw, int nk, int n, const int *rpw, int bm, const int *nk_w, int *kend, hipStream_t st)
{
    if (n < 1 || n > kMtSetsMax) return hipErrorInvalidValue;
    KendSets ks{};
    ks.n = n;
    ks.bm = bm;
    ks.tile0[0] = 0;
    for (int q = 0; q < n; q++) {
        ks.rpw[q] = rpw[q];
        ks.tile0[q + 1] = ks.tile0[q] + (rpw[q] * nw + bm - 1) / bm;
    }
    hipLaunchKernelGGL(k_kend_sets, dim3((ks.tile0[n] + 255) / 256), dim3(256), 0, st, nw, nk, ks, nk_w, kend);
    return hipGetLastError();
}

hipError_t launch_kend_tiles(int nw, int rows_per_tile, int nk, const int *nk_w, int *kend, hipStream_t st)
{
    const int ntile = (nw + rows_per_tile - 1) / rows_per_tile;
    hipLaunchKernelGGL(k_kend_tiles, dim3((ntile + 255) / 256), dim3(256), 0, st, nw, rows_per_tile, nk, nk_w, kend);
    return hipGetLastError();
}

hipError_t launch_copy_i32(const int *a, int *b, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy_i32, dim3((n + 255) / 256), dim3(256), 0, st, a, b, n);
    return hipGetLastError();
}

hipError_t launch_cells_fix(int compat, const void *lw64, int L, const Geo &g, int nk, double dk, double df,
                            double omega_i, void *A0, void *A1, long lda, const unsigned *flag_count,
                            const unsigned *flag_list, hipStream_t st, unsigned *row_peak)
{
    dim3 grid(1024);
#define GRTM_FIX(CP, BR)                                                                                           \
    hipLaunchKernelGGL((k_cells_fix<float, CP, BR>), grid, dim3(kCellThreads), 0, st, (const LayerW<double> *)lw64, \
                       L, g, nk, dk, df, omega_i, (float *)A0, (float *)A1, lda, flag_count, flag_list, row_peak)
    const int br = g.branch >= 1 && g.branch <= 4 ? g.branch : 0;
    if (compat) {
        if (br == 1) GRTM_FIX(true, 1);
        else if (br == 2) GRTM_FIX(true, 2);
        else if (br == 3) GRTM_FIX(true, 3);
        else GRTM_FIX(true, 0);  // compat refuses branch 4
    } else {
        if (br == 1) GRTM_FIX(false, 1);
        else if (br == 2) GRTM_FIX(false, 2);
        else if (br == 3) GRTM_FIX(false, 3);
        else if (br == 4) GRTM_FIX(false, 4);
        else GRTM_FIX(false, 0);
    }
#undef GRTM_FIX
    return hipGetLastError();
}

hipError_t launch_plant_nonfinite(const unsigned *cells, int n, int nk, void *A0, void *A1, long lda,
                                  unsigned *flag_count, unsigned *flag_list, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_plant_nonfinite, dim3((n + 255) / 256), dim3(256), 0, st, cells, n, nk, (float *)A0,
                       (float *)A1, lda, flag_count, flag_list);
    return hipGetLastError();
}

hipError_t launch_flag_cells(const void *lw64, int L, int r64, int nw, int nk, double dk, int hb, const float *poles,
                             float pole_widths, const int *nk_w, unsigned *flag_count, unsigned *flag_list,
                             hipStream_t st)
{
    if (nw <= r64 || L > kMaxLayers) return nw <= r64 ? hipSuccess : hipErrorInvalidValue;
    hipLaunchKernelGGL(k_flag_cells, dim3(nw - r64), dim3(256), 0, st, (const LayerW<double> *)lw64, L, r64, nk, dk,
                       hb, poles, pole_widths, nk_w, flag_count, flag_list);
    return hipGetLastError();
}

hipError_t launch_zero_u32(unsigned *p, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_green_finish(int nw, int nblk, const double *partial, double *raw, hipStream_t st)
{
    const int n = nw * 60;
    hipLaunchKernelGGL(k_green_finish, dim3((n + 255) / 256), dim3(256), 0, st, nw, nblk, partial, raw);
    return hipGetLastError();
}

hipError_t launch_bessel_tab(int prec, int mode, int nk, double dk, const double *dists, int nr, void *B0,
                             void *B1, long ldb, hipStream_t st, int n0, int k0)
{
    const size_t n = (size_t)nk * nr;
    dim3 grid((unsigned)((n + 255) / 256));
    if (prec == 0)
        hipLaunchKernelGGL(k_bessel_tab<double>, grid, dim3(256), 0, st, mode, n0, nk, k0, dk, dists, nr,
                           (double *)B0, (double *)B1, ldb);
    else
        hipLaunchKernelGGL(k_bessel_tab<float>, grid, dim3(256), 0, st, mode, n0, nk, k0, dk, dists, nr,
                           (float *)B0, (float *)B1, ldb);
    return hipGetLastError();
}

// One thread per (row, source q, component, receiver); receivers fastest, so
// a wave reads one tail row (broadcast) and consecutive Bt entries.  The tail
// sums are formed in FP64 relative to the contraction's sum (their extrema
// are those of the partial sums); the records' binomial weights C(n-1, i) /
// 2^(n-1) are the n - 1 rounds of pairwise averaging.
template <typename TO>
__global__ __launch_bounds__(256) void k_ptam(int nw, int nr, int nt, int n, double dk,
                                              const double *__restrict__ dists, const TO *__restrict__ At0,
                                              const TO *__restrict__ At1, const TO *__restrict__ Bt0,
                                              const TO *__restrict__ Bt1, TO *__restrict__ C0, TO *__restrict__ C1,
                                              unsigned *__restrict__ misses)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nw * 4 * nr) return;
    const int ir = int(t % nr);
    const int sel = int((t / nr) % 4), iw = int(t / ((size_t)nr * 4));
    const int comp = sel >> 1, q = sel & 1;  // comp 0: Z (J0), 1: R (J1)
    const TO *At = comp ? At1 : At0, *Bt = comp ? Bt1 : Bt0;
    TO *C = comp ? C1 : C0;
    const size_t row = (size_t)4 * iw + 2 * q;  // real part; the imaginary part follows
    const TO *are = At + row * nt, *aim = are + nt;
    const double want = ceil((0.5 * n + 2.0) * 2.0 * M_PI / (dists[ir] * dk));
    const int ntr = want < nt ? int(want) : nt;
    // T: the tail sum before term j; an extremum of the partial sums sits at
    // j - 1 when terms j - 1 and j have opposite signs
    double T[2] = {0.0, 0.0}, est[2] = {0.0, 0.0}, cw[2] = {1.0, 1.0}, prev[2] = {0.0, 0.0};
    int m[2] = {0, 0};
    for (int j = 0; j < ntr; j++) {
        const double b = double(Bt[(size_t)j * nr + ir]);
        const double cur[2] = {double(are[j]) * b, double(aim[j]) * b};
        for (int p = 0; p < 2; p++) {
            const double a = prev[p], c = cur[p];
            if (j > 0 && m[p] < n && ((a < 0.0 && c > 0.0) || (a > 0.0 && c < 0.0))) {
                est[p] += cw[p] * (T[p] - (a + c) * (a + c) / (8.0 * (c - a)));
                cw[p] = cw[p] * double(n - 1 - m[p]) / double(m[p] + 1);
                m[p]++;
            }
            T[p] += c;
            prev[p] = c;
        }
    }
    const double norm = ldexp(1.0, -(n - 1));
    const size_t ci = row * nr + ir;
    const double s0[2] = {double(C[ci]), double(C[ci + nr])};
    bool miss = false;
    for (int p = 0; p < 2; p++) {
        const double v = m[p] == n ? s0[p] + est[p] * norm : s0[p] + T[p];
        miss = miss || m[p] < n;
        C[ci + (size_t)p * nr] = TO(v);
    }
    if (miss && misses) atomicAdd(misses, 1u);
}

hipError_t launch_ptam(int prec, int nw, int nr, int nt, int n, double dk, const double *dists, const void *At0,
                       const void *At1, const void *Bt0, const void *Bt1, void *C0, void *C1, unsigned *misses,
                       hipStream_t st)
{
    const size_t nth = (size_t)nw * 4 * nr;
    dim3 grid((unsigned)((nth + 255) / 256));
    if (prec == 0)
        hipLaunchKernelGGL(k_ptam<double>, grid, dim3(256), 0, st, nw, nr, nt, n, dk, dists, (const double *)At0,
                           (const double *)At1, (const double *)Bt0, (const double *)Bt1, (double *)C0,
                           (double *)C1, misses);
    else
        hipLaunchKernelGGL(k_ptam<float>, grid, dim3(256), 0, st, nw, nr, nt, n, dk, dists, (const float *)At0,
                           (const float *)At1, (const float *)Bt0, (const float *)Bt1, (float *)C0, (float *)C1,
                           misses);
    return hipGetLastError();
}

hipError_t launch_gemm(int prec, int M, int N, int K, const void *A, long lda, const void *B, long ldb, void *Cm,
                       long ldc, hipStream_t st)
{
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    if (prec == 0)
        hipLaunchKernelGGL(k_gemm<double>, grid, dim3(256), 0, st, M, N, K, (const double *)A, lda,
                           (const double *)B, ldb, (double *)Cm, ldc);
    else
        hipLaunchKernelGGL(k_gemm<float>, grid, dim3(256), 0, st, M, N, K, (const float *)A, lda,
                           (const float *)B, ldb, (float *)Cm, ldc);
    return hipGetLastError();
}

// LDS of an IFFT workgroup of R traces: R staged spectra and one transform
// buffer, nw complex values each.
size_t post_ifft_lds_bytes(int prec, int nw, int R)
{
    return (size_t)(R + 1) * nw * (prec == 0 ? 16 : 8);
}


// Traces per IFFT workgroup: 2 adjacent receivers when nr is even and their
// LDS (with `extra` bytes of static LDS) fits a workgroup (160 KiB), else 1.
// Measured on the 256-receiver workloads: 2 per workgroup beat 1 (column
// traces 0.18 -> 0.15 ms) and 4 (0.23 ms: one workgroup per compute unit,
// four transforms in turn).
static int ifft_group(int prec, int nw, int nr, size_t extra)
{
    return (nr > 0 && nr % 2 == 0 && post_ifft_lds_bytes(prec, nw, 2) + extra <= 163840) ? 2 : 1;
}

template <typename K>
static hipError_t ifft_lds_attr(K kernel, size_t lds)
{
    if (lds <= 65536) return hipSuccess;  // within the default dynamic-LDS limit
    return hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_post_ifft_cols(int prec, const PostParams &pp, const MtCols *mc, int ncol, long ldc, int nr,
                                 const double *dists, void *traces, hipStream_t st)
{
    if (ncol < 1 || ncol > kMtColsMax) return hipErrorInvalidValue;
    const int ntr = ncol * nr;
    const int R = ifft_group(prec, pp.nw, nr, sizeof(double) * kMtTermsMax * 2);
    const size_t lds = post_ifft_lds_bytes(prec, pp.nw, R);
    const dim3 grid(xcd_grid(ntr / R));
#define GRTM_COLS(T, RR)                                                                                     \
    do {                                                                                                     \
        hipError_t e = ifft_lds_attr(k_post_ifft_cols<T, RR>, lds);                                         \
        if (e != hipSuccess) return e;                                                                       \
        hipLaunchKernelGGL((k_post_ifft_cols<T, RR>), grid, dim3(kIfftThreads), lds, st, pp, mc, ldc, nr, ntr, \
                           dists, (T *)traces);                                                              \
    } while (0)
    if (prec == 0) {
        if (R == 2) GRTM_COLS(double, 2);
        else GRTM_COLS(double, 1);
    } else {
        if (R == 2) GRTM_COLS(float, 2);
        else GRTM_COLS(float, 1);
    }
#undef GRTM_COLS
    return hipGetLastError();
}

hipError_t launch_post_ifft(int prec, const PostParams &pp, int layout, int ntr, const void *raw, const void *C0,
                            const void *C1, long ldc, int nr, void *traces, void *spec_post, hipStream_t st,
                            long cms)
{
    const int R = layout == 1 ? ifft_group(prec, pp.nw, nr, 0) : 1;
    const size_t lds = post_ifft_lds_bytes(prec, pp.nw, R);
    const dim3 grid(xcd_grid(ntr / R));
#define GRTM_POST(T, RR)                                                                                          \
    do {                                                                                                          \
        hipError_t e = ifft_lds_attr(k_post_ifft<T, RR>, lds);                                                   \
        if (e != hipSuccess) return e;                                                                            \
        hipLaunchKernelGGL((k_post_ifft<T, RR>), grid, dim3(kIfftThreads), lds, st, pp, layout, (const T *)raw,  \
                           (const T *)C0, (const T *)C1, ldc, nr, cms, ntr, (T *)traces, (T *)spec_post);         \
    } while (0)
    if (prec == 0) {
        if (R == 2) GRTM_POST(double, 2);
        else GRTM_POST(double, 1);
    } else {
        if (R == 2) GRTM_POST(float, 2);
        else GRTM_POST(float, 1);
    }
#undef GRTM_POST
    return hipGetLastError();
}

}  // namespace grtm
