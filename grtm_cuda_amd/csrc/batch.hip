// batch.hip -- recursion kernels of a batch of models (many models on one
// grid, groups of models per launch).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cellops.h"

namespace grtm {

// The cell's eight operand values into the group row's four component rows
// of A0 / A1 (RowLayout); a resident row stores only its first fext wavenumbers.
template <typename TO>
__device__ __forceinline__ void store_zr_rl(const TO v[8], int grow, int ik, TO *A0, TO *A1, const RowLayout &rl)
{
    size_t r0, st;
    if (rl.foff) {
        const int ext = rl.fext[grow];
        if (ik >= ext) return;
        r0 = size_t(rl.foff[grow]) + size_t(ik), st = size_t(ext);
    } else {
        r0 = size_t(4 * grow) * size_t(rl.lda) + size_t(ik), st = size_t(rl.lda);
    }
    for (int i = 0; i < 4; i++) {
        A0[r0 + i * st] = v[i];
        A1[r0 + i * st] = v[4 + i];
    }
}

__device__ __forceinline__ void flag_cell(const FlagList &fl, int grow, int nk, int ik)
{
    const unsigned slot = atomicAdd(fl.count, 1u);
    if (slot < fl.cap) fl.list[slot] = (unsigned long long)grow * (unsigned long long)nk + (unsigned long long)ik;
}

// The recursion of one cell with the receiver branch fixed at compile time
// (BR 1-4), or for BR = 0 the model's own branch, each branch's solve
// compiled separately (so the kernel's registers are the largest branch's,
// not the union a run-time branch would carry).
template <typename T, typename TO, bool COMPAT, int BR>
__device__ __forceinline__ bool cell_zr_br(const LayerW<T> *lw, const Geo &g, int iw, int ik, double dk, double df,
                                           double omega_i, TO v[8])
{
    if constexpr (BR != 0) {
        return cell_zr<T, TO, COMPAT, BR>(lw, g, iw, ik, dk, df, omega_i, v);
    } else {
        switch (g.branch) {
        case 1: return cell_zr<T, TO, COMPAT, 1>(lw, g, iw, ik, dk, df, omega_i, v);
        case 2: return cell_zr<T, TO, COMPAT, 2>(lw, g, iw, ik, dk, df, omega_i, v);
        case 3: return cell_zr<T, TO, COMPAT, 3>(lw, g, iw, ik, dk, df, omega_i, v);
        default: return cell_zr<T, TO, COMPAT, 4>(lw, g, iw, ik, dk, df, omega_i, v);
        }
    }
}

// ---------------------------------------------------------------- batches
// Recursion of a group of models in one launch: one block of BS threads per
// work item (model j of the group, row, k tile of BS wavenumbers), items
// ordered heaviest model first so the launch ends on light items.  Model
// m0 + j has its geometry and the offset of its layer table in md; its A
// operand rows are group rows j nw + iw (RowLayout).  Non-finite FP32 cells
// go to the flag list as (j nw + iw) nk + ik.  Adaptive grids: nk_w[group row] limits
// the wavenumbers solved and row_peak[group row] collects the row maxima; an
// item whose row / tile word carries kCheckItem solves the row's residual-check
// cells BS c .. BS c + BS - 1 (wavenumbers nk_w + stride j + stride / 2, j <
// nchk[group row]) and takes their largest |value| to tail_max[group row].
// BR: the receiver branch shared by every model of the launch's items, or 0
// (each item's model's own).
// dev_count (persistent form, learned grids formed on the device): the grid
// is the resident blocks, each taking items [dev_count[which],
// dev_count[which + 1]) `chunk` at a time from the counter *ctr (one counter
// serialises its atomics: small tiles take several per visit).
template <typename T, typename TO, bool COMPAT, int BS, int BR>
__global__ __launch_bounds__(BS, sizeof(T) == 4 ? 4 : 1) void k_cells_zr_batch(const ModelDesc *__restrict__ md, int m0,
                                                       const LayerW<T> *__restrict__ lw_all, int nw, int nk,
                                                       double dk, double df, double omega_i,
                                                       const unsigned long long *__restrict__ items,
                                                       TO *__restrict__ A0, TO *__restrict__ A1, RowLayout rl,
                                                       FlagList fl, const int *__restrict__ nk_w,
                                                       unsigned *__restrict__ row_peak,
                                                       const int *__restrict__ nchk, int stride,
                                                       unsigned *__restrict__ tail_max,
                                                       const unsigned *__restrict__ dev_count, int which,
                                                       unsigned *__restrict__ ctr, int chunk)
{
    __shared__ LayerW<T> lw[kMaxLayers];
    __shared__ unsigned item_sh;
    const unsigned lo_i = dev_count ? dev_count[which] : 0u, hi_i = dev_count ? dev_count[which + 1] : 0u;
    unsigned base = 0;
    for (int pass = 0;; pass++) {
        unsigned idx;
        if (dev_count) {
            const int c = pass % chunk;
            if (c == 0) {
                __syncthreads();  // the previous item's reads of lw and item_sh are done
                if (threadIdx.x == 0) item_sh = lo_i + atomicAdd(ctr, unsigned(chunk));
                __syncthreads();
                base = item_sh;
            }
            idx = base + unsigned(c);
            if (idx >= hi_i) return;  // uniform over the block: the list is drained
            if (c) __syncthreads();   // the previous item's reads of lw are done
        } else {
            if (pass) return;
            idx = blockIdx.x;
        }
        const unsigned long long e = items[idx];
        const int j = int(e >> 32);
        const unsigned lo = unsigned(e & 0xffffffffull);
        const bool chk = (lo & kCheckItem) != 0;
        const int iw = int((lo & ~kCheckItem) >> kTileBits), kb = int(lo & ((1u << kTileBits) - 1));
        const int grow = j * nw + iw;  // row of the group
        const int nkr = nk_w ? nk_w[grow] : nk;
        const int lim = chk ? nchk[grow] : nkr;  // cells (check cells) of the row
        if (kb * BS >= lim) continue;  // uniform over the block
        const Geo g = md[m0 + j].g;
        const size_t off = md[m0 + j].lw_off + (size_t)iw * g.L;
        for (int q = threadIdx.x; q < g.L; q += BS) lw[q] = lw_all[off + q];
        __syncthreads();
        const int jj = kb * BS + threadIdx.x;
        const int ik = chk ? nkr + stride * jj + stride / 2 : jj;
        unsigned mag = 0;
        if (jj < lim && ik < nk) {
            TO v[8];
            const bool ok = cell_zr_br<T, TO, COMPAT, BR>(lw, g, iw, ik, dk, df, omega_i, v);
            if (chk) {
                float m = 0.f;
                for (int i = 0; i < 8; i++) m = fmaxf(m, fabsf(float(v[i])));
                mag = __float_as_uint(isfinite(m) ? m : 3.0e38f);  // a non-finite check value fails the row
            } else {
                store_zr_rl(v, grow, ik, A0, A1, rl);
                mag = cell_mag_bits(v);
                if (!ok && fl.count) flag_cell(fl, grow, nk, ik);
            }
        }
        unsigned *dst = chk ? tail_max : row_peak;
        if (dst) {
            mag = wave_max_u32(mag);
            if ((threadIdx.x & 63) == 0 && mag) atomicMax(dst + grow, mag);
        }
    }
}

// Items of a group's learned grids formed on the device.  The group's item
// order is fixed by the host as a list of slots, each a group row with a
// phase (kCheckItem: its residual-check items, else its tiles of BS
// wavenumbers).  Every slot's first item is an exclusive scan of the slots'
// item counts, over blocks of 1024 slots in three launches: k_batch_count
// (each block's total), k_batch_scan_sums (one block: the totals' exclusive
// scan and the grand total), k_batch_offsets (each block's own scan plus its
// offset; seg_items[s], the first item of segment s whose slots start at
// seg_slots[s], nseg + 1 bounds).  k_batch_fill (one thread per slot) then
// writes the items (j << 32 | row << kTileBits | tile, with kCheckItem on
// check items).  A single block walking all the slots took 8.7 ms for the
// 2.1 M slots of a 1024-model batch.
__device__ __forceinline__ unsigned slot_items(int q, const unsigned *__restrict__ slots, int bs,
                                               const int *__restrict__ nk_pred, const int *__restrict__ nchk)
{
    const unsigned sl = slots[q];
    const int grow = int(sl & ~kCheckItem);
    const int n = (sl & kCheckItem) ? nchk[grow] : nk_pred[grow];
    return unsigned((n + bs - 1) / bs);
}

// Inclusive scan over a block of 1024 threads (16 waves): shuffles within
// each wave, then the wave totals (wsum: 16 words of LDS).  Returns the
// thread's inclusive prefix; *total gets the block's sum.
__device__ __forceinline__ unsigned block_scan_1024(unsigned v, unsigned *wsum, unsigned *total)
{
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned a = unsigned(__shfl_up(int(v), o, 64));
        if (lane >= o) v += a;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    unsigned run = 0;
    for (int w = 0; w < 16; w++) {
        if (w == wv) v += run;  // the earlier waves' totals
        run += wsum[w];
    }
    *total = run;
    return v;
}

__global__ __launch_bounds__(1024) void k_batch_count(int nslots, const unsigned *__restrict__ slots, int bs,
                                                      const int *__restrict__ nk_pred, const int *__restrict__ nchk,
                                                      unsigned *__restrict__ bsum)
{
    __shared__ unsigned wsum[16];
    const int q = blockIdx.x * 1024 + threadIdx.x;
    const unsigned c = q < nslots ? slot_items(q, slots, bs, nk_pred, nchk) : 0u;
    unsigned total;
    block_scan_1024(c, wsum, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// bsum[0 .. nb): block totals -> their exclusive prefixes; bsum[nb] = the sum.
__global__ __launch_bounds__(1024) void k_batch_scan_sums(int nb, unsigned *__restrict__ bsum)
{
    __shared__ unsigned wsum[16];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024, s0 = min(nb, t * per), s1 = min(nb, s0 + per);
    unsigned sum = 0;
    for (int q = s0; q < s1; q++) sum += bsum[q];
    unsigned total;
    unsigned run = block_scan_1024(sum, wsum, &total) - sum;
    __syncthreads();  // every thread has read its totals before any is overwritten
    for (int q = s0; q < s1; q++) {
        const unsigned v = bsum[q];
        bsum[q] = run;
        run += v;
    }
    if (t == 0) bsum[nb] = total;
}

__global__ __launch_bounds__(1024) void k_batch_offsets(int nslots, const unsigned *__restrict__ slots, int bs,
                                                        const int *__restrict__ nk_pred, const int *__restrict__ nchk,
                                                        const unsigned *__restrict__ bsum, int nb,
                                                        unsigned *__restrict__ slot_off, int nseg,
                                                        const unsigned *__restrict__ seg_slots,
                                                        unsigned *__restrict__ seg_items)
{
    __shared__ unsigned wsum[16], off[1024];
    const int q0 = blockIdx.x * 1024, q = q0 + threadIdx.x;
    const unsigned c = q < nslots ? slot_items(q, slots, bs, nk_pred, nchk) : 0u;
    unsigned total;
    const unsigned o = bsum[blockIdx.x] + block_scan_1024(c, wsum, &total) - c;
    off[threadIdx.x] = o;
    if (q < nslots) slot_off[q] = o;
    __syncthreads();
    // segment bounds: the first item of the bound's slot, or the grand total
    // for a bound at nslots (taken by block 0)
    for (int sg = threadIdx.x; sg <= nseg; sg += 1024) {
        const unsigned b = seg_slots[sg];
        if (b >= unsigned(nslots)) {
            if (blockIdx.x == 0) seg_items[sg] = bsum[nb];
        } else if (b >= unsigned(q0) && b < unsigned(q0) + 1024u) {
            seg_items[sg] = off[b - unsigned(q0)];
        }
    }
}

__global__ void k_batch_fill(int nslots, const unsigned *__restrict__ slots, int nw, int bs,
                             const int *__restrict__ nk_pred, const int *__restrict__ nchk,
                             const unsigned *__restrict__ slot_off, unsigned long long *__restrict__ items)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nslots) return;
    const unsigned sl = slots[q];
    const bool chk = (sl & kCheckItem) != 0;
    const int grow = int(sl & ~kCheckItem), j = grow / nw, iw = grow - j * nw;
    const int n = chk ? nchk[grow] : nk_pred[grow];
    const unsigned nt = unsigned((n + bs - 1) / bs), base = slot_off[q];
    const unsigned long long hi = (unsigned long long)j << 32;
    for (unsigned k = 0; k < nt; k++)
        items[base + k] = hi | ((unsigned(iw) << kTileBits) | k) | (chk ? kCheckItem : 0u);
}

// Rows of a group whose checked tail exceeds tol x row peak fall back to the
// dense grid: appended to the group's list (group row, then the first
// wavenumber still to solve at nrow + slot) and counted per model.
__global__ void k_fallback_select_batch(int nrow, int nw, int nk, double tol, const unsigned *__restrict__ row_peak,
                                        const unsigned *__restrict__ tail_max, int *__restrict__ nk_w,
                                        unsigned *__restrict__ count, unsigned *__restrict__ list,
                                        unsigned *__restrict__ model_count)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    const float pk = __uint_as_float(row_peak[r]), tm = __uint_as_float(tail_max[r]);
    if (nk_w[r] < nk && !(tm <= float(tol) * pk)) {
        const unsigned slot = atomicAdd(count, 1u);
        list[slot] = unsigned(r);
        list[nrow + slot] = unsigned(nk_w[r]);
        nk_w[r] = nk;
        atomicAdd(model_count + r / nw, 1u);
    }
}

// Dense completion of a group's fallback rows: wavenumbers [start, nk) of each
// listed row in the arithmetic T of that row (FP64 below its model's r64:
// F64ROWS true; FP32 above: F64ROWS false).  Grid (k blocks, row slots);
// slots stride over the list.  FP32 cells that are not finite go to the flag
// list as group row x nk + ik.
template <typename T, typename TO, bool COMPAT, bool F64ROWS>
__global__ __launch_bounds__(kCellThreads) void k_fallback_batch(const ModelDesc *__restrict__ md, int m0,
                                                                 const LayerW<T> *__restrict__ lw_all, int nw, int nk,
                                                                 double dk, double df, double omega_i,
                                                                 TO *__restrict__ A0, TO *__restrict__ A1,
                                                                 RowLayout rl, const unsigned *__restrict__ count,
                                                                 const unsigned *__restrict__ list, int nrow,
                                                                 FlagList fl, unsigned *__restrict__ row_peak)
{
    __shared__ LayerW<T> lw[kMaxLayers];
    const unsigned n = *count;
    for (unsigned slot = blockIdx.y; slot < n; slot += gridDim.y) {
        const int grow = int(list[slot]);
        const int j = grow / nw, iw = grow - j * nw;
        const ModelDesc d = md[m0 + j];
        if ((iw < d.r64) != F64ROWS) continue;  // uniform over the block
        const int start = int(list[nrow + slot]);
        __syncthreads();
        for (int q = threadIdx.x; q < d.g.L; q += blockDim.x) lw[q] = lw_all[d.lw_off + (size_t)iw * d.g.L + q];
        __syncthreads();
        for (int ik = start + blockIdx.x * blockDim.x + threadIdx.x; ik < nk; ik += gridDim.x * blockDim.x) {
            TO v[8];
            const bool ok = cell_zr<T, TO, COMPAT>(lw, d.g, iw, ik, dk, df, omega_i, v);
            store_zr_rl(v, grow, ik, A0, A1, rl);
            if (!ok && fl.count) flag_cell(fl, grow, nk, ik);
            // the row peak covers the completed tail (it sets the FP16 split's row scale)
            if (row_peak) {
                const unsigned m = cell_mag_bits(v);
                if (m) atomicMax(row_peak + grow, m);
            }
        }
    }
}

// FP64 re-solve of the flagged cells of a group.
template <typename TO, bool COMPAT>
__global__ __launch_bounds__(kCellThreads) void k_cells_fix_batch(const ModelDesc *__restrict__ md, int m0,
                                                                  const LayerW<double> *__restrict__ lw_all,
                                                                  int nw, int nk, double dk, double df,
                                                                  double omega_i, TO *__restrict__ A0,
                                                                  TO *__restrict__ A1, RowLayout rl, FlagList fl,
                                                                  unsigned *__restrict__ row_peak)
{
    const unsigned n = min(*fl.count, fl.cap);
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const unsigned long long cell = fl.list[t];
        const int grow = int(cell / (unsigned long long)nk), ik = int(cell % (unsigned long long)nk);
        const int j = grow / nw, iw = grow - j * nw;
        const Geo g = md[m0 + j].g;
        TO v[8];
        cell_zr<double, TO, COMPAT>(lw_all + md[m0 + j].lw_off + (size_t)iw * g.L, g, iw, ik, dk, df, omega_i, v);
        store_zr_rl(v, grow, ik, A0, A1, rl);
        if (row_peak) {
            const unsigned m = cell_mag_bits(v);
            if (m) atomicMax(row_peak + grow, m);
        }
    }
}

// Test hook: the listed cells (group row nk + k) get non-finite operands and
// join the flag list.
__global__ void k_plant_nonfinite_batch(const unsigned long long *__restrict__ cells, int n, int nk,
                                        float *__restrict__ A0, float *__restrict__ A1, RowLayout rl, FlagList fl)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const unsigned long long cell = cells[t];
    const int grow = int(cell / (unsigned long long)nk), ik = int(cell % (unsigned long long)nk);
    float v[8];
    for (int i = 0; i < 8; i++) v[i] = __builtin_nanf("");
    store_zr_rl(v, grow, ik, A0, A1, rl);
    flag_cell(fl, grow, nk, ik);
}

// A resident row's final extent cut to its stored length (counted in *over).
__global__ void k_clamp_extent(int nrow, int *__restrict__ nk_w, const int *__restrict__ fext,
                               unsigned *__restrict__ over)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    if (nk_w[r] > fext[r]) {
        nk_w[r] = fext[r];
        atomicAdd(over, 1u);
    }
}

// ---------------------------------------------------------------- launchers
// Blocks of a BS-thread kernel resident on the whole device, capped at ub.
static int resident_blocks_bs(const void *kernel, int bs, int ub)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, bs, 0) != hipSuccess || per < 1)
        per = 1, cus = cus > 0 ? cus : 256;
    return std::max(1, std::min(ub, per * cus));
}

size_t batch_items_words(int nslots) { return (size_t)nslots + (size_t)(nslots + 1023) / 1024 + 1; }

hipError_t launch_batch_items(int nslots, const unsigned *slots, int nw, int bs, const int *nk_pred, const int *nchk,
                              unsigned *slot_off, int nseg, const unsigned *seg_slots, unsigned *seg_items,
                              unsigned long long *items, hipStream_t st)
{
    if (nslots < 1 || nseg < 1 || nseg > 1023) return hipErrorInvalidValue;
    const int nb = (nslots + 1023) / 1024;
    unsigned *bsum = slot_off + nslots;  // nb + 1 words after the slots' offsets
    hipLaunchKernelGGL(k_batch_count, dim3(nb), dim3(1024), 0, st, nslots, slots, bs, nk_pred, nchk, bsum);
    hipLaunchKernelGGL(k_batch_scan_sums, dim3(1), dim3(1024), 0, st, nb, bsum);
    hipLaunchKernelGGL(k_batch_offsets, dim3(nb), dim3(1024), 0, st, nslots, slots, bs, nk_pred, nchk, bsum, nb,
                       slot_off, nseg, seg_slots, seg_items);
    hipLaunchKernelGGL(k_batch_fill, dim3((nslots + 255) / 256), dim3(256), 0, st, nslots, slots, nw, bs, nk_pred, nchk,
                       slot_off, items);
    return hipGetLastError();
}

hipError_t launch_cells_zr_batch(int prec, int compat, int bs, int br, const ModelDesc *md, int m0, const void *lw_all,
                                 int nw, int nk, double dk, double df, double omega_i,
                                 const unsigned long long *items, int nitems, void *A0, void *A1, RowLayout rl,
                                 FlagList fl, const int *nk_w, unsigned *row_peak, hipStream_t st, const int *nchk,
                                 int stride, unsigned *tail_max, const unsigned *dev_count, int which, unsigned *ctr,
                                 int chunk)
{
    if (nitems <= 0) return hipSuccess;
    if (dev_count) {
        if (!ctr || which < 0 || chunk < 1) return hipErrorInvalidValue;
        hipError_t e = launch_zero_u32(ctr, 1, st);
        if (e != hipSuccess) return e;
    }
#define GRTM_ZB(T, TO, CP, BSZ, BR, FC, FL)                                                                      \
    hipLaunchKernelGGL((k_cells_zr_batch<T, TO, CP, BSZ, BR>),                                                    \
                       dim3(dev_count ? resident_blocks_bs((const void *)k_cells_zr_batch<T, TO, CP, BSZ, BR>, BSZ, \
                                                           nitems)                                                 \
                                      : nitems),                                                                 \
                       dim3(BSZ), 0, st, md, m0, (const LayerW<T> *)lw_all, nw, nk, dk, df, omega_i, items, (TO *)A0, \
                       (TO *)A1, rl, FC, nk_w, row_peak, nchk, stride, tail_max, dev_count, which, ctr, chunk)
#define GRTM_ZB_BR(T, TO, CP, BSZ, FC, FL)                           \
    do {                                                             \
        if (br == 1) GRTM_ZB(T, TO, CP, BSZ, 1, FC, FL);             \
        else if (br == 2) GRTM_ZB(T, TO, CP, BSZ, 2, FC, FL);        \
        else if (br == 3) GRTM_ZB(T, TO, CP, BSZ, 3, FC, FL);        \
        else GRTM_ZB(T, TO, CP, BSZ, 4, FC, FL);                     \
    } while (0)
// br = 0 (every branch in one kernel) for the FP32-operand launches only
#define GRTM_ZB_BR0(T, TO, CP, BSZ, FC, FL)                          \
    do {                                                             \
        if (br == 0) GRTM_ZB(T, TO, CP, BSZ, 0, FC, FL);             \
        else GRTM_ZB_BR(T, TO, CP, BSZ, FC, FL);                     \
    } while (0)
#define GRTM_ZB_BS(T, TO, CP, FC, FL, BRM)                           \
    do {                                                             \
        if (bs == 64) BRM(T, TO, CP, 64, FC, FL);                    \
        else if (bs == 128) BRM(T, TO, CP, 128, FC, FL);             \
        else BRM(T, TO, CP, 256, FC, FL);                            \
    } while (0)
    if (bs != 64 && bs != 128 && bs != 256) return hipErrorInvalidValue;
    if (br < 0 || br > 4 || (br == 0 && prec == 0)) return hipErrorInvalidValue;
    const FlagList none{nullptr, nullptr, 0u};
    if (prec == 0) {
        if (compat) GRTM_ZB_BS(double, double, true, none, none, GRTM_ZB_BR);
        else GRTM_ZB_BS(double, double, false, none, none, GRTM_ZB_BR);
    } else if (prec == 1) {
        if (compat) GRTM_ZB_BS(float, float, true, fl, fl, GRTM_ZB_BR0);
        else GRTM_ZB_BS(float, float, false, fl, fl, GRTM_ZB_BR0);
    } else {
        if (compat) GRTM_ZB_BS(double, float, true, none, none, GRTM_ZB_BR0);
        else GRTM_ZB_BS(double, float, false, none, none, GRTM_ZB_BR0);
    }
#undef GRTM_ZB_BS
#undef GRTM_ZB_BR0
#undef GRTM_ZB_BR
#undef GRTM_ZB
    return hipGetLastError();
}

hipError_t launch_fallback_batch(int prec, int compat, const ModelDesc *md, int m0, const void *lw64, const void *lw32,
                                 int nrow, int nw, int nk, double dk, double df, double omega_i, double tol,
                                 unsigned *row_peak, const unsigned *tail_max, int *nk_w, unsigned *count,
                                 unsigned *list, unsigned *model_count, void *A0, void *A1, RowLayout rl,
                                 FlagList fl, hipStream_t st, bool raise_peaks)
{
    hipLaunchKernelGGL(k_fallback_select_batch, dim3((nrow + 255) / 256), dim3(256), 0, st, nrow, nw, nk, tol, row_peak,
                       tail_max, nk_w, count, list, model_count);
    const dim3 grid((nk + kCellThreads - 1) / kCellThreads, 16);
#define GRTM_FBB(T, TO, CP, F64, LWP, FLL)                                                                       \
    hipLaunchKernelGGL((k_fallback_batch<T, TO, CP, F64>), grid, dim3(kCellThreads), 0, st, md, m0,              \
                       (const LayerW<T> *)LWP, nw, nk, dk, df, omega_i, (TO *)A0, (TO *)A1, rl, count, list, nrow,  \
                       FLL, raise_peaks ? row_peak : nullptr)
    const FlagList none{nullptr, nullptr, 0u};
    if (prec == 0) {
        if (compat) GRTM_FBB(double, double, true, true, lw64, none);
        else GRTM_FBB(double, double, false, true, lw64, none);
    } else if (compat) {
        GRTM_FBB(double, float, true, true, lw64, none);
        GRTM_FBB(float, float, true, false, lw32, fl);
    } else {
        GRTM_FBB(double, float, false, true, lw64, none);
        GRTM_FBB(float, float, false, false, lw32, fl);
    }
#undef GRTM_FBB
    return hipGetLastError();
}

hipError_t launch_cells_fix_batch(int compat, const ModelDesc *md, int m0, const void *lw64, int nw, int nk, double dk,
                                  double df, double omega_i, void *A0, void *A1, RowLayout rl, FlagList fl,
                                  hipStream_t st, unsigned *row_peak)
{
    dim3 grid(1024);
    if (compat)
        hipLaunchKernelGGL((k_cells_fix_batch<float, true>), grid, dim3(kCellThreads), 0, st, md, m0,
                           (const LayerW<double> *)lw64, nw, nk, dk, df, omega_i, (float *)A0, (float *)A1, rl, fl,
                           row_peak);
    else
        hipLaunchKernelGGL((k_cells_fix_batch<float, false>), grid, dim3(kCellThreads), 0, st, md, m0,
                           (const LayerW<double> *)lw64, nw, nk, dk, df, omega_i, (float *)A0, (float *)A1, rl, fl,
                           row_peak);
    return hipGetLastError();
}

hipError_t launch_plant_nonfinite_batch(const unsigned long long *cells, int n, int nk, void *A0, void *A1,
                                        RowLayout rl, FlagList fl, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_plant_nonfinite_batch, dim3((n + 255) / 256), dim3(256), 0, st, cells, n, nk, (float *)A0,
                       (float *)A1, rl, fl);
    return hipGetLastError();
}

hipError_t launch_clamp_extent(int nrow, int *nk_w, const int *fext, unsigned *over, hipStream_t st)
{
    hipLaunchKernelGGL(k_clamp_extent, dim3((nrow + 255) / 256), dim3(256), 0, st, nrow, nk_w, fext, over);
    return hipGetLastError();
}

}  // namespace grtm
