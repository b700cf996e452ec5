// batch.hip -- recursion kernels of a batch of models (many models on one
// grid, groups of models per launch).
#include <hip/hip_runtime.h>

#include "cellops.h"

namespace grtm {

// ---------------------------------------------------------------- batches
// Recursion of a group of models in one launch: one block of BS threads per
// work item (model j of the group, row, k tile of BS wavenumbers), items
// ordered heaviest model first so the launch ends on light items.  Model
// m0 + j has its geometry and the offset of its layer table in md; its A
// operand rows are group rows j nw + iw.  Non-finite FP32 cells go to the
// flag list as (j nw + iw) nk + ik.  Adaptive grids: nk_w[group row] limits
// the wavenumbers solved and row_peak[group row] collects the row maxima.
// BR: the receiver branch shared by every model of the launch's items.
template <typename T, typename TO, bool COMPAT, int BS, int BR>
__global__ __launch_bounds__(BS, sizeof(T) == 4 ? 4 : 1) void k_cells_zr_batch(const ModelDesc *__restrict__ md, int m0,
                                                       const LayerW<T> *__restrict__ lw_all, int nw, int nk,
                                                       double dk, double df, double omega_i,
                                                       const unsigned long long *__restrict__ items,
                                                       TO *__restrict__ A0, TO *__restrict__ A1, long lda,
                                                       unsigned *__restrict__ flag_count,
                                                       unsigned *__restrict__ flag_list,
                                                       const int *__restrict__ nk_w,
                                                       unsigned *__restrict__ row_peak)
{
    __shared__ LayerW<T> lw[kMaxLayers];
    const unsigned long long e = items[blockIdx.x];
    const int j = int(e >> 32);
    const unsigned lo = unsigned(e & 0xffffffffull);
    const int iw = int(lo >> kTileBits), kb = int(lo & ((1u << kTileBits) - 1));
    const int grow = j * nw + iw;  // row of the group
    const int nkr = nk_w ? nk_w[grow] : nk;
    if (kb * BS >= nkr) return;  // uniform over the block
    const Geo g = md[m0 + j].g;
    const size_t off = md[m0 + j].lw_off + (size_t)iw * g.L;
    for (int q = threadIdx.x; q < g.L; q += BS) lw[q] = lw_all[off + q];
    __syncthreads();
    const int ik = kb * BS + threadIdx.x;
    unsigned mag = 0;
    if (ik < nkr) {
        TO v[8];
        const bool ok = cell_zr<T, TO, COMPAT, BR>(lw, g, iw, ik, dk, df, omega_i, v);
        store_zr(v, grow, ik, A0, A1, lda);
        mag = cell_mag_bits(v);
        if (!ok && flag_count) {
            const unsigned slot = atomicAdd(flag_count, 1u);
            flag_list[slot] = unsigned(grow) * unsigned(nk) + unsigned(ik);
        }
    }
    if (row_peak) {
        mag = wave_max_u32(mag);
        if ((threadIdx.x & 63) == 0 && mag) atomicMax(row_peak + grow, mag);
    }
}

// FP64 re-solve of the flagged cells of a group.
template <typename TO, bool COMPAT>
__global__ __launch_bounds__(kCellThreads) void k_cells_fix_batch(const ModelDesc *__restrict__ md, int m0,
                                                                  const LayerW<double> *__restrict__ lw_all,
                                                                  int nw, int nk, double dk, double df,
                                                                  double omega_i, TO *__restrict__ A0,
                                                                  TO *__restrict__ A1, long lda,
                                                                  const unsigned *__restrict__ flag_count,
                                                                  const unsigned *__restrict__ flag_list)
{
    const unsigned n = *flag_count;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const unsigned cell = flag_list[t];
        const int grow = int(cell / unsigned(nk)), ik = int(cell % unsigned(nk));
        const int j = grow / nw, iw = grow - j * nw;
        const Geo g = md[m0 + j].g;
        TO v[8];
        cell_zr<double, TO, COMPAT>(lw_all + md[m0 + j].lw_off + (size_t)iw * g.L, g, iw, ik, dk, df, omega_i, v);
        store_zr(v, grow, ik, A0, A1, lda);
    }
}

// ---------------------------------------------------------------- launchers
hipError_t launch_cells_zr_batch(int prec, int compat, int bs, int br, const ModelDesc *md, int m0, const void *lw_all,
                                 int nw, int nk, double dk, double df, double omega_i,
                                 const unsigned long long *items, int nitems, void *A0, void *A1, long lda,
                                 unsigned *flag_count, unsigned *flag_list, const int *nk_w, unsigned *row_peak,
                                 hipStream_t st)
{
    if (nitems <= 0) return hipSuccess;
#define GRTM_ZB(T, TO, CP, BSZ, BR, FC, FL)                                                                      \
    hipLaunchKernelGGL((k_cells_zr_batch<T, TO, CP, BSZ, BR>), dim3(nitems), dim3(BSZ), 0, st, md, m0,          \
                       (const LayerW<T> *)lw_all, nw, nk, dk, df, omega_i, items, (TO *)A0, (TO *)A1, lda, FC, FL, \
                       nk_w, row_peak)
#define GRTM_ZB_BR(T, TO, CP, BSZ, FC, FL)                           \
    do {                                                             \
        if (br == 1) GRTM_ZB(T, TO, CP, BSZ, 1, FC, FL);             \
        else if (br == 2) GRTM_ZB(T, TO, CP, BSZ, 2, FC, FL);        \
        else if (br == 3) GRTM_ZB(T, TO, CP, BSZ, 3, FC, FL);        \
        else GRTM_ZB(T, TO, CP, BSZ, 4, FC, FL);                     \
    } while (0)
#define GRTM_ZB_BS(T, TO, CP, FC, FL)                                \
    do {                                                             \
        if (bs == 64) GRTM_ZB_BR(T, TO, CP, 64, FC, FL);             \
        else if (bs == 128) GRTM_ZB_BR(T, TO, CP, 128, FC, FL);      \
        else GRTM_ZB_BR(T, TO, CP, 256, FC, FL);                     \
    } while (0)
    if (bs != 64 && bs != 128 && bs != 256) return hipErrorInvalidValue;
    if (br < 1 || br > 4) return hipErrorInvalidValue;
    if (prec == 0) {
        if (compat) GRTM_ZB_BS(double, double, true, nullptr, nullptr);
        else GRTM_ZB_BS(double, double, false, nullptr, nullptr);
    } else if (prec == 1) {
        if (compat) GRTM_ZB_BS(float, float, true, flag_count, flag_list);
        else GRTM_ZB_BS(float, float, false, flag_count, flag_list);
    } else {
        if (compat) GRTM_ZB_BS(double, float, true, nullptr, nullptr);
        else GRTM_ZB_BS(double, float, false, nullptr, nullptr);
    }
#undef GRTM_ZB_BS
#undef GRTM_ZB_BR
#undef GRTM_ZB
    return hipGetLastError();
}

hipError_t launch_cells_fix_batch(int compat, const ModelDesc *md, int m0, const void *lw64, int nw, int nk, double dk,
                                  double df, double omega_i, void *A0, void *A1, long lda, const unsigned *flag_count,
                                  const unsigned *flag_list, hipStream_t st)
{
    dim3 grid(1024);
    if (compat)
        hipLaunchKernelGGL((k_cells_fix_batch<float, true>), grid, dim3(kCellThreads), 0, st, md, m0,
                           (const LayerW<double> *)lw64, nw, nk, dk, df, omega_i, (float *)A0, (float *)A1, lda,
                           flag_count, flag_list);
    else
        hipLaunchKernelGGL((k_cells_fix_batch<float, false>), grid, dim3(kCellThreads), 0, st, md, m0,
                           (const LayerW<double> *)lw64, nw, nk, dk, df, omega_i, (float *)A0, (float *)A1, lda,
                           flag_count, flag_list);
    return hipGetLastError();
}

}  // namespace grtm
