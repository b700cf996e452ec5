// green.hip -- the drop-in path's cell kernel: per-(omega,k) recursion -> all
// 30 integrand columns of green()'s output at one distance, reduced over k
// per block (green.cu:1799-1903: set_pars .. get_int_val_*, and the host
// k-sum at 1878-1903, with the per-row ragged k counts of 2145-2194).
//
// Compiled in two arithmetics: FP64, and double-double for the lowest
// frequency rows, where the P and SV eigenvector columns become parallel
// (|w| ~ w_I) and an FP64 solve of the recursion is rounding noise (it loses
// about (k V / |w|)^4: 40% per cell at the first row of a 4096 s trace).  The
// layer constants stay FP64 in both; the double-double kernel widens them
// exactly in LDS.  The integrand and the k-sum run in FP64 from the cell
// coefficients in both.
#include <hip/hip_runtime.h>

#include "cell.h"
#include "cellops.h"
#include "integrand.h"
#include "kernels.h"

namespace grtm {

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ LayerW<T> cell_table(const LayerW<double> &l);
template <>
__device__ __forceinline__ LayerW<double> cell_table<double>(const LayerW<double> &l)
{
    return l;
}
template <>
__device__ __forceinline__ LayerW<dd> cell_table<dd>(const LayerW<double> &l)
{
    return widen_dd(l);
}

// Rows w_base .. w_base + gridDim.y - 1; block x covers wavenumbers
// [x kCellThreads, (x + 1) kCellThreads) of its row (those below nk_w[iw]).
// partial[(iw * nblk + x) * 60 + 2 c + {re, im}].
template <typename T, bool COMPAT, int SRC>
__global__ __launch_bounds__(kCellThreads) void k_cells_full(const LayerW<double> *__restrict__ lw_all,
                                                             const int *__restrict__ nk_w, int L, Geo g,
                                                             int w_base, double dk, double dist, double df,
                                                             double omega_i, double p0r_x, double p0r_y,
                                                             double p0v_x, double p0v_y, int nblk,
                                                             double *__restrict__ partial)
{
    __shared__ LayerW<T> lw[kMaxLayers];
    __shared__ double red[kCellThreads / 64][60];
    const int iw = w_base + int(blockIdx.y);
    for (int j = threadIdx.x; j < L; j += blockDim.x) lw[j] = cell_table<T>(lw_all[(size_t)iw * L + j]);
    __syncthreads();
    const int ik = blockIdx.x * blockDim.x + threadIdx.x;
    cx<double> col[30];
    for (int i = 0; i < 30; i++) col[i] = cx<double>{0, 0};
    if (ik < nk_w[iw]) {
        const double kd = dk * ik;
        const T k = T(kd);
        const cx<T> w = C<T>(T(2.0 * M_PI * df * iw), T(-omega_i));
        CellOut<T> c;
        cell_solve<T, COMPAT, true>(lw, g, k, C<T>(T(p0r_x), T(p0r_y)), C<T>(T(p0v_x), T(p0v_y)), c);
        Coef<T> p;
        coef_from_cell<T, true, COMPAT>(c, SRC, g.up, k, w, p);
        BesselSet bs;
        bessel_set(COMPAT ? 0 : 1, dist * kd, bs);
        integrand_cols<T>(COMPAT ? 1 : 0, SRC, p, kd, dist, bs, col);
        const double wt = (ik == 0) ? dk / 2.0 : dk;
        for (int i = 0; i < 30; i++) col[i] = col[i] * wt;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = 0; i < 30; i++) {
        double sx = wave_sum(col[i].x), sy = wave_sum(col[i].y);
        if (lane == 0) {
            red[wv][2 * i] = sx;
            red[wv][2 * i + 1] = sy;
        }
    }
    __syncthreads();
    if (threadIdx.x < 60) {
        double s = 0;
        for (int q = 0; q < kCellThreads / 64; q++) s += red[q][threadIdx.x];
        partial[((size_t)iw * nblk + blockIdx.x) * 60 + threadIdx.x] = s;
    }
}

hipError_t launch_cells_full(int compat, int src, int dd_rows, const void *lw, const int *nk_w, int L,
                             const Geo &g, int w_base, int nrows, int nblk, double dk, double dist, double df,
                             double omega_i, const double p0[4], double *partial, hipStream_t st)
{
    if (nrows <= 0) return hipSuccess;
    dim3 grid(nblk, nrows);
#define GRTM_FULL(T, CP, SR)                                                                          \
    hipLaunchKernelGGL((k_cells_full<T, CP, SR>), grid, dim3(kCellThreads), 0, st,                  \
                       (const LayerW<double> *)lw, nk_w, L, g, w_base, dk, dist, df, omega_i, p0[0], \
                       p0[1], p0[2], p0[3], nblk, partial)
#define GRTM_FULL_T(T)                       \
    if (compat) {                            \
        if (src == 0) GRTM_FULL(T, true, 0); \
        else GRTM_FULL(T, true, 1);          \
    } else {                                 \
        if (src == 0) GRTM_FULL(T, false, 0); \
        else GRTM_FULL(T, false, 1);         \
    }
    if (dd_rows) {
        GRTM_FULL_T(dd)
    } else {
        GRTM_FULL_T(double)
    }
#undef GRTM_FULL_T
#undef GRTM_FULL
    return hipGetLastError();
}

}  // namespace grtm
