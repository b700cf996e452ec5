// columns.hip -- recursion kernels of the displacement- and strain-column
// plans (tensor and single-force columns for many receivers).
#include <hip/hip_runtime.h>

#include "cellops.h"

namespace grtm {

// Work items as k_cells_zr: the list `tiles` (row << kTileBits | k tile,
// residual-check items flagged with kCheckItem) or every tile of rows
// [w0, w0 + gridDim.x / tpr); nk_w (optional) limits the wavenumbers of each
// row, row_peak (optional) collects the row maxima.  BR: the plan's receiver
// branch, fixed at compile time.
template <typename T, typename TO, bool COMPAT, bool DERIV, int BR>
__global__ __launch_bounds__(kCellThreads) void k_cells_mt(const LayerW<T> *__restrict__ lw_all, int L, Geo g,
                                                           int nk, double dk, double df, double omega_i, int w0,
                                                           int tpr, const unsigned *__restrict__ tiles, MtOps o,
                                                           long lda, unsigned *__restrict__ flag_count,
                                                           unsigned *__restrict__ flag_list,
                                                           const int *__restrict__ nk_w,
                                                           unsigned *__restrict__ row_peak,
                                                           const int *__restrict__ nchk, int stride,
                                                           unsigned *__restrict__ tail_max)
{
    __shared__ LayerW<T> lw[kMaxLayers];
    int iw, kb;
    bool chk = false;
    if (tiles) {
        const unsigned e = tiles[blockIdx.x];
        chk = (e & kCheckItem) != 0;
        iw = int((e & ~kCheckItem) >> kTileBits);
        kb = int(e & ((1u << kTileBits) - 1));
    } else {
        const int q = int(blockIdx.x) / tpr;
        iw = w0 + q;
        kb = int(blockIdx.x) - q * tpr;
    }
    const int nkr = nk_w ? nk_w[iw] : nk;
    const int lim = chk ? nchk[iw] : nkr;   // cells (check cells) of the row
    if (kb * kCellThreads >= lim) return;  // uniform over the block
    for (int j = threadIdx.x; j < L; j += blockDim.x) lw[j] = lw_all[(size_t)iw * L + j];
    __syncthreads();
    const int j = kb * kCellThreads + threadIdx.x;
    const int ik = chk ? nkr + stride * j + stride / 2 : j;
    unsigned mag = 0;
    if (j < lim && ik < nk) {
        // the cell's rows: stored (and flagged when not finite), or only their
        // largest |value| for a check cell
        bool ok = true;
        float m = 0.f;
        auto put = [&](int set, int jr, T vr, T vi) {
            const TO re = TO(vr), im = TO(vi);
            if (!chk) {
                TO *A = reinterpret_cast<TO *>(o.A[set]);
                const size_t row = 2 * ((size_t)o.rows[set] * iw + jr);
                A[row * lda + ik] = re;
                A[(row + 1) * lda + ik] = im;
            }
            const float fr = fabsf(float(re)), fi = fabsf(float(im));
            ok = ok && isfinite(fr) && isfinite(fi);
            m = fmaxf(m, fmaxf(fr, fi));
        };
        cell_mt_rows<T, COMPAT, DERIV, decltype(put) &, BR>(lw, g, iw, ik, dk, df, omega_i, put);
        if (chk) {
            mag = __float_as_uint(isfinite(m) ? m : 3.0e38f);  // a non-finite check value fails the row
        } else {
            mag = isfinite(m) ? __float_as_uint(m) : 0u;
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

// FP64 re-solve of the flagged cells (grid-stride over the list).
template <typename TO, bool COMPAT, bool DERIV>
__global__ __launch_bounds__(kCellThreads) void k_cells_mt_fix(const LayerW<double> *__restrict__ lw_all, int L,
                                                               Geo g, int nk, double dk, double df, double omega_i,
                                                               MtOps o, long lda,
                                                               const unsigned *__restrict__ flag_count,
                                                               const unsigned *__restrict__ flag_list)
{
    const unsigned n = *flag_count;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const unsigned cell = flag_list[t];
        const int iw = int(cell / unsigned(nk)), ik = int(cell % unsigned(nk));
        unsigned mag;
        cell_mt<double, TO, COMPAT, DERIV>(lw_all + (size_t)iw * L, g, iw, ik, dk, df, omega_i, o, lda, mag);
    }
}

// ---------------------------------------------------------------- launchers
hipError_t launch_cells_mt(int prec, int compat, int deriv, const void *lw, int L, const Geo &g, int nk, double dk,
                           double df, double omega_i, int w0, int nwr, const unsigned *tiles, int ntiles,
                           const MtOps &o, long lda, unsigned *flag_count, unsigned *flag_list, const int *nk_w,
                           unsigned *row_peak, hipStream_t st, const int *nchk, int stride, unsigned *tail_max)
{
    const int tpr = cells_tiles_per_row(nk);
    if (!tiles) ntiles = nwr * tpr;
    if (ntiles <= 0) return hipSuccess;
    if (deriv && compat) return hipErrorInvalidValue;
#define GRTM_MT_BR(T, TO, CP, DV, BR, FC, FL)                                                              \
    hipLaunchKernelGGL((k_cells_mt<T, TO, CP, DV, BR>), dim3(ntiles), dim3(kCellThreads), 0, st,            \
                       (const LayerW<T> *)lw, L, g, nk, dk, df, omega_i, w0, tpr, tiles, o, lda, FC, FL, nk_w, \
                       row_peak, nchk, stride, tail_max)
#define GRTM_MT(T, TO, CP, DV, FC, FL)                                                   \
    do {                                                                                 \
        switch (g.branch) {                                                              \
        case 1: GRTM_MT_BR(T, TO, CP, DV, 1, FC, FL); break;                             \
        case 2: GRTM_MT_BR(T, TO, CP, DV, 2, FC, FL); break;                             \
        case 3: GRTM_MT_BR(T, TO, CP, DV, 3, FC, FL); break;                             \
        default: GRTM_MT_BR(T, TO, CP, DV, 4, FC, FL); break;                            \
        }                                                                                \
    } while (0)
    if (prec == 0) {
        if (deriv) GRTM_MT(double, double, false, true, nullptr, nullptr);
        else if (compat) GRTM_MT(double, double, true, false, nullptr, nullptr);
        else GRTM_MT(double, double, false, false, nullptr, nullptr);
    } else if (prec == 1) {
        if (deriv) GRTM_MT(float, float, false, true, flag_count, flag_list);
        else if (compat) GRTM_MT(float, float, true, false, flag_count, flag_list);
        else GRTM_MT(float, float, false, false, flag_count, flag_list);
    } else {
        if (deriv) GRTM_MT(double, float, false, true, nullptr, nullptr);
        else if (compat) GRTM_MT(double, float, true, false, nullptr, nullptr);
        else GRTM_MT(double, float, false, false, nullptr, nullptr);
    }
#undef GRTM_MT
#undef GRTM_MT_BR
    return hipGetLastError();
}

hipError_t launch_cells_mt_fix(int compat, int deriv, const void *lw64, int L, const Geo &g, int nk, double dk,
                               double df, double omega_i, const MtOps &o, long lda, const unsigned *flag_count,
                               const unsigned *flag_list, hipStream_t st)
{
    dim3 grid(1024);
    if (deriv && compat) return hipErrorInvalidValue;
#define GRTM_FIX(CP, DV)                                                                                 \
    hipLaunchKernelGGL((k_cells_mt_fix<float, CP, DV>), grid, dim3(kCellThreads), 0, st,                 \
                       (const LayerW<double> *)lw64, L, g, nk, dk, df, omega_i, o, lda, flag_count, flag_list)
    if (deriv) GRTM_FIX(false, true);
    else if (compat) GRTM_FIX(true, false);
    else GRTM_FIX(false, false);
#undef GRTM_FIX
    return hipGetLastError();
}

}  // namespace grtm
