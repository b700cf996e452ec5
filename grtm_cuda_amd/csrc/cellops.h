// cellops.h -- device-side cell helpers shared by the recursion kernels:
// one (omega, k) cell solved and turned into the Z/R operand values or the
// displacement-column rows, their stores and magnitudes.
#pragma once
#include <hip/hip_runtime.h>

#include "cell.h"
#include "kernels.h"

namespace grtm {

constexpr int kCellThreads = 256;
constexpr int kTileBits = 12;  // work item = row << kTileBits | k tile

// ---------------------------------------------------------------- recursion
// One thread per (omega, k) cell; T = arithmetic of the recursion, TO = type
// of the A operands.  FP32 cells whose result is not finite are appended to
// a list (flag_count / flag_list) and re-solved in FP64 by k_cells_fix.
template <typename T, typename TO, bool COMPAT, int BR = 0>
__device__ __forceinline__ bool cell_zr(const LayerW<T> *lw, const Geo &g, int iwg, int ik, double dk, double df,
                                        double omega_i, TO v[8])
{
    const double kd = dk * ik;
    const T k = T(kd);
    const cx<T> w = C<T>(T(2.0 * M_PI * df * iwg), T(-omega_i));
    CellOut<T> c;
    cell_solve<T, COMPAT, false, BR>(lw, g, kd, C<T>(0, 0), C<T>(0, 0), c);
    cx<T> kze, kre, kzv, krv;
    // the coupling direction follows from the branch (above the source: 1, 2)
    zr_from_cell<T, COMPAT>(c, BR ? int(BR <= 2) : g.up, k, w, kze, kre, kzv, krv);
    const T wt = T(ik == 0 ? 0.5 * dk : dk);
    const T r[8] = {wt * kze.x, wt * kze.y, wt * kzv.x, wt * kzv.y, wt * kre.x, wt * kre.y, wt * krv.x, wt * krv.y};
    bool ok = true;
    for (int i = 0; i < 8; i++) {
        v[i] = TO(r[i]);
        ok = ok && isfinite(float(v[i]));
    }
    return ok;
}

template <typename TO>
__device__ __forceinline__ void store_zr(const TO v[8], int iw, int ik, TO *A0, TO *A1, long lda)
{
    const size_t r0 = (size_t)(4 * iw) * lda + ik;
    for (int i = 0; i < 4; i++) {
        A0[r0 + i * lda] = v[i];
        A1[r0 + i * lda] = v[4 + i];
    }
}

// Largest |value| of a cell, as the bit pattern of a non-negative float
// (monotone in the value, so atomicMax on the bits is a float max).
template <typename TO>
__device__ __forceinline__ unsigned cell_mag_bits(const TO v[8])
{
    float m = 0.f;
    for (int i = 0; i < 8; i++) m = fmaxf(m, fabsf(float(v[i])));
    return isfinite(m) ? __float_as_uint(m) : 0u;
}

// Largest value over the 64 lanes of a wave.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, unsigned(__shfl_xor(int(v), o, 64)));
    return v;
}

// ---------------------------------------------------------------- moment tensor
// Recursion of the displacement-column operands: one thread per (omega, k)
// cell, the rows of mt_rows_from_cell weighted by the k-quadrature;
// put(set, row, re, im) receives them.
template <typename T, bool COMPAT, bool DERIV, typename Put, int BR = 0>
__device__ __forceinline__ void cell_mt_rows(const LayerW<T> *lw, const Geo &g, int iw, int ik, double dk, double df,
                                             double omega_i, Put &&put)
{
    const double kd = dk * ik;
    const T k = T(kd);
    const cx<T> w = C<T>(T(2.0 * M_PI * df * iw), T(-omega_i));
    CellOut<T> c;
    cell_solve<T, COMPAT, DERIV, BR>(lw, g, kd, C<T>(0, 0), C<T>(0, 0), c);
    const T wt = T(ik == 0 ? 0.5 * dk : dk);
    mt_rows_from_cell<T, COMPAT, DERIV>(c, BR ? int(BR <= 2) : g.up, k, w,
                                        [&](int set, int j, cx<T> v) { put(set, j, wt * v.x, wt * v.y); });
}

// Stores a cell's rows to A_s[2 (rows_s iw + j) + {0 re, 1 im}][ik]; returns
// whether all are finite, and their largest |value| as float bits in mag.
template <typename T, typename TO, bool COMPAT, bool DERIV, int BR = 0>
__device__ __forceinline__ bool cell_mt(const LayerW<T> *lw, const Geo &g, int iw, int ik, double dk, double df,
                                        double omega_i, const MtOps &o, long lda, unsigned &mag)
{
    bool ok = true;
    float m = 0.f;
    auto put = [&](int set, int j, T vr, T vi) {
        const TO re = TO(vr), im = TO(vi);
        TO *A = reinterpret_cast<TO *>(o.A[set]);
        const size_t row = 2 * ((size_t)o.rows[set] * iw + j);
        A[row * lda + ik] = re;
        A[(row + 1) * lda + ik] = im;
        const float fr = fabsf(float(re)), fi = fabsf(float(im));
        ok = ok && isfinite(fr) && isfinite(fi);
        m = fmaxf(m, fmaxf(fr, fi));
    };
    cell_mt_rows<T, COMPAT, DERIV, decltype(put) &, BR>(lw, g, iw, ik, dk, df, omega_i, put);
    mag = isfinite(m) ? __float_as_uint(m) : 0u;
    return ok;
}

}  // namespace grtm
