// integrand.h -- the 30 integrand columns of the drop-in output
// (cusrc/popm/func/green.cu:1558-1621 single force, 1626-1716 tensor).
//
// Column layout follows the reference.  compat = 1 keeps the reference's
// column contents; compat = 0 fixes what the reference gets wrong:
//   single force: column 10 divides the b1 J1' term by r like column 12;
//   tensor: columns 2/3 use the radial coefficient a.[0], column 5 uses b1,
//   the d/dz block (10-19) is kept and the d/dr block goes to 20-29;
//   the d/dr columns use the exact J''.
#pragma once
#include "bessel.h"
#include "cell.h"

namespace grtm {

template <typename T>
GRTM_HD void integrand_cols(int compat, int src_type, const Coef<T> &p, double k, double dist,
                            const BesselSet &bs, cx<double> *c)
{
    auto D = [](cx<T> z) { return cx<double>{double(z.x), double(z.y)}; };
    const double r = dist;
    const double j0 = bs.j[0], j1 = bs.j[1], j2 = bs.j[2];
    const double j0d = bs.jd[0], j1d = bs.jd[1], j2d = bs.jd[2];
    const double j0dd = bs.jdd[0], j1dd = bs.jdd[1], j2dd = bs.jdd[2];
    const cx<double> a00 = D(p.a0[0]), a01 = D(p.a0[1]), a10 = D(p.a1[0]), a11 = D(p.a1[1]);
    const cx<double> a20 = D(p.a2[0]), a21 = D(p.a2[1]), b1 = D(p.b1), b2 = D(p.b2);
    const cx<double> a00t = D(p.a0t[0]), a01t = D(p.a0t[1]), a10t = D(p.a1t[0]), a11t = D(p.a1t[1]);
    const cx<double> a20t = D(p.a2t[0]), a21t = D(p.a2t[1]), b1t = D(p.b1t), b2t = D(p.b2t);
    for (int i = 0; i < 30; i++) c[i] = cx<double>{0, 0};
    if (src_type == 0) {
        c[0] = b1 * (j1 / r) + a10 * (k * j1d);
        c[1] = a00 * (k * j0d);
        c[2] = a10 * (j1 / r) + b1 * (k * j1d);
        c[3] = a11 * (k * j1);
        c[4] = a01 * (k * j0);
        c[5] = b1t * (j1 / r) + a10t * (k * j1d);
        c[6] = a00t * (k * j0d);
        c[7] = a10t * (j1 / r) + b1t * (k * j1d);
        c[8] = a11t * (k * j1);
        c[9] = a01t * (k * j0);
        c[10] = b1 * (-j1 / r / r) + b1 * (j1d * k / (compat ? 1.0 : r)) + a10 * (k * j1dd * k);
        c[11] = a00 * (k * j0dd * k);
        c[12] = a10 * (-j1 / r / r) + a10 * (j1d * k / r) + b1 * (k * j1dd * k);
        c[13] = a11 * (k * j1d * k);
        c[14] = a01 * (k * j0d * k);
        return;
    }
    const cx<double> a2q = compat ? a21 : a20, a0q = compat ? a01 : a00;
    const cx<double> b5 = compat ? b2 : b1;
    c[0] = b2 * (2.0 * j2 / r) + a20 * (k * j2d);
    c[1] = b1 * (j1 / r) + a10 * (k * j1d);
    c[2] = a2q * (k * j0d);
    c[3] = a0q * (k * j0d);
    c[4] = b2 * (k * j2d) + a20 * (2.0 * j2 / r);
    c[5] = b5 * (k * j1d) + a10 * (j1 / r);
    c[6] = a21 * (k * j2);
    c[7] = a11 * (k * j1);
    c[8] = a01 * (k * j0);
    c[9] = a21 * (k * j0);
    const int o = compat ? 10 : 20;
    c[o + 0] = b2 * (-2.0 * j2 / r / r) + b2 * (2.0 * j2d * k / r) + a20 * (k * j2dd * k);
    c[o + 1] = b1 * (-j1 / r / r) + b1 * (j1d * k / r) + a10 * (k * j1dd * k);
    c[o + 2] = a2q * (k * j0dd * k);
    c[o + 3] = a0q * (k * j0dd * k);
    c[o + 4] = b2 * (k * j2dd * k) + a20 * (-2.0 * j2 / r / r) + a20 * (2.0 * j2d * k / r);
    c[o + 5] = b5 * (k * j1dd * k) + a10 * (-j1 / r / r) + a10 * (j1d * k / r);
    c[o + 6] = a21 * (k * j2d * k);
    c[o + 7] = a11 * (k * j1d * k);
    c[o + 8] = a01 * (k * j0d * k);
    c[o + 9] = a21 * (k * j0d * k);
    if (!compat) {
        c[10] = b2t * (2.0 * j2 / r) + a20t * (k * j2d);
        c[11] = b1t * (j1 / r) + a10t * (k * j1d);
        c[12] = a20t * (k * j0d);
        c[13] = a00t * (k * j0d);
        c[14] = b2t * (k * j2d) + a20t * (2.0 * j2 / r);
        c[15] = b1t * (k * j1d) + a10t * (j1 / r);
        c[16] = a21t * (k * j2);
        c[17] = a11t * (k * j1);
        c[18] = a01t * (k * j0);
        c[19] = a21t * (k * j0);
    }
}

}  // namespace grtm
