// bessel.h -- Bessel functions for the wavenumber integrand.
//
// mode 0 ("reference"): the Abramowitz & Stegun 9.4.1-9.4.6 polynomial forms,
//   integer-order J_n by upward recurrence above n / Miller's downward
//   recurrence below, and the derivative formulas the reference uses
//   (cusrc/popm/func/green.cu:1418-1554, including its "J''" formula at 1552).
// mode 1 ("physics"): libm / device-libm j0, j1, jn with exact derivative
//   identities.
#pragma once
#include "cx.h"

namespace grtm {

GRTM_HD double as_j0(double x)
{
    const double a0 = 1.00000000, a1 = -2.24999970, a2 = 1.26562080, a3 = -0.31638660, a4 = 0.04444790,
                 a5 = -0.00394440, a6 = 0.00021000;
    const double b0 = 0.79788456, b1 = -0.00000077, b2 = -0.00552740, b3 = -0.00009512, b4 = 0.00137237,
                 b5 = -0.00072805, b6 = 0.00014476;
    const double c0 = -0.78539816, c1 = -0.04166397, c2 = -0.00003954, c3 = 0.00262573, c4 = -0.00054125,
                 c5 = -0.00029333, c6 = 0.00013558;
    double ax = fabs(x);
    if (ax <= 3) {
        double y = (ax / 3.0) * (ax / 3.0);
        return a0 + y * (a1 + y * (a2 + y * (a3 + y * (a4 + y * (a5 + y * a6)))));
    }
    double t = 3.0 / ax;
    double f = b0 + t * (b1 + t * (b2 + t * (b3 + t * (b4 + t * (b5 + t * b6)))));
    double th = ax + c0 + t * (c1 + t * (c2 + t * (c3 + t * (c4 + t * (c5 + t * c6)))));
    return f * cos(th) / sqrt(ax);
}

GRTM_HD double as_j1(double x)
{
    const double a0 = 0.50000000, a1 = -0.56249985, a2 = 0.21093573, a3 = -0.03954289, a4 = 0.00443319,
                 a5 = -0.00031761, a6 = 0.00001109;
    const double b0 = 0.79788456, b1 = 0.00000156, b2 = 0.01659667, b3 = 0.00017105, b4 = -0.00249511,
                 b5 = 0.00113653, b6 = -0.00020033;
    const double c0 = -2.35619449, c1 = 0.12499612, c2 = 0.00005650, c3 = -0.00637879, c4 = 0.00074348,
                 c5 = 0.00079824, c6 = -0.00029166;
    double ax = fabs(x);
    if (ax <= 3.0) {
        double y = (ax / 3.0) * (ax / 3.0);
        return x * (a0 + y * (a1 + y * (a2 + y * (a3 + y * (a4 + y * (a5 + y * a6))))));
    }
    double t = 3.0 / ax;
    double f = b0 + t * (b1 + t * (b2 + t * (b3 + t * (b4 + t * (b5 + t * b6)))));
    double th = ax + c0 + t * (c1 + t * (c2 + t * (c3 + t * (c4 + t * (c5 + t * c6)))));
    return (x >= 0 ? 1.0 : -1.0) * f * cos(th) / sqrt(ax);
}

GRTM_HD double as_jn(double x, int n)
{
    const int iacc = 40;
    const double bigno = 1e10, bigni = 1e-10;
    int nn = n;
    double res;
    if (n == 0) {
        res = as_j0(x);
    } else if (n == 1) {
        res = as_j1(x);
    } else {
        if (n < 0) n = -n;
        double ax = fabs(x);
        if (ax == 0) {
            res = 0.0;
        } else if (ax > n) {
            double tox = 2.0 / ax, bjm = as_j0(ax), bj = as_j1(ax);
            for (int j = 0; j < n - 1; j++) {
                double bjp = (double)(j + 1) * tox * bj - bjm;
                bjm = bj;
                bj = bjp;
            }
            res = bj;
        } else {
            double tox = 2.0 / ax, bjp = 0.0, bj = 1.0, sum = 0.0;
            int m = 2 * ((n + (int)(sqrt((double)(iacc * n)))) / 2), jsum = 0;
            res = 0.0;
            for (int j = m; j >= 1; j--) {
                double bjm = (double)j * tox * bj - bjp;
                bjp = bj;
                bj = bjm;
                if (fabs(bj) > bigno) {
                    bj *= bigni;
                    bjp *= bigni;
                    res *= bigni;
                    sum *= bigni;
                }
                if (jsum != 0) sum += bj;
                jsum = 1 - jsum;
                if (j == n) res = bjp;
            }
            sum = 2.0 * sum - bj;
            res = res / sum;
        }
        if (x < 0.0 && (n % 2) == 1) res = -res;
    }
    if (nn < 0 && (n % 2) == 1) res = -res;
    return res;
}

GRTM_HD double as_djn(double x, int n)
{
    const double bigni = 1e-10;
    if (fabs(x) < bigni) return (as_jn(bigni, n) - as_jn(0, n)) / bigni;
    return (double)n * as_jn(x, n) / x - as_jn(x, n + 1);
}

GRTM_HD double as_ddjn(double x, int n)
{
    const double bigni = 1e-10;
    if (fabs(x) < bigni) return (as_djn(bigni, n) - as_djn(0, n)) / bigni;
    return (double)n * as_jn(x, n - 1) / x - as_jn(x, n) * (n + n * n + x * x) / x / x;
}

// J_m(x) for |x| below 1e-3 from the first two series terms
// (x/2)^m / m! (1 - (x/2)^2 / (m+1)); exact at x = 0, where a forward
// recurrence in m would divide by x.
GRTM_HD double small_jn(int m, double x)
{
    const double h = 0.5 * x;
    double t = 1.0;
    for (int i = 1; i <= m; i++) t *= h / i;
    return t * (1.0 - h * h / (m + 1));
}

// Accurate integer-order J_n from the C / device math library.
GRTM_HD double lib_jn(int n, double x)
{
    int m = n < 0 ? -n : n;
    double v;
    if (fabs(x) < 1e-3 && m >= 2) {
        v = small_jn(m, x);
    } else {
#ifdef __HIP_DEVICE_COMPILE__
        v = (m == 0) ? ::j0(x) : (m == 1) ? ::j1(x) : ::jn(m, x);
#else
        v = ::jn(m, x);
#endif
    }
    return (n < 0 && (m & 1)) ? -v : v;
}

// Values J_n, J_n', J_n'' (mode 1) or the reference's three (mode 0) for n = 0..2.
struct BesselSet {
    double j[3], jd[3], jdd[3];
};

GRTM_HD void bessel_set(int mode, double x, BesselSet &b)
{
    if (mode == 0) {
        for (int n = 0; n < 3; n++) {
            b.j[n] = as_jn(x, n);
            b.jd[n] = as_djn(x, n);
            b.jdd[n] = as_ddjn(x, n);
        }
        return;
    }
    double J[5];  // J_0..J_4
    for (int n = 0; n < 5; n++) J[n] = lib_jn(n, x);
    auto Jn = [&](int n) { return n < 0 ? ((n & 1) ? -J[-n] : J[-n]) : J[n]; };
    for (int n = 0; n < 3; n++) {
        b.j[n] = J[n];
        b.jd[n] = 0.5 * (Jn(n - 1) - Jn(n + 1));
        b.jdd[n] = 0.25 * (Jn(n - 2) - 2.0 * J[n] + Jn(n + 2));
    }
}

}  // namespace grtm
