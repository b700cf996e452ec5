// dd.h -- double-double arithmetic: a real number as an unevaluated sum
// hi + lo of two FP64 values with |lo| <= ulp(hi) / 2, about 106 bits of
// significand.  Host+device, so the cell solve can run in it on the GPU and
// in host self-tests.
//
// Used for the lowest frequency rows, where the P and SV eigenvector columns
// of the layer matrices become parallel (|w| -> w_I, k V / |w| ~ 1e4) and an
// FP64 solve of the reflection/transmission recursion loses about
// (k V / |w|)^4 of relative accuracy: the FP64 answer there is rounding noise
// in any formulation.
//
// Methods (published): error-free transformations two-sum (Knuth) and
// two-product by fused multiply-add; double-double addition, product and
// quotient as in Hida, Li and Bailey, "Algorithms for quad-double precision
// floating point arithmetic" (ARITH-15, 2001); square root by one Newton
// step on the FP64 root; exp by reduction modulo ln 2, a Taylor series of
// the argument scaled by 2^-9 and nine squarings of (1 + x); sin and cos by
// reduction modulo pi/2 and Taylor series on [-pi/4, pi/4].  The constants
// ln 2, pi/2, 2 pi and the series coefficients 1/n! are their correctly
// rounded double-double values.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#ifndef GRTM_HD
#define GRTM_HD __host__ __device__ __forceinline__
#endif

namespace grtm {

// The FP64 overloads stay visible beside the double-double ones declared
// below (an unqualified call in this namespace would otherwise stop at them).
using ::copysign;
using ::exp;
using ::fabs;
using ::fma;
using ::hypot;
using ::sqrt;

// Every product and sum below must round on its own: a fused multiply-add
// formed from them (the GPU backend contracts FP64 products into later sums,
// even products with other uses) breaks the error-free transformations.
#define DD_NO_CONTRACT _Pragma("clang fp contract(off)")

struct dd {
    double hi, lo;
    dd() = default;  // trivial: LDS arrays of records holding dd are allowed
    GRTM_HD constexpr dd(double h) : hi(h), lo(0.0) {}
    GRTM_HD constexpr dd(double h, double l) : hi(h), lo(l) {}
    GRTM_HD explicit operator double() const { return hi + lo; }
    GRTM_HD explicit operator float() const { return float(hi + lo); }
};

namespace ddx {
GRTM_HD dd two_sum(double a, double b)
{
    DD_NO_CONTRACT
    const double s = a + b;
    const double bb = s - a;
    return dd(s, (a - (s - bb)) + (b - bb));
}
GRTM_HD dd quick_two_sum(double a, double b)
{
    DD_NO_CONTRACT
    const double s = a + b;
    return dd(s, b - (s - a));
}
GRTM_HD dd two_prod(double a, double b)
{
    DD_NO_CONTRACT
    const double p = a * b;
    return dd(p, ::fma(a, b, -p));
}
}  // namespace ddx

GRTM_HD dd operator+(dd a, dd b)
{
    DD_NO_CONTRACT
    dd s = ddx::two_sum(a.hi, b.hi);
    const dd t = ddx::two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = ddx::quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return ddx::quick_two_sum(s.hi, s.lo);
}
GRTM_HD dd operator-(dd a) { return dd(-a.hi, -a.lo); }
GRTM_HD dd operator-(dd a, dd b) { return a + (-b); }
GRTM_HD dd operator*(dd a, dd b)
{
    DD_NO_CONTRACT
    dd p = ddx::two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return ddx::quick_two_sum(p.hi, p.lo);
}
GRTM_HD dd operator/(dd a, dd b)
{
    DD_NO_CONTRACT
    const double q1 = a.hi / b.hi;
    dd r = a - b * dd(q1);
    const double q2 = r.hi / b.hi;
    r = r - b * dd(q2);
    const double q3 = r.hi / b.hi;
    return ddx::quick_two_sum(q1, q2) + dd(q3);
}
GRTM_HD dd &operator+=(dd &a, dd b) { return a = a + b; }
GRTM_HD dd &operator-=(dd &a, dd b) { return a = a - b; }
GRTM_HD dd &operator*=(dd &a, dd b) { return a = a * b; }

GRTM_HD bool operator<(dd a, dd b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
GRTM_HD bool operator>(dd a, dd b) { return b < a; }
GRTM_HD bool operator<=(dd a, dd b) { return !(b < a); }
GRTM_HD bool operator>=(dd a, dd b) { return !(a < b); }
GRTM_HD bool operator==(dd a, dd b) { return a.hi == b.hi && a.lo == b.lo; }
GRTM_HD bool operator!=(dd a, dd b) { return !(a == b); }

GRTM_HD dd fabs(dd a) { return a.hi < 0.0 || (a.hi == 0.0 && a.lo < 0.0) ? -a : a; }
GRTM_HD dd copysign(dd a, dd s) { return (s.hi < 0.0) != (a.hi < 0.0) ? -a : a; }
GRTM_HD dd fma(dd a, dd b, dd c) { return a * b + c; }
GRTM_HD dd ldexp_dd(dd a, int e) { return dd(::ldexp(a.hi, e), ::ldexp(a.lo, e)); }

GRTM_HD dd sqrt(dd a)
{
    DD_NO_CONTRACT
    if (!(a.hi > 0.0)) return dd(0.0);
    const double s = ::sqrt(a.hi);
    const dd p = ddx::two_prod(s, s);
    const double e = ((a.hi - p.hi) - p.lo + a.lo) / (2.0 * s);
    return ddx::quick_two_sum(s, e);
}
// estimates (the depth-cut criterion): the FP64 root suffices
GRTM_HD dd sqrt_est(dd a) { return dd(::sqrt(a.hi)); }
GRTM_HD dd hypot(dd a, dd b) { return sqrt(a * a + b * b); }

// 1 / n! for n = 0..28 as correctly rounded double-double values
__device__ __constant__ static const double kInvFact[29][2] = {
    {1.0, 0.0},
    {1.0, 0.0},
    {0.5, 0.0},
    {0.16666666666666666, 9.25185853854297e-18},
    {0.041666666666666664, 2.3129646346357427e-18},
    {0.008333333333333333, 1.1564823173178714e-19},
    {0.001388888888888889, -5.300543954373577e-20},
    {0.0001984126984126984, 1.7209558293420705e-22},
    {2.48015873015873e-05, 2.1511947866775882e-23},
    {2.7557319223985893e-06, -1.858393274046472e-22},
    {2.755731922398589e-07, 2.3767714622250297e-23},
    {2.505210838544172e-08, -1.448814070935912e-24},
    {2.08767569878681e-09, -1.20734505911326e-25},
    {1.6059043836821613e-10, 1.2585294588752098e-26},
    {1.1470745597729725e-11, 2.0655512752830745e-28},
    {7.647163731819816e-13, 7.03872877733453e-30},
    {4.779477332387385e-14, 4.399205485834081e-31},
    {2.8114572543455206e-15, 1.6508842730861433e-31},
    {1.5619206968586225e-16, 1.1910679660273754e-32},
    {8.22063524662433e-18, 2.2141894119604265e-34},
    {4.110317623312165e-19, 1.4412973378659527e-36},
    {1.9572941063391263e-20, -1.3643503830087908e-36},
    {8.896791392450574e-22, -7.911402614872376e-38},
    {3.868170170630684e-23, -8.843177655482344e-40},
    {1.6117375710961184e-24, -3.6846573564509766e-41},
    {6.446950284384474e-26, -1.9330404233703465e-42},
    {2.4795962632247976e-27, -1.2953730964765229e-43},
    {9.183689863795546e-29, 1.4303150396787322e-45},
    {3.279889237069838e-30, 1.5117542744029879e-46},
};
static const double kInvFactHost[29][2] = {
    {1.0, 0.0},
    {1.0, 0.0},
    {0.5, 0.0},
    {0.16666666666666666, 9.25185853854297e-18},
    {0.041666666666666664, 2.3129646346357427e-18},
    {0.008333333333333333, 1.1564823173178714e-19},
    {0.001388888888888889, -5.300543954373577e-20},
    {0.0001984126984126984, 1.7209558293420705e-22},
    {2.48015873015873e-05, 2.1511947866775882e-23},
    {2.7557319223985893e-06, -1.858393274046472e-22},
    {2.755731922398589e-07, 2.3767714622250297e-23},
    {2.505210838544172e-08, -1.448814070935912e-24},
    {2.08767569878681e-09, -1.20734505911326e-25},
    {1.6059043836821613e-10, 1.2585294588752098e-26},
    {1.1470745597729725e-11, 2.0655512752830745e-28},
    {7.647163731819816e-13, 7.03872877733453e-30},
    {4.779477332387385e-14, 4.399205485834081e-31},
    {2.8114572543455206e-15, 1.6508842730861433e-31},
    {1.5619206968586225e-16, 1.1910679660273754e-32},
    {8.22063524662433e-18, 2.2141894119604265e-34},
    {4.110317623312165e-19, 1.4412973378659527e-36},
    {1.9572941063391263e-20, -1.3643503830087908e-36},
    {8.896791392450574e-22, -7.911402614872376e-38},
    {3.868170170630684e-23, -8.843177655482344e-40},
    {1.6117375710961184e-24, -3.6846573564509766e-41},
    {6.446950284384474e-26, -1.9330404233703465e-42},
    {2.4795962632247976e-27, -1.2953730964765229e-43},
    {9.183689863795546e-29, 1.4303150396787322e-45},
    {3.279889237069838e-30, 1.5117542744029879e-46},
};
GRTM_HD dd inv_fact(int n)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return dd(kInvFact[n][0], kInvFact[n][1]);
#else
    return dd(kInvFactHost[n][0], kInvFactHost[n][1]);
#endif
}

GRTM_HD dd exp(dd a)
{
    DD_NO_CONTRACT
    constexpr double ln2_hi = 6.931471805599452862e-01, ln2_lo = 2.319046813846299558e-17;
    if (a.hi < -745.0) return dd(0.0);
    if (a.hi > 709.0) return dd(HUGE_VAL);
    const double m = ::rint(a.hi / ln2_hi);
    // r = (a - m ln2) / 512, |r| <= 6.8e-4
    const dd r = ldexp_dd(a - dd(ln2_hi, ln2_lo) * dd(m), -9);
    // e^r - 1 = r (1 + r (1/2! + r (1/3! + ... + r / 11!))); r^12 / 12! < 1e-46
    dd p = inv_fact(11);
    for (int n = 10; n >= 1; n--) p = p * r + inv_fact(n);
    dd e = p * r;
    // (1 + e)^2 - 1 = 2 e + e^2, nine times: e^{512 r} - 1
    for (int i = 0; i < 9; i++) e = ldexp_dd(e, 1) + e * e;
    return ldexp_dd(e + dd(1.0), int(m));
}

GRTM_HD void sincos_t(dd a, dd *s, dd *c)
{
    DD_NO_CONTRACT
    constexpr double pi2_hi = 6.283185307179586232e+00, pi2_lo = 2.449293598294706414e-16;
    constexpr double pio2_hi = 1.570796326794896558e+00, pio2_lo = 6.123233995736766036e-17;
    // a - n 2 pi, then - j pi/2 (|t| <= pi/4)
    const double n = ::rint(a.hi / pi2_hi);
    dd t = a - dd(pi2_hi, pi2_lo) * dd(n);
    const double jq = ::rint(t.hi / pio2_hi);
    t = t - dd(pio2_hi, pio2_lo) * dd(jq);
    const int j = int(jq) & 3;
    const dd u = t * t;
    // sin t = t sum_m (-1)^m u^m / (2m+1)!  (m <= 13: t^27 / 27!),
    // cos t =   sum_m (-1)^m u^m / (2m)!    (m <= 14: t^28 / 28!), by Horner in u
    dd ss = inv_fact(27), cc = inv_fact(28);
    for (int m = 12; m >= 0; m--) ss = inv_fact(2 * m + 1) - u * ss;
    for (int m = 13; m >= 0; m--) cc = inv_fact(2 * m) - u * cc;
    const dd sn = t * ss;
    switch (j) {
    case 0: *s = sn, *c = cc; break;
    case 1: *s = cc, *c = -sn; break;
    case 2: *s = -sn, *c = -cc; break;
    default: *s = -cc, *c = sn; break;
    }
}

}  // namespace grtm
