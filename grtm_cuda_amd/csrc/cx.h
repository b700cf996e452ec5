// cx.h -- small complex and 2x2-complex algebra for the GRT kernels.
// Host+device so the same arithmetic can be exercised by CPU self-tests.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define GRTM_HD __host__ __device__ __forceinline__

namespace grtm {

template <typename T>
struct cx {
    T x, y;
};

template <typename T> GRTM_HD cx<T> C(T a, T b) { return cx<T>{a, b}; }
template <typename T> GRTM_HD cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> GRTM_HD cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> GRTM_HD cx<T> operator-(cx<T> a) { return {-a.x, -a.y}; }
template <typename T> GRTM_HD cx<T> operator*(cx<T> a, cx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> GRTM_HD cx<T> operator*(T s, cx<T> a) { return {s * a.x, s * a.y}; }
template <typename T> GRTM_HD cx<T> operator*(cx<T> a, T s) { return {s * a.x, s * a.y}; }
template <typename T> GRTM_HD cx<T> conj(cx<T> a) { return {a.x, -a.y}; }
// c + a b and c - a b as four fused multiply-adds each (a b + d e costs one
// complex product and one cfma: 8 instructions instead of 10)
template <typename T> GRTM_HD cx<T> cfma(cx<T> a, cx<T> b, cx<T> c)
{
    return {fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y))};
}
template <typename T> GRTM_HD cx<T> cfms(cx<T> a, cx<T> b, cx<T> c)
{
    return {fma(-a.x, b.x, fma(a.y, b.y, c.x)), fma(-a.x, b.y, fma(-a.y, b.x, c.y))};
}
template <typename T> GRTM_HD T norm2(cx<T> a) { return a.x * a.x + a.y * a.y; }

// Hardware reciprocals on the device (v_rcp_f32: 1 ulp; v_rcp_f64: about
// 2^-26, for a Newton start); exact division in host builds.
GRTM_HD float frcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
GRTM_HD double drcp_est(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcp(x);
#else
    return 1.0 / x;
#endif
}
GRTM_HD float frsq(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}
GRTM_HD float fsqrt_est(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}

// Square root for estimates: the hardware FP32 instruction (about 1 ulp,
// no correctly rounded sequence); FP64 keeps the library root.
GRTM_HD float sqrt_est(float x) { return fsqrt_est(x); }
GRTM_HD double sqrt_est(double x) { return ::sqrt(x); }

// Reciprocal with Smith-style scaling (robust against over/underflow of |z|^2).
template <typename T> GRTM_HD cx<T> recip(cx<T> b)
{
    if (fabs(b.x) >= fabs(b.y)) {
        T r = b.y / b.x;
        T d = T(1) / (b.x + b.y * r);
        return {d, -r * d};
    } else {
        T r = b.x / b.y;
        T d = T(1) / (b.x * r + b.y);
        return {r * d, -d};
    }
}
// FP32: the same scaling with the hardware reciprocal (1 ulp) in place of
// two IEEE divides, and no branch.
template <> GRTM_HD cx<float> recip(cx<float> b)
{
    const bool xb = fabsf(b.x) >= fabsf(b.y);
    const float p = xb ? b.x : b.y, q = xb ? b.y : b.x;
    const float r = q * frcp(p);
    const float d = frcp(p + q * r);
    return xb ? cx<float>{d, -r * d} : cx<float>{r * d, -d};
}
template <typename T> GRTM_HD cx<T> operator/(cx<T> a, cx<T> b) { return a * recip(b); }

// Principal square root (Re >= 0), computed without cancellation.
template <typename T> GRTM_HD cx<T> csqrt(cx<T> z)
{
    T a = z.x, b = z.y;
    T m = hypot(a, b);
    if (m == T(0)) return {T(0), T(0)};
    if (a >= T(0)) {
        T t = sqrt(T(0.5) * (m + a));
        return {t, b / (T(2) * t)};
    }
    T t = sqrt(T(0.5) * (m - a));
    return {fabs(b) / (T(2) * t), copysign(t, b)};
}

// Hardware transcendentals for FP32 (device): 2^x (v_exp_f32, 1 ulp), and
// sin / cos of 2 pi x for x in revolutions (v_sin_f32 / v_cos_f32: 1.3e-7
// absolute on [-1/2, 1/2], measured); libm in host builds.
GRTM_HD float fexp2(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_exp2f(x);
#else
    return exp2f(x);
#endif
}
GRTM_HD float fsin_rev(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sinf(x);
#else
    return sinf(6.2831853071795865f * x);
#endif
}
GRTM_HD float fcos_rev(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_cosf(x);
#else
    return cosf(6.2831853071795865f * x);
#endif
}

// Principal square root of an FP64 value to about 1e-13 relative without FP64
// divides or square roots: an FP32 estimate (hardware square root and
// reciprocal square root, about 2^-21) refined by one Newton step s = (s0 + z / s0) / 2 in
// FP64, z / s0 = z conj(s0) / |s0|^2 with a reciprocal refined by one Newton
// step (the square root's step needs about 2^-45).  The
// step keeps the estimate's branch (Re >= 0).  |z| must lie well inside the
// FP32 range (the vertical wavenumbers: 1e-12 < |z| < 1e18).
GRTM_HD cx<double> csqrt_nr(cx<double> z)
{
    const float a = float(z.x), b = float(z.y);
    const float m = fsqrt_est(a * a + b * b);
    const float h = 0.5f * (m + fabsf(a));
    const float rh = frsq(h);  // t = sqrt(h) = h / sqrt(h), u = b / (2 t)
    const float t = h * rh;
    const float u = 0.5f * b * rh;
    const double s0x = a >= 0.f ? double(t) : double(fabsf(u));
    const double s0y = a >= 0.f ? double(u) : double(copysignf(t, b));
    const double n = s0x * s0x + s0y * s0y;
    double inv = drcp_est(n);
    inv = inv * (2.0 - n * inv);  // about 2^-52 from the 2^-26 estimate
    const double qx = (z.x * s0x + z.y * s0y) * inv, qy = (z.y * s0x - z.x * s0y) * inv;
    return {0.5 * (s0x + qx), 0.5 * (s0y + qy)};
}

GRTM_HD void sincos_t(double a, double *s, double *c) { sincos(a, s, c); }
GRTM_HD void sincos_t(float a, float *s, float *c) { sincosf(a, s, c); }

template <typename T> GRTM_HD cx<T> cexp(cx<T> z)
{
    T e = exp(z.x);
    T s, c;
    sincos_t(z.y, &s, &c);
    return {e * c, e * s};
}

template <typename T> GRTM_HD cx<T> clog(cx<T> z) { return {log(hypot(z.x, z.y)), atan2(z.y, z.x)}; }

// 2x2 complex matrix [[a, b], [c, d]]
template <typename T>
struct m2 {
    cx<T> a, b, c, d;
};

template <typename T> GRTM_HD m2<T> operator*(const m2<T> &p, const m2<T> &q)
{
    return {cfma(p.b, q.c, p.a * q.a), cfma(p.b, q.d, p.a * q.b), cfma(p.d, q.c, p.c * q.a), cfma(p.d, q.d, p.c * q.b)};
}
// r + p q and r - p q in fused multiply-adds
template <typename T> GRTM_HD m2<T> mmadd(const m2<T> &p, const m2<T> &q, const m2<T> &r)
{
    return {cfma(p.b, q.c, cfma(p.a, q.a, r.a)), cfma(p.b, q.d, cfma(p.a, q.b, r.b)),
            cfma(p.d, q.c, cfma(p.c, q.a, r.c)), cfma(p.d, q.d, cfma(p.c, q.b, r.d))};
}
template <typename T> GRTM_HD m2<T> mmsub(const m2<T> &p, const m2<T> &q, const m2<T> &r)
{
    return {cfms(p.b, q.c, cfms(p.a, q.a, r.a)), cfms(p.b, q.d, cfms(p.a, q.b, r.b)),
            cfms(p.d, q.c, cfms(p.c, q.a, r.c)), cfms(p.d, q.d, cfms(p.c, q.b, r.d))};
}
template <typename T> GRTM_HD m2<T> operator+(const m2<T> &p, const m2<T> &q)
{
    return {p.a + q.a, p.b + q.b, p.c + q.c, p.d + q.d};
}
template <typename T> GRTM_HD m2<T> operator-(const m2<T> &p, const m2<T> &q)
{
    return {p.a - q.a, p.b - q.b, p.c - q.c, p.d - q.d};
}
template <typename T> GRTM_HD m2<T> mscale(T s, const m2<T> &p) { return {s * p.a, s * p.b, s * p.c, s * p.d}; }
template <typename T> GRTM_HD m2<T> ident()
{
    return {C<T>(1, 0), C<T>(0, 0), C<T>(0, 0), C<T>(1, 0)};
}
// diag(l) * p * diag(r)
template <typename T> GRTM_HD m2<T> dmd(cx<T> l0, cx<T> l1, const m2<T> &p, cx<T> r0, cx<T> r1)
{
    return {l0 * p.a * r0, l0 * p.b * r1, l1 * p.c * r0, l1 * p.d * r1};
}
// diag(l) * p
template <typename T> GRTM_HD m2<T> dm(cx<T> l0, cx<T> l1, const m2<T> &p)
{
    return {l0 * p.a, l0 * p.b, l1 * p.c, l1 * p.d};
}
// p * diag(r)
template <typename T> GRTM_HD m2<T> md(const m2<T> &p, cx<T> r0, cx<T> r1)
{
    return {p.a * r0, p.b * r1, p.c * r0, p.d * r1};
}
// Closed-form inverse adj(p)/det(p).
template <typename T> GRTM_HD m2<T> minv(const m2<T> &p)
{
    cx<T> di = recip(cfms(p.b, p.c, p.a * p.d));
    return {p.d * di, -(p.b * di), -(p.c * di), p.a * di};
}
template <typename T> GRTM_HD cx<T> mdet(const m2<T> &p) { return cfms(p.b, p.c, p.a * p.d); }
// p * v for a 2-vector
template <typename T> GRTM_HD void mv(const m2<T> &p, cx<T> v0, cx<T> v1, cx<T> &o0, cx<T> &o1)
{
    o0 = cfma(p.b, v1, p.a * v0);
    o1 = cfma(p.d, v1, p.c * v0);
}

}  // namespace grtm
