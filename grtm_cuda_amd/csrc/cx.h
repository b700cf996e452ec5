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
template <typename T> GRTM_HD T norm2(cx<T> a) { return a.x * a.x + a.y * a.y; }

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
    return {p.a * q.a + p.b * q.c, p.a * q.b + p.b * q.d, p.c * q.a + p.d * q.c, p.c * q.b + p.d * q.d};
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
// p * diag(r)
template <typename T> GRTM_HD m2<T> md(const m2<T> &p, cx<T> r0, cx<T> r1)
{
    return {p.a * r0, p.b * r1, p.c * r0, p.d * r1};
}
// Closed-form inverse adj(p)/det(p).
template <typename T> GRTM_HD m2<T> minv(const m2<T> &p)
{
    cx<T> di = recip(p.a * p.d - p.b * p.c);
    return {p.d * di, -(p.b * di), -(p.c * di), p.a * di};
}
template <typename T> GRTM_HD cx<T> mdet(const m2<T> &p) { return p.a * p.d - p.b * p.c; }
// p * v for a 2-vector
template <typename T> GRTM_HD void mv(const m2<T> &p, cx<T> v0, cx<T> v1, cx<T> &o0, cx<T> &o1)
{
    o0 = p.a * v0 + p.b * v1;
    o1 = p.c * v0 + p.d * v1;
}

}  // namespace grtm
