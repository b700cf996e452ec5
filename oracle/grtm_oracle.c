/*
 * grtm_oracle.c -- FP64 CPU restatement of the reference's Green's-function
 * path.  TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or
 * shipped).  Every function cites the part of cusrc/popm/func/green.cu it
 * restates.  Complex semantics follow cusrc/scilib/cucomplex.hpp:
 * sqrt = polar(sqrt|z|, arg/2) (:945-958), log = (log|z|, atan2) (:925-930),
 * exp = e^re (cos, sin) (:963-985).
 *
 * Arithmetic type.  The default build computes in FP64, as the reference does
 * (float32 is double, cusrc/scilib/mathematics.h:26).  Built with -DOR_QUAD
 * (libgrtm_oracle_q.so) every cell, k-sum, post-processing and inverse-FFT
 * value is computed in IEEE binary128 (__float128, 113-bit significand)
 * instead: the same restatement with 2^-60 smaller rounding, which decides
 * what the FP64 answer should be where the formulation is ill-conditioned
 * (the lowest frequency rows).  The grid (k spacing, per-row k counts,
 * frequencies) and the input parsing stay in FP64 in both builds, so both
 * evaluate the same (w, k) points as the reference.  The API takes and
 * returns double in both builds.
 */
#include "grtm_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef OR_QUAD
#include <quadmath.h>
typedef __float128 or_real;
typedef __complex128 cd;
#define CRE crealq
#define CIM cimagq
#define CABS cabsq
#define CARG cargq
#define CONJ conjq
#define RSQRT sqrtq
#define RCOS cosq
#define RSIN sinq
#define REXP expq
#define RLOG logq
#define RFABS fabsq
#define RJN jnq
#define RPOW2(e) ldexpq((__float128)1, (e))
#define R_PI M_PIq
static inline cd MKC(or_real re, or_real im)
{
    cd z;
    __real__ z = re;
    __imag__ z = im;
    return z;
}
#else
typedef double or_real;
typedef double complex cd;
#define CRE creal
#define CIM cimag
#define CABS cabs
#define CARG carg
#define CONJ conj
#define RSQRT sqrt
#define RCOS cos
#define RSIN sin
#define REXP exp
#define RLOG log
#define RFABS fabs
#define RJN jn
#define RPOW2(e) ldexp(1.0, (e))
#define R_PI 3.1415926535897932384626433
#define MKC CMPLX
#endif

/* OR_PI: the FP64 constant of the grid (green.cu:2065-2080, 2146);
 * R_PI: pi in the arithmetic type, for cell values and the transforms. */
#define OR_PI 3.1415926535897932384626433
#define OR_PI2 (OR_PI * 2.0)
#define R_PI2 (R_PI * 2)

/* ------------------------------------------------------------------ */
/* complex helpers (cucomplex.hpp semantics)                           */
/* ------------------------------------------------------------------ */
static cd c_sqrt(cd z)
{
    or_real m = RSQRT(CABS(z));
    or_real t = CARG(z) * 0.5;
    return MKC(m * RCOS(t), m * RSIN(t));
}
static cd c_log(cd z) { return MKC(RLOG(CABS(z)), CARG(z)); }
static cd c_exp(cd z)
{
    or_real e = REXP(CRE(z));
    return MKC(e * RCOS(CIM(z)), e * RSIN(CIM(z)));
}

/* 2x2 helpers (green.cu:388-423) */
static void m2_inv(cd a[2][2])
{
    cd d = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    cd b00 = a[0][0], b01 = a[0][1], b10 = a[1][0], b11 = a[1][1];
    a[0][0] = b11 / d;
    a[1][1] = b00 / d;
    a[0][1] = -b01 / d;
    a[1][0] = -b10 / d;
}
static void m2_mul(cd a[2][2], cd b[2][2], cd c[2][2])
{
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) c[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j];
}

/* ------------------------------------------------------------------ */
/* per-cell state (field meanings of cusrc/popm/popm.h:16-136)        */
/* ------------------------------------------------------------------ */
typedef struct { or_real h; cd a, b, u, r, v, x; } OPar;
typedef struct { cd tdd1[2][2], tuu1[2][2], rdu1[2][2], rud1[2][2], tdd0, tuu0, rdu0, rud0; } OMrt;
typedef struct { cd rud1[2][2], rdu1[2][2], tu1[2][2], td1[2][2], rud0, rdu0, tu0, td0; } OGrt;
typedef struct {
    cd su0[2], su1[2], su2[2], sd0[2], sd1[2], sd2[2];
    cd shu1, shd1, shu2, shd2, gshu, gshd;
    cd gu0[2], gu1[2], gd0[2], gd1[2];
} OSrc;
typedef struct { cd a0[2], a1[2], a2[2], b1, b2, a0t[2], a1t[2], a2t[2], b1t, b2t; } OPpar;

/* Dispersed velocity v0*(1 + ln(w/2pi)/(pi Q) + i/(2Q))  (green.cu:177-180) */
static cd disp_vel(cd w, or_real v0, or_real q)
{
    cd f = 1.0 + c_log(w / R_PI2) / (R_PI * q) + I / (2.0 * q);
    return v0 * f;
}

/* The same dispersion in FP64 for the grid (cal_kmax), as the reference
 * forms it; identical to disp_vel in the FP64 build. */
static double complex disp_vel_d(double complex w, double v0, double q)
{
    double complex f = 1.0 + CMPLX(log(cabs(w / OR_PI2)), carg(w / OR_PI2)) / (OR_PI * q) + I / (2.0 * q);
    return v0 * f;
}

void or_dispersion(double wr, double wi, double v0, double q, double *o)
{
    cd z = disp_vel(MKC(wr, wi), v0, q);
    o[0] = CRE(z);
    o[1] = CIM(z);
}

/* set_pars vertical wavenumbers (green.cu:184-194) */
static void set_par(OPar *p, or_real k, cd w)
{
    cd wa = w / p->a, wb = w / p->b;
    p->r = c_sqrt(k * k - wa * wa);
    p->v = c_sqrt(k * k - wb * wb);
    if (CRE(p->r) < 0) p->r = -p->r;
    if (CRE(p->v) < 0) p->v = -p->v;
    p->x = k * k + p->v * p->v;
}

/* P-SV eigenvector matrix E (green.cu:196-214) */
static void psv_e(const OPar *p, or_real k, cd w, cd E[4][4])
{
    E[0][0] = p->a * k / w;
    E[0][1] = p->b * p->v / w;
    E[0][2] = E[0][0];
    E[0][3] = E[0][1];
    E[1][0] = p->a * p->r / w;
    E[1][1] = p->b * k / w;
    E[1][2] = -E[1][0];
    E[1][3] = -E[1][1];
    E[2][0] = -2.0 * p->a * p->u * k * p->r / w;
    E[2][1] = -p->b * p->u * p->x / w;
    E[2][2] = -E[2][0];
    E[2][3] = -E[2][1];
    E[3][0] = -p->a * p->u * p->x / w;
    E[3][1] = -2.0 * p->b * p->u * k * p->v / w;
    E[3][2] = E[3][0];
    E[3][3] = E[3][1];
}

/* Solve A X = B (4x4 complex) by Gaussian elimination with partial pivoting.
 * Stands in for cublasZmatinvBatched + cublasZgemmBatched (green.cu:1813-1816). */
static void solve4(cd A[4][4], cd B[4][4], cd X[4][4])
{
    cd a[4][4], b[4][4];
    memcpy(a, A, sizeof(a));
    memcpy(b, B, sizeof(b));
    for (int c = 0; c < 4; c++) {
        int p = c;
        or_real best = CABS(a[c][c]);
        for (int r = c + 1; r < 4; r++)
            if (CABS(a[r][c]) > best) { best = CABS(a[r][c]); p = r; }
        if (p != c) {
            for (int j = 0; j < 4; j++) {
                cd t = a[c][j]; a[c][j] = a[p][j]; a[p][j] = t;
                t = b[c][j]; b[c][j] = b[p][j]; b[p][j] = t;
            }
        }
        for (int r = c + 1; r < 4; r++) {
            cd f = a[r][c] / a[c][c];
            for (int j = c; j < 4; j++) a[r][j] -= f * a[c][j];
            for (int j = 0; j < 4; j++) b[r][j] -= f * b[c][j];
        }
    }
    for (int j = 0; j < 4; j++) {
        for (int r = 3; r >= 0; r--) {
            cd s = b[r][j];
            for (int q = r + 1; q < 4; q++) s -= a[r][q] * X[q][j];
            X[r][j] = s / a[r][r];
        }
    }
}

/* 0: solve E1 M = E2 directly; 1: form E1^-1 then multiply (the reference's
 * cuBLAS sequence). */
static int g_inverse_mode = 0;
void or_set_inverse_mode(int m) { g_inverse_mode = m; }

/* Test switch: the reference's SH factors of the tensor source in physics
 * mode too (a planted mistake for the reciprocity tests). */
static int g_tensor_sh_reference = 0;
void or_set_tensor_sh_reference(int on) { g_tensor_sh_reference = on; }

/* Interface between layer j (above) and j+1 (below):
 * SH (green.cu:339-353) and P-SV (green.cu:235-262 + 1813-1817 + 359-387). */
static void interface_rt(const OPar *pu, const OPar *pd, or_real k, cd w, OMrt *m)
{
    cd wa = pu->v * pu->u, wb = pd->v * pd->u;
    m->tdd0 = 2.0 * wa / (wa + wb);
    m->rdu0 = (wa - wb) / (wa + wb);
    m->rud0 = -m->rdu0;
    m->tuu0 = 2.0 * wb / (wa + wb);

    cd Eu[4][4], Ed[4][4], E1[4][4], E2[4][4], M[4][4];
    psv_e(pu, k, w, Eu);
    psv_e(pd, k, w, Ed);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 2; j++) {
            E1[i][j] = Ed[i][j];
            E1[i][j + 2] = -Eu[i][j + 2];
            E2[i][j] = Eu[i][j];
            E2[i][j + 2] = -Ed[i][j + 2];
        }
    if (g_inverse_mode) {
        cd Inv[4][4], Id[4][4];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) Id[i][j] = (i == j) ? 1.0 : 0.0;
        solve4(E1, Id, Inv);
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                cd s = 0;
                for (int q = 0; q < 4; q++) s += Inv[i][q] * E2[q][j];
                M[i][j] = s;
            }
    } else {
        solve4(E1, E2, M);
    }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            m->tdd1[i][j] = M[i][j];
            m->tuu1[i][j] = M[i + 2][j + 2];
            m->rdu1[i][j] = M[i + 2][j];
            m->rud1[i][j] = M[i][j + 2];
        }
}

static void load_lay(const double *l, OPar *p)
{
    p->h = l[0];
    p->a = MKC(l[1], l[2]);
    p->b = MKC(l[3], l[4]);
    p->u = MKC(l[5], l[6]);
}

void or_interface_rt(double k, double wr, double wi, const double *lu,
                     const double *ld, double *out)
{
    OPar pu, pd;
    double tu[7] = {0, lu[0], lu[1], lu[2], lu[3], lu[4], lu[5]};
    double td[7] = {0, ld[0], ld[1], ld[2], ld[3], ld[4], ld[5]};
    load_lay(tu, &pu);
    load_lay(td, &pd);
    cd w = MKC(wr, wi);
    set_par(&pu, k, w);
    set_par(&pd, k, w);
    OMrt m;
    interface_rt(&pu, &pd, k, w, &m);
    cd *blk[4] = {&m.tdd1[0][0], &m.rud1[0][0], &m.rdu1[0][0], &m.tuu1[0][0]};
    int o = 0;
    for (int b = 0; b < 4; b++)
        for (int e = 0; e < 4; e++) {
            out[o++] = CRE(blk[b][e]);
            out[o++] = CIM(blk[b][e]);
        }
    cd sh[4] = {m.tdd0, m.rud0, m.rdu0, m.tuu0};
    for (int e = 0; e < 4; e++) {
        out[o++] = CRE(sh[e]);
        out[o++] = CIM(sh[e]);
    }
}


/* Generalized R/T recursion over layers (green.cu:424-613). */
static void grt_sweeps(const OPar *par, const OMrt *mrt, int L, int s, or_real k, cd w,
                       OGrt *grt)
{
    cd E[4][4], a[2][2], b[2][2], exu[2];
    /* free surface: R_ud^0 = -(E21)^-1 E22, SH R^0 = 1 (green.cu:449-468) */
    psv_e(&par[0], k, w, E);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            a[i][j] = E[i + 2][j];
            b[i][j] = E[i + 2][j + 2];
        }
    m2_inv(a);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) grt[0].rud1[i][j] = -(a[i][0] * b[0][j] + a[i][1] * b[1][j]);
    grt[0].rud0 = 1.0;

    /* downward sweep, free surface -> source layer (green.cu:471-520) */
    for (int lay = 0; lay < s; lay++) {
        exu[0] = c_exp(-par[lay].r * par[lay].h);
        exu[1] = c_exp(-par[lay].v * par[lay].h);
        const OMrt *t = &mrt[lay];
        cd tem1 = grt[lay].rud0 * exu[1];
        cd tem2 = t->rdu0 * exu[1];
        cd tem3 = t->tdd0 * exu[1];
        grt[lay + 1].tu0 = t->tuu0 / (1. - tem1 * tem2);
        grt[lay + 1].rud0 = t->rud0 + tem3 * tem1 * grt[lay + 1].tu0;
        for (int i = 0; i < 2; i++) {
            for (int j = 0; j < 2; j++) {
                a[i][j] = -(t->rdu1[i][0] * exu[0] * grt[lay].rud1[0][j] +
                            t->rdu1[i][1] * exu[1] * grt[lay].rud1[1][j]) * exu[j];
                b[i][j] = (t->tdd1[i][0] * exu[0] * grt[lay].rud1[0][j] +
                           t->tdd1[i][1] * exu[1] * grt[lay].rud1[1][j]) * exu[j];
            }
            a[i][i] = 1.0 + a[i][i];
        }
        m2_inv(a);
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                grt[lay + 1].tu1[i][j] = a[i][0] * t->tuu1[0][j] + a[i][1] * t->tuu1[1][j];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                grt[lay + 1].rud1[i][j] = t->rud1[i][j] + b[i][0] * grt[lay + 1].tu1[0][j] +
                                          b[i][1] * grt[lay + 1].tu1[1][j];
    }
    /* half-space (green.cu:522-531) */
    {
        const OMrt *t = &mrt[L - 2];
        grt[L - 1].rdu0 = t->rdu0;
        grt[L - 1].td0 = t->tdd0;
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
                grt[L - 1].td1[i][j] = t->tdd1[i][j];
                grt[L - 1].rdu1[i][j] = t->rdu1[i][j];
            }
    }
    /* upward sweep, half-space -> source layer (green.cu:532-594) */
    for (int lay = L - 3; lay >= s; lay--) {
        exu[0] = c_exp(-par[lay + 1].r * par[lay + 1].h);
        exu[1] = c_exp(-par[lay + 1].v * par[lay + 1].h);
        const OMrt *t = &mrt[lay];
        cd tem1 = t->rud0 * exu[1];
        cd tem2 = grt[lay + 2].rdu0 * exu[1];
        cd tem3 = t->tuu0 * exu[1];
        grt[lay + 1].td0 = t->tdd0 / (1.0 - tem1 * tem2);
        grt[lay + 1].rdu0 = t->rdu0 + tem3 * tem2 * grt[lay + 1].td0;
        for (int i = 0; i < 2; i++) {
            for (int j = 0; j < 2; j++) {
                a[i][j] = -(t->rud1[i][0] * exu[0] * grt[lay + 2].rdu1[0][j] +
                            t->rud1[i][1] * exu[1] * grt[lay + 2].rdu1[1][j]) * exu[j];
                b[i][j] = (t->tuu1[i][0] * exu[0] * grt[lay + 2].rdu1[0][j] +
                           t->tuu1[i][1] * exu[1] * grt[lay + 2].rdu1[1][j]) * exu[j];
            }
            a[i][i] = 1. + a[i][i];
        }
        m2_inv(a);
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                grt[lay + 1].td1[i][j] = a[i][0] * t->tdd1[0][j] + a[i][1] * t->tdd1[1][j];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                grt[lay + 1].rdu1[i][j] = t->rdu1[i][j] + b[i][0] * grt[lay + 1].td1[0][j] +
                                          b[i][1] * grt[lay + 1].td1[1][j];
    }
}

/* Source terms.  Single force: green.cu:1271-1332; tensor: green.cu:1334-1413. */
static void source_terms(int compat, int src_type, const OPar *sp, const OGrt *grt, int s, or_real zs,
                         or_real k, cd w, OSrc *src)
{
    cd sexd[2], sexu[2], wru0[2][2], wrd0[2][2], wru1[2][2], wrd1[2][2];
    sexd[0] = c_exp(-sp->r * (sp->h - zs));
    sexd[1] = c_exp(-sp->v * (sp->h - zs));
    sexu[0] = c_exp(-sp->r * zs);
    sexu[1] = c_exp(-sp->v * zs);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            cd id = (i == j) ? 1.0 : 0.0;
            cd erd = sexd[i] * grt[s + 1].rdu1[i][j] * sexd[j];
            cd eru = sexu[i] * grt[s].rud1[i][j] * sexu[j];
            wru0[i][j] = erd - id;
            wrd0[i][j] = id - eru;
            wru1[i][j] = erd + id;
            wrd1[i][j] = id + eru;
        }
    cd ssh = 1. / (2. * sp->u * sp->v);
    if (src_type == 0) {
        cd tmp1 = sp->b / (2.0 * sp->u * sp->a * w * sp->r * sp->v);
        cd f00 = sp->b * sp->v * sp->r * tmp1;
        cd f01 = -sp->a * k * sp->r * tmp1;
        cd f10 = k * sp->b * sp->v * tmp1;
        cd f11 = -sp->a * sp->r * sp->v * tmp1;
        for (int i = 0; i < 2; i++) {
            src->gu0[i] = wru0[i][0] * f00 + wru0[i][1] * f01;
            src->gd0[i] = wrd0[i][0] * f00 + wrd0[i][1] * f01;
            src->gu1[i] = wru1[i][0] * f10 + wru1[i][1] * f11;
            src->gd1[i] = wrd1[i][0] * f10 + wrd1[i][1] * f11;
        }
        src->gshu = (1. + sexd[1] * grt[s + 1].rdu0 * sexd[1]) * ssh;
        src->gshd = (sexu[1] * grt[s].rud0 * sexu[1] + 1.) * ssh;
    } else {
        cd tmp1 = sp->b / (2.0 * sp->u * sp->a * w);
        cd s2[2], s1[2], s0[2];
        s2[0] = tmp1 * k / sp->r * k * sp->b;
        s2[1] = -k * tmp1 / sp->r * sp->a * sp->r;
        s1[0] = tmp1 / sp->v * (-2.0) * k * sp->v * sp->b;
        s1[1] = tmp1 / sp->v * (k * k + sp->v * sp->v) * sp->a;
        s0[0] = tmp1 * sp->r * sp->b;
        s0[1] = tmp1 * sp->a * (-k);
        for (int i = 0; i < 2; i++) {
            src->su0[i] = wru1[i][0] * s0[0] + wru1[i][1] * s0[1];
            src->sd0[i] = wrd1[i][0] * s0[0] + wrd1[i][1] * s0[1];
            src->su1[i] = wru0[i][0] * s1[0] + wru0[i][1] * s1[1];
            src->sd1[i] = wrd0[i][0] * s1[0] + wrd0[i][1] * s1[1];
            src->su2[i] = wru1[i][0] * s2[0] + wru1[i][1] * s2[1];
            src->sd2[i] = wrd1[i][0] * s2[0] + wrd1[i][1] * s2[1];
        }
        /* SH: the reference (green.cu:1408-1411) gives both orders the
         * symmetric factor 1 + R above the source and the antisymmetric
         * R - 1 below.  Physics mode: n = 1 (a jump in displacement, like
         * the P-SV s1) is antisymmetric on both sides, 1 - R above; n = 2
         * (a jump in traction, like s0 and s2) symmetric, 1 + R below. */
        cd erd0 = sexd[1] * grt[s + 1].rdu0 * sexd[1], eru0 = sexu[1] * grt[s].rud0 * sexu[1];
        int ref_sh = compat || g_tensor_sh_reference;
        src->shu1 = (ref_sh ? 1. + erd0 : 1. - erd0) * sp->v * ssh;
        src->shd1 = (eru0 - 1.) * sp->v * ssh;
        src->shu2 = (1. + erd0) * k * ssh;
        src->shd2 = (ref_sh ? eru0 - 1. : 1. + eru0) * k * ssh;
    }
}

static void diag_mul(cd m[2][2], const cd *dg, cd y[2][2], cd out[2][2])
{
    /* out = m * diag(dg) * y */
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) out[i][j] = m[i][0] * dg[0] * y[0][j] + m[i][1] * dg[1] * y[1][j];
}

/* Receiver propagation + source coupling.  green.cu:615-939 (single force) and
 * 942-1267 (tensor) share everything but the final coupling.  compat keeps the
 * reference's choices: Q4 (rexut from the first cell, :674-675), Q5 (extra v in
 * u0t, :720) and the r<s chain with yu0t formed from the already multiplied
 * yu0 (:842).  The r>s branch is physics mode only (see the check below). */
static int ypars(int compat, int src_type, const OPar *par, const OGrt *grt, const OSrc *src,
                 int L, int s, or_real zs, int rl, or_real zr, or_real k, cd w, cd par0_r,
                 cd par0_v, OPpar *pp)
{
    cd sexd[2], sexu[2], rexd[2], rexu[2], rexdt[2], rexut[2];
    cd tmp1[2][2], tmp2[2][2], yu_psv[2][2], yd_psv[2][2], y_sh;
    cd E[4][4], u[2][2], d[2][2], ut[2][2], dt[2][2];
    cd yu1[2][2], yd1[2][2], yu1t[2][2], yd1t[2][2], ch[2][2];
    cd u0 = 0, d0 = 0, u0t = 0, d0t = 0, yu0 = 0, yd0 = 0, yu0t = 0, yd0t = 0, ex[2];
    const OPar *sp = &par[s], *rp = &par[rl];
    memset(yu1, 0, sizeof(yu1));
    memset(yd1, 0, sizeof(yd1));
    memset(yu1t, 0, sizeof(yu1t));
    memset(yd1t, 0, sizeof(yd1t));

    int up;
    if ((rl == s && zr < zs) || rl < s) up = 1;
    else if ((rl == s && zr > zs) || rl > s) up = 0;
    else up = -1; /* zr == zs in the source layer: the reference writes nothing */
    if (up < 0) {
        if (compat) return OR_E_UNSUPPORTED;
        up = 1;
    }
    /* The reference's receiver search (green.cu:2092-2101, Q1) compares the
     * source's depth inside its layer, which is at most h[s], with each
     * thickness, so it stops at a layer <= s: its r > s code (green.cu:858-915)
     * is never reached from an input, and compat mode has no such geometry. */
    if (compat && rl > s) return OR_E_UNSUPPORTED;

    sexd[0] = c_exp(-sp->r * (sp->h - zs));
    sexd[1] = c_exp(-sp->v * (sp->h - zs));
    sexu[0] = c_exp(-sp->r * zs);
    sexu[1] = c_exp(-sp->v * zs);
    rexd[0] = c_exp(-rp->r * zr);
    rexd[1] = c_exp(-rp->v * zr);
    rexdt[0] = -rp->r * c_exp(-rp->r * zr);
    rexdt[1] = -rp->v * c_exp(-rp->v * zr);
    if (L - 1 == rl) {
        rexu[0] = rexu[1] = rexut[0] = rexut[1] = 0.0;
    } else {
        rexu[0] = c_exp(-rp->r * (rp->h - zr));
        rexu[1] = c_exp(-rp->v * (rp->h - zr));
        rexut[0] = rexu[0] * (compat ? par0_r : rp->r);
        rexut[1] = rexu[1] * (compat ? par0_v : rp->v);
    }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            tmp1[i][j] = sexd[i] * grt[s + 1].rdu1[i][j] * sexd[j];
            tmp2[i][j] = sexu[i] * grt[s].rud1[i][j] * sexu[j];
        }
    m2_mul(tmp1, tmp2, yu_psv);
    m2_mul(tmp2, tmp1, yd_psv);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            cd id = (i == j) ? 1.0 : 0.0;
            yu_psv[i][j] = id - yu_psv[i][j];
            yd_psv[i][j] = id - yd_psv[i][j];
        }
    m2_inv(yu_psv);
    m2_inv(yd_psv);
    y_sh = 1. / (1. - sexd[1] * grt[s + 1].rdu0 * sexd[1] * sexu[1] * grt[s].rud0 * sexu[1]);
    psv_e(rp, k, w, E);

    if (s == rl && zr <= zs) { /* green.cu:709-743 */
        rexd[0] = c_exp(-rp->r * zr);
        rexd[1] = c_exp(-rp->v * zr);
        rexu[0] = c_exp(-rp->r * (zs - zr));
        rexu[1] = c_exp(-rp->v * (zs - zr));
        rexdt[0] = -rp->r * c_exp(-rp->r * zr);
        rexdt[1] = -rp->v * c_exp(-rp->v * zr);
        rexut[0] = rp->r * c_exp(-rp->r * (zs - zr));
        rexut[1] = rp->v * c_exp(-rp->v * (zs - zr));
        u0 = rexd[1] * grt[s].rud0 * sexu[1] + rexu[1];
        u0t = rexdt[1] * grt[s].rud0 * sexu[1] + rexut[1] * (compat ? rp->v : 1.0);
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
                u[i][j] = E[i][0] * rexd[0] * grt[s].rud1[0][j] * sexu[j] +
                          E[i][1] * rexd[1] * grt[s].rud1[1][j] * sexu[j] + E[i][j + 2] * rexu[j];
                ut[i][j] = E[i][0] * rexdt[0] * grt[s].rud1[0][j] * sexu[j] +
                           E[i][1] * rexdt[1] * grt[s].rud1[1][j] * sexu[j] + E[i][j + 2] * rexut[j];
            }
        yu0 = u0 * y_sh;
        yu0t = u0t * y_sh;
        m2_mul(u, yu_psv, yu1);
        m2_mul(ut, yu_psv, yu1t);
    } else if (s == rl && zr > zs) { /* green.cu:744-774 */
        rexd[0] = c_exp(-sp->r * (zr - zs));
        rexd[1] = c_exp(-sp->v * (zr - zs));
        rexu[0] = c_exp(-sp->r * (sp->h - zr));
        rexu[1] = c_exp(-sp->v * (sp->h - zr));
        rexdt[0] = -rp->r * rexd[0];
        rexdt[1] = -rp->v * rexd[1];
        rexut[0] = rp->r * rexu[0];
        rexut[1] = rp->v * rexu[1];
        d0 = rexd[1] + rexu[1] * grt[s + 1].rdu0 * sexd[1];
        d0t = rexdt[1] + rexut[1] * grt[s + 1].rdu0 * sexd[1];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
                d[i][j] = E[i][j] * rexd[j] + E[i][2] * rexu[0] * grt[s + 1].rdu1[0][j] * sexd[j] +
                          E[i][3] * rexu[1] * grt[s + 1].rdu1[1][j] * sexd[j];
                dt[i][j] = E[i][j] * rexdt[j] + E[i][2] * rexut[0] * grt[s + 1].rdu1[0][j] * sexd[j] +
                           E[i][3] * rexut[1] * grt[s + 1].rdu1[1][j] * sexd[j];
            }
        yd0 = d0 * y_sh;
        yd0t = d0t * y_sh;
        m2_mul(d, yd_psv, yd1);
        m2_mul(dt, yd_psv, yd1t);
    } else if (rl < s) { /* green.cu:775-857 */
        ex[0] = c_exp(-rp->r * rp->h);
        ex[1] = c_exp(-rp->v * rp->h);
        u0 = rexd[1] * grt[rl].rud0 * ex[1] + rexu[1];
        u0t = rexdt[1] * grt[rl].rud0 * ex[1] + rexut[1];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
                u[i][j] = E[i][0] * rexd[0] * grt[rl].rud1[0][j] * ex[j] +
                          E[i][1] * rexd[1] * grt[rl].rud1[1][j] * ex[j] + E[i][j + 2] * rexu[j];
                ut[i][j] = E[i][0] * rexdt[0] * grt[rl].rud1[0][j] * ex[j] +
                           E[i][1] * rexdt[1] * grt[rl].rud1[1][j] * ex[j] + E[i][j + 2] * rexut[j];
            }
        /* chain T_u[r+1] L_{r+1} ... T_u[s] sexu (I - ERER)^-1 */
        for (int lay = s - 1; lay >= rl; lay--) {
            if (lay == s - 1) {
                yu0 = grt[lay + 1].tu0 * sexu[1] * y_sh;
                diag_mul((cd(*)[2])grt[lay + 1].tu1, sexu, yu_psv, yu1);
            } else {
                ex[0] = c_exp(-par[lay + 1].r * par[lay + 1].h);
                ex[1] = c_exp(-par[lay + 1].v * par[lay + 1].h);
                yu0 = grt[lay + 1].tu0 * ex[1] * yu0;
                diag_mul((cd(*)[2])grt[lay + 1].tu1, ex, yu1, ch);
                memcpy(yu1, ch, sizeof(ch));
            }
        }
        cd chain0 = yu0;
        yu0 = u0 * chain0;
        yu0t = u0t * (compat ? yu0 : chain0);
        memcpy(ch, yu1, sizeof(ch));
        m2_mul(u, ch, yu1);
        m2_mul(ut, ch, yu1t);
    } else { /* rl > s (physics mode only), green.cu:858-915 */
        ex[0] = c_exp(-rp->r * rp->h);
        ex[1] = c_exp(-rp->v * rp->h);
        int has_below = (rl + 1 <= L - 1);
        cd rd0 = has_below ? grt[rl + 1].rdu0 : 0.0;
        d0 = rexu[1] * rd0 * ex[1] + rexd[1];
        d0t = rexut[1] * rd0 * ex[1] + rexdt[1];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
                cd r0j = has_below ? grt[rl + 1].rdu1[0][j] : 0.0;
                cd r1j = has_below ? grt[rl + 1].rdu1[1][j] : 0.0;
                d[i][j] = E[i][j] * rexd[j] + E[i][2] * rexu[0] * r0j * ex[j] +
                          E[i][3] * rexu[1] * r1j * ex[j];
                dt[i][j] = E[i][j] * rexdt[j] + E[i][2] * rexut[0] * r0j * ex[j] +
                           E[i][3] * rexut[1] * r1j * ex[j];
            }
        yd0 = grt[s + 1].td0 * sexd[1] * y_sh;
        diag_mul((cd(*)[2])grt[s + 1].td1, sexd, yd_psv, yd1);
        for (int lay = s + 2; lay <= rl; lay++) {
            ex[0] = c_exp(-par[lay - 1].r * par[lay - 1].h);
            ex[1] = c_exp(-par[lay - 1].v * par[lay - 1].h);
            yd0 = grt[lay].td0 * ex[1] * yd0;
            diag_mul((cd(*)[2])grt[lay].td1, ex, yd1, ch);
            memcpy(yd1, ch, sizeof(ch));
        }
        cd chain0 = yd0;
        yd0 = d0 * chain0;
        yd0t = d0t * chain0;
        memcpy(ch, yd1, sizeof(ch));
        m2_mul(d, ch, yd1);
        m2_mul(dt, ch, yd1t);
    }

    memset(pp, 0, sizeof(*pp));
    cd(*Y)[2] = up ? yu1 : yd1;
    cd(*Yt)[2] = up ? yu1t : yd1t;
    cd y0 = up ? yu0 : yd0, y0t = up ? yu0t : yd0t;
    if (src_type == 0) { /* green.cu:918-938 */
        const cd *g0 = up ? src->gu0 : src->gd0, *g1 = up ? src->gu1 : src->gd1;
        cd gsh = up ? src->gshu : src->gshd;
        pp->b1 = y0 * gsh;
        pp->b1t = y0t * gsh;
        for (int i = 0; i < 2; i++) {
            pp->a0[i] = Y[i][0] * g0[0] + Y[i][1] * g0[1];
            pp->a1[i] = Y[i][0] * g1[0] + Y[i][1] * g1[1];
            pp->a0t[i] = Yt[i][0] * g0[0] + Yt[i][1] * g0[1];
            pp->a1t[i] = Yt[i][0] * g1[0] + Yt[i][1] * g1[1];
        }
    } else { /* green.cu:1239-1266 */
        const cd *s0 = up ? src->su0 : src->sd0, *s1 = up ? src->su1 : src->sd1,
                 *s2 = up ? src->su2 : src->sd2;
        cd sh1 = up ? src->shu1 : src->shd1, sh2 = up ? src->shu2 : src->shd2;
        pp->b1 = y0 * sh1;
        pp->b2 = y0 * sh2;
        pp->b1t = y0t * sh1;
        pp->b2t = y0t * sh2;
        for (int i = 0; i < 2; i++) {
            pp->a0[i] = Y[i][0] * s0[0] + Y[i][1] * s0[1];
            pp->a1[i] = Y[i][0] * s1[0] + Y[i][1] * s1[1];
            pp->a2[i] = Y[i][0] * s2[0] + Y[i][1] * s2[1];
            pp->a0t[i] = Yt[i][0] * s0[0] + Yt[i][1] * s0[1];
            pp->a1t[i] = Yt[i][0] * s1[0] + Yt[i][1] * s1[1];
            pp->a2t[i] = Yt[i][0] * s2[0] + Yt[i][1] * s2[1];
        }
    }
    return OR_OK;
}

/* Full per-cell chain: set_pars .. get_ypars (green.cu:1799-1855). */
static int cell_ppar(int compat, int src_type, or_real k, cd w, int L, const OPar *lay0,
                     int s, or_real zs, int rl, or_real zr, cd par0_r, cd par0_v, OPpar *pp)
{
    OPar par[OR_MAX_LAYERS];
    OMrt mrt[OR_MAX_LAYERS];
    OGrt grt[OR_MAX_LAYERS + 1];
    OSrc src;
    memset(grt, 0, sizeof(grt));
    for (int j = 0; j < L; j++) {
        par[j] = lay0[j];
        set_par(&par[j], k, w);
    }
    for (int j = 0; j < L - 1; j++) interface_rt(&par[j], &par[j + 1], k, w, &mrt[j]);
    grt_sweeps(par, mrt, L, s, k, w, grt);
    source_terms(compat, src_type, &par[s], grt, s, zs, k, w, &src);
    return ypars(compat, src_type, par, grt, &src, L, s, zs, rl, zr, k, w, par0_r, par0_v, pp);
}

int or_cell(int compat, int src_type, double k, double wr, double wi, int n_layer,
            const double *lay, int s_layer, double zs, int r_layer, double zr,
            const double *par0_rv, double *out)
{
    OPar l0[OR_MAX_LAYERS];
    if (n_layer < 2 || n_layer > OR_MAX_LAYERS) return OR_E_INPUT;
    for (int j = 0; j < n_layer; j++) load_lay(lay + 7 * j, &l0[j]);
    cd p0r = par0_rv ? MKC(par0_rv[0], par0_rv[1]) : 0.0;
    cd p0v = par0_rv ? MKC(par0_rv[2], par0_rv[3]) : 0.0;
    OPpar pp;
    int st = cell_ppar(compat, src_type, k, MKC(wr, wi), n_layer, l0, s_layer, zs, r_layer, zr,
                       p0r, p0v, &pp);
    if (st) return st;
    cd v[18] = {pp.a0[0], pp.a0[1], pp.a1[0], pp.a1[1], pp.a2[0], pp.a2[1], pp.b1, pp.b2,
                pp.a0t[0], pp.a0t[1], pp.a1t[0], pp.a1t[1], pp.a2t[0], pp.a2t[1], pp.b1t, pp.b2t,
                0, 0};
    for (int i = 0; i < 18; i++) {
        out[2 * i] = CRE(v[i]);
        out[2 * i + 1] = CIM(v[i]);
    }
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* Bessel family                                                       */
/* ------------------------------------------------------------------ */
/* Abramowitz & Stegun 9.4.1-9.4.6 polynomial forms, as used by the reference
 * (green.cu:1418-1465).  The coefficients are the published table values. */
static or_real as_j0(or_real x)
{
    static const or_real a[7] = {1.00000000, -2.24999970, 1.26562080, -0.31638660,
                                0.04444790, -0.00394440, 0.00021000};
    static const or_real b[7] = {0.79788456, -0.00000077, -0.00552740, -0.00009512,
                                0.00137237, -0.00072805, 0.00014476};
    static const or_real c[7] = {-0.78539816, -0.04166397, -0.00003954, 0.00262573,
                                -0.00054125, -0.00029333, 0.00013558};
    or_real ax = RFABS(x);
    if (ax <= 3) {
        or_real y = (ax / 3.0) * (ax / 3.0);
        return a[0] + y * (a[1] + y * (a[2] + y * (a[3] + y * (a[4] + y * (a[5] + y * a[6])))));
    }
    or_real t = 3.0 / ax;
    or_real f = b[0] + t * (b[1] + t * (b[2] + t * (b[3] + t * (b[4] + t * (b[5] + t * b[6])))));
    or_real th = ax + c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * (c[4] + t * (c[5] + t * c[6])))));
    return f * RCOS(th) / RSQRT(ax);
}
static or_real as_j1(or_real x)
{
    static const or_real a[7] = {0.50000000, -0.56249985, 0.21093573, -0.03954289,
                                0.00443319, -0.00031761, 0.00001109};
    static const or_real b[7] = {0.79788456, 0.00000156, 0.01659667, 0.00017105,
                                -0.00249511, 0.00113653, -0.00020033};
    static const or_real c[7] = {-2.35619449, 0.12499612, 0.00005650, -0.00637879,
                                0.00074348, 0.00079824, -0.00029166};
    or_real ax = RFABS(x);
    if (ax <= 3.0) {
        or_real y = (ax / 3.0) * (ax / 3.0);
        return x * (a[0] + y * (a[1] + y * (a[2] + y * (a[3] + y * (a[4] + y * (a[5] + y * a[6]))))));
    }
    or_real t = 3.0 / ax;
    or_real f = b[0] + t * (b[1] + t * (b[2] + t * (b[3] + t * (b[4] + t * (b[5] + t * b[6])))));
    or_real th = ax + c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * (c[4] + t * (c[5] + t * c[6])))));
    return (x >= 0 ? 1.0 : -1.0) * f * RCOS(th) / RSQRT(ax);
}
/* J_n for integer n: upward recurrence above n, Miller's downward recurrence
 * with renormalisation below (green.cu:1467-1529; iacc 40, bigno 1e10). */
static or_real ref_jn(or_real x, int n)
{
    const int iacc = 40;
    const or_real bigno = 1e10, bigni = 1e-10;
    int nn = n;
    or_real res;
    if (n == 0) res = as_j0(x);
    else if (n == 1) res = as_j1(x);
    else {
        if (n < 0) n = -n;
        or_real ax = RFABS(x);
        if (ax == 0) res = 0.0;
        else if (ax > n) {
            or_real tox = 2.0 / ax, bjm = as_j0(ax), bj = as_j1(ax), bjp;
            for (int j = 0; j < n - 1; j++) {
                bjp = (or_real)(j + 1) * tox * bj - bjm;
                bjm = bj;
                bj = bjp;
            }
            res = bj;
        } else {
            or_real tox = 2.0 / ax, bjm, bjp = 0.0, bj = 1.0, sum = 0.0;
            int m = 2 * ((n + (int)(RSQRT((or_real)(iacc * n)))) / 2), jsum = 0;
            res = 0.0;
            for (int j = m; j >= 1; j--) {
                bjm = (or_real)j * tox * bj - bjp;
                bjp = bj;
                bj = bjm;
                if (RFABS(bj) > bigno) {
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
static or_real ref_djn(or_real x, int n) /* green.cu:1530-1541 */
{
    const or_real bigni = 1e-10;
    if (RFABS(x) < bigni) return (ref_jn(bigni, n) - ref_jn(0, n)) / bigni;
    return (or_real)n * ref_jn(x, n) / x - ref_jn(x, n + 1);
}
static or_real ref_ddjn(or_real x, int n) /* green.cu:1543-1554 (Q7: not J'') */
{
    const or_real bigni = 1e-10;
    if (RFABS(x) < bigni) return (ref_djn(bigni, n) - ref_djn(0, n)) / bigni;
    return (or_real)n * ref_jn(x, n - 1) / x - ref_jn(x, n) * (n + n * n + x * x) / x / x;
}
/* physics mode: libm jn and exact derivative identities */
static or_real phy_jn(or_real x, int n) { return RJN(n, x); }
static or_real phy_djn(or_real x, int n) { return 0.5 * (RJN(n - 1, x) - RJN(n + 1, x)); }
static or_real phy_ddjn(or_real x, int n)
{
    return 0.25 * (RJN(n - 2, x) - 2.0 * RJN(n, x) + RJN(n + 2, x));
}

static or_real bes(int mode, int which, int n, or_real x)
{
    if (mode == 0) return which == 0 ? ref_jn(x, n) : which == 1 ? ref_djn(x, n) : ref_ddjn(x, n);
    return which == 0 ? phy_jn(x, n) : which == 1 ? phy_djn(x, n) : phy_ddjn(x, n);
}
double or_bessel(int mode, int which, int n, double x) { return (double)bes(mode, which, n, x); }

/* Integrand columns (green.cu:1558-1621 single, 1626-1716 tensor). */
static void integrand(int compat, int src_type, const OPpar *p, or_real k, or_real dist, cd *c)
{
    or_real x = dist * k;
    int m = compat ? 0 : 1;
    or_real j0 = bes(m, 0, 0, x), j1 = bes(m, 0, 1, x), j2 = bes(m, 0, 2, x);
    or_real j0d = bes(m, 1, 0, x), j1d = bes(m, 1, 1, x), j2d = bes(m, 1, 2, x);
    or_real j0dd = bes(m, 2, 0, x), j1dd = bes(m, 2, 1, x), j2dd = bes(m, 2, 2, x);
    or_real r = dist;
    for (int i = 0; i < 30; i++) c[i] = 0.0;
    if (src_type == 0) {
        c[0] = p->b1 * j1 / r + p->a1[0] * k * j1d;
        c[1] = p->a0[0] * k * j0d;
        c[2] = p->a1[0] * j1 / r + p->b1 * k * j1d;
        c[3] = p->a1[1] * k * j1;
        c[4] = p->a0[1] * k * j0;
        c[5] = p->b1t * j1 / r + p->a1t[0] * k * j1d;
        c[6] = p->a0t[0] * k * j0d;
        c[7] = p->a1t[0] * j1 / r + p->b1t * k * j1d;
        c[8] = p->a1t[1] * k * j1;
        c[9] = p->a0t[1] * k * j0;
        /* d/dr; the reference omits 1/r on the b1 J1' term of column 10 */
        c[10] = -p->b1 * j1 / r / r + p->b1 * j1d * k / (compat ? 1.0 : r) + p->a1[0] * k * j1dd * k;
        c[11] = p->a0[0] * k * j0dd * k;
        c[12] = -p->a1[0] * j1 / r / r + p->a1[0] * j1d * k / r + p->b1 * k * j1dd * k;
        c[13] = p->a1[1] * k * j1d * k;
        c[14] = p->a0[1] * k * j0d * k;
        return;
    }
    /* tensor; Q9: compat takes a.[1] in the R n=0 columns and b2 in column 5 */
    int q = compat ? 1 : 0;
    cd b5 = compat ? p->b2 : p->b1;
    cd a2q = q ? p->a2[1] : p->a2[0], a0q = q ? p->a0[1] : p->a0[0];
    c[0] = p->b2 * 2.0 * j2 / r + p->a2[0] * k * j2d;
    c[1] = p->b1 * j1 / r + p->a1[0] * k * j1d;
    c[2] = a2q * k * j0d;
    c[3] = a0q * k * j0d;
    c[4] = p->b2 * k * j2d + 2.0 * p->a2[0] * j2 / r;
    c[5] = b5 * k * j1d + p->a1[0] * j1 / r;
    c[6] = p->a2[1] * k * j2;
    c[7] = p->a1[1] * k * j1;
    c[8] = p->a0[1] * k * j0;
    c[9] = p->a2[1] * k * j0;
    /* d/dr block: columns 10-19 in compat (the d/dz block is overwritten,
     * Q8), columns 20-29 in physics mode with the d/dz block kept at 10-19 */
    int o = compat ? 10 : 20;
    cd b5r = compat ? p->b2 : p->b1;
    c[o + 0] = -p->b2 * 2.0 * j2 / r / r + p->b2 * 2.0 * j2d * k / r + p->a2[0] * k * j2dd * k;
    c[o + 1] = -p->b1 * j1 / r / r + p->b1 * j1d * k / r + p->a1[0] * k * j1dd * k;
    c[o + 2] = a2q * k * j0dd * k;
    c[o + 3] = a0q * k * j0dd * k;
    c[o + 4] = p->b2 * k * j2dd * k - 2.0 * p->a2[0] * j2 / r / r + 2.0 * p->a2[0] * j2d * k / r;
    c[o + 5] = b5r * k * j1dd * k - p->a1[0] * j1 / r / r + p->a1[0] * j1d * k / r;
    c[o + 6] = p->a2[1] * k * j2d * k;
    c[o + 7] = p->a1[1] * k * j1d * k;
    c[o + 8] = p->a0[1] * k * j0d * k;
    c[o + 9] = p->a2[1] * k * j0d * k;
    if (!compat) {
        c[10] = p->b2t * 2.0 * j2 / r + p->a2t[0] * k * j2d;
        c[11] = p->b1t * j1 / r + p->a1t[0] * k * j1d;
        c[12] = p->a2t[0] * k * j0d;
        c[13] = p->a0t[0] * k * j0d;
        c[14] = p->b2t * k * j2d + 2.0 * p->a2t[0] * j2 / r;
        c[15] = p->b1t * k * j1d + p->a1t[0] * j1 / r;
        c[16] = p->a2t[1] * k * j2;
        c[17] = p->a1t[1] * k * j1;
        c[18] = p->a0t[1] * k * j0;
        c[19] = p->a2t[1] * k * j0;
    }
}

/* ------------------------------------------------------------------ */
/* model preparation: input.txt parsing (green.cu:1989-2101)           */
/* ------------------------------------------------------------------ */
/* Build the layer stack from layer tops, inserting the source interface
 * (green.cu:2001-2049) and locating source/receiver (green.cu:2081-2101). */
static int prepare_stack(int compat, int n_in, const double *top, const double *vp,
                         const double *vs, const double *rho, const double *qp,
                         const double *qs, double hs, double hr, or_run *run)
{
    if (n_in < 2 || n_in + 1 > OR_MAX_LAYERS) return OR_E_INPUT;
    int cnt = 0;
    run->split_done = 0;
#define PUSH(T, I)                                                                             \
    do {                                                                                       \
        run->h[cnt] = (T);                                                                     \
        run->a0[cnt] = vp[I];                                                                  \
        run->b0[cnt] = vs[I];                                                                  \
        run->rho[cnt] = rho[I];                                                                \
        run->qp[cnt] = qp[I];                                                                  \
        run->qs[cnt] = qs[I];                                                                  \
        cnt++;                                                                                 \
    } while (0)
    PUSH(top[0], 0);
    for (int i = 1; i < n_in; i++) {
        if (top[i - 1] < hs && top[i] > hs) {
            PUSH(hs, i - 1);
            run->split_done = 1;
        }
        PUSH(top[i], i);
    }
    if (!run->split_done) {
        /* Q2: the reference still counts an extra (uninitialised) layer */
        if (compat) return OR_E_UNSUPPORTED;
        if (hs > top[n_in - 1]) { /* source inside the half-space: split it */
            PUSH(hs, n_in - 1);
            run->split_done = 1;
        }
    }
#undef PUSH
    run->n_layer = cnt;
    int L = cnt;
    double tops[OR_MAX_LAYERS];
    for (int i = 0; i < L; i++) tops[i] = run->h[i];
    for (int i = 0; i < L - 1; i++) run->h[i] = run->h[i + 1] - run->h[i];
    run->hs = hs;
    run->hr = hr;
    run->s_layer = -1;
    run->r_layer = -1;
    if (!compat) { /* physics mode: locate from the layer tops */
        int s = 0, r = 0;
        for (int i = 0; i < L; i++) {
            if (tops[i] < hs) s = i;
            if (tops[i] < hr) r = i;
        }
        run->s_layer = s;
        run->zs = hs - tops[s];
        run->r_layer = r;
        run->zr = hr - tops[r];
        return s >= L - 1 ? OR_E_UNSUPPORTED : OR_OK;
    }
    double zs = hs, zr = hr;
    for (int i = 0; i < L; i++) {
        if (zs <= run->h[i] || i == L - 1) {
            run->s_layer = i;
            run->zs = zs;
            break;
        }
        zs -= run->h[i];
    }
    for (int i = 0; i < L; i++) {
        double key = compat ? zs : zr; /* Q1: the reference tests the source depth */
        if (key <= run->h[i] || (!compat && i == L - 1)) {
            run->r_layer = i;
            run->zr = zr;
            break;
        }
        zr -= run->h[i];
    }
    if (run->s_layer < 0 || run->r_layer < 0) return OR_E_INPUT;
    if (run->s_layer >= L - 1) return OR_E_UNSUPPORTED;
    return OR_OK;
}

int or_parse_input(const char *path, int compat, or_run *run)
{
    FILE *f = fopen(path, "r");
    if (!f) return OR_E_IO;
    memset(run, 0, sizeof(*run));
    double hs, hr;
    int n;
    double top[OR_MAX_LAYERS], vp[OR_MAX_LAYERS], vs[OR_MAX_LAYERS], rho[OR_MAX_LAYERS],
        qp[OR_MAX_LAYERS], qs[OR_MAX_LAYERS];
    int st = OR_OK;
    if (fscanf(f, "%lf", &hs) != 1 || fscanf(f, "%lf", &hr) != 1 || fscanf(f, "%d", &n) != 1) {
        st = OR_E_PARSE;
        goto done;
    }
    if (n < 2 || n + 1 > OR_MAX_LAYERS) {
        st = OR_E_INPUT;
        goto done;
    }
    for (int i = 0; i < n; i++) {
        if (fscanf(f, " %lf , %lf , %lf , %lf , %lf , %lf", &top[i], &vp[i], &vs[i], &rho[i], &qp[i],
                   &qs[i]) != 6) {
            st = OR_E_PARSE;
            goto done;
        }
    }
    if (fscanf(f, "%lf", &run->dt) != 1 || fscanf(f, "%lf", &run->dist) != 1 ||
        fscanf(f, "%d", &run->nf1) != 1 || fscanf(f, "%d", &run->nf2) != 1 ||
        fscanf(f, "%lf", &run->twin) != 1 || fscanf(f, "%d", &run->log2n) != 1 ||
        fscanf(f, "%lf", &run->t0) != 1 || fscanf(f, "%lf", &run->taper) != 1 ||
        fscanf(f, "%d", &run->n_ptam) != 1 || fscanf(f, "%d", &run->src_type) != 1) {
        st = OR_E_PARSE;
        goto done;
    }
    if (run->log2n < 2 || run->log2n > 24 || !(run->dt > 0) || !(run->dist > 0)) {
        st = OR_E_INPUT;
        goto done;
    }
    {
        double dt = run->dt, dist = run->dist, twin = run->twin, t0 = run->t0, taper = run->taper;
        int nf1 = run->nf1, nf2 = run->nf2, l2 = run->log2n, np = run->n_ptam, stp = run->src_type;
        st = prepare_stack(compat, n, top, vp, vs, rho, qp, qs, hs, hr, run);
        run->dt = dt;
        run->dist = dist;
        run->twin = twin;
        run->t0 = t0;
        run->taper = taper;
        run->nf1 = nf1;
        run->nf2 = nf2;
        run->log2n = l2;
        run->n_ptam = np;
        run->src_type = stp;
    }
done:
    fclose(f);
    return st;
}

/* cal_kmax (green.cu:108-162).  b = dispersed S velocities at the grid's
 * current frequency; w is the frequency the estimate is evaluated at. */
static double cal_kmax(double complex w, const double complex *b, const or_run *run, int nf)
{
    int s = run->s_layer, rl = run->r_layer;
    const double *h = run->h;
    double hs = run->hs, hr = run->hr, zs = run->zs, zr = run->zr;
    double pmin = 0, dzmax = 0, kmax, t;
#define MX(x, y) (((x) > (y)) ? (x) : (y))
#define MN(x, y) (((x) < (y)) ? (x) : (y))
#define AB(x) (((x) > 0) ? (x) : (-(x)))
    if (rl == s - 1) {
        pmin = MX(cabs(w / b[s]), cabs(w / b[s - 1]));
        dzmax = MN(zs, h[rl] - zr);
    } else if (rl < s - 1) {
        pmin = MX(cabs(w / b[s]), cabs(w / b[s - 1]));
        dzmax = MN(zs, hs - hr);
        for (int lay = rl; lay <= s - 2; lay++) {
            t = cabs(w / b[lay + 1]);
            if (pmin < t) pmin = t;
            if (dzmax > h[lay]) dzmax = h[lay];
        }
    } else if (rl == s) {
        pmin = cabs(w / b[s]);
        dzmax = AB(hr - hs);
    } else if (rl == s + 1) {
        pmin = MX(cabs(w / b[s]), cabs(w / b[s + 1]));
        dzmax = MN(h[s] - zs, zr);
    } else {
        pmin = MX(cabs(w / b[s]), cabs(w / b[s + 1]));
        dzmax = MN(h[s] - zs, AB(hr - hs));
        for (int lay = s + 1; lay < rl - 1; lay++) {
            t = cabs(w / b[lay]);
            if (pmin < t) pmin = t;
            if (dzmax > h[lay]) dzmax = h[lay];
        }
    }
    if (dzmax < 0.2) {
        if (nf < 10) kmax = MN(pmin + 2.0, 20. * pmin);
        else kmax = MN(pmin + 2.0, 2.0 * pmin);
    } else {
        kmax = sqrt((3.0 / dzmax) * (3.0 / dzmax) + pmin * pmin);
    }
    if ((kmax - pmin) < 0.5) kmax = pmin + 0.5;
#undef MX
#undef MN
#undef AB
    return kmax;
}

typedef struct {
    int nw;         /* Nw = N/2 */
    double df, omega_i, dk;
} grid_consts;

/* Grid constants (green.cu:2063-2080); omega_I = 2/T is the reference's. */
static void ref_consts(const or_run *run, grid_consts *g)
{
    int nh = 1 << (run->log2n - 1);
    int ns = nh * 2;
    double twin = ns * run->dt;
    double vmax = 0, vmin = 100000;
    for (int i = 0; i < run->n_layer; i++) {
        if (run->a0[i] > vmax) vmax = run->a0[i];
        if (run->b0[i] < vmin) vmin = run->b0[i];
    }
    double vavg = (vmax + vmin) * 0.5, maxd = 0;
    double Lw = (vmax * twin + maxd +
                 vmax / vavg * sqrt(maxd * maxd + (run->hs - run->hr) * (run->hs - run->hr)) + 100);
    Lw = Lw * 2;
    g->nw = nh;
    g->df = 1 / (ns * run->dt);
    g->omega_i = 2.0 / twin;
    g->dk = 2 * OR_PI / Lw;
}

static void layer_disp(const or_run *run, int compat, cd w, OPar *l0)
{
    for (int i = 0; i < run->n_layer; i++) {
        l0[i].h = run->h[i];
        l0[i].a = disp_vel(w, run->a0[i], run->qp[i]);
        l0[i].b = disp_vel(w, run->b0[i], run->qs[i]);
        /* Q3: the reference keeps mu real and unattenuated (green.cu:2047-2049) */
        l0[i].u = compat ? (cd)(run->b0[i] * run->b0[i] * run->rho[i]) : run->rho[i] * l0[i].b * l0[i].b;
    }
}

static int nk_for(const or_run *run, int compat, const grid_consts *g, int nw)
{
    double complex b[OR_MAX_LAYERS];
    double complex w = CMPLX(2 * OR_PI * g->df * nw, -g->omega_i);
    (void)compat;
    for (int i = 0; i < run->n_layer; i++) b[i] = disp_vel_d(w, run->b0[i], run->qs[i]);
    double km;
    if (nw < 3) {
        double complex wt = OR_PI2 * g->df * 3.0 - g->omega_i * I;
        km = cal_kmax(wt, b, run, nw);
    } else {
        km = cal_kmax(w, b, run, nw);
    }
    km = 1.1 * km;
    return (int)(km / g->dk) + 2;
}

int or_ref_grid(const or_run *run, int compat, double *dk, double *omega_i, int *nk_per_w,
                long long *nkw)
{
    grid_consts g;
    ref_consts(run, &g);
    if (dk) *dk = g.dk;
    if (omega_i) *omega_i = g.omega_i;
    long long tot = 0;
    for (int nw = 0; nw < g.nw; nw++) {
        int nk = nk_for(run, compat, &g, nw);
        if (nk_per_w) nk_per_w[nw] = nk;
        tot += nk;
    }
    if (nkw) *nkw = tot;
    return OR_OK;
}

/* k-sum of one frequency row (green.cu:1878-1903): dk * (f_0/2 + sum f_k). */
static int green_row_raw(const or_run *run, int compat, const grid_consts *g, int nw, cd par0_r,
                         cd par0_v, cd *sums)
{
    OPar l0[OR_MAX_LAYERS];
    cd w = MKC(2 * OR_PI * g->df * nw, -g->omega_i);
    layer_disp(run, compat, w, l0);
    int nk = nk_for(run, compat, g, nw);
    for (int c = 0; c < 30; c++) sums[c] = 0;
    for (int ik = 0; ik < nk; ik++) {
        or_real k = g->dk * ik;
        OPpar pp;
        int st = cell_ppar(compat, run->src_type ? 1 : 0, k, w, run->n_layer, l0, run->s_layer,
                           run->zs, run->r_layer, run->zr, par0_r, par0_v, &pp);
        if (st) return st;
        cd c[30];
        integrand(compat, run->src_type ? 1 : 0, &pp, k, run->dist, c);
        or_real wt = (ik == 0) ? g->dk / 2.0 : g->dk;
        for (int q = 0; q < 30; q++) sums[q] += c[q] * wt;
    }
    return OR_OK;
}

/* Phase, cosine taper and scale of one row (green.cu:1905-1921). */
static void post_row(int nw_tot, int iw, or_real wr, or_real t0, or_real taper, or_real df, cd *v,
                     int ncol)
{
    int wc = (int)((or_real)nw_tot * (1 - taper));
    or_real phi = wr * t0;
    or_real st0 = 1;
    if (iw > wc) st0 = 0.5 * (1. + RCOS((iw - wc + 1) * R_PI / (nw_tot - wc + 1)));
    or_real coef = df * nw_tot / R_PI;
    for (int j = 0; j < ncol; j++) {
        v[j] = v[j] * MKC(RCOS(phi), RSIN(phi));
        v[j] = st0 * v[j] * coef;
    }
}

static void par0_rv(const or_run *run, int compat, const grid_consts *g, cd *r0, cd *v0)
{
    OPar l0[OR_MAX_LAYERS];
    cd w0 = MKC(0.0, -g->omega_i);
    layer_disp(run, compat, w0, l0);
    OPar p = l0[run->r_layer];
    set_par(&p, 0.0, w0);
    *r0 = p.r;
    *v0 = p.v;
}

int or_green_rows(const char *iname, int compat, const int *rows, int nrows, double *out,
                  int nthreads)
{
    or_run run;
    int st = or_parse_input(iname, compat, &run);
    if (st) return st;
    grid_consts g;
    ref_consts(&run, &g);
    cd r0, v0;
    par0_rv(&run, compat, &g, &r0, &v0);
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (int i = 0; i < nrows; i++) {
        int nw = rows[i];
        cd s[30];
        if (nw < 0 || nw >= g.nw) {
            err |= 1;
            continue;
        }
        if (green_row_raw(&run, compat, &g, nw, r0, v0, s)) {
            err |= 2;
            continue;
        }
        post_row(g.nw, nw, 2 * OR_PI * g.df * nw, run.t0, run.taper, g.df, s, 30);
        for (int c = 0; c < 30; c++) {
            out[i * 60 + 2 * c] = CRE(s[c]);
            out[i * 60 + 2 * c + 1] = CIM(s[c]);
        }
    }
    return err ? OR_E_INPUT : OR_OK;
}

/* radix-2 inverse DFT (e^{+i}), unnormalised, in place */
static void ifft_inplace(cd *x, int n)
{
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            cd t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        or_real ang = 2 * R_PI / len;
        for (int i = 0; i < n; i += len)
            for (int j = 0; j < len / 2; j++) {
                cd wv = MKC(RCOS(ang * j), RSIN(ang * j));
                cd u = x[i + j], v = x[i + j + len / 2] * wv;
                x[i + j] = u + v;
                x[i + j + len / 2] = u - v;
            }
    }
}

/* Hermitian completion + inverse FFT + time scaling (green.cu:1925-1938,
 * 2222-2237) on an already post-processed half spectrum. */
static void hermitian_ifft(int nw, const cd *half, or_real dt, or_real omega_i, or_real t0, cd *full,
                           or_real *trace)
{
    int n = 2 * nw;
    for (int i = 0; i < nw; i++) full[i] = half[i];
    for (int iw = 0; iw < nw; iw++) {
        cd invs = (nw - iw >= nw) ? 0.0 : half[nw - iw];
        full[nw + iw] = CONJ(invs);
    }
    cd *x = (cd *)malloc(sizeof(cd) * n);
    memcpy(x, full, sizeof(cd) * n);
    ifft_inplace(x, n);
    for (int j = 0; j < n; j++) {
        or_real v = CRE(x[j]) / (or_real)nw / 2.0;
        trace[j] = v * REXP((or_real)j * dt * omega_i) * REXP(t0 * omega_i);
    }
    free(x);
}

/* post_row + hermitian_ifft of one raw half spectrum in the arithmetic type;
 * spec_post (optional) gets the post-processed half spectrum. */
static void post_ifft(int nw, const cd *raw, double df, double omega_i, double dt, double t0,
                      double taper, cd *spec_post, or_real *trace)
{
    cd *half = (cd *)malloc(sizeof(cd) * nw), *full = (cd *)malloc(sizeof(cd) * 2 * nw);
    for (int i = 0; i < nw; i++) {
        half[i] = raw[i];
        post_row(nw, i, 2 * OR_PI * df * i, t0, taper, df, &half[i], 1);
        if (spec_post) spec_post[i] = half[i];
    }
    hermitian_ifft(nw, half, dt, omega_i, t0, full, trace);
    free(half);
    free(full);
}

void or_post_ifft(int nw, const double *spec_raw, double df, double omega_i, double dt, double t0,
                  double taper, double *spec_post, double *trace)
{
    if (nw < 1) return;
    cd *raw = (cd *)malloc(sizeof(cd) * nw), *sp = (cd *)malloc(sizeof(cd) * nw);
    or_real *tr = (or_real *)malloc(sizeof(or_real) * 2 * nw);
    for (int i = 0; i < nw; i++) raw[i] = MKC(spec_raw[2 * i], spec_raw[2 * i + 1]);
    post_ifft(nw, raw, df, omega_i, dt, t0, taper, sp, tr);
    for (int i = 0; i < nw; i++) {
        if (spec_post) {
            spec_post[2 * i] = (double)CRE(sp[i]);
            spec_post[2 * i + 1] = (double)CIM(sp[i]);
        }
    }
    for (int j = 0; j < 2 * nw; j++) trace[j] = (double)tr[j];
    free(raw);
    free(sp);
    free(tr);
}

int or_green_files(const char *iname, const char *oname_w, const char *oname_t, int compat,
                   int nthreads)
{
    or_run run;
    int st = or_parse_input(iname, compat, &run);
    if (st) return st;
    grid_consts g;
    ref_consts(&run, &g);
    int nw = g.nw, n = 2 * nw;
    cd r0, v0;
    par0_rv(&run, compat, &g, &r0, &v0);
    cd *half = (cd *)calloc((size_t)nw * 30, sizeof(cd));
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (int iw = 0; iw < nw; iw++) {
        cd s[30];
        if (green_row_raw(&run, compat, &g, iw, r0, v0, s)) {
            err |= 1;
            continue;
        }
        post_row(nw, iw, 2 * OR_PI * g.df * iw, run.t0, run.taper, g.df, s, 30);
        for (int c = 0; c < 30; c++) half[(size_t)c * nw + iw] = s[c];
    }
    if (err) {
        free(half);
        return OR_E_INPUT;
    }
    cd *full = (cd *)malloc(sizeof(cd) * (size_t)n * 30);
    or_real *tr = (or_real *)malloc(sizeof(or_real) * (size_t)n * 30);
    for (int c = 0; c < 30; c++)
        hermitian_ifft(nw, half + (size_t)c * nw, run.dt, g.omega_i, run.t0, full + (size_t)c * n,
                       tr + (size_t)c * n);
    FILE *ft = fopen(oname_t, "w");
    FILE *fw = fopen(oname_w, "w");
    if (!ft || !fw) {
        if (ft) fclose(ft);
        if (fw) fclose(fw);
        free(half);
        free(full);
        free(tr);
        return OR_E_IO;
    }
    for (int j = 0; j < n; j++) {
        for (int c = 0; c < 30; c++) fprintf(ft, "%.18e,", (double)tr[(size_t)c * n + j]);
        fprintf(ft, "\n");
    }
    for (int j = 0; j < n; j++) {
        for (int c = 0; c < 30; c++)
            fprintf(fw, "%.18e,%.18e,", (double)CRE(full[(size_t)c * n + j]),
                    (double)CIM(full[(size_t)c * n + j]));
        fprintf(fw, "\n");
    }
    fclose(ft);
    fclose(fw);
    free(half);
    free(full);
    free(tr);
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* dense-grid multi-receiver traces                                    */
/* ------------------------------------------------------------------ */
static int trace_run(int compat, const double *lay, int n_in, double src_depth, double rcv_depth,
                     or_run *run)
{
    double top[OR_MAX_LAYERS], vp[OR_MAX_LAYERS], vs[OR_MAX_LAYERS], rho[OR_MAX_LAYERS],
        qp[OR_MAX_LAYERS], qs[OR_MAX_LAYERS];
    if (n_in < 2 || n_in + 1 > OR_MAX_LAYERS) return OR_E_INPUT;
    double z = 0;
    for (int i = 0; i < n_in; i++) {
        top[i] = z;
        z += lay[6 * i];
        vp[i] = lay[6 * i + 1];
        vs[i] = lay[6 * i + 2];
        rho[i] = lay[6 * i + 3];
        qp[i] = lay[6 * i + 4];
        qs[i] = lay[6 * i + 5];
    }
    memset(run, 0, sizeof(*run));
    return prepare_stack(compat, n_in, top, vp, vs, rho, qp, qs, src_depth, rcv_depth, run);
}

/* Z and R k-kernels for a source kind from the cell coefficients:
 * explosion = s0 - s2 tensor terms (columns 8-9 / 3-2), vertical force =
 * single-force n=0 terms (columns 4 / 1).  R uses J0' = -J1. */
static void zr_kernels(int compat, int kind, const OPpar *p, or_real k, cd *kz, cd *kr)
{
    if (kind == 0) {
        *kz = (p->a0[1] - p->a2[1]) * k;
        *kr = compat ? -(p->a0[1] - p->a2[1]) * k : -(p->a0[0] - p->a2[0]) * k;
    } else {
        *kz = p->a0[1] * k;
        *kr = -p->a0[0] * k;
    }
}

/* kern[((iw - w_lo) * nk + ik) * 2 + {0: Kz, 1: Kr}] in the arithmetic type */
static int kernels_cd(int compat, const double *lay, int n_in, double src_depth, double rcv_depth,
                      int src_kind, int nw, double dt, int nk, double kmax, double omega_i, int w_lo,
                      int w_hi, cd *kern, int nthreads)
{
    or_run run;
    int st = trace_run(compat, lay, n_in, src_depth, rcv_depth, &run);
    if (st) return st;
    double df = 1.0 / (2 * nw * dt), dk = kmax / nk;
    grid_consts g = {nw, df, omega_i, dk};
    cd r0, v0;
    par0_rv(&run, compat, &g, &r0, &v0);
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (int iw = w_lo; iw < w_hi; iw++) {
        OPar l0[OR_MAX_LAYERS];
        cd w = MKC(2 * OR_PI * df * iw, -omega_i);
        layer_disp(&run, compat, w, l0);
        for (int ik = 0; ik < nk; ik++) {
            double k = dk * ik;
            OPpar pp;
            if (cell_ppar(compat, src_kind == 0 ? 1 : 0, k, w, run.n_layer, l0, run.s_layer, run.zs,
                          run.r_layer, run.zr, r0, v0, &pp)) {
                err |= 1;
                break;
            }
            cd *o = kern + ((size_t)(iw - w_lo) * nk + ik) * 2;
            zr_kernels(compat, src_kind, &pp, k, &o[0], &o[1]);
        }
    }
    return err ? OR_E_UNSUPPORTED : OR_OK;
}

int or_kernels(int compat, const double *lay, int n_in, double src_depth, double rcv_depth,
               int src_kind, int nw, double dt, int nk, double kmax, double omega_i, int w_lo,
               int w_hi, double *kern, int nthreads)
{
    size_t n = (size_t)(w_hi > w_lo ? w_hi - w_lo : 0) * nk * 2;
    cd *kc = (cd *)malloc(sizeof(cd) * n + 1);
    int st = kernels_cd(compat, lay, n_in, src_depth, rcv_depth, src_kind, nw, dt, nk, kmax, omega_i,
                        w_lo, w_hi, kc, nthreads);
    if (!st)
        for (size_t i = 0; i < n; i++) {
            kern[2 * i] = (double)CRE(kc[i]);
            kern[2 * i + 1] = (double)CIM(kc[i]);
        }
    free(kc);
    return st;
}

/* Peak-trough averaging (PTAM) of the tail of a discrete k-sum.  The
 * reference reads n_ptam (green.cu:2059) and leaves the method as an unused
 * sketch (integral.cu:38-88): keep adding grid terms beyond k_max, record the
 * partial sums at the peaks and troughs of their oscillation, and average
 * neighbouring records repeatedly (Zhang, Chen & Chang 2003).  Here, for one
 * real part: partial sums S_j = s0 + t_0 + ... + t_j; an extremum at j where
 * t_j and t_{j+1} have opposite signs, recorded as the vertex of the parabola
 * through S_{j-1}, S_j, S_{j+1}; the first n records averaged pairwise n - 1
 * times, i.e. weighted by C(n-1, i) / 2^(n-1).  Fewer than n records in the
 * nt terms: the plain extended sum.  *found: records seen (at most n). */
static or_real ptam_part(or_real s0, const or_real *t, int stride, int nt, int n, int *found)
{
    or_real T = 0, est = 0, c = 1, all = 0;
    const or_real norm = RPOW2(-(n - 1));
    int m = 0;
    for (int j = 0; j < nt; j++) all += t[(size_t)j * stride];
    for (int j = 0; j + 1 < nt && m < n; j++) {
        const or_real a = t[(size_t)j * stride], b = t[(size_t)(j + 1) * stride];
        T += a;
        if ((a < 0 && b > 0) || (a > 0 && b < 0)) {
            est += c * (T - (a + b) * (a + b) / (8 * (b - a)));
            c = c * (n - 1 - m) / (m + 1);
            m++;
        }
    }
    *found = m;
    return m == n ? s0 + est * norm : s0 + all;
}

/* Terms of the PTAM tail of receiver r: (n/2 + 2) periods of J(k r) on the
 * grid's dk, at most nt. */
static int ptam_terms(int n, double r, double dk, int nt)
{
    const double want = ceil((0.5 * n + 2.0) * 2.0 * OR_PI / (r * dk));
    return want < nt ? (int)want : nt;
}

int or_ptam(double s0_re, double s0_im, const double *terms, int nt, int n, double *out, int *found)
{
    if (nt < 0 || n < 2 || n > OR_PTAM_MAX) return OR_E_INPUT;
    or_real *t = (or_real *)malloc(sizeof(or_real) * 2 * (size_t)nt + 1);
    for (int i = 0; i < 2 * nt; i++) t[i] = terms[i];
    int f0, f1;
    out[0] = (double)ptam_part(s0_re, t, 2, nt, n, &f0);
    out[1] = (double)ptam_part(s0_im, t + 1, 2, nt, n, &f1);
    if (found) *found = f0 < f1 ? f0 : f1;
    free(t);
    return OR_OK;
}

int or_traces(int compat, const double *lay, int n_in, double src_depth, double rcv_depth,
              const double *dists, int nr, int src_kind, int nw, double dt, int nk, double kmax,
              double omega_i, double t0, double taper, int w_lo, int w_hi, double *traces,
              double *spec, int nthreads)
{
    return or_traces_ptam(compat, lay, n_in, src_depth, rcv_depth, dists, nr, src_kind, nw, dt, nk, kmax,
                          omega_i, t0, taper, w_lo, w_hi, 0, 0, traces, spec, nthreads);
}

int or_traces_ptam(int compat, const double *lay, int n_in, double src_depth, double rcv_depth,
                   const double *dists, int nr, int src_kind, int nw, double dt, int nk, double kmax,
                   double omega_i, double t0, double taper, int w_lo, int w_hi, int n_ptam, int nt,
                   double *traces, double *spec, int nthreads)
{
    if (w_lo < 0 || w_hi > nw || w_lo > w_hi || nr < 1 || nt < 0) return OR_E_INPUT;
    if (nt > 0 && (n_ptam < 2 || n_ptam > OR_PTAM_MAX)) return OR_E_INPUT;
    int nwr = w_hi - w_lo;
    const double dk = kmax / nk;
    const int nkt = nk + nt; /* the grid, then the PTAM tail on the same dk */
    cd *kern = (cd *)malloc(sizeof(cd) * 2 * (size_t)nwr * nkt + 1);
    int st = kernels_cd(compat, lay, n_in, src_depth, rcv_depth, src_kind, nw, dt, nkt, dk * nkt, omega_i,
                        w_lo, w_hi, kern, nthreads);
    if (st) {
        free(kern);
        return st;
    }
    double df = 1.0 / (2 * nw * dt);
    int m = compat ? 0 : 1;
    /* Bessel tables J0(k r), J1(k r) */
    or_real *bt = (or_real *)malloc(sizeof(or_real) * 2 * (size_t)nkt * nr);
    for (int ir = 0; ir < nr; ir++)
        for (int ik = 0; ik < nkt; ik++) {
            or_real x = dk * ik * dists[ir];
            bt[((size_t)ir * nkt + ik) * 2] = bes(m, 0, 0, x);
            bt[((size_t)ir * nkt + ik) * 2 + 1] = bes(m, 0, 1, x);
        }
    cd *raw = (cd *)calloc((size_t)nr * 2 * nw, sizeof(cd));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int ir = 0; ir < nr; ir++) {
        or_real *tail = (or_real *)malloc(sizeof(or_real) * 4 * (size_t)nt + 1);
        const int ntr = nt ? ptam_terms(n_ptam, dists[ir], dk, nt) : 0;
        for (int iw = w_lo; iw < w_hi; iw++) {
            cd sz = 0, sr = 0;
            const cd *kw = kern + (size_t)(iw - w_lo) * nkt * 2;
            const or_real *b = bt + (size_t)ir * nkt * 2;
            for (int ik = 0; ik < nk; ik++) {
                or_real wgt = (ik == 0) ? dk / 2.0 : dk;
                sz += kw[2 * ik] * (b[2 * ik] * wgt);
                sr += kw[2 * ik + 1] * (b[2 * ik + 1] * wgt);
            }
            if (ntr) {
                /* tail terms: Z re, Z im, R re, R im */
                for (int j = 0; j < ntr; j++) {
                    const int ik = nk + j;
                    const cd tz = kw[2 * ik] * (b[2 * ik] * dk), trr = kw[2 * ik + 1] * (b[2 * ik + 1] * dk);
                    tail[4 * j] = CRE(tz);
                    tail[4 * j + 1] = CIM(tz);
                    tail[4 * j + 2] = CRE(trr);
                    tail[4 * j + 3] = CIM(trr);
                }
                int f;
                sz = MKC(ptam_part(CRE(sz), tail, 4, ntr, n_ptam, &f), ptam_part(CIM(sz), tail + 1, 4, ntr, n_ptam, &f));
                sr = MKC(ptam_part(CRE(sr), tail + 2, 4, ntr, n_ptam, &f),
                         ptam_part(CIM(sr), tail + 3, 4, ntr, n_ptam, &f));
            }
            raw[(size_t)(ir * 2 + 0) * nw + iw] = sz;
            raw[(size_t)(ir * 2 + 1) * nw + iw] = sr;
        }
        free(tail);
    }
    cd *sp = (cd *)malloc(sizeof(cd) * nw);
    or_real *tr = (or_real *)malloc(sizeof(or_real) * 2 * nw);
    for (int t = 0; t < nr * 2; t++) {
        post_ifft(nw, raw + (size_t)t * nw, df, omega_i, dt, t0, taper, sp, tr);
        if (spec)
            for (int i = 0; i < nw; i++) {
                spec[((size_t)t * nw + i) * 2] = (double)CRE(sp[i]);
                spec[((size_t)t * nw + i) * 2 + 1] = (double)CIM(sp[i]);
            }
        if (traces)
            for (int j = 0; j < 2 * nw; j++) traces[(size_t)t * 2 * nw + j] = (double)tr[j];
    }
    free(sp);
    free(tr);
    free(raw);
    free(bt);
    free(kern);
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* dense-grid multi-receiver displacement columns                      */
/* ------------------------------------------------------------------ */
/* The displacement columns of both sources on the dense grid: tensor
 * columns 0-9 (green.cu:1651-1662) then single-force columns 0-4
 * (green.cu:1582-1586), each summed over k with the trapezoid weight
 * (green.cu:1887-1901) from the reference's integrand formulas with direct
 * Bessel values (J_n, d_besselj, J_n / r), then phase / taper / scale and
 * the inverse FFT (green.cu:1908-1938, 2208-2239).
 * traces: [15][nr][2 nw]. */
int or_columns(int compat, const double *lay, int n_in, double src_depth, double rcv_depth,
               const double *dists, int nr, int nw, double dt, int nk, double kmax, double omega_i,
               double t0, double taper, double *traces, int nthreads)
{
    return or_columns_all(compat, 0, lay, n_in, src_depth, rcv_depth, dists, nr, nw, dt, nk, kmax, omega_i, t0,
                          taper, traces, nthreads);
}

int or_columns_all(int compat, int all30, const double *lay, int n_in, double src_depth, double rcv_depth,
                   const double *dists, int nr, int nw, double dt, int nk, double kmax, double omega_i,
                   double t0, double taper, double *traces, int nthreads)
{
    if (nr < 1 || nw < 1 || nk < 1) return OR_E_INPUT;
    or_run run;
    int st = trace_run(compat, lay, n_in, src_depth, rcv_depth, &run);
    if (st) return st;
    double df = 1.0 / (2 * nw * dt), dk = kmax / nk;
    grid_consts g = {nw, df, omega_i, dk};
    cd r0, v0;
    par0_rv(&run, compat, &g, &r0, &v0);
    const int ncol = all30 ? 60 : OR_N_COLUMNS;
    cd *raw = (cd *)calloc((size_t)ncol * nr * nw, sizeof(cd));
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (int iw = 0; iw < nw; iw++) {
        OPar l0[OR_MAX_LAYERS];
        cd w = MKC(2 * OR_PI * df * iw, -omega_i);
        layer_disp(&run, compat, w, l0);
        for (int ik = 0; ik < nk && !err; ik++) {
            double k = dk * ik;
            or_real wgt = ik == 0 ? dk / 2.0 : dk;
            OPpar pt, pf;
            if (cell_ppar(compat, 1, k, w, run.n_layer, l0, run.s_layer, run.zs, run.r_layer, run.zr, r0, v0,
                          &pt) ||
                cell_ppar(compat, 0, k, w, run.n_layer, l0, run.s_layer, run.zs, run.r_layer, run.zr, r0, v0,
                          &pf)) {
                err |= 1;
                break;
            }
            for (int ir = 0; ir < nr; ir++) {
                cd ct[30], cf[30];
                integrand(compat, 1, &pt, k, dists[ir], ct);
                integrand(compat, 0, &pf, k, dists[ir], cf);
                const int nt = all30 ? 30 : 10, nf = all30 ? 30 : 5;
                for (int c = 0; c < nt; c++) raw[((size_t)c * nr + ir) * nw + iw] += ct[c] * wgt;
                for (int c = 0; c < nf; c++) raw[((size_t)(nt + c) * nr + ir) * nw + iw] += cf[c] * wgt;
            }
        }
    }
    if (err) {
        free(raw);
        return OR_E_UNSUPPORTED;
    }
    or_real *tr = (or_real *)malloc(sizeof(or_real) * 2 * nw);
    for (size_t t = 0; t < (size_t)ncol * nr; t++) {
        post_ifft(nw, raw + t * nw, df, omega_i, dt, t0, taper, NULL, tr);
        for (int j = 0; j < 2 * nw; j++) traces[t * 2 * nw + j] = (double)tr[j];
    }
    free(tr);
    free(raw);
    return OR_OK;
}
