// cell.h -- the per-(omega,k) generalized reflection/transmission solve.
//
// One call handles one wavenumber cell of one frequency.  Layer constants of
// the frequency come in as LayerW records; everything per cell lives in
// registers and the layers are streamed twice (half-space -> source, then
// free surface -> source), so no per-layer state is ever written to memory.
//
// Reference algorithm: cusrc/popm/func/green.cu set_pars (164-195), get_emat
// (196-233), interface R/T (235-262 + the cuBLAS 4x4 inverse/GEMM at
// 1813-1816 + 339-387), get_grt (424-613), get_src_* (1271-1413) and
// get_ypars_a_* (615-1267).  The 4x4 interface solve is replaced by a
// closed-form 2x2 block factorisation (see interface_xy below).
#pragma once
#include "cx.h"
#include "dd.h"

namespace grtm {

// Precision of the vertical wavenumbers and exponent arguments: FP64 for
// both FP32 and FP64 cells, double-double for double-double cells.
template <typename T>
struct WideOf {
    using type = double;
};
template <>
struct WideOf<dd> {
    using type = dd;
};

// Constants of one layer at one complex frequency w.
template <typename T>
struct LayerW {
    using W = typename WideOf<T>::type;
    cx<W> sa2, sb2;  // (w/alpha)^2, (w/beta)^2 in FP64 (vertical wavenumbers are formed in FP64)
    cx<T> sa, sb;  // w/alpha, w/beta (complex slownesses times w)
    cx<T> mg;      // mu (w/beta)^2
    cx<T> ca, cb;  // alpha/w, beta/w (normalisation of the E columns)
    cx<T> mu;      // shear modulus
    cx<T> a, b;    // alpha, beta
    W h;           // thickness
    W hrev, hlog;  // -h / (2 pi), -h log2(e): exponent factors of Lambda
};

// The FP64 layer constants of one frequency as double-double values (exact).
GRTM_HD LayerW<dd> widen_dd(const LayerW<double> &l)
{
    auto W = [](cx<double> z) { return cx<dd>{dd(z.x), dd(z.y)}; };
    LayerW<dd> o;
    o.sa2 = W(l.sa2), o.sb2 = W(l.sb2), o.sa = W(l.sa), o.sb = W(l.sb), o.mg = W(l.mg);
    o.ca = W(l.ca), o.cb = W(l.cb), o.mu = W(l.mu), o.a = W(l.a), o.b = W(l.b);
    o.h = dd(l.h), o.hrev = dd(l.hrev), o.hlog = dd(l.hlog);
    return o;
}

// Geometry of one (model, source) pair.
struct Geo {
    int L;       // layers including the half-space
    int s;       // source layer
    int rl;      // receiver layer
    int branch;  // 1: rl<s, 2: rl==s & zr<=zs, 3: rl==s & zr>zs, 4: rl>s
    int up;      // 1: receiver-above coupling, 0: receiver-below coupling
    int pad;
    double zs, zr;
};

// Layer quantities at wavenumber k.
template <typename T>
struct LQ {
    using W = typename WideOf<T>::type;
    cx<W> rd, vd;  // vertical wavenumbers in FP64 (exponent arguments)
    cx<T> r, v, x, mg, mu, sa, sb, ca, cb;
    W h;
};

template <typename T, typename W>
GRTM_HD cx<T> narrow(cx<W> z)
{
    return C<T>(T(z.x), T(z.y));
}

// e^{-nu z} with the argument formed in FP64.  For FP32 cells the phase is
// reduced to revolutions in [-1/2, 1/2] in FP64 first: Im(nu z) reaches
// hundreds of radians, and rounding it to FP32 would cost ~1e-5 of relative
// accuracy per factor.  The modulus is 2^(Re(-nu z) log2 e) with the product
// in FP64; exp, sin and cos are then the hardware FP32 instructions.
template <typename T, typename W>
GRTM_HD cx<T> cexp_nz(cx<W> nu, W z)
{
    const W ax = -(nu.x * z), ay = -(nu.y * z);
    if constexpr (sizeof(T) == sizeof(W)) {
        return narrow<T>(cexp(C<W>(ax, ay)));
    } else {
        const double rev = ay * 0.15915494309189533576888376337251;  // 1 / (2 pi)
        const float f = float(rev - rint(rev));
        const float e = fexp2(float(ax * 1.4426950408889634073599246810019));
        return C<T>(e * fcos_rev(f), e * fsin_rev(f));
    }
}

template <typename T>
GRTM_HD LQ<T> layer_q(const LayerW<T> &L, typename WideOf<T>::type kd)
{
    LQ<T> q;
    // vertical wavenumbers sqrt(k^2 - (w/V)^2) with Re >= 0 (green.cu:184-194),
    // in FP64 at the FP64 wavenumber (rounding k to FP32 first would shift the
    // abscissa by up to 6e-8 k, to which the phase of the kernel is sensitive
    // near k = w/V).  k^2 - Re (w/V)^2 cancels near k = Re w/V, but there
    // |Im (w/V)^2| = 2 |Re w/V| |Im w/V| bounds |z| from below: with Q
    // attenuation and the damping w_I, |Im / Re (w/V)| >= about w_I / w, so
    // the rounding of k^2 costs at most 1e-16 w / w_I (4e-13 at the highest
    // frequency of a 4096-sample trace) of |z|
    using W = typename WideOf<T>::type;
    const W k2 = kd * kd;
    const cx<W> zr = C<W>(k2 - L.sa2.x, -L.sa2.y);
    const cx<W> zv = C<W>(k2 - L.sb2.x, -L.sb2.y);
    if constexpr (sizeof(T) == 4) {
        // FP32 cells: FP32 estimate + FP64 Newton step (about 1e-13 relative)
        q.rd = csqrt_nr(zr);
        q.vd = csqrt_nr(zv);
    } else {
        q.rd = csqrt(zr);
        q.vd = csqrt(zv);
    }
    q.r = narrow<T>(q.rd);
    q.v = narrow<T>(q.vd);
    // k^2 + v^2 = 2 k^2 - (w/beta)^2, formed in FP64: it cancels near
    // k = w / (beta sqrt 2)
    q.x = narrow<T>(C<W>(W(2) * k2 - L.sb2.x, -L.sb2.y));
    q.mg = L.mg;
    q.mu = L.mu;
    q.sa = L.sa;
    q.sb = L.sb;
    q.ca = L.ca;
    q.cb = L.cb;
    q.h = L.h;
    return q;
}

// Interface between layer a, the side of the incident wave, and layer b.
// With E = [[P, P], [N, -N]] (E's rows reordered to (0,3 | 1,2); columns
// down | up; E^-1 = 1/2 [[P^-1, N^-1], [P^-1, -N^-1]]), continuity of
// displacement and traction for a wave incident from a, b's side reduced to
// its generalized reflection Q referenced at the interface, gives with
// X = P_a^-1 P_b, Y = N_a^-1 N_b and G = X (I + Q) + Y (I - Q)
//   generalized reflection  R = (X (I + Q) - Y (I - Q)) G^-1,
//   generalized transmission T = 2 G^-1.
// X and Y are formed times D = mu_a (w/beta_a)^2 r_a v_a, which clears both
// determinants (D / det Pt_a = -r_a, D / det Nt_a = v_a, up to the w^2 factor
// Pt_b and Nt_b share): R is unchanged by the common factor, and
// T = 2 D (D G)^-1.
// These are the quantities green.cu:471-594 forms from the interface
// coefficients (235-262, 339-387 and the cuBLAS 4x4 solve at 1813-1816):
// downward incidence (up sweep, a above b) gives R_du and T_d, upward
// incidence (down sweep, a below b) R_ud and T_u; with Q = 0 they are the
// interface coefficients themselves.
// The E columns are taken without their a/w, b/w normalisation: P = Pt D,
// N = Nt D with D = diag(a/w, b/w), Pt = [[k, v], [-mu x, -2 mu k v]],
// Nt = [[r, k], [-2 mu k r, -mu x]].  Every generalized R and T of the
// recursion is then the similarity transform D R D^-1 of the normalised one,
// so the normalisation returns once, on the source vector (src_scaled).
// adj(Pt_a) Pt_b and adj(Nt_a) Nt_b are written through dmu = mu_b - mu_a
// (mg = mu (w/beta)^2 from the frequency tables), so identical layers give
// off-diagonal zeros exactly and D X = D Y = D I to rounding (R = Q, T = I).
// Returns D X, D Y and 2 D.
template <typename T>
GRTM_HD void interface_xy(const LQ<T> &a, const LQ<T> &b, T k, m2<T> &X, m2<T> &Y, cx<T> &d2)
{
    const cx<T> dmu = b.mu - a.mu;
    const cx<T> k2dmu2 = (T(2) * k * k) * dmu;
    const cx<T> rv = a.r * a.v;
    const cx<T> t1 = a.mg + k2dmu2, t2 = b.mg - k2dmu2, t3 = t1 - b.mg;
    const cx<T> w = ((T(-2) * k) * dmu) * rv;
    X.a = rv * t2;
    X.b = w * b.v;
    X.c = k * (a.r * t3);
    X.d = (a.r * b.v) * t1;
    Y.a = (a.v * b.r) * t1;
    Y.b = k * (a.v * t3);
    Y.c = w * b.r;
    Y.d = X.a;
    d2 = T(2) * (a.mg * rv);
}

// One step of the recursion across an interface (X, Y, d2 of interface_xy)
// with b's side reflecting Q: F = X (I + Q), H = Y (I - Q) (entry by entry in
// fused multiply-adds, X + X Q and Y - Y Q), G = F + H.  R is kept as
// R = I - c P: F - H = G - 2 H, so R = (F - H) G^-1 = I - 2 H G^-1 with
// P = H adj(G) and c = 2 / det G.  P needs one 2x2 product where R needs G^-1
// and a product; c and the identity enter where R is used next (lrl_p,
// r_from_p).  T = d2 G^-1.  grt_p takes F and H (F = X, H = Y when nothing
// reflects on b's side, b the half-space).
template <typename T>
GRTM_HD void grt_p(const m2<T> &F, const m2<T> &H, cx<T> d2, m2<T> &P, cx<T> &c, m2<T> &Tr)
{
    const m2<T> G = F + H;
    const cx<T> di = recip(cfms(G.b, G.c, G.a * G.d));
    P = {cfms(H.b, G.c, H.a * G.d), cfms(H.a, G.b, H.b * G.a), cfms(H.d, G.c, H.c * G.d), cfms(H.c, G.b, H.d * G.a)};
    c = T(2) * di;
    const cx<T> dt = d2 * di;
    Tr = {G.d * dt, -(G.b * dt), -(G.c * dt), G.a * dt};
}
template <typename T>
GRTM_HD void grt_step_p(const m2<T> &X, const m2<T> &Y, cx<T> d2, const m2<T> &Q, m2<T> &P, cx<T> &c, m2<T> &Tr)
{
    const m2<T> F = {cfma(X.b, Q.c, cfma(X.a, Q.a, X.a)), cfma(X.a, Q.b, cfma(X.b, Q.d, X.b)),
                     cfma(X.d, Q.c, cfma(X.c, Q.a, X.c)), cfma(X.c, Q.b, cfma(X.d, Q.d, X.d))};
    const m2<T> H = {cfms(Y.b, Q.c, cfms(Y.a, Q.a, Y.a)), cfms(Y.a, Q.b, cfms(Y.b, Q.d, Y.b)),
                     cfms(Y.d, Q.c, cfms(Y.c, Q.a, Y.c)), cfms(Y.c, Q.b, cfms(Y.d, Q.d, Y.d))};
    grt_p(F, H, d2, P, c, Tr);
}
// R = I - c P.
template <typename T>
GRTM_HD m2<T> r_from_p(const m2<T> &P, cx<T> c)
{
    return {cfms(c, P.a, C<T>(1, 0)), -(c * P.b), -(c * P.c), cfms(c, P.d, C<T>(1, 0))};
}
// Lambda (I - c P) Lambda with Lambda = diag(e0, e1), R never formed; e11 =
// e1^2 for the SH term.
template <typename T>
GRTM_HD m2<T> lrl_p(cx<T> e0, cx<T> e1, const m2<T> &P, cx<T> c, cx<T> &e11)
{
    const cx<T> e00 = e0 * e0, e01 = e0 * e1;
    e11 = e1 * e1;
    const cx<T> c00 = c * e00, c01 = c * e01, c11 = c * e11;
    return {cfms(c00, P.a, e00), -(c01 * P.b), -(c01 * P.c), cfms(c11, P.d, e11)};
}
// SH (E_SH = [[1, 1], [-mu v, mu v]]): impedances wa = mu_a v_a, wb = mu_b v_b,
// q the generalized SH reflection of b's side.
template <typename T>
GRTM_HD void grt_step_sh(cx<T> wa, cx<T> wb, cx<T> q, cx<T> &r, cx<T> &t)
{
    const cx<T> p = cfma(wa, q, wa), m = cfms(wb, q, wb);  // wa (1 + q), wb (1 - q)
    const cx<T> gi = recip(p + m);
    r = (p - m) * gi;
    t = T(2) * wa * gi;
}

// Free-surface reflection R_ud^0 = -(E21)^-1 E22 (green.cu:449-468) in closed
// form, without the E column normalisation: E21 = -mu F, F = [[2kr, x],
// [x, 2kv]], E22 = diag(-1, 1) E21.  det F = 4 k^2 r v - x^2 (the Rayleigh
// function of the top layer) cancels near the layer's Rayleigh slowness,
// where a shallow source excites the surface waves strongly: it is formed in
// FP64 from the FP64 wavenumber, vertical wavenumbers and (w / beta)^2 (sb2).
template <typename T>
GRTM_HD m2<T> free_surface(const LQ<T> &q, typename WideOf<T>::type kd, cx<typename WideOf<T>::type> sb2)
{
    using W = typename WideOf<T>::type;
    const T k = T(kd);
    const T k2 = k * k;
    const cx<T> rv4 = (T(4) * k2) * (q.r * q.v);
    const cx<T> x2 = q.x * q.x;
    const cx<W> xd = C<W>(W(2) * kd * kd - sb2.x, -sb2.y);
    // det F in FP64, its reciprocal in the cell's precision
    const cx<T> nr = recip(narrow<T>(C<W>(W(4) * kd * kd, W(0)) * (q.rd * q.vd) - xd * xd));  // 1/det F
    const cx<T> dd = (x2 + rv4) * nr;
    const cx<T> k4x = (T(4) * k) * q.x * nr;
    return m2<T>{dd, k4x * q.v, -(k4x * q.r), -dd};
}

// The P-SV source vector in the normalisation-free basis: its components
// times a/w, b/w of the source layer.
template <typename T>
GRTM_HD void src_scaled(const LQ<T> &q, cx<T> &v0, cx<T> &v1)
{
    v0 = v0 * q.ca;
    v1 = v1 * q.cb;
}

// e^{-nu h} for a layer's thickness from its precomputed factors: the phase
// in revolutions nu_y (-h / (2 pi)) and the modulus exponent nu_x (-h log2 e)
// each in one FP64 product (as cexp_nz for FP32 cells).
template <typename T, typename W>
GRTM_HD cx<T> cexp_layer(cx<W> nu, const LayerW<T> &L)
{
    if constexpr (sizeof(T) == sizeof(W)) {
        return cexp_nz<T>(nu, L.h);
    } else {
        const double rev = nu.y * L.hrev;
        const float f = float(rev - rint(rev));
        const float e = fexp2(float(nu.x * L.hlog));
        return C<T>(e * fcos_rev(f), e * fsin_rev(f));
    }
}
template <typename T>
GRTM_HD void lam_layer(const LQ<T> &q, const LayerW<T> &L, cx<T> &e0, cx<T> &e1)
{
    e0 = cexp_layer<T>(q.rd, L);
    e1 = cexp_layer<T>(q.vd, L);
}

template <typename T, typename W>
GRTM_HD void lam(const LQ<T> &q, W h, cx<T> &e0, cx<T> &e1)
{
    e0 = cexp_nz<T>(q.rd, h);
    e1 = cexp_nz<T>(q.vd, h);
}

// d/dz of the SH receiver response above the source: the upgoing wave
// ru1 = e^{-v (d - zr)} (d the depth of the layer's bottom or of the source
// below zr, e_top = e^{-v d}) minus its reflection rud0 e_top e^{-v zr} at
// the top, times v, written as v ((ru1 - e_top rd1) + (1 - rud0) e_top rd1)
// with rd1 = e^{-v zr}.  At a free surface (zr = 0: ru1 and e_top are formed
// from the same argument, rd1 = 1, rud0 = 1; the traction-free condition)
// both brackets are exactly zero, where the direct difference of the two
// products leaves an FP32 rounding residue that is large beside the small
// true d/dz of T.  Every factor stays bounded.
template <typename T>
GRTM_HD cx<T> sh_dz_above(const LQ<T> &q, cx<T> rud0, cx<T> rd1, cx<T> e_top, cx<T> ru1)
{
    const cx<T> down = e_top * rd1;
    return q.v * ((ru1 - down) + (C<T>(1, 0) - rud0) * down);
}

// Result of the receiver-side solve for one cell.
template <typename T>
struct CellOut {
    m2<T> Y, Yt;        // P-SV receiver response (incl. source-layer reverberation)
    cx<T> y0, y0t;      // SH receiver response
    m2<T> ERd, ERu;     // sexd R_du sexd, sexu R_ud sexu at the source
    cx<T> erd0, eru0;   // SH analogues
    LQ<T> sq;           // source-layer quantities
    int status;         // 0 ok
};

// Depth truncation of the upward sweep.  With the receiver in or above the
// source layer the sweep only feeds R_du at the bottom of the source layer,
// where the stack below layer j enters through the two-way decay
// e^{-2 x_j} relative to the direct wave, x_j = sum of Re(nu_beta) h from the
// deeper of source and receiver (z0, depth in layer s) down to the bottom of
// layer j (S decays least; P and converted paths decay faster).  Once
// x_j > kDecayCut (e^{-36} = 2e-16 two-way, below FP64 rounding) layer j is
// taken as the half-space; j = s drops the sweep (R_du = 0).  Returns that
// layer, or L - 1.  The criterion runs in the cell's precision.
constexpr double kDecayCut = 18.0;
template <typename T>
GRTM_HD int sweep_bottom(const LayerW<T> *lw, int L, int s, typename WideOf<T>::type z0,
                         typename WideOf<T>::type kd)
{
    const T k = T(kd), cut = T(kDecayCut);
    T x = T(0);
    for (int j = s; j < L - 1; j++) {
        const cx<T> sb = lw[j].sb;
        const cx<T> z = (C<T>(k, T(0)) - sb) * (C<T>(k, T(0)) + sb);
        const T re = sqrt_est(T(0.5) * (sqrt_est(z.x * z.x + z.y * z.y) + z.x));  // Re sqrt(z)
        x = x + re * T(j == s ? lw[j].h - z0 : lw[j].h);
        if (x >= cut) return j;
    }
    return L - 1;
}

// compat: reproduce the reference's index choices (rexut from the first cell,
// the extra v in u0t, yu0t/yd0t products, the shifted r>s chain).
// BR: the receiver branch fixed at compile time (1-4), or 0 to read g.branch;
// a fixed branch lets the compiler drop the other branches' state (the
// surface-receiver solve needs 134 VGPRs instead of 155).
// NEED_T: also form the d/dz responses.  par0_r/par0_v: vertical wavenumbers
// of layer rl at the first cell (k = 0, w_0), used only in compat mode.
// bot_in: the upward sweep's bottom layer (operation counts); -1: from
// sweep_bottom when the receiver is at or above the source layer, else L - 1.
template <typename T, bool COMPAT, bool NEED_T, int BR = 0>
GRTM_HD void cell_solve(const LayerW<T> *lw, const Geo &g, typename WideOf<T>::type kd, cx<T> par0_r,
                        cx<T> par0_v, CellOut<T> &out, int bot_in = -1)
{
    const T k = T(kd);
    const int L = g.L, s = g.s, rl = g.rl;
    const int branch = BR ? BR : g.branch;
    using W = typename WideOf<T>::type;
    const W zs = W(g.zs), zr = W(g.zr);
    out.status = 0;

    // ---------------- upward sweep: half-space -> bottom of source layer
    // R_du at the bottom of s as I - cs1 Ps1 (grt_p), at the bottom of rl, T_d chain
    m2<T> Ps1, Rdu_r1, Md;
    cx<T> cs1, rdu0_s1, rdu0_r1, md0;
    Rdu_r1 = ident<T>();
    rdu0_r1 = C<T>(0, 0);
    Md = ident<T>();
    md0 = C<T>(1, 0);
    const int bot = bot_in >= 0 ? bot_in : branch != 4 ? sweep_bottom(lw, L, s, branch == 3 ? zr : zs, kd) : L - 1;
    if (bot == s) {
        Ps1 = ident<T>();  // R_du = 0
        cs1 = C<T>(1, 0);
        rdu0_s1 = C<T>(0, 0);
    } else {
        LQ<T> qd = layer_q(lw[bot], kd);  // layer below the interface
        LQ<T> qu = layer_q(lw[bot - 1], kd);
        m2<T> X, Y, P, Td;
        cx<T> c, rg0, td0, d2;
        interface_xy(qu, qd, k, X, Y, d2);
        grt_p(X, Y, d2, P, c, Td);  // nothing reflects below: F = X, H = Y
        grt_step_sh(qu.v * qu.mu, qd.v * qd.mu, C<T>(0, 0), rg0, td0);
        // grt[bot] = (Rg, Td)
        int j = bot;  // index of the grt entry just formed
        // lj0, lj1 = Lambda_jj (layer jj, just below the interface formed)
        auto visit = [&](int jj, const m2<T> &PJ, cx<T> cJ, cx<T> rg0J, const m2<T> &TdJ, cx<T> td0J,
                         cx<T> lj0, cx<T> lj1) {
            // jj: grt index formed (R_du = I - cJ PJ at interface (jj-1, jj), T_d into jj)
            if (jj == s + 1) {
                Ps1 = PJ;
                cs1 = cJ;
                rdu0_s1 = rg0J;
            }
            if (branch == 4) {
                if (jj == rl + 1) {
                    Rdu_r1 = r_from_p(PJ, cJ);
                    rdu0_r1 = rg0J;
                }
                if (!COMPAT) {
                    // chain T_d[rl] L_{rl-1} ... L_{s+1} T_d[s+1]
                    if (jj == rl) {
                        Md = TdJ;
                        md0 = td0J;
                    } else if (jj < rl && jj >= s + 1) {
                        Md = Md * dm(lj0, lj1, TdJ);
                        md0 = md0 * lj1 * td0J;
                    }
                } else {
                    // reference chain T_d[rl+1] L_{rl-2} T_d[rl] ... L_s T_d[s+2]
                    if (jj == rl + 1) {
                        Md = TdJ;
                        md0 = td0J;
                    } else if (jj <= rl && jj >= s + 2) {
                        LQ<T> q2 = layer_q(lw[jj - 2], kd);
                        cx<T> e0, e1;
                        lam(q2, lw[jj - 2].h, e0, e1);
                        Md = Md * dm(e0, e1, TdJ);
                        md0 = md0 * e1 * td0J;
                    }
                }
            }
        };
        visit(j, P, c, rg0, Td, td0, C<T>(0, 0), C<T>(0, 0));  // Lambda of the half-space is never used
        for (int lay = bot - 2; lay >= s; lay--) {
            // interface lay: between layer lay (above) and lay+1 (below, = qu)
            cx<T> e0, e1;
            lam_layer(qu, lw[lay + 1], e0, e1);
            qd = qu;
            qu = layer_q(lw[lay], kd);
            // green.cu:532-594: generalized R_du, T_d across interface lay with
            // the stack below reflecting Lambda R Lambda
            cx<T> e11;
            const m2<T> Q = lrl_p(e0, e1, P, c, e11);
            interface_xy(qu, qd, k, X, Y, d2);
            grt_step_p(X, Y, d2, Q, P, c, Td);
            grt_step_sh(qu.v * qu.mu, qd.v * qd.mu, e11 * rg0, rg0, td0);
            visit(lay + 1, P, c, rg0, Td, td0, e0, e1);
        }
    }

    // ---------------- downward sweep: free surface -> top of source layer
    m2<T> Ps, Rud_r, Mu;  // R_ud at the top of s as I - cs Ps
    cx<T> cs, rud0_s, rud0_r, mu0;
    Mu = ident<T>();
    mu0 = C<T>(1, 0);
    LQ<T> qr;  // receiver layer quantities
    {
        LQ<T> qu = layer_q(lw[0], kd);
        m2<T> Rg = free_surface(qu, kd, lw[0].sb2);
        cx<T> rg0 = C<T>(1, 0);
        if (rl == 0) qr = qu;
        Rud_r = Rg;
        rud0_r = rg0;
        // the free surface's R as I - c P with c = 1
        m2<T> P = {C<T>(1, 0) - Rg.a, -Rg.b, -Rg.c, C<T>(1, 0) - Rg.d};
        cx<T> c = C<T>(1, 0);
        for (int lay = 0; lay < s; lay++) {
            cx<T> e0, e1;
            lam_layer(qu, lw[lay], e0, e1);  // Lambda_lay
            LQ<T> qd = layer_q(lw[lay + 1], kd);
            // green.cu:471-520: generalized R_ud, T_u across interface lay
            // (incidence from below) with the stack above reflecting
            // Lambda R Lambda
            m2<T> X, Y, Tu;
            cx<T> tu0, d2, e11;
            const m2<T> Q = lrl_p(e0, e1, P, c, e11);
            interface_xy(qd, qu, k, X, Y, d2);
            grt_step_p(X, Y, d2, Q, P, c, Tu);
            grt_step_sh(qd.v * qd.mu, qu.v * qu.mu, e11 * rg0, rg0, tu0);
            const int jn = lay + 1;  // grt index formed
            if (branch == 1) {
                if (jn == rl) {
                    Rud_r = r_from_p(P, c);
                    rud0_r = rg0;
                }
                if (jn == rl + 1) {
                    Mu = Tu;
                    mu0 = tu0;
                } else if (jn > rl + 1) {
                    Mu = Mu * dm(e0, e1, Tu);
                    mu0 = mu0 * e1 * tu0;
                }
            }
            if (jn == rl) qr = qd;
            qu = qd;
        }
        Ps = P;
        cs = c;
        rud0_s = rg0;
        out.sq = qu;  // layer s
    }
    if (rl > s) qr = layer_q(lw[rl], kd);

    // ---------------- source layer reverberation (green.cu:654-702)
    const LQ<T> &sq = out.sq;
    const W hs_ = lw[s].h;
    cx<T> sexd0 = cexp_nz<T>(sq.rd, hs_ - zs), sexd1 = cexp_nz<T>(sq.vd, hs_ - zs);
    cx<T> sexu0 = cexp_nz<T>(sq.rd, zs), sexu1 = cexp_nz<T>(sq.vd, zs);
    {
        cx<T> t;
        out.ERd = lrl_p(sexd0, sexd1, Ps1, cs1, t);
        out.ERu = lrl_p(sexu0, sexu1, Ps, cs, t);
    }
    out.erd0 = sexd1 * rdu0_s1 * sexd1;
    out.eru0 = sexu1 * rud0_s * sexu1;
    const m2<T> yu_psv = minv(mmsub(out.ERd, out.ERu, ident<T>()));
    const m2<T> yd_psv = minv(mmsub(out.ERu, out.ERd, ident<T>()));
    const cx<T> y_sh = recip(C<T>(1, 0) - out.erd0 * out.eru0);

    // receiver-layer E rows 0,1 (displacements) without the column
    // normalisation: down cols [[k, v], [r, k]], up cols [[k, v], [-r, -k]]
    const m2<T> Ed{C<T>(k, 0), qr.v, qr.r, C<T>(k, 0)};
    const m2<T> Eu{C<T>(k, 0), qr.v, -qr.r, C<T>(-k, 0)};
    const W hr_ = lw[rl].h;
    const bool rl_half = (rl == L - 1);

    if (branch == 2) {  // green.cu:709-743
        cx<T> rd0 = cexp_nz<T>(qr.rd, zr), rd1 = cexp_nz<T>(qr.vd, zr);
        cx<T> ru0 = cexp_nz<T>(qr.rd, zs - zr), ru1 = cexp_nz<T>(qr.vd, zs - zr);
        const m2<T> Rud_s = r_from_p(Ps, cs);
        const m2<T> RS = dmd(rd0, rd1, Rud_s, sexu0, sexu1);
        const m2<T> u = mmadd(Ed, RS, md(Eu, ru0, ru1));
        const cx<T> u0 = rd1 * rud0_s * sexu1 + ru1;
        out.Y = u * yu_psv;
        out.y0 = u0 * y_sh;
        if (NEED_T) {
            const m2<T> RSt = dmd(-(qr.r * rd0), -(qr.v * rd1), Rud_s, sexu0, sexu1);
            const m2<T> ut = mmadd(Ed, RSt, md(Eu, qr.r * ru0, qr.v * ru1));
            cx<T> u0t = COMPAT ? -(qr.v * rd1) * rud0_s * sexu1 + qr.v * ru1 * qr.v
                               : sh_dz_above(qr, rud0_s, rd1, sexu1, ru1);
            out.Yt = ut * yu_psv;
            out.y0t = u0t * y_sh;
        }
    } else if (branch == 3) {  // green.cu:744-774
        cx<T> rd0 = cexp_nz<T>(qr.rd, zr - zs), rd1 = cexp_nz<T>(qr.vd, zr - zs);
        cx<T> ru0 = cexp_nz<T>(qr.rd, hr_ - zr), ru1 = cexp_nz<T>(qr.vd, hr_ - zr);
        const m2<T> Rdu_s1 = r_from_p(Ps1, cs1);
        const m2<T> RS = dmd(ru0, ru1, Rdu_s1, sexd0, sexd1);
        const m2<T> d = mmadd(Eu, RS, md(Ed, rd0, rd1));
        const cx<T> d0 = rd1 + ru1 * rdu0_s1 * sexd1;
        out.Y = d * yd_psv;
        out.y0 = d0 * y_sh;
        if (NEED_T) {
            const m2<T> RSt = dmd(qr.r * ru0, qr.v * ru1, Rdu_s1, sexd0, sexd1);
            const m2<T> dt = mmadd(Eu, RSt, md(Ed, -(qr.r * rd0), -(qr.v * rd1)));
            const cx<T> d0t = -(qr.v * rd1) + qr.v * ru1 * rdu0_s1 * sexd1;
            out.Yt = dt * yd_psv;
            out.y0t = d0t * y_sh;
        }
    } else if (branch == 1) {  // green.cu:775-857
        cx<T> ru0 = cexp_nz<T>(qr.rd, hr_ - zr), ru1 = cexp_nz<T>(qr.vd, hr_ - zr);
        // Lambda of the receiver layer: e^{-nu h} = e^{-nu zr} e^{-nu (h - zr)}
        cx<T> rd0, rd1, ex0, ex1, u0;
        m2<T> RS;
        if (!NEED_T && zr == W(0)) {
            // receiver on its layer's top (the free surface for rl = 0):
            // e^{-nu zr} = 1 and Lambda = ru, the general products below with
            // their unit factors dropped (two exponentials and six complex
            // products fewer per cell)
            RS = md(Rud_r, ru0, ru1);
            u0 = rud0_r * ru1 + ru1;
        } else {
            rd0 = cexp_nz<T>(qr.rd, zr);
            rd1 = cexp_nz<T>(qr.vd, zr);
            ex0 = rd0 * ru0;
            ex1 = rd1 * ru1;
            RS = dmd(rd0, rd1, Rud_r, ex0, ex1);
            u0 = rd1 * rud0_r * ex1 + ru1;
        }
        const m2<T> u = mmadd(Ed, RS, md(Eu, ru0, ru1));
        const m2<T> chain = Mu * dm(sexu0, sexu1, yu_psv);
        const cx<T> chain0 = mu0 * sexu1 * y_sh;
        out.Y = u * chain;
        out.y0 = u0 * chain0;
        if (NEED_T) {
            const cx<T> pr = COMPAT ? par0_r : qr.r, pv = COMPAT ? par0_v : qr.v;
            const m2<T> RSt = dmd(-(qr.r * rd0), -(qr.v * rd1), Rud_r, ex0, ex1);
            const m2<T> ut = mmadd(Ed, RSt, md(Eu, ru0 * pr, ru1 * pv));
            const cx<T> u0t = COMPAT ? -(qr.v * rd1) * rud0_r * ex1 + ru1 * pv : sh_dz_above(qr, rud0_r, rd1, ex1, ru1);
            out.Yt = ut * chain;
            out.y0t = COMPAT ? u0t * out.y0 : u0t * chain0;
        }
    } else {  // branch 4, rl > s (physics mode only): green.cu:858-915
        // The reference's receiver search stops at a layer <= s (it compares
        // the source's depth inside its layer with each thickness), so its
        // r > s code is never reached; the host refuses compat geometries here.
        cx<T> rd0 = cexp_nz<T>(qr.rd, zr), rd1 = cexp_nz<T>(qr.vd, zr);
        cx<T> ru0 = C<T>(0, 0), ru1 = C<T>(0, 0);
        if (!rl_half) {
            ru0 = cexp_nz<T>(qr.rd, hr_ - zr);
            ru1 = cexp_nz<T>(qr.vd, hr_ - zr);
        }
        // e^{-nu h} = e^{-nu zr} e^{-nu (h - zr)}; ex only enters times ru,
        // which is zero in the half-space
        const cx<T> ex0 = rd0 * ru0, ex1 = rd1 * ru1;
        const m2<T> RS = dmd(ru0, ru1, Rdu_r1, ex0, ex1);
        const m2<T> d = mmadd(Eu, RS, md(Ed, rd0, rd1));
        const cx<T> d0 = ru1 * rdu0_r1 * ex1 + rd1;
        const m2<T> chain = Md * dm(sexd0, sexd1, yd_psv);
        const cx<T> chain0 = md0 * sexd1 * y_sh;
        out.Y = d * chain;
        out.y0 = d0 * chain0;
        if (NEED_T) {
            const m2<T> RSt = dmd(ru0 * qr.r, ru1 * qr.v, Rdu_r1, ex0, ex1);
            const m2<T> dt = mmadd(Eu, RSt, md(Ed, -(qr.r * rd0), -(qr.v * rd1)));
            const cx<T> d0t = ru1 * qr.v * rdu0_r1 * ex1 - qr.v * rd1;
            out.Yt = dt * chain;
            out.y0t = d0t * chain0;
        }
    }
    if (!COMPAT && NEED_T && rl == 0 && zr == W(0)) {
        // Receiver on the free surface: the traction-free conditions give the
        // d/dz response exactly, d/dz (u_h, u_z) = (k u_z, -k (1 - 2 b^2/a^2) u_h)
        // and zero for SH; the general formula forms it as a difference of
        // up- and downgoing terms that cancels to O(k) at small k (in FP32
        // that cost the d/dz of T 1e-3).
        const cx<T> ba = qr.sa * recip(qr.sb);  // beta / alpha
        const cx<T> kg = k * (C<T>(1, 0) - T(2) * ba * ba);
        out.Yt = m2<T>{k * out.Y.c, k * out.Y.d, -(kg * out.Y.a), -(kg * out.Y.b)};
        out.y0t = C<T>(0, 0);
    }
}

// Kernel coefficients (popm_parameter a0,a1,a2,b1,b2 and their d/dz
// versions) from a solved cell.
template <typename T>
struct Coef {
    cx<T> a0[2], a1[2], a2[2], b1, b2;
    cx<T> a0t[2], a1t[2], a2t[2], b1t, b2t;
};

// Source vectors and coupling.  src_type 0: single force (green.cu:1271-1332,
// 918-938); otherwise moment tensor (green.cu:1334-1413, 1239-1266).
// The tensor's SH terms: the reference (green.cu:1408-1411) gives both orders
// the factor 1 + R above the source and R - 1 below.  Source-receiver
// reciprocity with the single force's strain fixes them: n = 1 is a jump in
// displacement (antisymmetric, like the P-SV s1: 1 - R above, R - 1 below)
// and n = 2 a jump in traction (symmetric, like s0 and s2: 1 + R on both
// sides).  COMPAT keeps the reference's factors.
template <typename T, bool NEED_T, bool COMPAT>
GRTM_HD void coef_from_cell(const CellOut<T> &c, int src_type, int up, T k, cx<T> w, Coef<T> &o)
{
    const LQ<T> &q = c.sq;
    const cx<T> b = w * q.cb;  // beta
    const m2<T> I2 = ident<T>();
    // wr?0 and wr?1 for the selected direction
    const m2<T> W0 = up ? (c.ERd - I2) : (I2 - c.ERu);
    const m2<T> W1 = up ? (c.ERd + I2) : (I2 + c.ERu);
    const cx<T> ssh = recip(T(2) * q.mu * q.v);
    if (src_type == 0) {
        // f0 = b/(2 mu a w r v) * (b v r, -a k r); f1 = (k b v, -a r v) * same
        const cx<T> tmp = b * recip(T(2) * q.mu * w * w * q.r * q.v) * q.sa;  // b/(2 mu a w r v)
        cx<T> f00 = b * q.v * q.r * tmp, f01 = -(k * (w * q.ca) * q.r * tmp);
        cx<T> f10 = k * b * q.v * tmp, f11 = -((w * q.ca) * q.r * q.v * tmp);
        src_scaled(q, f00, f01);
        src_scaled(q, f10, f11);
        cx<T> g00, g01, g10, g11;
        mv(W0, f00, f01, g00, g01);
        mv(W1, f10, f11, g10, g11);
        const cx<T> gsh = (up ? (C<T>(1, 0) + c.erd0) : (c.eru0 + C<T>(1, 0))) * ssh;
        o.b1 = c.y0 * gsh;
        mv(c.Y, g00, g01, o.a0[0], o.a0[1]);
        mv(c.Y, g10, g11, o.a1[0], o.a1[1]);
        o.a2[0] = o.a2[1] = o.b2 = C<T>(0, 0);
        if (NEED_T) {
            o.b1t = c.y0t * gsh;
            mv(c.Yt, g00, g01, o.a0t[0], o.a0t[1]);
            mv(c.Yt, g10, g11, o.a1t[0], o.a1t[1]);
            o.a2t[0] = o.a2t[1] = o.b2t = C<T>(0, 0);
        }
    } else {
        const cx<T> av = w * q.ca;  // alpha
        const cx<T> tmp = b * recip(T(2) * q.mu * av * w);  // b/(2 mu a w)
        cx<T> s20 = tmp * k * recip(q.r) * k * b, s21 = -(k * tmp * av);
        cx<T> s10 = T(-2) * k * b * tmp, s11 = tmp * recip(q.v) * q.x * av;
        cx<T> s00 = tmp * q.r * b, s01 = -(tmp * av * k);
        src_scaled(q, s00, s01);
        src_scaled(q, s10, s11);
        src_scaled(q, s20, s21);
        cx<T> u00, u01, u10, u11, u20, u21;
        mv(W1, s00, s01, u00, u01);
        mv(W0, s10, s11, u10, u11);
        mv(W1, s20, s21, u20, u21);
        const cx<T> one = C<T>(1, 0);
        const cx<T> sh1f = up ? (COMPAT ? one + c.erd0 : one - c.erd0) : (c.eru0 - one);
        const cx<T> sh2f = up ? (one + c.erd0) : (COMPAT ? c.eru0 - one : one + c.eru0);
        const cx<T> sh1 = sh1f * q.v * ssh, sh2 = sh2f * k * ssh;
        o.b1 = c.y0 * sh1;
        o.b2 = c.y0 * sh2;
        mv(c.Y, u00, u01, o.a0[0], o.a0[1]);
        mv(c.Y, u10, u11, o.a1[0], o.a1[1]);
        mv(c.Y, u20, u21, o.a2[0], o.a2[1]);
        if (NEED_T) {
            o.b1t = c.y0t * sh1;
            o.b2t = c.y0t * sh2;
            mv(c.Yt, u00, u01, o.a0t[0], o.a0t[1]);
            mv(c.Yt, u10, u11, o.a1t[0], o.a1t[1]);
            mv(c.Yt, u20, u21, o.a2t[0], o.a2t[1]);
        }
    }
}

// Z/R k-kernels of the n = 0 sources used by the trace path:
//  explosion      = s0 - s2 tensor source (only the P part survives:
//                   s0 - s2 = (-b (w/a)^2 tmp / r, 0), tmp = b/(2 mu a w)),
//                   Z = (a0[1]-a2[1]) k, R = -(a0[0]-a2[0]) k  [J0' = -J1]
//  vertical force = single-force f0 source, Z = a0[1] k, R = -a0[0] k.
// compat keeps the reference's index slip in the R column (a.[1], Q9).
template <typename T, bool COMPAT>
GRTM_HD void zr_from_cell(const CellOut<T> &c, int up, T k, cx<T> w, cx<T> &kz_ex, cx<T> &kr_ex,
                          cx<T> &kz_vf, cx<T> &kr_vf)
{
    const LQ<T> &q = c.sq;
    const m2<T> I2 = ident<T>();
    const m2<T> W0 = up ? (c.ERd - I2) : (I2 - c.ERu);
    const m2<T> W1 = up ? (c.ERd + I2) : (I2 + c.ERu);
    const cx<T> b = w * q.cb;  // beta
    const cx<T> av = w * q.ca;  // alpha
    // one reciprocal for the three denominators: R = 1 / (2 mu w alpha r v)
    const cx<T> rv = q.r * q.v;
    const cx<T> bR = b * recip((T(2) * q.mu * w) * av * rv);
    const cx<T> tmp = bR * rv;  // b / (2 mu alpha w)
    // explosion; the source vector carries the a/w normalisation of the
    // source layer (q.sa q.ca = 1): -b (w/alpha) tmp / r
    const cx<T> sx = -(b * q.sa * (bR * q.v));
    const cx<T> e0 = W1.a * sx, e1 = W1.c * sx;
    cx<T> a0, a1;
    mv(c.Y, e0, e1, a0, a1);
    kz_ex = a1 * k;
    kr_ex = COMPAT ? -(a1 * k) : -(a0 * k);
    // vertical force: f0 = (b^2/(2 mu a w), -b k/(2 mu w v))
    cx<T> f00 = b * tmp;
    cx<T> f01 = -(k * (bR * av * q.r));  // b k / (2 mu w v)
    src_scaled(q, f00, f01);
    cx<T> g0, g1, v0, v1;
    mv(W0, f00, f01, g0, g1);
    mv(c.Y, g0, g1, v0, v1);
    kz_vf = v1 * k;
    kr_vf = -(v0 * k);
}

// Displacement columns of both sources for many receivers: tensor columns
// 0-9 (green.cu:1651-1662) and single-force columns 0-4 (green.cu:1582-1586).
// With k J_n'(k r) = n J_n(k r) / r - k J_{n+1}(k r), the reference's
// d_besselj form (green.cu:1530-1541), every column is a sum of terms
// coef(w, k) J_m(k r) r^p, so the columns are assembled after the k-sum from
// contractions against J_m tables, the r^p factors applied per receiver.
// Displacement rows per Bessel order (complex; t: tensor, f: single force):
//   J0: t.a0[1] k, t.a2[1] k, f.a0[1] k
//   J1: t.a2[q] k, t.a0[q] k, t.a1[1] k, t.b1 + t.a1[0], f.a0[0] k, f.a1[1] k,
//       f.b1 + f.a1[0], and in compat t.b2 + t.a1[0]
//   J2: t.b2 + t.a2[0], t.a1[0] k, t.b5 k, t.a2[1] k, f.a1[0] k, f.b1 k
//   J3: t.a2[0] k, t.b2 k
// with q = 0 (compat: 1) and b5 = b1 (compat: b2), the index slips of the
// reference's columns 2, 3 and 5 (green.cu:1653-1657).
// Strain columns (physics mode): the d/dz columns are the same rows formed
// from the receiver-side d/dz coefficients (a.t, b.t; green.cu:1665-1691
// before the reference overwrites them), and d/dr of a term s r^-q S_m with
// S_m = sum_k A J_m(k r) is s (m - q) r^-(q+1) S_m - s r^-q sum_k (k A) J_{m+1}:
// every displacement row also enters the next order multiplied by k.  Set m
// then holds [displacement rows][their d/dz][k x displacement rows of set m-1]
// over J0..J4.
constexpr int kMtSetsMax = 5;
GRTM_HD constexpr int mt_disp_rows(int compat, int set)
{
    return set == 0 ? 3 : set == 1 ? (compat ? 8 : 7) : set == 2 ? 6 : set == 3 ? 2 : 0;
}
GRTM_HD constexpr int mt_sets(int deriv) { return deriv ? 5 : 4; }
GRTM_HD constexpr int mt_set_rows(int compat, int deriv, int set)
{
    return deriv ? 2 * mt_disp_rows(0, set) + (set > 0 ? mt_disp_rows(0, set - 1) : 0) : mt_disp_rows(compat, set);
}

template <typename T, bool COMPAT, typename Put>
GRTM_HD void mt_tensor_rows(const cx<T> *a0, const cx<T> *a1, const cx<T> *a2, cx<T> b1, cx<T> b2, T k, Put &&put)
{
    put(0, 0, a0[1] * k);
    put(0, 1, a2[1] * k);
    put(1, 0, (COMPAT ? a2[1] : a2[0]) * k);
    put(1, 1, (COMPAT ? a0[1] : a0[0]) * k);
    put(1, 2, a1[1] * k);
    put(1, 3, b1 + a1[0]);
    if (COMPAT) put(1, 7, b2 + a1[0]);
    put(2, 0, b2 + a2[0]);
    put(2, 1, a1[0] * k);
    put(2, 2, (COMPAT ? b2 : b1) * k);
    put(2, 3, a2[1] * k);
    put(3, 0, a2[0] * k);
    put(3, 1, b2 * k);
}

template <typename T, typename Put>
GRTM_HD void mt_force_rows(const cx<T> *a0, const cx<T> *a1, cx<T> b1, T k, Put &&put)
{
    put(0, 2, a0[1] * k);
    put(1, 4, a0[0] * k);
    put(1, 5, a1[1] * k);
    put(1, 6, b1 + a1[0]);
    put(2, 4, a1[0] * k);
    put(2, 5, b1 * k);
}

// put(set, row, value) receives a cell's rows; DERIV (physics mode only)
// adds the strain rows.  The cell must have been solved with NEED_T = DERIV.
template <typename T, bool COMPAT, bool DERIV, typename Put>
GRTM_HD void mt_rows_from_cell(const CellOut<T> &c, int up, T k, cx<T> w, Put &&put)
{
    static_assert(!(COMPAT && DERIV), "strain columns are physics-mode only");
    auto disp = [&](int set, int j, cx<T> v) {
        put(set, j, v);
        if (DERIV) put(set + 1, 2 * mt_disp_rows(0, set + 1) + j, k * v);
    };
    auto dz = [&](int set, int j, cx<T> v) { put(set, mt_disp_rows(0, set) + j, v); };
    {
        Coef<T> t;
        coef_from_cell<T, DERIV, COMPAT>(c, 1, up, k, w, t);
        mt_tensor_rows<T, COMPAT>(t.a0, t.a1, t.a2, t.b1, t.b2, k, disp);
        if constexpr (DERIV) mt_tensor_rows<T, COMPAT>(t.a0t, t.a1t, t.a2t, t.b1t, t.b2t, k, dz);
    }
    Coef<T> f;
    coef_from_cell<T, DERIV, COMPAT>(c, 0, up, k, w, f);
    mt_force_rows<T>(f.a0, f.a1, f.b1, k, disp);
    if constexpr (DERIV) mt_force_rows<T>(f.a0t, f.a1t, f.b1t, k, dz);
}

}  // namespace grtm
