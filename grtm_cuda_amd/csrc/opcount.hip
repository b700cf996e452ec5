// opcount.hip -- exact operation count of one recursion cell.
//
// Instantiates the device cell solver on a counting scalar type on the host:
// every add / sub / mul / div / sqrt executed by cell_solve + zr_from_cell
// (Z / R plans) or + mt_rows_from_cell (displacement-column plans) for a
// given layer stack and geometry is tallied, transcendental calls
// (exp, sin/cos pair, log, atan2, hypot) separately.  This is the algorithmic
// flop count the benchmark's roofline uses for the recursion kernel.
#include <hip/hip_runtime.h>

#include <cmath>

namespace grtm_count {
struct Tally {
    double flops = 0, trans = 0;
};
static thread_local Tally g_t;

struct Cnt {
    double v;
    Cnt() : v(0) {}
    Cnt(double x) : v(x) {}
};
inline Cnt operator+(Cnt a, Cnt b) { g_t.flops += 1; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { g_t.flops += 1; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { g_t.flops += 1; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { g_t.flops += 1; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline bool operator>=(Cnt a, Cnt b) { return a.v >= b.v; }
inline bool operator==(Cnt a, Cnt b) { return a.v == b.v; }
inline Cnt fabs(Cnt a) { return Cnt(std::fabs(a.v)); }
inline Cnt copysign(Cnt a, Cnt b) { return Cnt(std::copysign(a.v, b.v)); }
inline Cnt sqrt(Cnt a) { g_t.flops += 1; return Cnt(std::sqrt(a.v)); }
inline Cnt fma(Cnt a, Cnt b, Cnt c) { g_t.flops += 2; return Cnt(std::fma(a.v, b.v, c.v)); }
inline Cnt sqrt_est(Cnt a) { g_t.flops += 1; return Cnt(std::sqrt(a.v)); }
inline Cnt hypot(Cnt a, Cnt b) { g_t.trans += 1; return Cnt(std::hypot(a.v, b.v)); }
inline Cnt exp(Cnt a) { g_t.trans += 1; return Cnt(std::exp(a.v)); }
inline Cnt log(Cnt a) { g_t.trans += 1; return Cnt(std::log(a.v)); }
inline Cnt atan2(Cnt a, Cnt b) { g_t.trans += 1; return Cnt(std::atan2(a.v, b.v)); }
inline void sincos_t(Cnt a, Cnt *s, Cnt *c)
{
    g_t.trans += 1;
    s->v = std::sin(a.v);
    c->v = std::cos(a.v);
}
}  // namespace grtm_count

namespace grtm {
using grtm_count::copysign;
using grtm_count::exp;
using grtm_count::fabs;
using grtm_count::hypot;
using grtm_count::sincos_t;
using grtm_count::sqrt;
using grtm_count::fma;
using grtm_count::sqrt_est;
}  // namespace grtm

#include "cell.h"

// count the FP64 vertical-wavenumber work as well
template <>
struct grtm::WideOf<grtm_count::Cnt> {
    using type = grtm_count::Cnt;
};

static int count_cell(int kind, int L, int s, int rl, int branch, int up, int bot, double *out, bool top = false)
{
    using grtm::cx;
    using grtm::LayerW;
    using grtm_count::Cnt;
    if (L < 2 || L > 64 || s < 0 || s >= L - 1 || rl < 0 || rl >= L || !out) return -3;
    if (bot < 0) bot = L - 1;
    if (bot < s || bot > L - 1 || (branch == 4 && bot != L - 1)) return -3;
    LayerW<Cnt> lw[64];
    const double w_re = 3.0, w_im = -0.05;
    for (int j = 0; j < L; j++) {
        const double a = 5.0 + 0.3 * j, b = 3.0 + 0.15 * j;
        LayerW<Cnt> &q = lw[j];
        q.sa2 = cx<Cnt>{Cnt((w_re * w_re - w_im * w_im) / (a * a)), Cnt(2 * w_re * w_im / (a * a))};
        q.sb2 = cx<Cnt>{Cnt((w_re * w_re - w_im * w_im) / (b * b)), Cnt(2 * w_re * w_im / (b * b))};
        q.sa = cx<Cnt>{Cnt(w_re / a), Cnt(w_im / a)};
        q.sb = cx<Cnt>{Cnt(w_re / b), Cnt(w_im / b)};
        q.mg = cx<Cnt>{Cnt(2.7 * (w_re * w_re - w_im * w_im)), Cnt(2.7 * 2 * w_re * w_im)};
        q.ca = cx<Cnt>{Cnt(a / w_re), Cnt(0)};
        q.cb = cx<Cnt>{Cnt(b / w_re), Cnt(0)};
        q.mu = cx<Cnt>{Cnt(2.7 * b * b), Cnt(0)};
        q.a = cx<Cnt>{Cnt(a), Cnt(0)};
        q.b = cx<Cnt>{Cnt(b), Cnt(0)};
        q.h = Cnt(2.0 + j);
        q.hrev = Cnt(-(2.0 + j) / (2 * M_PI));
        q.hlog = Cnt(-(2.0 + j) * M_LOG2E);
    }
    grtm::Geo g;
    g.L = L;
    g.s = s;
    g.rl = rl;
    g.branch = branch;
    g.up = up;
    g.pad = 0;
    g.zs = 0.5 * (2.0 + s);
    // top: the receiver on its layer's top (the free surface for rl = 0)
    g.zr = top ? 0.0 : (rl == s) ? (branch == 2 ? 0.25 * (2.0 + s) : 0.75 * (2.0 + s)) : 0.3;
    // the depth-truncation criterion: one iteration per layer from the source
    // down to bot (all L - 1 - s when nothing is cut)
    double crit = 0;
    if (branch != 4) {
        grtm_count::g_t = grtm_count::Tally();
        (void)grtm::sweep_bottom(lw, L, s, Cnt(branch == 3 ? g.zr : g.zs), Cnt(0.7));  // k = 0.7: no layer is cut
        const int iters = bot < L - 1 ? bot - s + 1 : L - 1 - s;
        crit = grtm_count::g_t.flops / (L - 1 - s) * iters;
    }
    grtm_count::g_t = grtm_count::Tally();
    grtm::CellOut<Cnt> c;
    const cx<Cnt> z0{Cnt(0), Cnt(0)};
    if (kind == 2) grtm::cell_solve<Cnt, false, true>(lw, g, Cnt(0.7), z0, z0, c, bot);
    else grtm::cell_solve<Cnt, false, false>(lw, g, Cnt(0.7), z0, z0, c, bot);
    const cx<Cnt> w{Cnt(w_re), Cnt(w_im)};
    if (kind == 0) {
        cx<Cnt> a, b, d, e;
        grtm::zr_from_cell<Cnt, false>(c, up, Cnt(0.7), w, a, b, d, e);
    } else if (kind == 1) {
        grtm::mt_rows_from_cell<Cnt, false, false>(c, up, Cnt(0.7), w, [](int, int, cx<Cnt>) {});
    } else {
        grtm::mt_rows_from_cell<Cnt, false, true>(c, up, Cnt(0.7), w, [](int, int, cx<Cnt>) {});
    }
    out[0] = grtm_count::g_t.flops + crit;
    out[1] = grtm_count::g_t.trans;
    return 0;
}

namespace grtm {
int opcount_cell_bot(int kind, int L, int s, int rl, int branch, int up, int bot, double *out, bool top)
{
    return count_cell(kind, L, s, rl, branch, up, bot, out, top);
}
}  // namespace grtm

extern "C" int grtm_opcount_cell(int L, int s, int rl, int branch, int up, double *out)
{
    return count_cell(0, L, s, rl, branch, up, -1, out);
}

extern "C" int grtm_opcount_columns_cell(int L, int s, int rl, int branch, int up, int derivatives, double *out)
{
    return count_cell(derivatives ? 2 : 1, L, s, rl, branch, up, -1, out);
}
