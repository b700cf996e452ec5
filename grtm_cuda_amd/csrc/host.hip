// host.hip -- host driver and C ABI of libgrtm.so.
//
// Input handling follows cusrc/popm/func/green.cu:1989-2101 (input.txt
// parsing, source-layer insertion, thickness conversion, source/receiver
// layer search), the grid set-up follows green.cu:2063-2194 (incl. cal_kmax,
// green.cu:108-162), and the outputs follow green.cu:2242-2258.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/grtm.h"
#include "bessel.h"
#include "cell.h"
#include "kernels.h"

using grtm::Geo;
using grtm::LayerW;
using cd = std::complex<double>;

namespace {

thread_local int g_last_status = GRTM_OK;

#define HIPCHK(x)                                      \
    do {                                               \
        hipError_t e_ = (x);                           \
        if (e_ != hipSuccess) return GRTM_E_DEVICE;    \
    } while (0)

// ------------------------------------------------------------------ model
struct Stack {
    int L = 0;
    double h[grtm::kMaxLayers], a0[grtm::kMaxLayers], b0[grtm::kMaxLayers], rho[grtm::kMaxLayers];
    double qp[grtm::kMaxLayers], qs[grtm::kMaxLayers];
    double hs = 0, hr = 0;
    int s = -1, rl = -1;
    double zs = 0, zr = 0;
};

struct RunParams {
    double dt = 0, dist = 0, twin = 0, t0 = 0, taper = 0;
    int nf1 = 0, nf2 = 0, log2n = 0, n_ptam = 0, src_type = 0;
};

// Layer stack from layer tops (green.cu:2001-2049) and source / receiver
// location (green.cu:2081-2101).  compat: the receiver search keys on the
// source depth as the reference does, and a source that needs no split is
// rejected (the reference would leave its last layer uninitialised).
int prepare_stack(int compat, int n, const double *top, const double *vp, const double *vs, const double *rho,
                  const double *qp, const double *qs, double hs, double hr, Stack &st)
{
    if (n < 2 || n + 1 > grtm::kMaxLayers) return GRTM_E_INPUT;
    for (int i = 0; i < n; i++) {
        if (!(std::isfinite(top[i]) && vp[i] > 0 && vs[i] > 0 && rho[i] > 0 && qp[i] > 0 && qs[i] > 0))
            return GRTM_E_INPUT;
        if (i > 0 && !(top[i] > top[i - 1])) return GRTM_E_INPUT;
    }
    if (!(std::isfinite(hs) && std::isfinite(hr) && hs >= 0 && hr >= 0)) return GRTM_E_INPUT;
    int c = 0;
    bool split = false;
    auto push = [&](double t, int i) {
        st.h[c] = t;
        st.a0[c] = vp[i];
        st.b0[c] = vs[i];
        st.rho[c] = rho[i];
        st.qp[c] = qp[i];
        st.qs[c] = qs[i];
        c++;
    };
    push(top[0], 0);
    for (int i = 1; i < n; i++) {
        if (top[i - 1] < hs && top[i] > hs) {
            push(hs, i - 1);
            split = true;
        }
        push(top[i], i);
    }
    if (!split) {
        if (compat) return GRTM_E_UNSUPPORTED;
        if (hs > top[n - 1]) {  // source inside the half-space: split it
            push(hs, n - 1);
            split = true;
        }
    }
    st.L = c;
    double tops[grtm::kMaxLayers];
    for (int i = 0; i < c; i++) tops[i] = st.h[i];
    for (int i = 0; i < c - 1; i++) st.h[i] = st.h[i + 1] - st.h[i];
    st.hs = hs;
    st.hr = hr;
    st.s = st.rl = -1;
    if (!compat) {
        // physics mode: locate from the layer tops (the source sits at the
        // bottom of the layer above its interface, zs = h_s exactly)
        int s = 0, r = 0;
        for (int i = 0; i < c; i++) {
            if (tops[i] < hs) s = i;
            if (tops[i] < hr) r = i;
        }
        st.s = s;
        st.zs = hs - tops[s];
        st.rl = r;
        st.zr = hr - tops[r];
        if (st.s >= c - 1) return GRTM_E_UNSUPPORTED;
        return GRTM_OK;
    }
    double zs = hs, zr = hr;
    for (int i = 0; i < c; i++) {
        if (zs <= st.h[i] || i == c - 1) {
            st.s = i;
            st.zs = zs;
            break;
        }
        zs -= st.h[i];
    }
    for (int i = 0; i < c; i++) {
        const double key = compat ? zs : zr;
        if (key <= st.h[i] || (!compat && i == c - 1)) {
            st.rl = i;
            st.zr = zr;
            break;
        }
        zr -= st.h[i];
    }
    if (st.s < 0 || st.rl < 0) return GRTM_E_INPUT;
    if (st.s >= c - 1) return GRTM_E_UNSUPPORTED;
    return GRTM_OK;
}

int make_geo(const Stack &st, int compat, Geo &g)
{
    g.L = st.L;
    g.s = st.s;
    g.rl = st.rl;
    g.zs = st.zs;
    g.zr = st.zr;
    g.pad = 0;
    if (st.rl < st.s) g.branch = 1;
    else if (st.rl == st.s) g.branch = (st.zr <= st.zs) ? 2 : 3;
    else g.branch = 4;
    // coupling direction (green.cu:918-938)
    if ((st.rl == st.s && st.zr < st.zs) || st.rl < st.s) g.up = 1;
    else if ((st.rl == st.s && st.zr > st.zs) || st.rl > st.s) g.up = 0;
    else {
        if (compat) return GRTM_E_UNSUPPORTED;  // the reference writes no coefficients
        g.up = 1;
    }
    // the reference's search never places the receiver below the source layer
    if (compat && g.branch == 4) return GRTM_E_UNSUPPORTED;
    return GRTM_OK;
}

int parse_input(const char *path, int compat, Stack &st, RunParams &rp)
{
    FILE *f = fopen(path, "r");
    if (!f) return GRTM_E_IO;
    double hs, hr;
    int n;
    double top[grtm::kMaxLayers], vp[grtm::kMaxLayers], vs[grtm::kMaxLayers], rho[grtm::kMaxLayers];
    double qp[grtm::kMaxLayers], qs[grtm::kMaxLayers];
    int status = GRTM_OK;
    if (fscanf(f, "%lf", &hs) != 1 || fscanf(f, "%lf", &hr) != 1 || fscanf(f, "%d", &n) != 1) {
        status = GRTM_E_PARSE;
    } else if (n < 2 || n + 1 > grtm::kMaxLayers) {
        status = GRTM_E_INPUT;
    } else {
        for (int i = 0; i < n && status == GRTM_OK; i++)
            if (fscanf(f, " %lf , %lf , %lf , %lf , %lf , %lf", &top[i], &vp[i], &vs[i], &rho[i], &qp[i], &qs[i]) !=
                6)
                status = GRTM_E_PARSE;
        if (status == GRTM_OK &&
            (fscanf(f, "%lf", &rp.dt) != 1 || fscanf(f, "%lf", &rp.dist) != 1 || fscanf(f, "%d", &rp.nf1) != 1 ||
             fscanf(f, "%d", &rp.nf2) != 1 || fscanf(f, "%lf", &rp.twin) != 1 || fscanf(f, "%d", &rp.log2n) != 1 ||
             fscanf(f, "%lf", &rp.t0) != 1 || fscanf(f, "%lf", &rp.taper) != 1 ||
             fscanf(f, "%d", &rp.n_ptam) != 1 || fscanf(f, "%d", &rp.src_type) != 1))
            status = GRTM_E_PARSE;
    }
    fclose(f);
    if (status) return status;
    if (rp.log2n < 3 || rp.log2n > 14 || !(rp.dt > 0) || !(rp.dist > 0) || !std::isfinite(rp.t0) ||
        !(rp.taper >= 0 && rp.taper <= 1))
        return GRTM_E_INPUT;
    return prepare_stack(compat, n, top, vp, vs, rho, qp, qs, hs, hr, st);
}

// Dispersed velocity v0 (1 + ln(w/2pi)/(pi Q) + i/(2Q)) (green.cu:177-180).
cd disp_vel(cd w, double v0, double q)
{
    const cd lw = std::log(w / (2.0 * M_PI));
    return v0 * (1.0 + lw / (M_PI * q) + cd(0, 1) / (2.0 * q));
}

// Layer constants at one complex frequency.
void layer_w(const Stack &st, int compat, cd w, LayerW<double> *out)
{
    for (int j = 0; j < st.L; j++) {
        const cd a = disp_vel(w, st.a0[j], st.qp[j]);
        const cd b = disp_vel(w, st.b0[j], st.qs[j]);
        // the reference keeps mu = beta0^2 rho real (green.cu:2047-2049)
        const cd mu = compat ? cd(st.b0[j] * st.b0[j] * st.rho[j], 0) : st.rho[j] * b * b;
        const cd sa = w / a, sb = w / b;
        auto C = [](cd z) { return grtm::cx<double>{z.real(), z.imag()}; };
        LayerW<double> &q = out[j];
        q.sa2 = C(sa * sa);
        q.sb2 = C(sb * sb);
        q.sa = C(sa);
        q.sb = C(sb);
        q.mg = C(mu * sb * sb);
        q.ca = C(a / w);
        q.cb = C(b / w);
        q.mu = C(mu);
        q.a = C(a);
        q.b = C(b);
        q.h = st.h[j];
        q.hrev = -st.h[j] / (2 * M_PI);
        q.hlog = -st.h[j] * M_LOG2E;
    }
}

// cal_kmax (green.cu:108-162); b: S velocities dispersed at the grid's
// current frequency, w: the frequency at which the estimate is evaluated.
double cal_kmax(cd w, const cd *b, const Stack &st, int nf)
{
    const int s = st.s, rl = st.rl;
    const double *h = st.h;
    double pmin, dzmax, kmax, t;
    auto MX = [](double x, double y) { return x > y ? x : y; };
    auto MN = [](double x, double y) { return x < y ? x : y; };
    auto AB = [](double x) { return x > 0 ? x : -x; };
    if (rl == s - 1) {
        pmin = MX(std::abs(w / b[s]), std::abs(w / b[s - 1]));
        dzmax = MN(st.zs, h[rl] - st.zr);
    } else if (rl < s - 1) {
        pmin = MX(std::abs(w / b[s]), std::abs(w / b[s - 1]));
        dzmax = MN(st.zs, st.hs - st.hr);
        for (int lay = rl; lay <= s - 2; lay++) {
            t = std::abs(w / b[lay + 1]);
            if (pmin < t) pmin = t;
            if (dzmax > h[lay]) dzmax = h[lay];
        }
    } else if (rl == s) {
        pmin = std::abs(w / b[s]);
        dzmax = AB(st.hr - st.hs);
    } else if (rl == s + 1) {
        pmin = MX(std::abs(w / b[s]), std::abs(w / b[s + 1]));
        dzmax = MN(h[s] - st.zs, st.zr);
    } else {
        pmin = MX(std::abs(w / b[s]), std::abs(w / b[s + 1]));
        dzmax = MN(h[s] - st.zs, AB(st.hr - st.hs));
        for (int lay = s + 1; lay < rl - 1; lay++) {
            t = std::abs(w / b[lay]);
            if (pmin < t) pmin = t;
            if (dzmax > h[lay]) dzmax = h[lay];
        }
    }
    if (dzmax < 0.2) kmax = (nf < 10) ? MN(pmin + 2.0, 20. * pmin) : MN(pmin + 2.0, 2.0 * pmin);
    else kmax = std::sqrt((3.0 / dzmax) * (3.0 / dzmax) + pmin * pmin);
    if ((kmax - pmin) < 0.5) kmax = pmin + 0.5;
    return kmax;
}

// Device buffer that frees itself.
struct DBuf {
    void *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;  // owns device memory: no copies
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t bytes)
    {
        release();
        if (bytes == 0) bytes = 16;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            return GRTM_E_NOMEM;
        }
        n = bytes;
        return GRTM_OK;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

int device_ready()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return GRTM_E_DEVICE;
    return GRTM_OK;
}

// Element i of the post-processing tables of pp, in FP64: pv[2 nw] (phase,
// taper and scale of each frequency, green.cu:1905-1921), then tw[2 nw]
// (twiddles e^{i pi m / nw}), then ts[N] (e^{w_I t} / N, green.cu:2236-2237).
double post_table_value(const grtm::PostParams &pp, size_t i)
{
    const int M = pp.nw, N = 2 * M;
    if (i < (size_t)2 * M) {
        const int iw = int(i / 2);
        const double phi = 2.0 * M_PI * pp.df * iw * pp.t0;
        const double st0 = iw > pp.wc ? 0.5 * (1. + std::cos((iw - pp.wc + 1) * M_PI / (pp.nw - pp.wc + 1))) : 1.0;
        const double coef = pp.df * pp.nw / M_PI * st0;
        return (i & 1) ? std::sin(phi) * coef : std::cos(phi) * coef;
    }
    if (i < (size_t)4 * M) {
        const int m = int((i - 2 * M) / 2);
        return (i & 1) ? std::sin(M_PI * m / M) : std::cos(M_PI * m / M);
    }
    const int j = int(i - 4 * M);
    return std::exp(pp.omega_i * (j * pp.dt + pp.t0)) / N;
}

// The post-processing tables of pp (the PostParams fields pv, tw, ts), computed
// in FP64 and stored in buf in the trace precision (prec 0 FP64, else FP32).
int post_tables(grtm::PostParams &pp, int prec, DBuf &buf)
{
    const int M = pp.nw, N = 2 * M;
    std::vector<double> v((size_t)4 * M + N);
    for (size_t i = 0; i < v.size(); i++) v[i] = post_table_value(pp, i);
    const size_t es = prec == 0 ? sizeof(double) : sizeof(float);
    int st;
    if ((st = buf.alloc(v.size() * es))) return st;
    if (prec == 0) {
        HIPCHK(hipMemcpy(buf.p, v.data(), v.size() * es, hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(v.begin(), v.end());
        HIPCHK(hipMemcpy(buf.p, f.data(), f.size() * es, hipMemcpyHostToDevice));
    }
    const char *b = static_cast<const char *>(buf.p);
    pp.pv = b;
    pp.tw = b + (size_t)2 * M * es;
    pp.ts = b + (size_t)4 * M * es;
    return GRTM_OK;
}

// ------------------------------------------------------------------ green()
// Relative accuracy an FP64 solve of a frequency row loses where the P and SV
// columns of the layer matrices approach each other (|w| -> w_I): about
// (k V / |w|)^4 rounding units at the row's largest wavenumber kmax and the
// fastest velocity.  On the 2048-sample, 2 s demo model the estimate is 0.7
// at row 0 (the FP64 solve is 40% off per cell there), 1e-5 at row 5, and it
// overstates the measured loss of rows 1-8 about 100-fold.
double dd_row_loss(double kmax, double vmax, double wabs)
{
    const double q = kmax * vmax / wabs;
    return q * q * q * q * 0x1p-53;
}
constexpr double kDDRowLoss = 1e-5;

// ------------------------------------------------ no-allocation scratch
// Device memory of grtm_green_mem_noalloc: one fixed array per GPU, held
// until the process ends.  Host code never uses the array's name as an
// address: each device's copy is looked up once with hipGetSymbolAddress
// (with that device current) and that pointer is what kernels receive.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kScratchDevices];

// The current device's scratch address (looked up on the first call).
int scratch_ptr(void **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices) return GRTM_E_DEVICE;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GRTM_E_DEVICE;
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return GRTM_OK;
}

// Device memory of one drop-in run, carved from one block in 256-byte
// aligned pieces: layer tables [nw][L], per-row k counts, the cell kernel's
// per-block partial sums [nw][nblk][60], the summed spectra, the traces and
// post-processed spectra before their copy to the host, the post tables.
struct GreenLayout {
    size_t lw, nk, part, raw, tr, sp, pt, total;
};
GreenLayout green_layout(int nw, int L, int nblk)
{
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t N = 2 * (size_t)nw;
    GreenLayout o;
    size_t off = 0;
    o.lw = off, off += up((size_t)nw * L * sizeof(LayerW<double>));
    o.nk = off, off += up((size_t)nw * sizeof(int));
    o.part = off, off += up((size_t)nw * nblk * 60 * sizeof(double));
    o.raw = off, off += up((size_t)30 * nw * 2 * sizeof(double));
    o.tr = off, off += up((size_t)30 * N * sizeof(double));
    o.sp = off, off += up((size_t)30 * nw * 2 * sizeof(double));
    o.pt = off, off += up(((size_t)4 * nw + N) * sizeof(double));
    o.total = off;
    return o;
}

// Largest frequency count of a drop-in run (log2 N <= 14).
constexpr int kMaxRows = 8192;

// Host-to-device copy through a bounded buffer on the caller's stack, chunk by
// chunk; fill(i0, i1, buf) writes elements [i0, i1).  No heap memory.
template <typename E, int CHUNK, typename Fill>
int stage_h2d(E *dst, size_t n, hipStream_t s, Fill fill)
{
    E buf[CHUNK];
    for (size_t i0 = 0; i0 < n; i0 += CHUNK) {
        const size_t i1 = std::min(n, i0 + CHUNK);
        fill(i0, i1, buf);
        HIPCHK(hipMemcpyAsync(dst + i0, buf, (i1 - i0) * sizeof(E), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));  // buf is reused
    }
    return GRTM_OK;
}

// The whole drop-in run (green.cu:1975-2239) into host arrays in the layout
// of grtm_green_mem (spec: 30 x nw complex, traces: 30 x N; either NULL).
// get_dev: NULL allocates the run's device memory; otherwise it supplies
// `cap` bytes of device memory (the no-allocation entry point's scratch of the
// current device), and a run needing more is rejected with GRTM_E_TOO_LARGE
// before any device work.  stream: NULL uses a stream of the run's own
// (allocating form) or the null stream.  Host side: fixed-size arrays only
// (the layer tables and post tables are staged through bounded stack buffers).
int green_core(const char *iname, int compat, double *spec, double *traces, int *nw_out, int (*get_dev)(void **),
               size_t cap, hipStream_t user_stream)
{
    Stack st;
    RunParams rp;
    int status = parse_input(iname, compat, st, rp);
    if (status) return status;
    Geo g;
    status = make_geo(st, compat, g);
    if (status) return status;

    // grid constants (green.cu:2063-2080)
    const int nw = 1 << (rp.log2n - 1), N = 2 * nw;
    const double twin = N * rp.dt, df = 1.0 / (N * rp.dt), omega_i = 2.0 / twin;
    double vmax = 0, vmin = 100000;
    for (int i = 0; i < st.L; i++) {
        if (st.a0[i] > vmax) vmax = st.a0[i];
        if (st.b0[i] < vmin) vmin = st.b0[i];
    }
    const double vavg = (vmax + vmin) * 0.5;
    const double Lw = 2.0 * (vmax * twin + vmax / vavg * std::sqrt((st.hs - st.hr) * (st.hs - st.hr)) + 100);
    const double dk = 2 * M_PI / Lw;
    if (nw_out) *nw_out = nw;
    if (!spec && !traces) return GRTM_OK;

    // ragged k counts per frequency (green.cu:2145-2194)
    int nk_w[kMaxRows];
    int nk_max = 0;
    cd bvel[grtm::kMaxLayers];
    for (int iw = 0; iw < nw; iw++) {
        const cd w(2 * M_PI * df * iw, -omega_i);
        for (int j = 0; j < st.L; j++) bvel[j] = disp_vel(w, st.b0[j], st.qs[j]);
        double km = (iw < 3) ? cal_kmax(cd(2 * M_PI * df * 3.0, -omega_i), bvel, st, iw) : cal_kmax(w, bvel, st, iw);
        km = 1.1 * km;
        const double nkd = km / dk + 2.0;
        if (!(nkd < 2.0e9)) return GRTM_E_INPUT;
        nk_w[iw] = (int)(km / dk) + 2;
        if (nk_w[iw] > nk_max) nk_max = nk_w[iw];
    }
    const int nblk = (nk_max + 255) / 256;
    const GreenLayout lay = green_layout(nw, st.L, nblk);
    if (get_dev && lay.total > cap) return GRTM_E_TOO_LARGE;
    if ((status = device_ready())) return status;

    // vertical wavenumbers of the receiver layer in the first cell (k = 0, w_0),
    // used by the reference's d/dz columns (green.cu:674-675)
    double p0[4];
    {
        const cd w0(0.0, -omega_i);
        const cd a = disp_vel(w0, st.a0[st.rl], st.qp[st.rl]), b = disp_vel(w0, st.b0[st.rl], st.qs[st.rl]);
        cd r = std::sqrt(-(w0 / a) * (w0 / a)), v = std::sqrt(-(w0 / b) * (w0 / b));
        if (r.real() < 0) r = -r;
        if (v.real() < 0) v = -v;
        p0[0] = r.real(), p0[1] = r.imag(), p0[2] = v.real(), p0[3] = v.imag();
    }
    DBuf owned;
    void *dev = nullptr;
    if (get_dev) {
        if ((status = get_dev(&dev))) return status;
    } else {
        if ((status = owned.alloc(lay.total))) return status;
        dev = owned.p;
    }
    char *base = static_cast<char *>(dev);
    auto *d_lw = reinterpret_cast<LayerW<double> *>(base + lay.lw);
    auto *d_nk = reinterpret_cast<int *>(base + lay.nk);
    auto *d_part = reinterpret_cast<double *>(base + lay.part);
    auto *d_raw = reinterpret_cast<double *>(base + lay.raw);
    auto *d_tr = reinterpret_cast<double *>(base + lay.tr);
    auto *d_sp = reinterpret_cast<double *>(base + lay.sp);
    auto *d_pt = reinterpret_cast<double *>(base + lay.pt);

    // the caller's stream; else a stream of the run's own (allocating form) or
    // the null stream (no-allocation form)
    hipStream_t s = user_stream;
    struct SG {
        hipStream_t s = nullptr;
        ~SG()
        {
            if (s) (void)hipStreamDestroy(s);
        }
    } sg;
    if (!s && owned.p) {
        HIPCHK(hipStreamCreate(&s));
        sg.s = s;
    }

    // per-frequency layer tables, staged whole rows at a time
    const int L = st.L;
    {
        constexpr int kRowsPerChunk = 8;
        LayerW<double> buf[kRowsPerChunk * grtm::kMaxLayers];
        for (int r0 = 0; r0 < nw; r0 += kRowsPerChunk) {
            const int r1 = std::min(nw, r0 + kRowsPerChunk);
            for (int iw = r0; iw < r1; iw++)
                layer_w(st, compat, cd(2 * M_PI * df * iw, -omega_i), buf + (size_t)(iw - r0) * L);
            HIPCHK(hipMemcpyAsync(d_lw + (size_t)r0 * L, buf, (size_t)(r1 - r0) * L * sizeof(LayerW<double>),
                                  hipMemcpyHostToDevice, s));
            HIPCHK(hipStreamSynchronize(s));  // buf is reused
        }
    }
    HIPCHK(hipMemcpyAsync(d_nk, nk_w, (size_t)nw * sizeof(int), hipMemcpyHostToDevice, s));

    // the lowest rows, where an FP64 solve loses more than kDDRowLoss, in
    // double-double; the rest in FP64
    int n_dd = 0;
    while (n_dd < nw && dd_row_loss(nk_w[n_dd] * dk, vmax, std::abs(cd(2 * M_PI * df * n_dd, -omega_i))) > kDDRowLoss)
        n_dd++;
    HIPCHK(grtm::launch_cells_full(compat, rp.src_type ? 1 : 0, 1, d_lw, d_nk, L, g, 0, n_dd, nblk, dk, rp.dist, df,
                                   omega_i, p0, d_part, s));
    HIPCHK(grtm::launch_cells_full(compat, rp.src_type ? 1 : 0, 0, d_lw, d_nk, L, g, n_dd, nw - n_dd, nblk, dk,
                                   rp.dist, df, omega_i, p0, d_part, s));
    HIPCHK(grtm::launch_green_finish(nw, nblk, d_part, d_raw, s));

    grtm::PostParams pp{};
    pp.nw = nw;
    pp.wc = (int)((double)nw * (1 - rp.taper));
    pp.df = df;
    pp.dt = rp.dt;
    pp.t0 = rp.t0;
    pp.omega_i = omega_i;
    // post tables (as post_tables, FP64): pv [2 nw], tw [2 nw], ts [N]
    status = stage_h2d<double, 4096>(d_pt, (size_t)4 * nw + N, s, [&](size_t i0, size_t i1, double *buf) {
        for (size_t i = i0; i < i1; i++) buf[i - i0] = post_table_value(pp, i);
    });
    if (status) return status;
    pp.pv = d_pt;
    pp.tw = d_pt + 2 * nw;
    pp.ts = d_pt + 4 * nw;
    HIPCHK(grtm::launch_post_ifft(0, pp, 0, 30, d_raw, nullptr, nullptr, 0, 0, d_tr, d_sp, s));
    if (spec) HIPCHK(hipMemcpyAsync(spec, d_sp, (size_t)30 * nw * 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    if (traces) HIPCHK(hipMemcpyAsync(traces, d_tr, (size_t)30 * N * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return GRTM_OK;
}

int write_outputs(const char *oname_w, const char *oname_t, int nw, const std::vector<double> &spec,
                  const std::vector<double> &tr)
{
    const int N = 2 * nw;
    FILE *ft = fopen(oname_t, "w");
    if (!ft) return GRTM_E_IO;
    for (int j = 0; j < N; j++) {
        for (int c = 0; c < 30; c++) fprintf(ft, "%.18e,", tr[(size_t)c * N + j]);
        fprintf(ft, "\n");
    }
    if (fclose(ft) != 0) return GRTM_E_IO;
    FILE *fw = fopen(oname_w, "w");
    if (!fw) return GRTM_E_IO;
    for (int j = 0; j < N; j++) {
        for (int c = 0; c < 30; c++) {
            double re, im;
            if (j < nw) {
                re = spec[((size_t)c * nw + j) * 2];
                im = spec[((size_t)c * nw + j) * 2 + 1];
            } else if (j == nw) {
                re = im = 0.0;
            } else {  // conjugate mirror (green.cu:1925-1938)
                re = spec[((size_t)c * nw + (N - j)) * 2];
                im = -spec[((size_t)c * nw + (N - j)) * 2 + 1];
            }
            fprintf(fw, "%.18e,%.18e,", re, im);
        }
        fprintf(fw, "\n");
    }
    if (fclose(fw) != 0) return GRTM_E_IO;
    return GRTM_OK;
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

const char *grtm_version(void) { return "grtm-mi355x 0.1 (gfx950)"; }

const char *grtm_status_str(int s)
{
    switch (s) {
    case GRTM_OK: return "ok";
    case GRTM_E_IO: return "file i/o error";
    case GRTM_E_PARSE: return "malformed input";
    case GRTM_E_INPUT: return "invalid input values";
    case GRTM_E_UNSUPPORTED: return "configuration not supported in this mode";
    case GRTM_E_DEVICE: return "HIP device error";
    case GRTM_E_NOMEM: return "device allocation failed";
    case GRTM_E_TOO_LARGE: return "run exceeds the fixed scratch of the no-allocation entry point";
    default: return "unknown status";
    }
}

int grtm_last_status(void) { return g_last_status; }

int grtm_green_mem(const char *iname, int compat, double *spec, double *traces, int *nw_out)
{
    if (!iname) return GRTM_E_INPUT;
    return green_core(iname, compat, spec, traces, nw_out, nullptr, 0, nullptr);
}

int grtm_green_mem_noalloc(const char *iname, int compat, double *spec, double *traces, int *nw_out, void *stream)
{
    if (!iname) return GRTM_E_INPUT;
    return green_core(iname, compat, spec, traces, nw_out, scratch_ptr, kScratchBytes,
                      static_cast<hipStream_t>(stream));
}

int grtm_green(const char *iname, const char *oname_w, const char *oname_t, int compat)
{
    if (!iname || !oname_w || !oname_t) return GRTM_E_INPUT;
    int nw = 0;
    int st = green_core(iname, compat, nullptr, nullptr, &nw, nullptr, 0, nullptr);
    if (st) return st;
    std::vector<double> sp((size_t)30 * nw * 2), tr((size_t)30 * 2 * nw);
    st = green_core(iname, compat, sp.data(), tr.data(), &nw, nullptr, 0, nullptr);
    if (st) return st;
    return write_outputs(oname_w, oname_t, nw, sp, tr);
}

void green(char *iname, char *oname_w, char *oname_t) { g_last_status = grtm_green(iname, oname_w, oname_t, 1); }

}  // extern "C"

namespace {

// ------------------------------------------------------------ model set-up
// One model of a plan or batch: layer stack, geometry, per-frequency layer
// tables [nw x L] and (mixed precision) the rows solved in FP64.
struct ModelTables {
    Stack st;
    Geo g{};
    int rows64 = 0;
    std::vector<LayerW<double>> lw;
};

int model_tables(const double *layers, int n, double src_depth, double rcv_depth, int compat, int nw, double df,
                 double omega_i, bool mixed, ModelTables &mt)
{
    if (n < 2 || n + 1 > grtm::kMaxLayers || !layers) return GRTM_E_INPUT;
    std::vector<double> top(n), vp(n), vs(n), rho(n), qp(n), qs(n);
    double z = 0;
    for (int i = 0; i < n; i++) {
        const double *r = layers + 6 * i;
        top[i] = z;
        if (i < n - 1 && !(r[0] > 0)) return GRTM_E_INPUT;
        z += r[0];
        vp[i] = r[1], vs[i] = r[2], rho[i] = r[3], qp[i] = r[4], qs[i] = r[5];
    }
    int status = prepare_stack(compat, n, top.data(), vp.data(), vs.data(), rho.data(), qp.data(), qs.data(),
                               src_depth, rcv_depth, mt.st);
    if (status) return status;
    if ((status = make_geo(mt.st, compat, mt.g))) return status;
    const int L = mt.st.L;
    mt.lw.resize((size_t)nw * L);
    for (int iw = 0; iw < nw; iw++) layer_w(mt.st, compat, cd(2 * M_PI * df * iw, -omega_i), &mt.lw[(size_t)iw * L]);
    if (mixed) {
        // Low frequencies are solved in FP64: as w -> 0 the P and S columns of
        // the eigenvector basis become parallel and the recursion loses
        // ~(k beta / |w|)^2 of relative accuracy, with the integrand significant
        // up to k ~ |w| / beta + O(1 / dz).  Rows with |w| < 2.5 beta_min / dz.
        double bmin = 1e30;
        for (int j = 0; j < L; j++) bmin = std::min(bmin, mt.st.b0[j]);
        const double dz = std::max(std::fabs(src_depth - rcv_depth), 1e-3);
        const double wc = 2.5 * bmin / dz;
        int r64 = 0;
        while (r64 < nw && std::hypot(2 * M_PI * df * r64, omega_i) < wc) r64++;
        mt.rows64 = std::min(nw, std::max(r64, 4));
    }
    return GRTM_OK;
}

std::vector<LayerW<float>> narrow_tables(const std::vector<LayerW<double>> &lw)
{
    std::vector<LayerW<float>> lf(lw.size());
    auto F = [](grtm::cx<double> z) { return grtm::cx<float>{float(z.x), float(z.y)}; };
    for (size_t i = 0; i < lw.size(); i++) {
        const LayerW<double> &d = lw[i];
        lf[i] = LayerW<float>{d.sa2, d.sb2, F(d.sa), F(d.sb), F(d.mg), F(d.ca), F(d.cb), F(d.mu),
                              F(d.a), F(d.b), d.h, d.hrev, d.hlog};
    }
    return lf;
}

grtm::PostParams post_params(const grtm_grid *grid, int nw, double df, double omega_i)
{
    grtm::PostParams pp{};
    pp.nw = nw;
    pp.wc = (int)((double)nw * (1 - grid->taper));
    pp.df = df;
    pp.dt = grid->dt;
    pp.t0 = grid->t0;
    pp.omega_i = omega_i;
    return pp;
}

bool grid_ok(const grtm_grid *grid, int precision)
{
    if (grid->nw < 4 || (grid->nw & (grid->nw - 1)) || grid->nw > 8192 || grid->nk < 1 || !(grid->kmax > 0) ||
        !(grid->dt > 0) || !(grid->taper >= 0 && grid->taper <= 1))
        return false;
    return true;
}

}  // namespace

// ============================================================ trace plans
struct grtm_ctx {
    int device = 0;
};

struct grtm_plan {
    grtm_ctx *ctx = nullptr;
    int prec = 0, compat = 0;
    int nw = 0, N = 0, nk = 0, nr = 0, L = 0;
    double dk = 0, df = 0, omega_i = 0;
    Geo g{};
    grtm::PostParams pp{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int rows64 = 0;  // FP32 plans: frequency rows [0, rows64) are solved in FP64
    int gemm_mode = 0;  // 0 FP64 MFMA, 1 FP32 MFMA, 2 BF16x3 MFMA
    bool pair = false;  // FP32: both contractions in one streamed launch
    int kc = 0;         // its split-K slice length
    // split-BF16 pair contraction (3 or 6 products, gemm_mode 1): the Bessel
    // tables as bf16 planes S0 / S1 of row length ldk
    int bsplit = 0;
    long ldk = 0;
    DBuf lw, lw64, flags, dists, A0, A1, B0, B1, C0, C1, tr, ws, S0, S1;
    DBuf plant;      // test hook: cells given non-finite FP32 operands at every execution
    int nplant = 0;
    // FP64 cells next to branch points (within flag_hb cells of Re(w/V) of
    // every layer; < 0: none) and predicted pole clusters (within flag_pw
    // half widths; 0: none) of the FP32 rows (grtm_plan_set_flag64); poles:
    // the network's clusters of the last execution, [nw][8][2] km^-1
    int flag_hb = -1;
    float flag_pw = 0.f;
    DBuf poles;
    // peak-trough averaging of the k-sums' tails (grtm_plan_set_ptam): ptam_nt
    // wavenumbers beyond the grid per row, their weighted kernels At and
    // Bessel values Bt, and the count of sums that took the extended sum
    int ptam_n = 0, ptam_nt = 0;
    DBuf At0, At1, Bt0, Bt1, ptam_miss;
    DBuf ptab;  // post-processing tables of pp
    // adaptive wavenumber grid
    bool adaptive = false;
    int chk_stride = 8, max_chk = 0;
    double chk_tol = 1e-5;
    hipStream_t last_stream = nullptr;  // stream of the last grtm_plan_execute
    std::vector<int> nk_pred_h, nchk_h;
    DBuf nk_pred, nk_w, nchk, counters, fb_list, kend;  // counters: [row peak nw][tail max nw][fb count]
    DBuf tiles;             // recursion work items of the predicted grid, ordered by row
    int ntiles = 0, ntiles64 = 0;  // all items; items of rows [0, rows64)
    // learned grid built on the device at every execution (grtm_plan_set_learned):
    // the network, the model's predictor inputs, the device item bounds
    // {0, end of rows < rows64, end} and the persistent launches' item counters
    grtm_mlp *net = nullptr;
    grtm::LearnGrid lg{};
    grtm::LearnModel lm{};
    DBuf lmodel, icounts, ictr, ioff;
    int tiles_cap = 0;       // capacity of tiles (items)
    // mixed learned plans: rows [0, rows64) take the analytic cut (network
    // free, set up once: nk_pred / nchk and the item list low_tiles), so their
    // FP64 launch runs on the side stream beside the network and the item
    // kernels; ev_fork / ev_join bracket it
    DBuf low_tiles;
    int n_low_tiles = 0;
    bool low_apart = false;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool pred_stale = false;  // nk_pred_h / nchk_h behind the device's last execution
    // the job's model, kept for the predictor's inputs
    std::vector<double> layers_h;
    double src_depth = 0, rcv_depth = 0;
    int rows64_rule = 0;  // the default FP64-row split (also for FP64 plans: the learned floor)
    double bessel_ms = 0;
    // kind 0: explosion and vertical-force Z / R traces; kind 1: the
    // displacement columns, whose operands, tables and outputs are the four
    // contractions against J0..J3 laid end to end in A0, B0 and C0
    int kind = 0;
    int deriv = 0;   // column plans: with the d/dz and d/dr columns (physics mode)
    int nsets = 0;   // contractions of a column plan (J0..J3, or J0..J4 with deriv)
    int ncols = 0;   // columns per receiver (15, or 45 with deriv)
    grtm::MtOps mto{};
    grtm::MtCols mtc{};  // host copy; the kernels read mtc_d
    DBuf mtc_d;
    int ntr() const { return kind ? ncols * nr : 4 * nr; }
    // timing: events around the main recursion launch (the FP32 rows of a
    // mixed plan, all rows of an FP64 plan), recorded when set
    hipEvent_t ev_main0 = nullptr, ev_main1 = nullptr;
    hipEvent_t ev_pred0 = nullptr, ev_pred1 = nullptr;  // around a learned grid's formation
    int mt_M(int q) const { return 2 * mto.rows[q] * nw; }
    ~grtm_plan()
    {
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side) (void)hipStreamDestroy(side);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

size_t esize(int prec) { return prec == GRTM_PREC_FP64 ? sizeof(double) : sizeof(float); }

// Stream-K workspace of an FP32 pair contraction of M rows a side, enough for
// the split-BF16 kernel too when bsplit is set.
size_t pair_sk_bytes(int bsplit, int M, int N, int K)
{
    const int M2[2] = {M, M};
    size_t b = grtm::gemm_streamk_workspace_bytes(2, M2, N, K);
    if (bsplit) b = std::max(b, grtm::gemm_bsplit_workspace_bytes(M, N, K));
    return b;
}

// Split-precision contraction (GRTM_PREC_BF16X3 / BF16X6 / FP16X3 with a
// pair layout): the two Bessel tables' 16-bit planes, planes x nr x ldk values each
// (slab-major by 32 wavenumbers), ldk the wavenumber count rounded up to 32.
int split_alloc(int bsplit, int nk, int nr, DBuf &S0, DBuf &S1, long &ldk)
{
    ldk = (nk + 31) / 32 * 32;
    const size_t bytes = (size_t)grtm::split_planes(bsplit) * nr * ldk * 2;
    int st;
    if ((st = S0.alloc(bytes)) || (st = S1.alloc(bytes))) return st;
    return GRTM_OK;
}

// Displacement-column plans: the same sequence as the Z / R plans (FP64 rows,
// FP32 rows, residual check, dense fallback, FP64 re-solve of non-finite
// cells, tile extents), on the column rows.
int run_cells_mt(grtm_plan *p, hipStream_t s)
{
    const bool ad = p->adaptive;
    int *nkw = ad ? p->nk_w.as<int>() : nullptr;
    unsigned *rowpk = ad ? p->counters.as<unsigned>() : nullptr;
    unsigned *tailmx = ad ? rowpk + p->nw : nullptr;
    unsigned *fbcnt = ad ? rowpk + 2 * p->nw : nullptr;
    if (ad) {
        HIPCHK(grtm::launch_copy_i32(p->nk_pred.as<int>(), nkw, p->nw, s));
        HIPCHK(grtm::launch_zero_u32(rowpk, 2 * p->nw + 1, s));
    }
    const bool mixed = p->prec != GRTM_PREC_FP64;
    const int r64 = mixed ? p->rows64 : p->nw;
    const unsigned *t_lo = ad ? p->tiles.as<unsigned>() : nullptr;
    const unsigned *t_hi = ad ? t_lo + p->ntiles64 : nullptr;
    const int n_lo = ad ? p->ntiles64 : 0, n_hi = ad ? p->ntiles - p->ntiles64 : 0;
    unsigned *cnt = mixed ? p->flags.as<unsigned>() : nullptr, *list = mixed ? cnt + 1 : nullptr;
    // learned grid: the item lists carry the residual check beyond the cut
    // (FP64 on the FP64 rows) after each launch's tiles
    const int *nchk = ad ? p->nchk.as<int>() : nullptr;
    if (!mixed) {
        if (p->ev_main0) HIPCHK(hipEventRecord(p->ev_main0, s));
        HIPCHK(grtm::launch_cells_mt(0, p->compat, p->deriv, p->lw.p, p->L, p->g, p->nk, p->dk, p->df, p->omega_i, 0,
                                     p->nw, t_lo, n_lo, p->mto, p->nk, nullptr, nullptr, nkw, rowpk, s, nchk,
                                     p->chk_stride, tailmx));
        if (p->ev_main1) HIPCHK(hipEventRecord(p->ev_main1, s));
    } else {
        HIPCHK(grtm::launch_zero_u32(cnt, 1, s));
        HIPCHK(grtm::launch_cells_mt(2, p->compat, p->deriv, p->lw64.p, p->L, p->g, p->nk, p->dk, p->df, p->omega_i,
                                     0, p->rows64, t_lo, n_lo, p->mto, p->nk, nullptr, nullptr, nkw, rowpk, s, nchk,
                                     p->chk_stride, tailmx));
        if (p->ev_main0) HIPCHK(hipEventRecord(p->ev_main0, s));
        HIPCHK(grtm::launch_cells_mt(1, p->compat, p->deriv, p->lw.p, p->L, p->g, p->nk, p->dk, p->df, p->omega_i,
                                     p->rows64, p->nw - p->rows64, t_hi, n_hi, p->mto, p->nk, cnt, list, nkw, rowpk,
                                     s, nchk, p->chk_stride, tailmx));
        if (p->ev_main1) HIPCHK(hipEventRecord(p->ev_main1, s));
    }
    if (ad) {
        const void *lwf64 = mixed ? p->lw64.p : p->lw.p;
        HIPCHK(grtm::launch_fallback(!mixed, p->compat, lwf64, mixed ? p->lw.p : nullptr, r64, p->L, p->g, p->nk,
                                     p->nw, p->dk, p->df, p->omega_i, p->chk_tol, rowpk, tailmx, nkw, fbcnt,
                                     p->fb_list.as<unsigned>(), nullptr, nullptr, p->nk, cnt, list, 0, s, &p->mto,
                                     p->deriv));
    }
    if (mixed)
        HIPCHK(grtm::launch_cells_mt_fix(p->compat, p->deriv, p->lw64.p, p->L, p->g, p->nk, p->dk, p->df,
                                         p->omega_i, p->mto, p->nk, cnt, list, s));
    if (ad && p->pair) {
        int rpw[grtm::kMtSetsMax];
        for (int q = 0; q < p->nsets; q++) rpw[q] = 2 * p->mto.rows[q];
        HIPCHK(grtm::launch_kend_sets(p->nw, p->nk, p->nsets, rpw, 128, nkw, p->kend.as<int>(), s));
    }
    return GRTM_OK;
}

int run_cells_core(grtm_plan *p, hipStream_t s);

// The plan's recursion, then the PTAM tail's cells (FP64 recursion, operands
// in the plan's type) when set.
int run_cells(grtm_plan *p, hipStream_t s)
{
    int st;
    if ((st = run_cells_core(p, s))) return st;
    if (!p->ptam_n) return GRTM_OK;
    const bool f64 = p->prec == GRTM_PREC_FP64;
    HIPCHK(grtm::launch_cells_zr(f64 ? 0 : 2, p->compat, f64 ? p->lw.p : p->lw64.p, p->L, p->g, p->ptam_nt, p->dk,
                                 p->df, p->omega_i, 0, p->nw, nullptr, 0, p->At0.p, p->At1.p, p->ptam_nt, nullptr,
                                 nullptr, nullptr, nullptr, s, nullptr, 0, nullptr, nullptr, 0, nullptr, p->nk));
    return GRTM_OK;
}

int run_cells_core(grtm_plan *p, hipStream_t s)
{
    if (p->kind == 1) return run_cells_mt(p, s);
    const bool ad = p->adaptive;
    const bool learned = ad && p->net;
    const bool f16 = p->bsplit == grtm::kSplitF16x3;
    int *nkw = ad ? p->nk_w.as<int>() : nullptr;
    // row peaks: the adaptive grid's check, and the FP16 split's row scales
    unsigned *rowpk = (ad || f16) ? p->counters.as<unsigned>() : nullptr;
    unsigned *tailmx = ad ? rowpk + p->nw : nullptr;
    unsigned *fbcnt = ad ? rowpk + 2 * p->nw : nullptr;
    if (f16 && !ad) HIPCHK(grtm::launch_zero_u32(rowpk, p->nw, s));
    const int r64 = p->prec == GRTM_PREC_FP64 ? p->nw : p->rows64;
    const bool apart = learned && p->low_apart;
    if (apart) {
        // the FP64 rows on the side stream: their state, then their launch
        // over the set-up list, while the network and the item kernels run
        HIPCHK(hipEventRecord(p->ev_fork, s));
        HIPCHK(hipStreamWaitEvent(p->side, p->ev_fork, 0));
        HIPCHK(grtm::launch_rows_init(p->rows64, p->nk_pred.as<int>(), nkw, rowpk, tailmx, p->side));
        HIPCHK(grtm::launch_cells_zr(2, p->compat, p->lw64.p, p->L, p->g, p->nk, p->dk, p->df, p->omega_i, 0,
                                     p->rows64, p->low_tiles.as<unsigned>(), p->n_low_tiles, p->A0.p, p->A1.p, p->nk,
                                     nullptr, nullptr, nkw, rowpk, p->side, p->nchk.as<int>(), p->chk_stride, tailmx,
                                     nullptr, 0, nullptr, 0, false));
        HIPCHK(hipEventRecord(p->ev_join, p->side));
    }
    if (learned) {
        // the learned grid of this execution: network -> per-row cut and
        // check count -> recursion items, all on the stream
        int st;
        if (p->ev_pred0) HIPCHK(hipEventRecord(p->ev_pred0, s));
        if ((st = grtm::learned_grid(p->net, p->lmodel.as<grtm::LearnModel>(), 1, p->lg, p->nk_pred.as<int>(),
                                     p->nchk.as<int>(), s, p->flag_pw > 0.f ? p->poles.as<float>() : nullptr)))
            return st;
        // the item kernels also start the execution's state: nk_w = the cut,
        // zero row peaks, tail maxima, fallback and flag counts, item counters
        grtm::GridInit gi{};
        gi.zero[0] = fbcnt;
        gi.zero[1] = p->prec != GRTM_PREC_FP64 ? p->flags.as<unsigned>() : nullptr;
        gi.zero[2] = p->ictr.as<unsigned>();
        gi.zero[3] = p->ictr.as<unsigned>() + 1;
        gi.nk_w = nkw;
        gi.row_peak = rowpk;
        gi.tail_max = tailmx;
        HIPCHK(grtm::launch_grid_items(p->nw, r64, p->nk_pred.as<int>(), p->nchk.as<int>(), p->tiles.as<unsigned>(),
                                       p->ioff.as<unsigned>(), p->icounts.as<unsigned>(), s, gi, apart));
        if (p->ev_pred1) HIPCHK(hipEventRecord(p->ev_pred1, s));
        p->pred_stale = true;
    } else if (ad) {
        HIPCHK(grtm::launch_copy_i32(p->nk_pred.as<int>(), nkw, p->nw, s));
        HIPCHK(grtm::launch_zero_u32(rowpk, 2 * p->nw + 1, s));
    }
    // work items: the predicted grid's list (rows [0, r64) first), or every
    // tile; a learned grid's list and counts are on the device (persistent
    // launches, the item count's upper bound is the list's capacity)
    const unsigned *t_lo = ad ? p->tiles.as<unsigned>() : nullptr;
    const unsigned *t_hi = ad ? (learned ? t_lo : t_lo + p->ntiles64) : nullptr;
    const int n_lo = learned ? p->tiles_cap : ad ? p->ntiles64 : 0;
    const int n_hi = learned ? p->tiles_cap : ad ? p->ntiles - p->ntiles64 : 0;
    const unsigned *dcnt = learned ? p->icounts.as<unsigned>() : nullptr;
    unsigned *ctr = learned ? p->ictr.as<unsigned>() : nullptr;
    // adaptive: the item lists carry the residual check beyond the cut (FP64 on
    // the FP64 rows) after each launch's tiles
    const int *nchk = ad ? p->nchk.as<int>() : nullptr;
    if (p->prec == GRTM_PREC_FP64) {
        if (p->ev_main0) HIPCHK(hipEventRecord(p->ev_main0, s));
        HIPCHK(grtm::launch_cells_zr(0, p->compat, p->lw.p, p->L, p->g, p->nk, p->dk, p->df, p->omega_i, 0, p->nw,
                                     t_lo, n_lo, p->A0.p, p->A1.p, p->nk, nullptr, nullptr, nkw, rowpk, s, nchk,
                                     p->chk_stride, tailmx, dcnt, 0, ctr, 0, learned));
        if (p->ev_main1) HIPCHK(hipEventRecord(p->ev_main1, s));
    } else {
        unsigned *cnt = p->flags.as<unsigned>(), *list = cnt + 1;
        if (!learned) HIPCHK(grtm::launch_zero_u32(cnt, 1, s));
        if (p->flag_hb >= 0 || p->flag_pw > 0.f)
            HIPCHK(grtm::launch_flag_cells(p->lw64.p, p->L, p->rows64, p->nw, p->nk, p->dk, p->flag_hb,
                                           learned && p->flag_pw > 0.f ? p->poles.as<float>() : nullptr,
                                           p->flag_pw, nkw, cnt, list, s));
        if (!apart)
            HIPCHK(grtm::launch_cells_zr(2, p->compat, p->lw64.p, p->L, p->g, p->nk, p->dk, p->df, p->omega_i, 0,
                                         p->rows64, t_lo, n_lo, p->A0.p, p->A1.p, p->nk, nullptr, nullptr, nkw, rowpk,
                                         s, nchk, p->chk_stride, tailmx, dcnt, 0, ctr, 0, learned));
        if (p->ev_main0) HIPCHK(hipEventRecord(p->ev_main0, s));
        HIPCHK(grtm::launch_cells_zr(1, p->compat, p->lw.p, p->L, p->g, p->nk, p->dk, p->df, p->omega_i, p->rows64,
                                     p->nw - p->rows64, t_hi, n_hi, p->A0.p, p->A1.p, p->nk, cnt, list, nkw, rowpk, s,
                                     nchk, p->chk_stride, tailmx, dcnt, 1, ctr ? ctr + 1 : nullptr, 0, learned));
        if (p->ev_main1) HIPCHK(hipEventRecord(p->ev_main1, s));
        if (apart) HIPCHK(hipStreamWaitEvent(s, p->ev_join, 0));  // the FP64 rows are in
        if (p->nplant)
            HIPCHK(grtm::launch_plant_nonfinite(p->plant.as<unsigned>(), p->nplant, p->nk, p->A0.p, p->A1.p, p->nk,
                                                cnt, list, s));
    }
    const bool mixed = p->prec != GRTM_PREC_FP64;
    unsigned *cnt = mixed ? p->flags.as<unsigned>() : nullptr, *list = mixed ? cnt + 1 : nullptr;
    if (ad) {
        // dense fallback of the rows whose check failed (each row in its own
        // precision); the selection also forms the contraction tiles' K extents
        const void *lwf64 = mixed ? p->lw64.p : p->lw.p;
        HIPCHK(grtm::launch_fallback(!mixed, p->compat, lwf64, mixed ? p->lw.p : nullptr, r64, p->L, p->g, p->nk,
                                     p->nw, p->dk, p->df, p->omega_i, p->chk_tol, rowpk, tailmx, nkw, fbcnt,
                                     p->fb_list.as<unsigned>(), p->A0.p, p->A1.p, p->nk, cnt, list, 0, s, nullptr, 0,
                                     p->kend.as<int>(), (p->gemm_mode == 0 && !p->pair) ? 16 : 32));
    }
    // FP64 re-solve of the non-finite FP32 cells (main grid and fallback rows)
    if (mixed)
        HIPCHK(grtm::launch_cells_fix(p->compat, p->lw64.p, p->L, p->g, p->nk, p->dk, p->df, p->omega_i, p->A0.p,
                                      p->A1.p, p->nk, cnt, list, s, f16 ? rowpk : nullptr));
    return GRTM_OK;
}
int run_gemm_mt(grtm_plan *p, hipStream_t s)
{
    const size_t es = esize(p->prec);
    int M[grtm::kMtSetsMax], rpw[grtm::kMtSetsMax];
    const void *A[grtm::kMtSetsMax], *B[grtm::kMtSetsMax];
    void *C[grtm::kMtSetsMax];
    for (int q = 0; q < p->nsets; q++) {
        M[q] = p->mt_M(q);
        rpw[q] = 2 * p->mto.rows[q];
        A[q] = p->mto.A[q];
        B[q] = (const unsigned char *)p->B0.p + (size_t)q * p->nk * p->nr * es;
        C[q] = const_cast<void *>(p->mtc.C[q]);
    }
    const int *nkw = p->adaptive ? p->nk_w.as<int>() : nullptr;
    if (p->pair) {
        // adaptive: per-contraction tile extents laid end to end in kend
        const int *kend[grtm::kMtSetsMax];
        int t0 = 0;
        for (int q = 0; q < p->nsets; q++) {
            kend[q] = p->adaptive ? p->kend.as<int>() + t0 : nullptr;
            t0 += (M[q] + 127) / 128;
        }
        HIPCHK(grtm::launch_gemm_set(p->gemm_mode, p->nsets, M, rpw, p->nr, p->nk, p->kc, A, p->nk, B, p->nr, C,
                                     p->ws.p, kend, nkw, s));
        return GRTM_OK;
    }
    // unaligned shapes and BF16x3: one contraction per set, dense grid only
    // (grtm_plan_set_adaptive refuses these plans)
    for (int q = 0; q < p->nsets; q++)
        HIPCHK(grtm::launch_gemm_mfma(p->gemm_mode, M[q], p->nr, p->nk, A[q], p->nk, B[q], p->nr, C[q], p->ws.p,
                                      nullptr, nullptr, s));
    return GRTM_OK;
}

int run_gemm_core(grtm_plan *p, hipStream_t s);

// The contraction, then the PTAM continuation of every sum when set.
int run_gemm(grtm_plan *p, hipStream_t s)
{
    int st;
    if ((st = run_gemm_core(p, s))) return st;
    if (!p->ptam_n) return GRTM_OK;
    HIPCHK(grtm::launch_zero_u32(p->ptam_miss.as<unsigned>(), 1, s));
    HIPCHK(grtm::launch_ptam(p->prec == GRTM_PREC_FP64 ? 0 : 1, p->nw, p->nr, p->ptam_nt, p->ptam_n, p->dk,
                             p->dists.as<double>(), p->At0.p, p->At1.p, p->Bt0.p, p->Bt1.p, p->C0.p, p->C1.p,
                             p->ptam_miss.as<unsigned>(), s));
    return GRTM_OK;
}

int run_gemm_core(grtm_plan *p, hipStream_t s)
{
    if (p->kind == 1) return run_gemm_mt(p, s);
    const int M = 4 * p->nw;
    const int *kend = p->adaptive ? p->kend.as<int>() : nullptr;
    const int *nkw = p->adaptive ? p->nk_w.as<int>() : nullptr;
    if (p->bsplit) {
        HIPCHK(grtm::launch_gemm_pair_bsplit(p->bsplit, M, p->nr, p->nk, p->A0.as<float>(), p->A1.as<float>(), p->nk,
                                             p->S0.p, p->S1.p, p->ldk, p->C0.as<float>(), p->C1.as<float>(),
                                             p->ws.as<float>(), kend, nkw, s, nullptr, nullptr,
                                             p->bsplit == grtm::kSplitF16x3 ? p->counters.as<unsigned>() : nullptr));
        return GRTM_OK;
    }
    if (p->pair && p->gemm_mode == 1) {
        HIPCHK(grtm::launch_gemm_pair_f32(M, p->nr, p->nk, p->kc, p->A0.as<float>(), p->A1.as<float>(), p->nk,
                                          p->B0.as<float>(), p->B1.as<float>(), p->nr, p->C0.as<float>(),
                                          p->C1.as<float>(), p->ws.as<float>(), kend, nkw, s));
        return GRTM_OK;
    }
    if (p->pair && p->gemm_mode == 0) {
        HIPCHK(grtm::launch_gemm_pair_f64(M, p->nr, p->nk, p->kc, p->A0.as<double>(), p->A1.as<double>(), p->nk,
                                          p->B0.as<double>(), p->B1.as<double>(), p->nr, p->C0.as<double>(),
                                          p->C1.as<double>(), p->ws.as<double>(), kend, nkw, s));
        return GRTM_OK;
    }
    HIPCHK(grtm::launch_gemm_mfma(p->gemm_mode, M, p->nr, p->nk, p->A0.p, p->nk, p->B0.p, p->nr, p->C0.p, p->ws.p,
                                  kend, nkw, s));
    HIPCHK(grtm::launch_gemm_mfma(p->gemm_mode, M, p->nr, p->nk, p->A1.p, p->nk, p->B1.p, p->nr, p->C1.p, p->ws.p,
                                  kend, nkw, s));
    return GRTM_OK;
}
int run_ifft(grtm_plan *p, hipStream_t s)
{
    if (p->kind == 1) {
        HIPCHK(grtm::launch_post_ifft_cols(p->prec == GRTM_PREC_FP64 ? 0 : 1, p->pp, p->mtc_d.as<grtm::MtCols>(),
                                           p->ncols, p->nr, p->nr, p->dists.as<double>(), p->tr.p, s));
        return GRTM_OK;
    }
    HIPCHK(grtm::launch_post_ifft(p->prec == GRTM_PREC_FP64 ? 0 : 1, p->pp, 1, 4 * p->nr, nullptr, p->C0.p, p->C1.p,
                                  p->nr, p->nr, p->tr.p, nullptr, s));
    return GRTM_OK;
}

// Terms of each column over the contraction outputs (rows of
// mt_rows_from_cell).  Displacement columns: tensor 0-9, then single force
// 0-4; k J_n' = n J_n / r - k J_{n+1} turns, e.g., tensor column 0,
// b2 2 J2 / r + a2[0] k J2', into (2 / r) (b2 + a2[0]) J2 - a2[0] k J3.
// With deriv (physics mode) each source's d/dz columns follow its
// displacement columns (the same terms on the d/dz rows), then its d/dr
// columns: a term s r^-q S_m gives s (m - q) r^-(q+1) S_m - s r^-q S'_{m+1}
// with S'_{m+1} the k x row in the next order; output order tensor 0-29
// (as the reference's physics-mode columns), single force 30-44 (its 0-14).
grtm::MtCols mt_columns(int compat, int deriv)
{
    struct Tm {
        int set, row, rpow;
        double scale;
    };
    std::vector<std::vector<Tm>> base(15);
    auto term = [&](int c, int set, int row, double scale, int rpow) { base[c].push_back({set, row, rpow, scale}); };
    term(0, 2, 0, 2.0, 1), term(0, 3, 0, -1.0, 0);                  // R, n = 2
    term(1, 1, 3, 1.0, 1), term(1, 2, 1, -1.0, 0);                  // R, n = 1
    term(2, 1, 0, -1.0, 0);                                         // R, n = 0 (s2)
    term(3, 1, 1, -1.0, 0);                                         // R, n = 0 (s0)
    term(4, 2, 0, 2.0, 1), term(4, 3, 1, -1.0, 0);                  // T, n = 2
    term(5, 1, compat ? 7 : 3, 1.0, 1), term(5, 2, 2, -1.0, 0);     // T, n = 1
    term(6, 2, 3, 1.0, 0);                                          // Z, n = 2
    term(7, 1, 2, 1.0, 0);                                          // Z, n = 1
    term(8, 0, 0, 1.0, 0);                                          // Z, n = 0 (s0)
    term(9, 0, 1, 1.0, 0);                                          // Z, n = 0 (s2)
    term(10, 1, 6, 1.0, 1), term(10, 2, 4, -1.0, 0);                // horizontal force R
    term(11, 1, 4, -1.0, 0);                                        // vertical force R
    term(12, 1, 6, 1.0, 1), term(12, 2, 5, -1.0, 0);                // horizontal force T
    term(13, 1, 5, 1.0, 0);                                         // horizontal force Z
    term(14, 0, 2, 1.0, 0);                                         // vertical force Z
    std::vector<std::vector<Tm>> cols;
    if (!deriv) {
        cols = base;
    } else {
        auto dz = [&](const std::vector<Tm> &c) {
            std::vector<Tm> o = c;
            for (Tm &t : o) t.row += grtm::mt_disp_rows(0, t.set);
            return o;
        };
        auto dr = [&](const std::vector<Tm> &c) {
            std::vector<Tm> o;
            for (const Tm &t : c) {
                if (t.set != t.rpow) o.push_back({t.set, t.row, t.rpow + 1, t.scale * (t.set - t.rpow)});
                o.push_back({t.set + 1, 2 * grtm::mt_disp_rows(0, t.set + 1) + t.row, t.rpow, -t.scale});
            }
            return o;
        };
        for (int blk = 0; blk < 3; blk++)
            for (int c = 0; c < 10; c++) cols.push_back(blk == 0 ? base[c] : blk == 1 ? dz(base[c]) : dr(base[c]));
        for (int blk = 0; blk < 3; blk++)
            for (int c = 10; c < 15; c++)
                cols.push_back(blk == 0 ? base[c] : blk == 1 ? dz(base[c]) : dr(base[c]));
    }
    grtm::MtCols mc{};
    for (size_t c = 0; c < cols.size(); c++) {
        mc.nterm[c] = (int)cols[c].size();
        for (size_t u = 0; u < cols[c].size(); u++)
            mc.term[c][u] = grtm::MtTerm{cols[c][u].set, cols[c][u].row, cols[c][u].rpow, 0, cols[c][u].scale};
    }
    for (int q = 0; q < grtm::mt_sets(deriv); q++) mc.rows[q] = grtm::mt_set_rows(compat, deriv, q);
    return mc;
}

}  // namespace

extern "C" {

int grtm_create(grtm_ctx **ctx, int device)
{
    if (!ctx) return GRTM_E_INPUT;
    *ctx = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return GRTM_E_DEVICE;
    if (device < 0 || device >= n) return GRTM_E_INPUT;
    HIPCHK(hipSetDevice(device));
    grtm_ctx *c = new grtm_ctx;
    c->device = device;
    *ctx = c;
    return GRTM_OK;
}

void grtm_destroy(grtm_ctx *ctx) { delete ctx; }

static int plan_create(grtm_ctx *ctx, const grtm_job *job, const grtm_grid *grid, int precision, int compat,
                       int kind, int deriv, grtm_plan **plan_out)
{
    if (!ctx || !job || !grid || !plan_out) return GRTM_E_INPUT;
    *plan_out = nullptr;
    if (precision != GRTM_PREC_FP64 && precision != GRTM_PREC_FP32 && precision != GRTM_PREC_BF16X3 &&
        precision != GRTM_PREC_BF16X6 && precision != GRTM_PREC_FP16X3)
        return GRTM_E_INPUT;
    if (precision == GRTM_PREC_FP16X3 && kind != 0) return GRTM_E_UNSUPPORTED;  // Z / R plans only
    if (job->n_layers < 2 || job->n_layers + 1 > grtm::kMaxLayers || !job->layers || job->nr < 1 || !job->dists)
        return GRTM_E_INPUT;
    if (grid->nw < 4 || (grid->nw & (grid->nw - 1)) || grid->nw > 8192 || grid->nk < 1 || !(grid->kmax > 0) ||
        !(grid->dt > 0) || !(grid->taper >= 0 && grid->taper <= 1))
        return GRTM_E_INPUT;
    if (precision == GRTM_PREC_FP64 && grid->nw > 4096) return GRTM_E_UNSUPPORTED;  // IFFT LDS limit
    if (!grtm::cells_tile_limit(grid->nw, grid->nk)) return GRTM_E_UNSUPPORTED;
    for (int i = 0; i < job->nr; i++)
        if (!(job->dists[i] > 0)) return GRTM_E_INPUT;
    HIPCHK(hipSetDevice(ctx->device));
    const int nw = grid->nw;
    const double T = 2.0 * nw * grid->dt;
    const double df = 1.0 / T, omega_i = grid->omega_i > 0 ? grid->omega_i : M_PI / T;
    ModelTables mt;
    int status = model_tables(job->layers, job->n_layers, job->src_depth, job->rcv_depth, compat, nw, df, omega_i,
                              true, mt);
    if (status) return status;
    grtm_plan *p = new grtm_plan;
    struct Guard {
        grtm_plan *&p;
        ~Guard() { delete p; }
    } guard{p};
    p->layers_h.assign(job->layers, job->layers + 6 * (size_t)job->n_layers);
    p->src_depth = job->src_depth;
    p->rcv_depth = job->rcv_depth;
    p->rows64_rule = mt.rows64;
    p->g = mt.g;
    p->ctx = ctx;
    p->prec = precision;
    p->gemm_mode = precision == GRTM_PREC_FP64 ? 0 : precision == GRTM_PREC_BF16X3 ? 2 : 1;
    p->compat = compat;
    p->nw = nw;
    p->N = 2 * nw;
    p->nk = grid->nk;
    p->nr = job->nr;
    p->L = mt.st.L;
    p->dk = grid->kmax / grid->nk;
    p->df = df;
    p->omega_i = omega_i;
    p->pp = post_params(grid, nw, df, omega_i);
    if ((status = post_tables(p->pp, precision == GRTM_PREC_FP64 ? 0 : 1, p->ptab))) return status;
    p->kind = kind;
    if (kind == 1) {
        p->deriv = deriv ? 1 : 0;
        p->nsets = grtm::mt_sets(p->deriv);
        p->ncols = p->deriv ? grtm::kMtColsMax : 15;
    }
    HIPCHK(hipStreamCreate(&p->stream));
    p->own_stream = true;

    // layer tables
    const std::vector<LayerW<double>> &lw = mt.lw;
    const size_t es = esize(precision);
    if (precision == 0) {
        if ((status = p->lw.alloc(lw.size() * sizeof(LayerW<double>)))) return status;
        HIPCHK(hipMemcpy(p->lw.p, lw.data(), lw.size() * sizeof(LayerW<double>), hipMemcpyHostToDevice));
    } else {
        p->rows64 = mt.rows64;
        if ((status = p->lw64.alloc(lw.size() * sizeof(LayerW<double>)))) return status;
        HIPCHK(hipMemcpy(p->lw64.p, lw.data(), lw.size() * sizeof(LayerW<double>), hipMemcpyHostToDevice));
        if ((status = p->flags.alloc(sizeof(unsigned) * ((size_t)p->nw * p->nk + 1)))) return status;
        const std::vector<LayerW<float>> lf = narrow_tables(lw);
        if ((status = p->lw.alloc(lf.size() * sizeof(LayerW<float>)))) return status;
        HIPCHK(hipMemcpy(p->lw.p, lf.data(), lf.size() * sizeof(LayerW<float>), hipMemcpyHostToDevice));
    }
    if ((status = p->dists.alloc(sizeof(double) * p->nr))) return status;
    if ((status = p->tr.alloc((size_t)p->ntr() * p->N * es))) return status;
    if (kind == 1) {
        // four contractions, operands / tables / outputs end to end
        p->mtc = mt_columns(compat, p->deriv);
        size_t rows = 0, off[grtm::kMtSetsMax];
        int M[grtm::kMtSetsMax];
        for (int q = 0; q < p->nsets; q++) {
            p->mto.rows[q] = p->mtc.rows[q];
            M[q] = p->mt_M(q);
            off[q] = rows;
            rows += (size_t)M[q];
        }
        if ((status = p->A0.alloc(rows * p->nk * es))) return status;
        if ((status = p->B0.alloc((size_t)p->nsets * p->nk * p->nr * es))) return status;
        if ((status = p->C0.alloc(rows * p->nr * es))) return status;
        bool aligned = p->gemm_mode != 2 && p->nk % 4 == 0 && p->nr % 4 == 0;
        for (int q = 0; q < p->nsets; q++) {
            p->mto.A[q] = (unsigned char *)p->A0.p + off[q] * p->nk * es;
            p->mtc.C[q] = (unsigned char *)p->C0.p + off[q] * p->nr * es;
            const void *Bq = (unsigned char *)p->B0.p + (size_t)q * p->nk * p->nr * es;
            aligned = aligned && grtm::gemm_pair_ok(p->nk, p->nr, p->mto.A[q], p->mto.A[q], Bq, Bq);
        }
        size_t wsb = 0;
        for (int q = 0; q < p->nsets; q++)
            wsb = std::max(wsb, grtm::gemm_workspace_bytes(p->gemm_mode, M[q], p->nr, p->nk));
        if (aligned) {
            p->pair = true;
            p->kc = grtm::gemm_set_slice(p->gemm_mode, p->nsets, M, p->nr, p->nk);
            wsb = std::max(wsb, grtm::gemm_set_workspace_bytes(p->nsets, M, p->nr, p->nk, p->kc, es));
        }
        if ((status = p->ws.alloc(wsb))) return status;
        if ((status = p->mtc_d.alloc(sizeof(grtm::MtCols)))) return status;
        HIPCHK(hipMemcpy(p->mtc_d.p, &p->mtc, sizeof(grtm::MtCols), hipMemcpyHostToDevice));
    } else {
        const size_t arows = (size_t)4 * p->nw;
        if ((status = p->A0.alloc(arows * p->nk * es))) return status;
        if ((status = p->A1.alloc(arows * p->nk * es))) return status;
        if ((status = p->B0.alloc((size_t)p->nk * p->nr * es))) return status;
        if ((status = p->B1.alloc((size_t)p->nk * p->nr * es))) return status;
        if ((status = p->C0.alloc(arows * p->nr * es))) return status;
        if ((status = p->C1.alloc(arows * p->nr * es))) return status;
        const bool pair_ok = grtm::gemm_pair_ok(p->nk, p->nr, p->A0.p, p->A1.p, p->B0.p, p->B1.p);
        if (pair_ok && (precision == GRTM_PREC_BF16X3 || precision == GRTM_PREC_BF16X6)) {
            p->gemm_mode = 1;
            p->bsplit = precision == GRTM_PREC_BF16X3 ? grtm::kSplitBf16x3 : grtm::kSplitBf16x6;
        }
        if (precision == GRTM_PREC_FP16X3) {
            // the FP16 split runs on the pair layout only; its row scales come
            // from the recursion's per-row peaks, kept on dense grids too
            if (!pair_ok) return GRTM_E_UNSUPPORTED;
            p->bsplit = grtm::kSplitF16x3;
            if ((status = p->counters.alloc(sizeof(unsigned) * (2 * p->nw + 1)))) return status;
        }
        size_t wsb = grtm::gemm_workspace_bytes(p->gemm_mode, (int)arows, p->nr, p->nk);
        if (p->gemm_mode != 2 && pair_ok) {
            p->pair = true;
            p->kc = grtm::gemm_pair_slice(p->gemm_mode, (int)arows, p->nr, p->nk, p->nk);
            wsb = std::max(wsb, grtm::gemm_pair_workspace_bytes((int)arows, p->nr, p->nk, p->kc) * es / 4);
        }
        if (p->bsplit) {
            wsb = std::max(wsb, pair_sk_bytes(p->bsplit, (int)arows, p->nr, p->nk));
            if ((status = split_alloc(p->bsplit, p->nk, p->nr, p->S0, p->S1, p->ldk))) return status;
        }
        if ((status = p->ws.alloc(wsb))) return status;
    }
    HIPCHK(hipMemcpy(p->dists.p, job->dists, sizeof(double) * p->nr, hipMemcpyHostToDevice));

    // Bessel tables (planning work: they depend on the grid and distances only)
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, p->stream));
    if (kind == 1) {
        const size_t tab = (size_t)p->nk * p->nr * es;
        unsigned char *b = (unsigned char *)p->B0.p;
        for (int n0 = 0; n0 < p->nsets; n0 += 2)
            HIPCHK(grtm::launch_bessel_tab(precision == GRTM_PREC_FP64 ? 0 : 1, compat ? 0 : 1, p->nk, p->dk,
                                           p->dists.as<double>(), p->nr, b + n0 * tab,
                                           n0 + 1 < p->nsets ? b + (n0 + 1) * tab : nullptr, p->nr, p->stream, n0));
    } else {
        HIPCHK(grtm::launch_bessel_tab(precision == GRTM_PREC_FP64 ? 0 : 1, compat ? 0 : 1, p->nk, p->dk,
                                       p->dists.as<double>(), p->nr, p->B0.p, p->B1.p, p->nr, p->stream));
        if (p->bsplit) {
            HIPCHK(grtm::launch_split_table(p->bsplit, p->nk, p->nr, p->B0.as<float>(), p->nr, p->S0.p, p->ldk,
                                            p->stream));
            HIPCHK(grtm::launch_split_table(p->bsplit, p->nk, p->nr, p->B1.as<float>(), p->nr, p->S1.p, p->ldk,
                                            p->stream));
        }
    }
    HIPCHK(hipEventRecord(e1, p->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    p->bessel_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *plan_out = p;
    p = nullptr;  // released from the guard
    return GRTM_OK;
}

int grtm_plan_traces(grtm_ctx *ctx, const grtm_job *job, const grtm_grid *grid, int precision, int compat,
                     grtm_plan **plan_out)
{
    return plan_create(ctx, job, grid, precision, compat, 0, 0, plan_out);
}

int grtm_plan_columns(grtm_ctx *ctx, const grtm_job *job, const grtm_grid *grid, int precision, int compat,
                      int derivatives, grtm_plan **plan_out)
{
    if (plan_out) *plan_out = nullptr;
    if (derivatives && compat) return GRTM_E_UNSUPPORTED;  // the reference's d/dr columns are not J'' (Q7, Q8)
    return plan_create(ctx, job, grid, precision, compat, 1, derivatives, plan_out);
}

int grtm_plan_execute(grtm_plan *p, void *stream)
{
    if (!p) return GRTM_E_INPUT;
    hipStream_t s = stream ? (hipStream_t)stream : p->stream;
    p->last_stream = s;
    int st;
    if ((st = run_cells(p, s))) return st;
    if ((st = run_gemm(p, s))) return st;
    return run_ifft(p, s);
}

int grtm_plan_sync(grtm_plan *p)
{
    if (!p) return GRTM_E_INPUT;
    HIPCHK(hipStreamSynchronize(p->stream));
    if (p->last_stream && p->last_stream != p->stream) HIPCHK(hipStreamSynchronize(p->last_stream));
    return GRTM_OK;
}

int grtm_plan_traces_out(grtm_plan *p, double *traces)
{
    if (!p || !traces) return GRTM_E_INPUT;
    const size_t n = (size_t)p->ntr() * p->N;
    HIPCHK(hipStreamSynchronize(p->stream));
    if (p->prec == GRTM_PREC_FP64) {
        HIPCHK(hipMemcpy(traces, p->tr.p, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> f(n);
        HIPCHK(hipMemcpy(f.data(), p->tr.p, n * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) traces[i] = f[i];
    }
    return GRTM_OK;
}

int grtm_plan_kernels_out(grtm_plan *p, int q, int w_lo, int w_hi, double *kern)
{
    if (!p || !kern || q < 0 || q > 1 || w_lo < 0 || w_hi > p->nw || w_lo >= w_hi) return GRTM_E_INPUT;
    if (p->kind != 0) return GRTM_E_UNSUPPORTED;
    HIPCHK(hipStreamSynchronize(p->stream));
    const size_t es = esize(p->prec);
    std::vector<unsigned char> buf((size_t)p->nk * es);
    for (int iw = w_lo; iw < w_hi; iw++) {
        for (int part = 0; part < 2; part++) {        // 0: Z (A0), 1: R (A1)
            for (int ri = 0; ri < 2; ri++) {           // re, im
                const DBuf &A = part ? p->A1 : p->A0;
                const size_t row = (size_t)4 * iw + 2 * q + ri;
                HIPCHK(hipMemcpy(buf.data(), (const unsigned char *)A.p + row * p->nk * es, p->nk * es,
                                 hipMemcpyDeviceToHost));
                for (int ik = 0; ik < p->nk; ik++) {
                    double v = p->prec == GRTM_PREC_FP64 ? ((const double *)buf.data())[ik]
                                                         : ((const float *)buf.data())[ik];
                    const double wt = ik == 0 ? 0.5 * p->dk : p->dk;
                    kern[((size_t)(iw - w_lo) * p->nk + ik) * 4 + part * 2 + ri] = v / wt;
                }
            }
        }
    }
    return GRTM_OK;
}

int grtm_plan_timing(grtm_plan *p, int reps, double *ms)
{
    if (!p || !ms || reps < 1) return GRTM_E_INPUT;
    hipEvent_t ev[8];
    for (int i = 0; i < 8; i++) HIPCHK(hipEventCreate(&ev[i]));
    double acc[6] = {0, 0, 0, 0, 0, 0};
    int st = GRTM_OK;
    p->ev_main0 = ev[4];
    p->ev_main1 = ev[5];
    const bool learned = p->adaptive && p->net && p->kind == 0;
    if (learned) p->ev_pred0 = ev[6], p->ev_pred1 = ev[7];
    for (int r = 0; r < reps && !st; r++) {
        if (hipEventRecord(ev[0], p->stream) != hipSuccess) { st = GRTM_E_DEVICE; break; }
        if ((st = run_cells(p, p->stream))) break;
        if (hipEventRecord(ev[1], p->stream) != hipSuccess) { st = GRTM_E_DEVICE; break; }
        if ((st = run_gemm(p, p->stream))) break;
        if (hipEventRecord(ev[2], p->stream) != hipSuccess) { st = GRTM_E_DEVICE; break; }
        if ((st = run_ifft(p, p->stream))) break;
        if (hipEventRecord(ev[3], p->stream) != hipSuccess || hipEventSynchronize(ev[3]) != hipSuccess) {
            st = GRTM_E_DEVICE;
            break;
        }
        float a, b, c, d, e;
        if (hipEventElapsedTime(&a, ev[0], ev[3]) != hipSuccess || hipEventElapsedTime(&b, ev[0], ev[1]) != hipSuccess ||
            hipEventElapsedTime(&c, ev[1], ev[2]) != hipSuccess || hipEventElapsedTime(&d, ev[2], ev[3]) != hipSuccess ||
            hipEventElapsedTime(&e, ev[4], ev[5]) != hipSuccess) {
            st = GRTM_E_DEVICE;
            break;
        }
        float f = 0;
        if (learned && hipEventElapsedTime(&f, ev[6], ev[7]) != hipSuccess) {
            st = GRTM_E_DEVICE;
            break;
        }
        acc[0] += a, acc[1] += b, acc[2] += c, acc[3] += d, acc[4] += e, acc[5] += f;
    }
    p->ev_main0 = p->ev_main1 = p->ev_pred0 = p->ev_pred1 = nullptr;
    for (int i = 0; i < 8; i++) (void)hipEventDestroy(ev[i]);
    if (st) return st;
    for (int i = 0; i < 4; i++) ms[i] = acc[i] / reps;
    ms[4] = p->bessel_ms;
    ms[5] = acc[4] / reps;
    ms[6] = acc[5] / reps;
    return GRTM_OK;
}

// Recursion work items of the predicted grid, rows solved in FP64 first.
static int build_items(grtm_plan *p)
{
    const int r64 = p->prec == GRTM_PREC_FP64 ? p->nw : p->rows64;
    std::vector<unsigned> items;
    int n64 = 0, st;
    // each launch's residual-check items follow its tiles, so the check cells
    // run in the launch's tail
    auto checks = [&](int w_lo, int w_hi) {
        for (int iw = w_lo; iw < w_hi; iw++)
            for (int c = 0; c < grtm::cells_tiles_per_row(p->nchk_h[iw]); c++)
                items.push_back(grtm::cells_tile(iw, c) | grtm::kCheckItem);
    };
    for (int iw = 0; iw < p->nw; iw++) {
        if (iw == r64) {
            checks(0, r64);
            n64 = (int)items.size();
        }
        const int nt = grtm::cells_tiles_per_row(p->nk_pred_h[iw]);
        for (int kb = 0; kb < nt; kb++) items.push_back(grtm::cells_tile(iw, kb));
    }
    if (r64 >= p->nw) {
        checks(0, p->nw);
        n64 = (int)items.size();
    } else {
        checks(r64, p->nw);
    }
    if ((st = p->tiles.alloc(sizeof(unsigned) * std::max<size_t>(1, items.size())))) return st;
    if (!items.empty())
        HIPCHK(hipMemcpy(p->tiles.p, items.data(), sizeof(unsigned) * items.size(), hipMemcpyHostToDevice));
    p->ntiles = (int)items.size();
    p->ntiles64 = n64;
    return GRTM_OK;
}

static void row_cut(int nk, double kcut, double dk, int stride, double window, int &n, int &nc);

// Mixed learned plans: rows [0, rows64), solved in FP64, take the analytic
// cut kc = min(1.15 w / Vs_min + ln(1e7) / dz, kmax) (the floor the network's
// cut keeps on those rows: on the 256-receiver model the network's cut is
// below it on every such row) with the residual check over the window beyond
// it.  Their nk_pred / nchk and item list (tiles, then check items) are set
// here once, so their launch needs nothing from the network and runs on the
// side stream beside it.
static int set_low_rows(grtm_plan *p)
{
    p->low_apart = false;
    p->lg.low_fixed = 0;
    if (p->prec == GRTM_PREC_FP64 || !p->net || p->rows64 < 1) return GRTM_OK;
    // no execution may still read the buffers rewritten below
    HIPCHK(hipStreamSynchronize(p->stream));
    if (p->last_stream && p->last_stream != p->stream) HIPCHK(hipStreamSynchronize(p->last_stream));
    if (p->side) HIPCHK(hipStreamSynchronize(p->side));
    if (!p->side) {
        HIPCHK(hipStreamCreate(&p->side));
        HIPCHK(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming));
    }
    const int r64 = p->rows64;
    std::vector<int> n(r64), nc(r64);
    std::vector<unsigned> items;
    const double dz = std::max(std::fabs(p->lm.src - p->lm.rcv), 0.5);
    for (int iw = 0; iw < r64; iw++) {
        const double w = 2.0 * M_PI * iw / (2.0 * p->lg.nw * p->lg.dt);
        const double kc = std::min(1.15 * w / p->lm.vs_min + std::log(1e7) / dz, p->lg.kmax);
        row_cut(p->nk, kc, p->dk, p->chk_stride, p->lm.window, n[iw], nc[iw]);
        for (int kb = 0; kb < grtm::cells_tiles_per_row(n[iw]); kb++) items.push_back(grtm::cells_tile(iw, kb));
    }
    for (int iw = 0; iw < r64; iw++)
        for (int c = 0; c < grtm::cells_tiles_per_row(nc[iw]); c++)
            items.push_back(grtm::cells_tile(iw, c) | grtm::kCheckItem);
    int st;
    const size_t bytes = sizeof(unsigned) * std::max<size_t>(1, items.size());
    if (p->low_tiles.n < bytes && (st = p->low_tiles.alloc(bytes))) return st;
    if (!items.empty())
        HIPCHK(hipMemcpy(p->low_tiles.p, items.data(), sizeof(unsigned) * items.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p->nk_pred.p, n.data(), sizeof(int) * r64, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p->nchk.p, nc.data(), sizeof(int) * r64, hipMemcpyHostToDevice));
    if ((int)p->nk_pred_h.size() == p->nw && (int)p->nchk_h.size() == p->nw)
        for (int iw = 0; iw < r64; iw++) p->nk_pred_h[iw] = n[iw], p->nchk_h[iw] = nc[iw];
    p->n_low_tiles = (int)items.size();
    p->lg.low_fixed = 1;
    p->low_apart = true;
    return GRTM_OK;
}

// The learned grid's predictor inputs to the device (the floor rows follow the
// plan's FP64 split).
static int upload_learn_model(grtm_plan *p)
{
    p->lm.floor_rows = p->prec == GRTM_PREC_FP64 ? p->rows64_rule : p->rows64;
    HIPCHK(hipMemcpy(p->lmodel.p, &p->lm, sizeof(grtm::LearnModel), hipMemcpyHostToDevice));
    return set_low_rows(p);
}

int grtm_plan_set_ptam(grtm_plan *p, int n_ptam, int tail)
{
    if (!p || n_ptam < 0 || n_ptam == 1 || n_ptam > GRTM_PTAM_MAX || tail < 0) return GRTM_E_INPUT;
    if (p->kind != 0 || p->adaptive) return GRTM_E_UNSUPPORTED;
    HIPCHK(hipStreamSynchronize(p->stream));
    if (p->last_stream && p->last_stream != p->stream) HIPCHK(hipStreamSynchronize(p->last_stream));
    if (n_ptam == 0) {
        p->ptam_n = p->ptam_nt = 0;
        return GRTM_OK;
    }
    std::vector<double> d(p->nr);
    HIPCHK(hipMemcpy(d.data(), p->dists.p, sizeof(double) * p->nr, hipMemcpyDeviceToHost));
    const double rmin = *std::min_element(d.begin(), d.end());
    if (!(rmin > 0)) return GRTM_E_INPUT;
    // (n/2 + 2) periods of J(k r) at the nearest receiver, at most 4 grids long
    const double want = std::ceil((0.5 * n_ptam + 2.0) * 2.0 * M_PI / (rmin * p->dk));
    const int nt = tail ? tail : (int)std::min<double>(want, 4.0 * p->nk);
    if (nt < 2 || nt > 4 * p->nk || !grtm::cells_tile_limit(p->nw, nt)) return GRTM_E_UNSUPPORTED;
    const size_t es = esize(p->prec);
    int st;
    if ((st = p->At0.alloc((size_t)4 * p->nw * nt * es)) || (st = p->At1.alloc((size_t)4 * p->nw * nt * es)) ||
        (st = p->Bt0.alloc((size_t)nt * p->nr * es)) || (st = p->Bt1.alloc((size_t)nt * p->nr * es)) ||
        (st = p->ptam_miss.alloc(sizeof(unsigned))))
        return st;
    HIPCHK(grtm::launch_bessel_tab(p->prec == GRTM_PREC_FP64 ? 0 : 1, p->compat ? 0 : 1, nt, p->dk,
                                   p->dists.as<double>(), p->nr, p->Bt0.p, p->Bt1.p, p->nr, p->stream, 0, p->nk));
    HIPCHK(grtm::launch_zero_u32(p->ptam_miss.as<unsigned>(), 1, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    p->ptam_n = n_ptam;
    p->ptam_nt = nt;
    return GRTM_OK;
}

int grtm_plan_ptam_stats(grtm_plan *p, double *out)
{
    if (!p || !out) return GRTM_E_INPUT;
    out[0] = p->ptam_n;
    out[1] = p->ptam_nt;
    out[2] = 0;
    if (p->ptam_n) {
        unsigned m = 0;
        HIPCHK(hipStreamSynchronize(p->stream));
        if (p->last_stream && p->last_stream != p->stream) HIPCHK(hipStreamSynchronize(p->last_stream));
        HIPCHK(hipMemcpy(&m, p->ptam_miss.p, sizeof(unsigned), hipMemcpyDeviceToHost));
        out[2] = m;
    }
    out[3] = (double)p->nw * 4 * p->nr;
    return GRTM_OK;
}

// Room in the flag list for every cell of the grid once (a non-finite FP32
// result), the planted cells and the cells of grtm_plan_set_flag64.
static int alloc_flags(grtm_plan *p, int nplant, int hb, float pw)
{
    const size_t rows = (size_t)p->nw;  // any rows64 set later
    const size_t per_row = (hb >= 0 ? (size_t)2 * p->L * (2 * hb + 1) : 0) + (pw > 0.f ? 8 * (2 * 64 + 2) : 0);
    return p->flags.alloc(sizeof(unsigned) * ((size_t)p->nw * p->nk + 1 + nplant + rows * per_row));
}

int grtm_plan_plant_nonfinite(grtm_plan *p, const int *cells, int n)
{
    if (!p || n < 0 || n > GRTM_MAX_PLANT || (n && !cells)) return GRTM_E_INPUT;
    if (p->kind != 0 || p->prec == GRTM_PREC_FP64) return GRTM_E_UNSUPPORTED;
    std::vector<unsigned> c(std::max(n, 1));
    for (int i = 0; i < n; i++) {
        const int iw = cells[2 * i], ik = cells[2 * i + 1];
        if (iw < p->rows64 || iw >= p->nw || ik < 0 || ik >= p->nk) return GRTM_E_INPUT;
        c[i] = unsigned(iw) * unsigned(p->nk) + unsigned(ik);
    }
    HIPCHK(hipStreamSynchronize(p->stream));
    if (p->last_stream && p->last_stream != p->stream) HIPCHK(hipStreamSynchronize(p->last_stream));
    int st;
    if ((st = alloc_flags(p, n, p->flag_hb, p->flag_pw))) return st;
    if ((st = p->plant.alloc(sizeof(unsigned) * c.size()))) return st;
    HIPCHK(hipMemcpy(p->plant.p, c.data(), sizeof(unsigned) * c.size(), hipMemcpyHostToDevice));
    p->nplant = n;
    return GRTM_OK;
}

int grtm_plan_set_flag64(grtm_plan *p, int branch_cells, double pole_widths)
{
    if (!p || branch_cells > GRTM_MAX_FLAG_CELLS || !(pole_widths >= 0 && pole_widths <= 8)) return GRTM_E_INPUT;
    if (p->kind != 0 || p->prec == GRTM_PREC_FP64 || p->ptam_n) return GRTM_E_UNSUPPORTED;
    if (pole_widths > 0 && !p->net) return GRTM_E_UNSUPPORTED;  // clusters come from the network
    const int hb = branch_cells < 0 ? -1 : branch_cells;
    HIPCHK(hipStreamSynchronize(p->stream));
    if (p->last_stream && p->last_stream != p->stream) HIPCHK(hipStreamSynchronize(p->last_stream));
    int st;
    if (pole_widths > 0 && !p->poles.p && (st = p->poles.alloc(sizeof(float) * 16 * (size_t)p->nw))) return st;
    if ((st = alloc_flags(p, p->nplant, hb, float(pole_widths)))) return st;
    p->flag_hb = hb;
    p->flag_pw = float(pole_widths);
    return GRTM_OK;
}

int grtm_plan_flag_list(grtm_plan *p, unsigned *cells, int cap)
{
    if (!p || cap < 0 || (cap && !cells)) return GRTM_E_INPUT;
    if (p->kind != 0 || p->prec == GRTM_PREC_FP64) return GRTM_E_UNSUPPORTED;
    HIPCHK(hipStreamSynchronize(p->stream));
    if (p->last_stream && p->last_stream != p->stream) HIPCHK(hipStreamSynchronize(p->last_stream));
    unsigned c = 0;
    HIPCHK(hipMemcpy(&c, p->flags.p, sizeof(unsigned), hipMemcpyDeviceToHost));
    const size_t n = std::min<size_t>(c, (size_t)cap);
    if (n) HIPCHK(hipMemcpy(cells, p->flags.as<unsigned>() + 1, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
    return (int)std::min<unsigned>(c, (unsigned)INT_MAX);
}

int grtm_plan_set_rows64(grtm_plan *p, int rows64)
{
    if (!p || p->prec == GRTM_PREC_FP64 || rows64 < 0 || rows64 > p->nw) return GRTM_E_INPUT;
    p->rows64 = rows64;
    if (p->adaptive && p->net) return upload_learn_model(p);
    return p->adaptive ? build_items(p) : GRTM_OK;
}

// Cut n and residual-check cells nc of one frequency row.  Check cell j
// sits at wavenumber n + stride j + stride / 2 and must lie below nk; a row
// cut below nk with no check cell (no room for one, or a zero window) is
// kept dense, since its tail would otherwise be neither checked nor solved.
static void row_cut(int nk, double kcut, double dk, int stride, double window, int &n, int &nc)
{
    n = (int)std::min<double>(nk, std::max(1.0, std::ceil(kcut / dk)));
    const int room = nk - n;
    const int fit = room > stride / 2 ? (room - stride / 2 - 1) / stride + 1 : 0;
    nc = std::max(0, std::min(fit, (int)std::ceil(window / (dk * stride))));
    if (room > 0 && nc == 0) n = nk;
}

int grtm_plan_set_adaptive(grtm_plan *p, const double *kcut, double window, int stride, double tol)
{
    if (!p) return GRTM_E_INPUT;
    if (p->ptam_n && kcut) return GRTM_E_UNSUPPORTED;  // the PTAM tail continues the dense grid
    // column plans: only the streamed contraction set skips the rows' tails
    if (p->kind == 1 && !p->pair) return GRTM_E_UNSUPPORTED;
    int Mq[grtm::kMtSetsMax], ntile = (p->nw + 15) / 16;
    if (p->kind == 1) {
        ntile = 0;
        for (int q = 0; q < p->nsets; q++) {
            Mq[q] = p->mt_M(q);
            ntile += (Mq[q] + 127) / 128;
        }
    }
    p->net = nullptr;  // a host cut replaces a learned grid
    p->pred_stale = false;
    if (!kcut) {
        p->adaptive = false;
        if (p->pair)
            p->kc = p->kind == 1 ? grtm::gemm_set_slice(p->gemm_mode, p->nsets, Mq, p->nr, p->nk)
                                 : grtm::gemm_pair_slice(p->gemm_mode, 4 * p->nw, p->nr, p->nk, p->nk);
        return GRTM_OK;
    }
    if (!(window >= 0) || stride < 1 || !(tol > 0)) return GRTM_E_INPUT;
    for (int iw = 0; iw < p->nw; iw++)
        if (!(kcut[iw] >= 0)) return GRTM_E_INPUT;  // validated before any plan state changes
    p->nk_pred_h.assign(p->nw, 0);
    p->nchk_h.assign(p->nw, 0);
    int mx = 0;
    for (int iw = 0; iw < p->nw; iw++) {
        int n, nc;
        row_cut(p->nk, kcut[iw], p->dk, stride, window, n, nc);
        p->nk_pred_h[iw] = n;
        p->nchk_h[iw] = nc;
        mx = std::max(mx, p->nchk_h[iw]);
    }
    int st;
    if (!p->nk_pred.p) {
        if ((st = p->nk_pred.alloc(sizeof(int) * p->nw)) || (st = p->nk_w.alloc(sizeof(int) * p->nw)) ||
            (st = p->nchk.alloc(sizeof(int) * p->nw)) || (st = p->counters.alloc(sizeof(unsigned) * (2 * p->nw + 1))) ||
            (st = p->fb_list.alloc(sizeof(unsigned) * 2 * p->nw)) || (st = p->kend.alloc(sizeof(int) * ntile)))
            return st;
    }
    HIPCHK(hipMemcpy(p->nk_pred.p, p->nk_pred_h.data(), sizeof(int) * p->nw, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p->nchk.p, p->nchk_h.data(), sizeof(int) * p->nw, hipMemcpyHostToDevice));
    if ((st = build_items(p))) return st;
    if (p->kind == 1) {
        // slices from the mean K extent of the frequency rows
        double ksum = 0, rows = 0;
        for (int iw = 0; iw < p->nw; iw++) ksum += std::min(p->nk, (p->nk_pred_h[iw] + 31) / 32 * 32);
        for (int q = 0; q < p->nsets; q++) rows += Mq[q];
        const int kc = grtm::gemm_pair_slice(p->gemm_mode, (int)(rows / 2), p->nr, p->nk, ksum / p->nw);
        size_t need = grtm::gemm_set_workspace_bytes(p->nsets, Mq, p->nr, p->nk, kc, esize(p->prec));
        if (p->gemm_mode == 1) need = std::max(need, grtm::gemm_streamk_workspace_bytes(p->nsets, Mq, p->nr, p->nk));
        if (need > p->ws.n && (st = p->ws.alloc(need))) return st;
        p->kc = kc;
    } else if (p->pair) {
        // contraction slices from the mean K extent of the 32-row tiles
        double ksum = 0;
        int nt = 0;
        for (int r0 = 0; r0 < p->nw; r0 += 32, nt++) {
            int m = 0;
            for (int r = r0; r < std::min(p->nw, r0 + 32); r++) m = std::max(m, p->nk_pred_h[r]);
            ksum += std::min(p->nk, (m + 31) / 32 * 32);
        }
        const int kc = grtm::gemm_pair_slice(p->gemm_mode, 4 * p->nw, p->nr, p->nk, ksum / std::max(1, nt));
        size_t need = grtm::gemm_pair_workspace_bytes(4 * p->nw, p->nr, p->nk, kc) * esize(p->prec) / 4;
        if (p->gemm_mode == 1) need = std::max(need, pair_sk_bytes(p->bsplit, 4 * p->nw, p->nr, p->nk));
        if (need > p->ws.n && (st = p->ws.alloc(need))) return st;
        p->kc = kc;
    }
    p->max_chk = mx;
    p->chk_stride = stride;
    p->chk_tol = tol;
    p->adaptive = true;
    return GRTM_OK;
}

// The device-built grid of the last execution into nk_pred_h / nchk_h.
static int pred_to_host(grtm_plan *p)
{
    if (!p->net || !p->pred_stale) return GRTM_OK;
    HIPCHK(hipStreamSynchronize(p->stream));
    if (p->last_stream && p->last_stream != p->stream) HIPCHK(hipStreamSynchronize(p->last_stream));
    HIPCHK(hipMemcpy(p->nk_pred_h.data(), p->nk_pred.p, sizeof(int) * p->nw, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(p->nchk_h.data(), p->nchk.p, sizeof(int) * p->nw, hipMemcpyDeviceToHost));
    p->pred_stale = false;
    return GRTM_OK;
}

int grtm_plan_set_learned(grtm_plan *p, grtm_mlp *net, double lmargin, double window, int stride, double tol)
{
    if (!p) return GRTM_E_INPUT;
    if (!net) return grtm_plan_set_adaptive(p, nullptr, 0.0, 1, 1.0);
    if (p->kind != 0 || !p->pair || p->gemm_mode == 2 || p->ptam_n) return GRTM_E_UNSUPPORTED;
    if (!(window >= 0) || stride < 1 || !(tol > 0) || !std::isfinite(lmargin)) return GRTM_E_INPUT;
    const int n = (int)p->layers_h.size() / 6;
    if (n > grtm::kLearnSlots) return GRTM_E_UNSUPPORTED;  // the network's 30 layer slots
    int st;
    // per-row buffers as grtm_plan_set_adaptive, and the device item list
    const int tpr = grtm::cells_tiles_per_row(p->nk);
    const int cap = p->nw * (2 * tpr + 1);
    if (!p->nk_pred.p) {
        if ((st = p->nk_pred.alloc(sizeof(int) * p->nw)) || (st = p->nk_w.alloc(sizeof(int) * p->nw)) ||
            (st = p->nchk.alloc(sizeof(int) * p->nw)) || (st = p->counters.alloc(sizeof(unsigned) * (2 * p->nw + 1))) ||
            (st = p->fb_list.alloc(sizeof(unsigned) * 2 * p->nw)) || (st = p->kend.alloc(sizeof(int) * ((p->nw + 15) / 16))))
            return st;
    }
    if (p->tiles.n < sizeof(unsigned) * (size_t)cap && (st = p->tiles.alloc(sizeof(unsigned) * (size_t)cap)))
        return st;
    if (!p->lmodel.p && ((st = p->lmodel.alloc(sizeof(grtm::LearnModel))) ||
                         (st = p->icounts.alloc(3 * sizeof(unsigned))) || (st = p->ictr.alloc(2 * sizeof(unsigned))) ||
                         (st = p->ioff.alloc(2 * sizeof(unsigned) * p->nw))))
        return st;
    if (p->gemm_mode == 1) {
        const size_t need = pair_sk_bytes(p->bsplit, 4 * p->nw, p->nr, p->nk);
        if (need > p->ws.n && (st = p->ws.alloc(need))) return st;
    } else {
        p->kc = grtm::gemm_pair_slice(p->gemm_mode, 4 * p->nw, p->nr, p->nk, p->nk);
    }
    grtm::LearnModel &m = p->lm;
    m = grtm::LearnModel{};
    m.n = n;
    double vs_min = 1e30;
    for (int j = 0; j < n; j++) {
        for (int c = 0; c < 6; c++) m.rows[j][c] = p->layers_h[(size_t)6 * j + c];
        vs_min = std::min(vs_min, m.rows[j][2]);
    }
    m.src = p->src_depth;
    m.rcv = p->rcv_depth;
    m.vs_min = vs_min;
    m.window = window;
    p->lg = grtm::LearnGrid{p->nw, p->nk, stride, 0, p->pp.dt, p->dk, p->dk * p->nk, lmargin};
    p->tiles_cap = cap;
    p->chk_stride = stride;
    p->chk_tol = tol;
    p->net = net;
    p->adaptive = true;
    if ((st = upload_learn_model(p))) return st;
    // one evaluation now, so the host's view of the grid (work, stats) is set
    p->nk_pred_h.assign(p->nw, 0);
    p->nchk_h.assign(p->nw, 0);
    if ((st = grtm::learned_grid(net, p->lmodel.as<grtm::LearnModel>(), 1, p->lg, p->nk_pred.as<int>(),
                                 p->nchk.as<int>(), p->stream))) {
        p->net = nullptr;
        p->adaptive = false;
        return st;
    }
    p->pred_stale = true;
    return pred_to_host(p);
}

int grtm_plan_adaptive_rows(grtm_plan *p, int *cut, int *checks)
{
    if (!p || !cut || !checks) return GRTM_E_INPUT;
    if (!p->adaptive) {
        for (int iw = 0; iw < p->nw; iw++) cut[iw] = p->nk, checks[iw] = 0;
        return GRTM_OK;
    }
    int st;
    if ((st = pred_to_host(p))) return st;
    memcpy(cut, p->nk_pred_h.data(), sizeof(int) * p->nw);
    memcpy(checks, p->nchk_h.data(), sizeof(int) * p->nw);
    return GRTM_OK;
}

int grtm_plan_row_labels(grtm_plan *p, double tau, float *out)
{
    if (!p || !out || !(tau > 0) || p->adaptive) return GRTM_E_INPUT;
    if (p->kind != 0) return GRTM_E_UNSUPPORTED;
    if ((size_t)p->nk * sizeof(float) > 160 * 1024) return GRTM_E_UNSUPPORTED;
    DBuf d;
    int st;
    if ((st = d.alloc(sizeof(float) * 26 * (size_t)p->nw))) return st;
    HIPCHK(grtm::launch_row_labels(p->prec == GRTM_PREC_FP64 ? 0 : 1, p->nw, p->nk, p->nk, p->A0.p, p->A1.p, tau,
                                   d.as<float>(), p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    HIPCHK(hipMemcpy(out, d.p, sizeof(float) * 26 * (size_t)p->nw, hipMemcpyDeviceToHost));
    return GRTM_OK;
}

int grtm_plan_adaptive_stats(grtm_plan *p, double *out)
{
    if (!p || !out) return GRTM_E_INPUT;
    out[0] = out[1] = out[2] = out[3] = 0;
    if (!p->adaptive) return GRTM_OK;
    int st;
    if ((st = pred_to_host(p))) return st;
    HIPCHK(hipStreamSynchronize(p->stream));
    unsigned fb = 0;
    HIPCHK(hipMemcpy(&fb, p->counters.as<unsigned>() + 2 * p->nw, sizeof(unsigned), hipMemcpyDeviceToHost));
    std::vector<int> nkw(p->nw);
    HIPCHK(hipMemcpy(nkw.data(), p->nk_w.p, sizeof(int) * p->nw, hipMemcpyDeviceToHost));
    double pred = 0, fin = 0, chk = 0;
    for (int iw = 0; iw < p->nw; iw++) {
        pred += p->nk_pred_h[iw];
        fin += nkw[iw];
        chk += p->nchk_h[iw];
    }
    out[0] = pred;  // cells solved by the predicted grid
    out[1] = chk;   // residual-check cells
    out[2] = fin;   // cells after dense fallback
    out[3] = fb;    // rows that fell back
    return GRTM_OK;
}

int grtm_plan_work(grtm_plan *p, double *w)
{
    if (!p || !w) return GRTM_E_INPUT;
    {
        const int st = pred_to_host(p);
        if (st) return st;
    }
    const double cells = (double)p->nw * p->nk;
    w[0] = cells;
    w[1] = cells * p->L;
    double oc[2] = {0, 0};
    if (p->kind == 0) {
        w[2] = 2.0 * 2.0 * (4.0 * p->nw) * p->nk * p->nr;  // two contractions, 2 flop per MAC
        w[3] = 4.0 * p->nr * (8.0 * p->nw + 4.0 * p->N);   // complex64 in + fp32 out per trace component
        w[4] = 2.0 * p->nr;                                 // (model, source, distance) traces, Z+R each
        if (grtm::opcount_cell_bot(0, p->L, p->g.s, p->g.rl, p->g.branch, p->g.up, -1, oc, p->g.zr == 0.0))
            return GRTM_E_INPUT;
    } else {
        double rows = 0, terms = 0;
        for (int q = 0; q < p->nsets; q++) rows += p->mt_M(q);
        for (int c = 0; c < p->ncols; c++) terms += p->mtc.nterm[c];
        w[2] = 2.0 * rows * p->nk * p->nr;                         // the contractions of the set
        w[3] = p->nr * (terms * 8.0 * p->nw + p->ncols * 4.0 * p->N);  // complex64 terms in, fp32 out
        w[4] = (double)p->ncols * p->nr;                           // column traces
        if (grtm::opcount_cell_bot(p->deriv ? 2 : 1, p->L, p->g.s, p->g.rl, p->g.branch, p->g.up, -1, oc,
                                   p->g.zr == 0.0))
            return GRTM_E_INPUT;
    }
    w[5] = oc[0];  // flops per cell
    w[6] = oc[1];  // transcendental calls per cell
    w[7] = p->prec == GRTM_PREC_FP64 ? p->nw : p->rows64;  // rows solved in FP64
    // the main launch's residual-check cells (learned grid)
    const bool chk_main = p->adaptive;
    auto nchk_row = [&](int iw) {
        int n = 0;
        for (int j = 0; j < p->nchk_h[iw]; j++) n += p->nk_pred_h[iw] + p->chk_stride * j + p->chk_stride / 2 < p->nk;
        return n;
    };
    {
        // cells of the main recursion launch: rows [r0, nw) of the predicted or
        // dense grid (and their check cells)
        const int r0 = p->prec == GRTM_PREC_FP64 ? 0 : p->rows64;
        double mc = 0;
        for (int iw = r0; iw < p->nw; iw++) mc += (p->adaptive ? p->nk_pred_h[iw] : p->nk) + (chk_main ? nchk_row(iw) : 0);
        w[9] = mc;
    }
    {
        // mean flops per cell of the main launch with the depth truncation:
        // the upward sweep's bottom layer of every cell (sweep_bottom on the
        // FP64 tables), each bottom counted on the counting type
        const int r0 = p->prec == GRTM_PREC_FP64 ? 0 : p->rows64;
        const bool cut = p->g.branch != 4;
        std::vector<double> nbot(p->L, 0.0);
        if (cut) {
            std::vector<LayerW<double>> lw((size_t)p->nw * p->L);
            const DBuf &t64 = p->prec == GRTM_PREC_FP64 ? p->lw : p->lw64;
            HIPCHK(hipMemcpy(lw.data(), t64.p, lw.size() * sizeof(LayerW<double>), hipMemcpyDeviceToHost));
            const double z0 = p->g.branch == 3 ? p->g.zr : p->g.zs;
            for (int iw = r0; iw < p->nw; iw++) {
                const LayerW<double> *lr = lw.data() + (size_t)iw * p->L;
                const int nk = p->adaptive ? p->nk_pred_h[iw] : p->nk;
                for (int ik = 0; ik < nk; ik++) nbot[grtm::sweep_bottom(lr, p->L, p->g.s, z0, p->dk * ik)] += 1;
                if (chk_main)
                    for (int j = 0; j < p->nchk_h[iw]; j++) {
                        const int ik = nk + p->chk_stride * j + p->chk_stride / 2;
                        if (ik < p->nk) nbot[grtm::sweep_bottom(lr, p->L, p->g.s, z0, p->dk * ik)] += 1;
                    }
            }
        } else {
            nbot[p->L - 1] = w[9];
        }
        double fl = 0;
        for (int b = 0; b < p->L; b++) {
            if (nbot[b] == 0) continue;
            double ob[2];
            if (grtm::opcount_cell_bot(p->kind == 0 ? 0 : (p->deriv ? 2 : 1), p->L, p->g.s, p->g.rl, p->g.branch,
                                       p->g.up, b, ob, p->g.zr == 0.0))
                return GRTM_E_INPUT;
            fl += nbot[b] * ob[0];
        }
        w[10] = w[9] > 0 ? fl / w[9] : oc[0];
    }
    w[8] = 0;  // cells re-solved in FP64 after a non-finite FP32 result (last execution)
    if (p->prec != GRTM_PREC_FP64) {
        unsigned c = 0;
        HIPCHK(hipStreamSynchronize(p->stream));
        HIPCHK(hipMemcpy(&c, p->flags.p, sizeof(unsigned), hipMemcpyDeviceToHost));
        w[8] = c;
    }
    return GRTM_OK;
}

void grtm_plan_destroy(grtm_plan *p) { delete p; }

// ------------------------------------------------------------ self-tests
int grtm_selftest_gemm_cut(int prec, int M, int N, int K, const double *A, const double *B, const int *nk_w, double *C)
{
    if ((prec != 0 && prec != 1 && prec != grtm::kSplitBf16x3 && prec != grtm::kSplitBf16x6 &&
         prec != grtm::kSplitF16x3) ||
        M < 4 || M % 4 || N < 1 || K < 1 || !A || !B || !nk_w || !C)
        return GRTM_E_INPUT;
    const int nw = M / 4;
    double ksum = 0;
    for (int i = 0; i < nw; i++) {
        if (nk_w[i] < 0 || nk_w[i] > K) return GRTM_E_INPUT;
        ksum += nk_w[i];
    }
    int st;
    if ((st = device_ready())) return st;
    const long lda = (K + 3) / 4 * 4, ldb = (N + 3) / 4 * 4;
    const size_t es = prec == 0 ? sizeof(double) : sizeof(float);
    std::vector<double> a((size_t)M * lda, 0.0), b((size_t)K * ldb, 0.0);
    for (int i = 0; i < M; i++)
        for (int k = 0; k < K; k++) a[(size_t)i * lda + k] = A[(size_t)i * K + k];
    for (int k = 0; k < K; k++)
        for (int j = 0; j < N; j++) b[(size_t)k * ldb + j] = B[(size_t)k * N + j];
    const int bsplit = prec > 1 ? prec : 0;
    const int kc = grtm::gemm_pair_slice(prec ? 1 : 0, M, N, K, std::max(1.0, ksum / nw));
    size_t wsb = grtm::gemm_pair_workspace_bytes(M, N, K, kc) * es / 4;
    if (prec) wsb = std::max(wsb, pair_sk_bytes(bsplit, M, N, K));
    DBuf dA, dB, dC, dP, dn, dk, dS, dpk;
    long ldk = 0;
    if ((st = dA.alloc(a.size() * es)) || (st = dB.alloc(b.size() * es)) || (st = dC.alloc((size_t)2 * M * N * es)) ||
        (st = dP.alloc(wsb)) || (st = dn.alloc(sizeof(int) * nw)) || (st = dk.alloc(sizeof(int) * ((nw + 31) / 32))))
        return st;
    if (bsplit) {
        ldk = (K + 31) / 32 * 32;
        if ((st = dS.alloc((size_t)grtm::split_planes(bsplit) * N * ldk * 2))) return st;
    }
    if (bsplit == grtm::kSplitF16x3) {
        // the FP16 split's row scales: the peak |A| of each frequency row's
        // four operand rows below its cut, as float bits (as the recursion keeps them)
        std::vector<unsigned> pk(nw, 0u);
        for (int i = 0; i < nw; i++) {
            float m = 0.f;
            for (int q = 0; q < 4; q++)
                for (int k = 0; k < nk_w[i]; k++) m = std::max(m, std::fabs(float(A[(size_t)(4 * i + q) * K + k])));
            std::memcpy(&pk[i], &m, sizeof(float));
        }
        if ((st = dpk.alloc(sizeof(unsigned) * nw))) return st;
        HIPCHK(hipMemcpy(dpk.p, pk.data(), sizeof(unsigned) * nw, hipMemcpyHostToDevice));
    }
    if (prec == 0) {
        HIPCHK(hipMemcpy(dA.p, a.data(), a.size() * es, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dB.p, b.data(), b.size() * es, hipMemcpyHostToDevice));
    } else {
        const std::vector<float> af(a.begin(), a.end()), bf(b.begin(), b.end());
        HIPCHK(hipMemcpy(dA.p, af.data(), af.size() * es, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dB.p, bf.data(), bf.size() * es, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(dn.p, nk_w, sizeof(int) * nw, hipMemcpyHostToDevice));
    // tile extents as the plans form them: 32 frequency rows (128 operand rows) per tile
    HIPCHK(grtm::launch_kend_tiles(nw, 32, K, dn.as<int>(), dk.as<int>(), nullptr));
    if (prec == 0) {
        double *c0 = dC.as<double>(), *c1 = c0 + (size_t)M * N;
        HIPCHK(grtm::launch_gemm_pair_f64(M, N, K, kc, dA.as<double>(), dA.as<double>(), lda, dB.as<double>(),
                                          dB.as<double>(), ldb, c0, c1, dP.as<double>(), dk.as<int>(), dn.as<int>(),
                                          nullptr));
    } else if (bsplit) {
        // the split planes of B, one table for both halves
        float *c0 = dC.as<float>(), *c1 = c0 + (size_t)M * N;
        HIPCHK(grtm::launch_split_table(bsplit, K, N, dB.as<float>(), ldb, dS.p, ldk, nullptr));
        HIPCHK(grtm::launch_gemm_pair_bsplit(bsplit, M, N, K, dA.as<float>(), dA.as<float>(), lda, dS.p, dS.p, ldk, c0,
                                             c1, dP.as<float>(), dk.as<int>(), dn.as<int>(), nullptr, nullptr,
                                             nullptr, dpk.as<unsigned>()));
    } else {
        float *c0 = dC.as<float>(), *c1 = c0 + (size_t)M * N;
        HIPCHK(grtm::launch_gemm_pair_f32(M, N, K, kc, dA.as<float>(), dA.as<float>(), lda, dB.as<float>(),
                                          dB.as<float>(), ldb, c0, c1, dP.as<float>(), dk.as<int>(), dn.as<int>(),
                                          nullptr));
    }
    HIPCHK(hipDeviceSynchronize());
    std::vector<double> f((size_t)2 * M * N);
    if (prec == 0) {
        HIPCHK(hipMemcpy(f.data(), dC.p, f.size() * es, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> g32(f.size());
        HIPCHK(hipMemcpy(g32.data(), dC.p, g32.size() * es, hipMemcpyDeviceToHost));
        std::copy(g32.begin(), g32.end(), f.begin());
    }
    for (size_t i = 0; i < (size_t)M * N; i++) {
        if (f[i] != f[(size_t)M * N + i]) return GRTM_E_DEVICE;  // the two halves must agree bit for bit
        C[i] = f[i];
    }
    return GRTM_OK;
}

int grtm_selftest_gemm(int mode, int M, int N, int K, const double *A, const double *B, double *C)
{
    if (mode < -1 || mode > 4 || M < 1 || N < 1 || K < 1 || !A || !B || !C) return GRTM_E_INPUT;
    int st;
    if ((st = device_ready())) return st;
    if (mode == 3) {
        // streamed FP32 pair kernel on (A, B) twice; padded leading dimensions
        const long lda = (K + 3) / 4 * 4, ldb = (N + 3) / 4 * 4;
        std::vector<float> a((size_t)M * lda, 0.f), b((size_t)K * ldb, 0.f);
        for (int i = 0; i < M; i++)
            for (int k = 0; k < K; k++) a[(size_t)i * lda + k] = (float)A[(size_t)i * K + k];
        for (int k = 0; k < K; k++)
            for (int j = 0; j < N; j++) b[(size_t)k * ldb + j] = (float)B[(size_t)k * N + j];
        const int kc = grtm::gemm_pair_slice(1, M, N, K, K);
        DBuf dA, dB, dC, dP;
        if ((st = dA.alloc(a.size() * 4)) || (st = dB.alloc(b.size() * 4)) || (st = dC.alloc((size_t)2 * M * N * 4)) ||
            (st = dP.alloc(grtm::gemm_pair_workspace_bytes(M, N, K, kc))))
            return st;
        HIPCHK(hipMemcpy(dA.p, a.data(), a.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dB.p, b.data(), b.size() * 4, hipMemcpyHostToDevice));
        float *c0 = dC.as<float>(), *c1 = c0 + (size_t)M * N;
        HIPCHK(grtm::launch_gemm_pair_f32(M, N, K, kc, dA.as<float>(), dA.as<float>(), lda, dB.as<float>(),
                                          dB.as<float>(), ldb, c0, c1, dP.as<float>(), nullptr, nullptr, nullptr));
        HIPCHK(hipDeviceSynchronize());
        std::vector<float> f((size_t)2 * M * N);
        HIPCHK(hipMemcpy(f.data(), dC.p, f.size() * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < (size_t)M * N; i++) {
            if (f[i] != f[(size_t)M * N + i]) return GRTM_E_DEVICE;  // the two halves must agree bit for bit
            C[i] = f[i];
        }
        return GRTM_OK;
    }
    if (mode == 4) {
        // streamed FP64 pair kernel on (A, B) twice; padded leading dimensions
        const long lda = (K + 3) / 4 * 4, ldb = (N + 3) / 4 * 4;
        std::vector<double> a((size_t)M * lda, 0.0), b((size_t)K * ldb, 0.0);
        for (int i = 0; i < M; i++)
            for (int k = 0; k < K; k++) a[(size_t)i * lda + k] = A[(size_t)i * K + k];
        for (int k = 0; k < K; k++)
            for (int j = 0; j < N; j++) b[(size_t)k * ldb + j] = B[(size_t)k * N + j];
        const int kc = grtm::gemm_pair_slice(0, M, N, K, K);
        DBuf dA, dB, dC, dP;
        if ((st = dA.alloc(a.size() * 8)) || (st = dB.alloc(b.size() * 8)) || (st = dC.alloc((size_t)2 * M * N * 8)) ||
            (st = dP.alloc(2 * grtm::gemm_pair_workspace_bytes(M, N, K, kc))))
            return st;
        HIPCHK(hipMemcpy(dA.p, a.data(), a.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dB.p, b.data(), b.size() * 8, hipMemcpyHostToDevice));
        double *c0 = dC.as<double>(), *c1 = c0 + (size_t)M * N;
        HIPCHK(grtm::launch_gemm_pair_f64(M, N, K, kc, dA.as<double>(), dA.as<double>(), lda, dB.as<double>(),
                                          dB.as<double>(), ldb, c0, c1, dP.as<double>(), nullptr, nullptr, nullptr));
        HIPCHK(hipDeviceSynchronize());
        std::vector<double> f((size_t)2 * M * N);
        HIPCHK(hipMemcpy(f.data(), dC.p, f.size() * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < (size_t)M * N; i++) {
            if (f[i] != f[(size_t)M * N + i]) return GRTM_E_DEVICE;
            C[i] = f[i];
        }
        return GRTM_OK;
    }
    const size_t es = (mode == -1 || mode == 0) ? 8 : 4;
    DBuf dA, dB, dC, dP;
    if ((st = dA.alloc((size_t)M * K * es)) || (st = dB.alloc((size_t)K * N * es)) || (st = dC.alloc((size_t)M * N * es)))
        return st;
    if (mode >= 0 && (st = dP.alloc(grtm::gemm_workspace_bytes(mode, M, N, K)))) return st;
    auto up = [&](DBuf &d, const double *h, size_t n) -> int {
        if (es == 8) {
            HIPCHK(hipMemcpy(d.p, h, n * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(h, h + n);
            HIPCHK(hipMemcpy(d.p, f.data(), n * 4, hipMemcpyHostToDevice));
        }
        return GRTM_OK;
    };
    if ((st = up(dA, A, (size_t)M * K)) || (st = up(dB, B, (size_t)K * N))) return st;
    if (mode == -1) HIPCHK(grtm::launch_gemm(0, M, N, K, dA.p, K, dB.p, N, dC.p, N, nullptr));
    else HIPCHK(grtm::launch_gemm_mfma(mode, M, N, K, dA.p, K, dB.p, N, dC.p, dP.p, nullptr, nullptr, nullptr));
    HIPCHK(hipDeviceSynchronize());
    if (es == 8) {
        HIPCHK(hipMemcpy(C, dC.p, (size_t)M * N * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> f((size_t)M * N);
        HIPCHK(hipMemcpy(f.data(), dC.p, f.size() * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < f.size(); i++) C[i] = f[i];
    }
    return GRTM_OK;
}

int grtm_selftest_mfma_probe(int l0, double *out)
{
    if (l0 < 0 || l0 > 63 || !out) return GRTM_E_INPUT;
    int st;
    if ((st = device_ready())) return st;
    DBuf d;
    if ((st = d.alloc(64 * 4 * sizeof(double)))) return st;
    HIPCHK(grtm::launch_probe_f64(l0, d.as<double>(), nullptr));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(out, d.p, 64 * 4 * sizeof(double), hipMemcpyDeviceToHost));
    return GRTM_OK;
}

int grtm_selftest_copy(int n_int32)
{
    if (n_int32 < 1) return GRTM_E_INPUT;
    int st;
    if ((st = device_ready())) return st;
    DBuf a, b;
    if ((st = a.alloc(sizeof(int) * (size_t)n_int32)) || (st = b.alloc(sizeof(int) * (size_t)n_int32))) return st;
    HIPCHK(grtm::launch_zero_u32(a.as<unsigned>(), n_int32, nullptr));
    HIPCHK(grtm::launch_copy_i32(a.as<int>(), b.as<int>(), n_int32, nullptr));
    HIPCHK(hipDeviceSynchronize());
    return GRTM_OK;
}

}  // extern "C"

namespace {
// The device cell solve compiled for the host in arithmetic T (double or
// double-double; the layer constants are formed in FP64 either way).
template <typename T>
int selftest_cell_t(int compat, int src_type, double k, double wr, double wi, int n_layer, const double *lay,
                    int s_layer, double zs, int r_layer, double zr, const double *par0_rv, double *out)
{
    if (n_layer < 2 || n_layer > grtm::kMaxLayers || !lay || !out) return GRTM_E_INPUT;
    const cd w(wr, wi);
    std::vector<LayerW<double>> lw(n_layer);
    for (int j = 0; j < n_layer; j++) {
        const double *l = lay + 7 * j;
        const cd a(l[1], l[2]), b(l[3], l[4]), mu(l[5], l[6]);
        auto C = [](cd z) { return grtm::cx<double>{z.real(), z.imag()}; };
        lw[j].sa2 = C((w / a) * (w / a));
        lw[j].sb2 = C((w / b) * (w / b));
        lw[j].sa = C(w / a);
        lw[j].sb = C(w / b);
        lw[j].mg = C(mu * (w / b) * (w / b));
        lw[j].ca = C(a / w);
        lw[j].cb = C(b / w);
        lw[j].mu = C(mu);
        lw[j].a = C(a);
        lw[j].b = C(b);
        lw[j].h = l[0];
        lw[j].hrev = -l[0] / (2 * M_PI);
        lw[j].hlog = -l[0] * M_LOG2E;
    }
    std::vector<LayerW<T>> lt(n_layer);
    for (int j = 0; j < n_layer; j++) {
        if constexpr (std::is_same<T, double>::value) lt[j] = lw[j];
        else lt[j] = grtm::widen_dd(lw[j]);
    }
    Stack st;
    st.L = n_layer;
    st.s = s_layer;
    st.rl = r_layer;
    st.zs = zs;
    st.zr = zr;
    Geo g;
    int status = make_geo(st, compat, g);
    if (status) return status;
    grtm::cx<T> p0r{T(0), T(0)}, p0v{T(0), T(0)};
    if (par0_rv) p0r = {T(par0_rv[0]), T(par0_rv[1])}, p0v = {T(par0_rv[2]), T(par0_rv[3])};
    grtm::CellOut<T> c;
    grtm::Coef<T> p;
    const T kt = T(k);
    if (compat) grtm::cell_solve<T, true, true>(lt.data(), g, kt, p0r, p0v, c);
    else grtm::cell_solve<T, false, true>(lt.data(), g, kt, p0r, p0v, c);
    if (compat) grtm::coef_from_cell<T, true, true>(c, src_type, g.up, kt, grtm::cx<T>{T(wr), T(wi)}, p);
    else grtm::coef_from_cell<T, true, false>(c, src_type, g.up, kt, grtm::cx<T>{T(wr), T(wi)}, p);
    const grtm::cx<T> z{T(0), T(0)};
    const grtm::cx<T> v[18] = {p.a0[0], p.a0[1], p.a1[0], p.a1[1], p.a2[0], p.a2[1], p.b1, p.b2, p.a0t[0],
                               p.a0t[1], p.a1t[0], p.a1t[1], p.a2t[0], p.a2t[1], p.b1t, p.b2t, z, z};
    for (int i = 0; i < 18; i++) out[2 * i] = double(v[i].x), out[2 * i + 1] = double(v[i].y);
    return GRTM_OK;
}
}  // namespace

extern "C" {
int grtm_selftest_cell(int compat, int src_type, double k, double wr, double wi, int n_layer, const double *lay,
                       int s_layer, double zs, int r_layer, double zr, const double *par0_rv, double *out)
{
    return selftest_cell_t<double>(compat, src_type, k, wr, wi, n_layer, lay, s_layer, zs, r_layer, zr, par0_rv,
                                   out);
}

int grtm_selftest_cell_dd(int compat, int src_type, double k, double wr, double wi, int n_layer, const double *lay,
                          int s_layer, double zs, int r_layer, double zr, const double *par0_rv, double *out)
{
    return selftest_cell_t<grtm::dd>(compat, src_type, k, wr, wi, n_layer, lay, s_layer, zs, r_layer, zr, par0_rv,
                                     out);
}

}  // extern "C"

// ============================================================ batches
struct grtm_batch {
    grtm_ctx *ctx = nullptr;
    int prec = 0, compat = 0, gemm_mode = 0;
    int nm = 0, nw = 0, N = 0, nk = 0, nr = 0;
    double dk = 0, df = 0, omega_i = 0;
    grtm::PostParams pp{};
    hipStream_t stream = nullptr;
    std::vector<int> L, rows64;  // per model
    double flops = 0;            // recursion flops of one execution (operation count)
    std::vector<grtm::ModelDesc> md_h;
    DBuf md, lw64, lw32, dists, B0, B1, A0, A1, C0, C1, ws, flags, tr, items;
    int bsplit = 0;  // split-BF16 pair contraction (3 or 6 products) on the planes S0 / S1
    long ldk = 0;
    DBuf S0, S1;
    unsigned flag_cap = 0;  // flag list entries (flags: counter, then 64-bit cells)
    // test hook: cells (group row * nk + k) given non-finite FP32 operands, per
    // group, from the (model, row, k) triples in plant_h
    DBuf plant;
    std::vector<size_t> plant_off;  // per group: first entry in plant, then the end
    std::vector<int> plant_h;
    // resident layout (learned FP32 batches): one group of every model, each
    // frequency row's operands stored to its final extent (foff, fext), which
    // an eager pass fixes; rows that would exceed it are cut and counted
    bool resident = false;
    int nb_grouped = 0;  // models per group of the grouped layout
    DBuf foff, fext, ext_over;
    size_t resident_floats = 0;  // entries of A0 (and of A1)
    DBuf ptab;  // post-processing tables of pp
    int nb = 0;       // models per group
    int bs = 256;     // recursion tile: wavenumbers per block
    int kc = 0;       // streamed FP32 contraction: K slice
    bool pair = false;
    int tuned = 0;    // autotune candidates timed
    double tune_ms = 0;
    struct Group {
        int m0, n;
        // item ranges of the FP64 rows (pass 0) and FP32 rows (pass 1), one per
        // receiver branch 1-4 (the recursion is compiled per branch)
        size_t off[2][4], cnt[2][4];
    };
    std::vector<Group> groups;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t nodes = 0;
    // adaptive wavenumber grids, one per model ([model][row] arrays)
    bool adaptive = false;
    int chk_stride = 8;
    double chk_tol = 3e-6;
    std::vector<int> nk_pred_h, nchk_h, max_chk;
    DBuf nk_pred, nk_w, nchk, row_peak, tail_max, fb_count, fb_list, kend;
    DBuf fb_gcount;  // fallback rows listed per group (one counter per group)
    // learned grids formed on the device inside the graph (grtm_batch_set_learned):
    // the network, every model's predictor inputs, and per group its item
    // slots (group row | kCheckItem), the slots' first items and the
    // segments' item bounds (segment = pass * 4 + branch - 1)
    grtm_mlp *net = nullptr;
    grtm::LearnGrid lg{};
    DBuf lmodels, lslots, lslot_off, lseg_slots, lseg_items, lctr, litems;
    hipEvent_t ev_pred0 = nullptr, ev_pred1 = nullptr;  // around a group's grid formation (timing)
    std::vector<size_t> lslot0;     // per group: first slot in lslots
    std::vector<int> lnslots;       // per group: slots
    std::vector<unsigned> lseg_h;   // per group: 9 segment slot bounds
    int litems_cap = 0;             // item capacity of one group
    int lchunk = 1;                 // items a persistent block takes per visit of its counter (autotuned)
    bool pred_stale = false;
    // the models, kept for the predictor's inputs
    std::vector<std::vector<double>> layers_h;
    std::vector<double> src_h, rcv_h;
    std::vector<int> rows64_rule;
    ~grtm_batch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

bool batch_mixed(const grtm_batch *b) { return b->prec != GRTM_PREC_FP64; }

// Flag list capacity of a batch: every cell of a group, at most 16M entries.
constexpr size_t kBatchFlagCap = size_t(1) << 24;
int batch_alloc_flags(grtm_batch *b, size_t extra)
{
    if (!batch_mixed(b)) return GRTM_OK;
    const size_t cap = std::min<size_t>((size_t)b->nb * b->nw * b->nk, kBatchFlagCap) + extra;
    b->flag_cap = (unsigned)cap;
    return b->flags.alloc(8 + sizeof(unsigned long long) * cap);
}
grtm::FlagList batch_fl(const grtm_batch *b)
{
    if (!batch_mixed(b)) return grtm::FlagList{nullptr, nullptr, 0u};
    return grtm::FlagList{b->flags.as<unsigned>(),
                          reinterpret_cast<unsigned long long *>(static_cast<char *>(b->flags.p) + 8), b->flag_cap};
}
grtm::RowLayout batch_rl(const grtm_batch *b)
{
    if (b->resident) return grtm::RowLayout{b->foff.as<long long>(), b->fext.as<int>(), 0};
    return grtm::RowLayout{nullptr, nullptr, (long)b->nk};
}

// Frequency rows per K-extent tile of the contraction, and the offset of a
// group's tile extents (each group's tiles start at its own first row).
int batch_rpt(const grtm_batch *b) { return (b->gemm_mode == 0 && !b->pair) ? 16 : 32; }
size_t batch_kend_per_group(const grtm_batch *b) { return ((size_t)b->nb * b->nw + 15) / 16; }
size_t batch_kend_off(const grtm_batch *b, const grtm_batch::Group &g)
{
    return (size_t)(g.m0 / b->nb) * batch_kend_per_group(b);
}

// Work items of every group for recursion tiles of bs wavenumbers; within a
// group the models are taken in decreasing layer count.  Only groups
// [g0, g1) are built (the autotuner times group 0 alone).
int batch_items(grtm_batch *b, int bs, int g0, int g1)
{
    const int tpr = (b->nk + bs - 1) / bs;
    if (tpr > 4096) return GRTM_E_UNSUPPORTED;
    // tiles of a row: all, or those below the row's predicted cut
    auto ntile = [&](int m, int iw) { return b->adaptive ? (b->nk_pred_h[(size_t)m * b->nw + iw] + bs - 1) / bs : tpr; };
    std::vector<unsigned long long> it;
    for (int gi = g0; gi < g1; gi++) {
        grtm_batch::Group &g = b->groups[gi];
        std::vector<int> order(g.n);
        for (int j = 0; j < g.n; j++) order[j] = j;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return b->L[g.m0 + x] > b->L[g.m0 + y]; });
        for (int pass = 0; pass < 2; pass++)
            for (int br = 1; br <= 4; br++) {
                const size_t start = it.size();
                for (int j : order) {
                    if (b->md_h[g.m0 + j].g.branch != br) continue;
                    const int r64 = batch_mixed(b) ? b->rows64[g.m0 + j] : b->nw;
                    const int lo = pass == 0 ? 0 : r64, hi = pass == 0 ? r64 : b->nw;
                    for (int iw = lo; iw < hi; iw++)
                        for (int kb = 0, nt = ntile(g.m0 + j, iw); kb < nt; kb++)
                            it.push_back(((unsigned long long)j << 32) | grtm::cells_tile(iw, kb));
                }
                // learned grids: the residual-check items of these rows follow
                // their tiles, so the check cells run in the launch's tail
                if (b->adaptive)
                    for (int j : order) {
                        if (b->md_h[g.m0 + j].g.branch != br) continue;
                        const int r64 = batch_mixed(b) ? b->rows64[g.m0 + j] : b->nw;
                        const int lo = pass == 0 ? 0 : r64, hi = pass == 0 ? r64 : b->nw;
                        for (int iw = lo; iw < hi; iw++) {
                            const int nc = b->nchk_h[(size_t)(g.m0 + j) * b->nw + iw];
                            for (int c = 0; c * bs < nc; c++)
                                it.push_back(((unsigned long long)j << 32) | grtm::cells_tile(iw, c) | grtm::kCheckItem);
                        }
                    }
                g.off[pass][br - 1] = start;
                g.cnt[pass][br - 1] = it.size() - start;
            }
    }
    int st;
    if ((st = b->items.alloc(sizeof(unsigned long long) * std::max<size_t>(1, it.size())))) return st;
    if (!it.empty()) HIPCHK(hipMemcpy(b->items.p, it.data(), sizeof(unsigned long long) * it.size(), hipMemcpyHostToDevice));
    b->bs = bs;
    return GRTM_OK;
}

int batch_cells(grtm_batch *b, const grtm_batch::Group &g, hipStream_t s)
{
    const grtm::ModelDesc *md = b->md.as<grtm::ModelDesc>();
    const size_t r0 = (size_t)g.m0 * b->nw, nrow = (size_t)g.n * b->nw;
    const int gi = int(&g - b->groups.data());
    const bool learned = b->adaptive && b->net;
    // a learned grid: the group's network rows, cuts and items, formed here
    const unsigned long long *it = learned ? b->litems.as<unsigned long long>() : b->items.as<unsigned long long>();
    const unsigned *seg = learned ? b->lseg_items.as<unsigned>() + 9 * gi : nullptr;
    if (learned) {
        int st;
        if (b->ev_pred0) HIPCHK(hipEventRecord(b->ev_pred0, s));
        if ((st = grtm::learned_grid(b->net, b->lmodels.as<grtm::LearnModel>() + g.m0, g.n, b->lg,
                                     b->nk_pred.as<int>() + r0, b->nchk.as<int>() + r0, s)))
            return st;
        HIPCHK(grtm::launch_batch_items(b->lnslots[gi], b->lslots.as<unsigned>() + b->lslot0[gi], b->nw, b->bs,
                                        b->nk_pred.as<int>() + r0, b->nchk.as<int>() + r0, b->lslot_off.as<unsigned>(),
                                        8, b->lseg_slots.as<unsigned>() + 9 * gi, b->lseg_items.as<unsigned>() + 9 * gi,
                                        b->litems.as<unsigned long long>(), s));
        if (b->ev_pred1) HIPCHK(hipEventRecord(b->ev_pred1, s));
        b->pred_stale = true;
    }
    // item range of (pass, branch): a host list's range, or a device-built segment
    auto items_of = [&](int pass, int br, const unsigned long long *&p, int &n, const unsigned *&dc, int &which,
                        unsigned *&ctr) {
        const int sg = pass * 4 + br - 1;
        if (learned) {
            const unsigned *bounds = b->lseg_h.data() + 9 * gi;
            p = it, n = bounds[sg + 1] > bounds[sg] ? b->litems_cap : 0, dc = seg, which = sg;
            ctr = b->lctr.as<unsigned>() + sg;
        } else {
            p = it + g.off[pass][br - 1], n = (int)g.cnt[pass][br - 1], dc = nullptr, which = 0, ctr = nullptr;
        }
    };
    int *nkw = b->adaptive ? b->nk_w.as<int>() + r0 : nullptr;
    // row peaks: the adaptive grid's check, and the FP16 split's row scales
    const bool f16 = b->bsplit == grtm::kSplitF16x3;
    unsigned *rowpk = (b->adaptive || f16) ? b->row_peak.as<unsigned>() + r0 : nullptr;
    if (f16 && !b->adaptive) HIPCHK(grtm::launch_zero_u32(rowpk, (int)nrow, s));
    if (b->adaptive) {
        HIPCHK(grtm::launch_copy_i32(b->nk_pred.as<int>() + r0, nkw, (int)nrow, s));
        HIPCHK(grtm::launch_zero_u32(rowpk, (int)nrow, s));
        HIPCHK(grtm::launch_zero_u32(b->tail_max.as<unsigned>() + r0, (int)nrow, s));
        HIPCHK(grtm::launch_zero_u32(b->fb_count.as<unsigned>() + g.m0, g.n, s));
    }
    // learned grids: the item lists carry the residual-check cells (FP64 on the FP64 rows)
    const int *nchk = b->adaptive ? b->nchk.as<int>() + r0 : nullptr;
    unsigned *tailmx = b->adaptive ? b->tail_max.as<unsigned>() + r0 : nullptr;
    const unsigned long long *ip;
    const unsigned *dc;
    unsigned *ctr;
    int ni, which;
    const grtm::RowLayout rl = batch_rl(b);
    const grtm::FlagList fl = batch_fl(b);
    // resident layout: one launch per pass for every receiver branch (the
    // pass's slots are ordered by branch, then by decreasing layer count)
    auto launch_pass = [&](int pass, int prec, const void *lw, grtm::FlagList f) -> int {
        if (b->resident) {
            const unsigned *bounds = b->lseg_h.data() + 9 * gi;
            if (bounds[pass + 1] > bounds[pass])
                HIPCHK(grtm::launch_cells_zr_batch(prec, b->compat, b->bs, 0, md, g.m0, lw, b->nw, b->nk, b->dk,
                                                   b->df, b->omega_i, it, b->litems_cap, b->A0.p, b->A1.p, rl, f, nkw,
                                                   rowpk, s, nchk, b->chk_stride, tailmx, seg, pass,
                                                   b->lctr.as<unsigned>() + pass, b->lchunk));
            return GRTM_OK;
        }
        for (int br = 1; br <= 4; br++) {
            items_of(pass, br, ip, ni, dc, which, ctr);
            if (ni)
                HIPCHK(grtm::launch_cells_zr_batch(prec, b->compat, b->bs, br, md, g.m0, lw, b->nw, b->nk, b->dk,
                                                   b->df, b->omega_i, ip, ni, b->A0.p, b->A1.p, rl, f, nkw, rowpk, s,
                                                   nchk, b->chk_stride, tailmx, dc, which, ctr, b->lchunk));
        }
        return GRTM_OK;
    };
    const grtm::FlagList none{nullptr, nullptr, 0u};
    int pst;
    if (!batch_mixed(b)) {
        if ((pst = launch_pass(0, 0, b->lw64.p, none))) return pst;
    } else {
        HIPCHK(grtm::launch_zero_u32(fl.count, 1, s));
        if ((pst = launch_pass(0, 2, b->lw64.p, none)) || (pst = launch_pass(1, 1, b->lw32.p, fl))) return pst;
        if (!b->plant_off.empty() && b->plant_off[gi + 1] > b->plant_off[gi])
            HIPCHK(grtm::launch_plant_nonfinite_batch(b->plant.as<unsigned long long>() + b->plant_off[gi],
                                                      int(b->plant_off[gi + 1] - b->plant_off[gi]), b->nk, b->A0.p,
                                                      b->A1.p, rl, fl, s));
    }
    if (!b->adaptive) {
        if (fl.count)
            HIPCHK(grtm::launch_cells_fix_batch(b->compat, md, g.m0, b->lw64.p, b->nw, b->nk, b->dk, b->df,
                                                b->omega_i, b->A0.p, b->A1.p, rl, fl, s, f16 ? rowpk : nullptr));
        return GRTM_OK;
    }
    // dense fallback of the group's rows whose check failed (each row in its
    // own precision), then the group's tile extents
    unsigned *gcount = b->fb_gcount.as<unsigned>() + gi;
    HIPCHK(grtm::launch_zero_u32(gcount, 1, s));
    HIPCHK(grtm::launch_fallback_batch(batch_mixed(b) ? 1 : 0, b->compat, md, g.m0, b->lw64.p, b->lw32.p, (int)nrow,
                                       b->nw, b->nk, b->dk, b->df, b->omega_i, b->chk_tol, rowpk, tailmx, nkw, gcount,
                                       b->fb_list.as<unsigned>() + 2 * r0, b->fb_count.as<unsigned>() + g.m0, b->A0.p,
                                       b->A1.p, rl, fl, s, f16));
    if (fl.count)
        HIPCHK(grtm::launch_cells_fix_batch(b->compat, md, g.m0, b->lw64.p, b->nw, b->nk, b->dk, b->df, b->omega_i,
                                            b->A0.p, b->A1.p, rl, fl, s, f16 ? rowpk : nullptr));
    if (b->resident) HIPCHK(grtm::launch_clamp_extent((int)nrow, nkw, b->fext.as<int>(), b->ext_over.as<unsigned>(), s));
    HIPCHK(grtm::launch_kend_tiles((int)nrow, batch_rpt(b), b->nk, nkw, b->kend.as<int>() + batch_kend_off(b, g), s));
    return GRTM_OK;
}

int batch_gemm(grtm_batch *b, const grtm_batch::Group &g, hipStream_t s)
{
    const int M = 4 * b->nw * g.n;
    const size_t r0 = (size_t)g.m0 * b->nw;
    const int *kend = b->adaptive ? b->kend.as<int>() + batch_kend_off(b, g) : nullptr;
    const int *nkw = b->adaptive ? b->nk_w.as<int>() + r0 : nullptr;
    if (b->bsplit) {
        HIPCHK(grtm::launch_gemm_pair_bsplit(b->bsplit, M, b->nr, b->nk, b->A0.as<float>(), b->A1.as<float>(), b->nk,
                                             b->S0.p, b->S1.p, b->ldk, b->C0.as<float>(), b->C1.as<float>(),
                                             b->ws.as<float>(), kend, nkw, s,
                                             b->resident ? b->foff.as<long long>() : nullptr,
                                             b->resident ? b->fext.as<int>() : nullptr,
                                             b->bsplit == grtm::kSplitF16x3 ? b->row_peak.as<unsigned>() + r0
                                                                             : nullptr));
        return GRTM_OK;
    }
    if (b->pair && b->gemm_mode == 1) {
        HIPCHK(grtm::launch_gemm_pair_f32(M, b->nr, b->nk, b->kc, b->A0.as<float>(), b->A1.as<float>(), b->nk,
                                          b->B0.as<float>(), b->B1.as<float>(), b->nr, b->C0.as<float>(),
                                          b->C1.as<float>(), b->ws.as<float>(), kend, nkw, s,
                                          b->resident ? b->foff.as<long long>() : nullptr,
                                          b->resident ? b->fext.as<int>() : nullptr));
        return GRTM_OK;
    }
    if (b->pair && b->gemm_mode == 0) {
        HIPCHK(grtm::launch_gemm_pair_f64(M, b->nr, b->nk, b->kc, b->A0.as<double>(), b->A1.as<double>(), b->nk,
                                          b->B0.as<double>(), b->B1.as<double>(), b->nr, b->C0.as<double>(),
                                          b->C1.as<double>(), b->ws.as<double>(), kend, nkw, s));
        return GRTM_OK;
    }
    HIPCHK(grtm::launch_gemm_mfma(b->gemm_mode, M, b->nr, b->nk, b->A0.p, b->nk, b->B0.p, b->nr, b->C0.p, b->ws.p,
                                  kend, nkw, s));
    HIPCHK(grtm::launch_gemm_mfma(b->gemm_mode, M, b->nr, b->nk, b->A1.p, b->nk, b->B1.p, b->nr, b->C1.p, b->ws.p,
                                  kend, nkw, s));
    return GRTM_OK;
}

int batch_group(grtm_batch *b, const grtm_batch::Group &g, hipStream_t s)
{
    int st;
    if ((st = batch_cells(b, g, s)) || (st = batch_gemm(b, g, s))) return st;
    const size_t es = esize(b->prec);
    unsigned char *out = (unsigned char *)b->tr.p + (size_t)g.m0 * 4 * b->nr * b->N * es;
    HIPCHK(grtm::launch_post_ifft(batch_mixed(b) ? 1 : 0, b->pp, 1, 4 * b->nr * g.n, nullptr, b->C0.p, b->C1.p,
                                  b->nr, b->nr, out, nullptr, s, (long)4 * b->nw * b->nr));
    return GRTM_OK;
}

int batch_all(grtm_batch *b, hipStream_t s)
{
    int st;
    for (const grtm_batch::Group &g : b->groups)
        if ((st = batch_group(b, g, s))) return st;
    return GRTM_OK;
}

// mean ms of fn() over reps launches on the batch's stream (one warm-up)
template <typename F>
int time_ms(grtm_batch *b, int reps, F fn, double &ms)
{
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int st = fn();
    if (!st) {
        HIPCHK(hipEventRecord(e0, b->stream));
        for (int r = 0; r < reps && !st; r++) st = fn();
        HIPCHK(hipEventRecord(e1, b->stream));
        HIPCHK(hipEventSynchronize(e1));
        float f = 0;
        HIPCHK(hipEventElapsedTime(&f, e0, e1));
        ms = f / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

// Autotuner: recursion tile and contraction slice, timed on group 0.
int batch_autotune(grtm_batch *b)
{
    const auto t0 = std::chrono::steady_clock::now();
    int st, best_bs = 256;
    double best = 1e30;
    for (int bs : {64, 128, 256}) {
        if ((st = batch_items(b, bs, 0, 1))) return st;
        double ms = 0;
        if ((st = time_ms(b, 2, [&] { return batch_cells(b, b->groups[0], b->stream); }, ms))) return st;
        b->tuned++;
        if (ms < best) best = ms, best_bs = bs;
    }
    b->bs = best_bs;
    if (b->pair) {
        const int M = 4 * b->nw * b->groups[0].n;
        const int k0 = grtm::gemm_pair_slice(b->gemm_mode, M, b->nr, b->nk, b->nk);
        int best_kc = k0;
        best = 1e30;
        for (int kc : {k0 / 2, k0, 2 * k0}) {
            if (kc < 256 || kc % 32 || (kc > b->nk && kc / 2 >= b->nk)) continue;
            if (grtm::gemm_pair_workspace_bytes(M, b->nr, b->nk, kc) * esize(b->prec) / 4 > b->ws.n) continue;
            b->kc = kc;
            double ms = 0;
            if ((st = time_ms(b, 2, [&] { return batch_gemm(b, b->groups[0], b->stream); }, ms))) return st;
            b->tuned++;
            if (ms < best) best = ms, best_kc = kc;
        }
        b->kc = best_kc;
    }
    b->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return GRTM_OK;
}

// The planted cells (test hook) as per-group lists of group row nk + k, and
// room for them in the flag list.
int batch_upload_plant(grtm_batch *b)
{
    const size_t n = b->plant_h.size() / 3;
    std::vector<std::vector<unsigned long long>> per(b->groups.size());
    for (size_t i = 0; i < n; i++) {
        const int m = b->plant_h[3 * i], iw = b->plant_h[3 * i + 1], ik = b->plant_h[3 * i + 2];
        size_t gi = 0;
        while (gi + 1 < b->groups.size() && m >= b->groups[gi + 1].m0) gi++;
        const int j = m - b->groups[gi].m0;
        per[gi].push_back((unsigned long long)(j * b->nw + iw) * (unsigned long long)b->nk + (unsigned long long)ik);
    }
    std::vector<unsigned long long> all;
    std::vector<size_t> off(1, 0);
    for (const auto &v : per) {
        all.insert(all.end(), v.begin(), v.end());
        off.push_back(all.size());
    }
    int st;
    if ((st = batch_alloc_flags(b, n))) return st;
    if ((st = b->plant.alloc(sizeof(unsigned long long) * std::max<size_t>(1, all.size())))) return st;
    if (!all.empty())
        HIPCHK(hipMemcpy(b->plant.p, all.data(), sizeof(unsigned long long) * all.size(), hipMemcpyHostToDevice));
    b->plant_off = n ? off : std::vector<size_t>();
    return GRTM_OK;
}

// The learned grids' slot order of every group: per pass and receiver branch
// the group's models, heaviest first, with their rows' tiles, then their rows'
// check cells.  Segment s = pass * 4 + branch - 1 of a group spans slots
// [lseg_h[9 g + s], lseg_h[9 g + s + 1]).  A resident batch has one segment
// per pass (one launch runs every branch): bounds [0, end of pass 0, total...].
int batch_learned_slots(grtm_batch *b)
{
    int st;
    const int ng = (int)b->groups.size();
    std::vector<unsigned> slots, segs(9 * (size_t)ng);
    b->lslot0.assign(ng, 0);
    b->lnslots.assign(ng, 0);
    int maxslots = 0;
    for (int gi = 0; gi < ng; gi++) {
        const grtm_batch::Group &g = b->groups[gi];
        b->lslot0[gi] = slots.size();
        std::vector<int> order(g.n);
        for (int j = 0; j < g.n; j++) order[j] = j;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->L[g.m0 + x] > b->L[g.m0 + y]; });
        for (int pass = 0; pass < 2; pass++) {
            if (b->resident) segs[9 * (size_t)gi + pass] = unsigned(slots.size() - b->lslot0[gi]);
            for (int br = 1; br <= 4; br++) {
                if (!b->resident) segs[9 * (size_t)gi + pass * 4 + br - 1] = unsigned(slots.size() - b->lslot0[gi]);
                for (int phase = 0; phase < 2; phase++)
                    for (int j : order) {
                        if (b->md_h[g.m0 + j].g.branch != br) continue;
                        const int r64 = batch_mixed(b) ? b->rows64[g.m0 + j] : b->nw;
                        const int lo = pass == 0 ? 0 : r64, hi = pass == 0 ? r64 : b->nw;
                        for (int iw = lo; iw < hi; iw++)
                            slots.push_back(unsigned(j * b->nw + iw) | (phase ? grtm::kCheckItem : 0u));
                    }
            }
        }
        b->lnslots[gi] = int(slots.size() - b->lslot0[gi]);
        for (int q = b->resident ? 2 : 8; q <= 8; q++) segs[9 * (size_t)gi + q] = unsigned(b->lnslots[gi]);
        maxslots = std::max(maxslots, b->lnslots[gi]);
    }
    const int tpb = (b->nk + b->bs - 1) / b->bs;  // items of one slot at most
    const size_t cap = (size_t)maxslots * tpb;
    if (cap > 0x7fffffffu || (b->resident && (size_t)b->nm * b->nw >= (size_t(1) << 31) / 4))
        return GRTM_E_UNSUPPORTED;
    if ((st = b->lslots.alloc(sizeof(unsigned) * std::max<size_t>(1, slots.size()))) ||
        (st = b->lslot_off.alloc(sizeof(unsigned) * grtm::batch_items_words(std::max(1, maxslots)))) ||
        (st = b->lseg_slots.alloc(sizeof(unsigned) * segs.size())) ||
        (st = b->lseg_items.alloc(sizeof(unsigned) * segs.size())) || (st = b->lctr.alloc(sizeof(unsigned) * 8)) ||
        (st = b->litems.alloc(sizeof(unsigned long long) * std::max<size_t>(1, cap))))
        return st;
    if (!slots.empty()) HIPCHK(hipMemcpy(b->lslots.p, slots.data(), sizeof(unsigned) * slots.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->lseg_slots.p, segs.data(), sizeof(unsigned) * segs.size(), hipMemcpyHostToDevice));
    b->lseg_h = segs;
    b->litems_cap = (int)cap;
    return GRTM_OK;
}

// The device-built grids of the last execution into nk_pred_h / nchk_h.
int batch_pred_to_host(grtm_batch *b)
{
    if (!b->net || !b->pred_stale) return GRTM_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    const size_t n = (size_t)b->nm * b->nw;
    HIPCHK(hipMemcpy(b->nk_pred_h.data(), b->nk_pred.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(b->nchk_h.data(), b->nchk.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    b->pred_stale = false;
    return GRTM_OK;
}

// Back to the grouped layout (dense operand rows for nb_grouped models per
// group); the caller rebuilds items or slots and captures the graph again.
int batch_make_grouped(grtm_batch *b)
{
    if (!b->resident) return GRTM_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->resident = false;
    b->resident_floats = 0;
    b->foff.release();
    b->fext.release();
    b->ext_over.release();
    b->A0.release();
    b->A1.release();
    b->C0.release();
    b->C1.release();
    b->nb = b->nb_grouped;
    b->groups.clear();
    for (int m0 = 0; m0 < b->nm; m0 += b->nb) {
        grtm_batch::Group g{};
        g.m0 = m0;
        g.n = std::min(b->nb, b->nm - m0);
        b->groups.push_back(g);
    }
    const size_t arows = (size_t)4 * b->nw * b->nb, es = esize(b->prec);
    int gmax = 0;
    for (const auto &gr : b->groups) gmax = std::max(gmax, gr.n);
    const int M2[2] = {4 * b->nw * gmax, 4 * b->nw * gmax};
    size_t wsb = std::max(grtm::gemm_workspace_bytes(b->gemm_mode, (int)arows, b->nr, b->nk),
                          grtm::gemm_pair_workspace_bytes((int)arows, b->nr, b->nk, std::max(256, b->kc / 2)) * es / 4);
    if (b->gemm_mode == 1) wsb = std::max(wsb, pair_sk_bytes(b->bsplit, M2[0], b->nr, b->nk));
    int st;
    if ((st = b->A0.alloc(arows * b->nk * es)) || (st = b->A1.alloc(arows * b->nk * es)) ||
        (st = b->C0.alloc(arows * b->nr * es)) || (st = b->C1.alloc(arows * b->nr * es)) ||
        (st = b->ws.alloc(wsb)) || (st = b->fb_gcount.alloc(sizeof(unsigned) * b->groups.size())) ||
        (st = b->kend.alloc(sizeof(int) * b->groups.size() * batch_kend_per_group(b))))
        return st;
    return batch_upload_plant(b);
}

// The resident layout for a learned FP32 batch whose grids the last pass
// formed: row r keeps fext[r] = its final extent rounded up to 4 wavenumbers
// (nk_w after the fallback), one group of every model, the branches of a
// pass in one recursion launch.  It stays grouped when the layout does not
// fit the device with 4 GiB to spare.
int batch_make_resident(grtm_batch *b)
{
    if (b->resident || !b->adaptive || !b->net || !batch_mixed(b) || b->gemm_mode != 1 || !b->pair) return GRTM_OK;
    const size_t n = (size_t)b->nm * b->nw, nrows = 4 * n;
    if (nrows >= (size_t(1) << 31)) return GRTM_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    std::vector<int> nkw(n), ext(n);
    std::vector<long long> off(n);
    HIPCHK(hipMemcpy(nkw.data(), b->nk_w.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    size_t tot = 0;
    for (size_t r = 0; r < n; r++) {
        ext[r] = std::max(4, (std::min(nkw[r], b->nk) + 3) / 4 * 4);
        off[r] = (long long)tot;
        tot += 4 * (size_t)ext[r];
    }
    const size_t ws_need = pair_sk_bytes(b->bsplit, (int)nrows, b->nr, b->nk);
    const size_t a_bytes = (tot + 64) * sizeof(float);  // 64 entries past the last row (16-byte loads)
    const size_t c_bytes = nrows * b->nr * sizeof(float);
    const int tpb = (b->nk + b->bs - 1) / b->bs;
    const size_t need = 2 * a_bytes + 2 * c_bytes + ws_need + n * (sizeof(long long) + sizeof(int)) +
                        2 * n * tpb * sizeof(unsigned long long) + (kBatchFlagCap + 1) * sizeof(unsigned long long);
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    const size_t held = b->A0.n + b->A1.n + b->C0.n + b->C1.n + b->ws.n + b->litems.n + b->flags.n;
    if (need + (size_t(4) << 30) > free_b + held) return GRTM_OK;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->A0.release();
    b->A1.release();
    b->C0.release();
    b->C1.release();
    b->ws.release();
    b->litems.release();
    int st;
    if ((st = b->A0.alloc(a_bytes)) || (st = b->A1.alloc(a_bytes)) || (st = b->C0.alloc(c_bytes)) ||
        (st = b->C1.alloc(c_bytes)) || (st = b->ws.alloc(ws_need)) ||
        (st = b->foff.alloc(sizeof(long long) * n)) || (st = b->fext.alloc(sizeof(int) * n)) ||
        (st = b->ext_over.alloc(sizeof(unsigned))))
        return st;
    HIPCHK(hipMemcpy(b->foff.p, off.data(), sizeof(long long) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->fext.p, ext.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(b->ext_over.p, 0, sizeof(unsigned)));  // set-up, outside any graph
    b->resident = true;
    b->resident_floats = tot;
    b->nb = b->nm;
    b->groups.clear();
    grtm_batch::Group g{};
    g.m0 = 0;
    g.n = b->nm;
    b->groups.push_back(g);
    if ((st = b->fb_gcount.alloc(sizeof(unsigned))) ||
        (st = b->kend.alloc(sizeof(int) * batch_kend_per_group(b))) || (st = batch_upload_plant(b)) ||
        (st = batch_learned_slots(b)))
        return st;
    // the resident layout's own eager pass (the host's view of the grids)
    if ((st = batch_group(b, b->groups[0], b->stream))) return st;
    return batch_pred_to_host(b);
}

int batch_capture(grtm_batch *b)
{
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int st = batch_all(b, b->stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    HIPCHK(e);
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    size_t n = 0;
    HIPCHK(hipGraphGetNodes(b->graph, nullptr, &n));
    b->nodes = n;
    return GRTM_OK;
}

}  // namespace

extern "C" {

int grtm_batch_create(grtm_ctx *ctx, const grtm_model *models, int n_models, const double *dists, int nr,
                      const grtm_grid *grid, int precision, int compat, int flags, grtm_batch **out)
{
    if (!ctx || !models || n_models < 1 || !dists || nr < 1 || !grid || !out) return GRTM_E_INPUT;
    *out = nullptr;
    if (precision != GRTM_PREC_FP64 && precision != GRTM_PREC_FP32 && precision != GRTM_PREC_BF16X3 &&
        precision != GRTM_PREC_BF16X6 && precision != GRTM_PREC_FP16X3)
        return GRTM_E_INPUT;
    if (!grid_ok(grid, precision)) return GRTM_E_INPUT;
    if (precision == GRTM_PREC_FP64 && grid->nw > 4096) return GRTM_E_UNSUPPORTED;
    if (!grtm::cells_tile_limit(grid->nw, grid->nk) || (grid->nk + 63) / 64 > 4096) return GRTM_E_UNSUPPORTED;
    for (int i = 0; i < nr; i++)
        if (!(dists[i] > 0)) return GRTM_E_INPUT;
    HIPCHK(hipSetDevice(ctx->device));
    grtm_batch *b = new grtm_batch;
    struct Guard {
        grtm_batch *&b;
        ~Guard() { delete b; }
    } guard{b};
    b->ctx = ctx;
    b->prec = precision;
    b->compat = compat;
    b->gemm_mode = precision == GRTM_PREC_FP64 ? 0 : precision == GRTM_PREC_BF16X3 ? 2 : 1;
    b->nm = n_models;
    b->nw = grid->nw;
    b->N = 2 * grid->nw;
    b->nk = grid->nk;
    b->nr = nr;
    b->dk = grid->kmax / grid->nk;
    const double T = b->N * grid->dt;
    b->df = 1.0 / T;
    b->omega_i = grid->omega_i > 0 ? grid->omega_i : M_PI / T;
    b->pp = post_params(grid, b->nw, b->df, b->omega_i);
    if (int e = post_tables(b->pp, precision == GRTM_PREC_FP64 ? 0 : 1, b->ptab)) return e;
    HIPCHK(hipStreamCreate(&b->stream));

    // models: descriptors and layer tables, concatenated
    int st;
    std::vector<LayerW<double>> all64;
    b->L.resize(n_models);
    b->rows64.resize(n_models);
    b->md_h.resize(n_models);
    for (int m = 0; m < n_models; m++) {
        ModelTables mt;
        const grtm_model &gm = models[m];
        if ((st = model_tables(gm.layers, gm.n_layers, gm.src_depth, gm.rcv_depth, compat, b->nw, b->df, b->omega_i,
                               true, mt)))
            return st;
        b->L[m] = mt.st.L;
        b->rows64[m] = mt.rows64;
        b->layers_h.emplace_back(gm.layers, gm.layers + 6 * (size_t)gm.n_layers);
        b->src_h.push_back(gm.src_depth);
        b->rcv_h.push_back(gm.rcv_depth);
        b->rows64_rule.push_back(mt.rows64);
        b->md_h[m].g = mt.g;
        b->md_h[m].lw_off = (long)all64.size();
        b->md_h[m].r64 = precision != GRTM_PREC_FP64 ? mt.rows64 : b->nw;
        all64.insert(all64.end(), mt.lw.begin(), mt.lw.end());
        double oc[2] = {0, 0};
        if (grtm::opcount_cell_bot(0, mt.st.L, mt.g.s, mt.g.rl, mt.g.branch, mt.g.up, -1, oc, mt.g.zr == 0.0))
            return GRTM_E_INPUT;
        b->flops += oc[0] * b->nw * b->nk;
    }
    if ((st = b->md.alloc(sizeof(grtm::ModelDesc) * n_models)) ||
        (st = b->lw64.alloc(sizeof(LayerW<double>) * all64.size())))
        return st;
    HIPCHK(hipMemcpy(b->md.p, b->md_h.data(), sizeof(grtm::ModelDesc) * n_models, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->lw64.p, all64.data(), sizeof(LayerW<double>) * all64.size(), hipMemcpyHostToDevice));
    if (precision != GRTM_PREC_FP64) {
        const std::vector<LayerW<float>> lf = narrow_tables(all64);
        if ((st = b->lw32.alloc(sizeof(LayerW<float>) * lf.size()))) return st;
        HIPCHK(hipMemcpy(b->lw32.p, lf.data(), sizeof(LayerW<float>) * lf.size(), hipMemcpyHostToDevice));
    }
    std::vector<LayerW<double>>().swap(all64);

    // group size: device bytes per model of the group buffers within a budget
    const size_t es = esize(precision);
    const size_t per_model = 2 * (size_t)4 * b->nw * ((size_t)b->nk + b->nr) * es +
                             (precision != GRTM_PREC_FP64 ? (size_t)b->nw * b->nk * 4 : 0);
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    const size_t budget = std::min<size_t>((size_t)24 << 30, free_b / 4);
    size_t nb = std::max<size_t>(1, budget / per_model);
    nb = std::min<size_t>(nb, 64);
    nb = std::min<size_t>(nb, (size_t)n_models);
    b->nb = b->nb_grouped = (int)nb;
    for (int m0 = 0; m0 < n_models; m0 += b->nb)
        b->groups.push_back({m0, std::min(b->nb, n_models - m0), 0, 0, 0, 0});

    const size_t arows = (size_t)4 * b->nw * b->nb;
    if ((st = b->dists.alloc(sizeof(double) * nr)) || (st = b->B0.alloc((size_t)b->nk * nr * es)) ||
        (st = b->B1.alloc((size_t)b->nk * nr * es)) || (st = b->A0.alloc(arows * b->nk * es)) ||
        (st = b->A1.alloc(arows * b->nk * es)) || (st = b->C0.alloc(arows * nr * es)) ||
        (st = b->C1.alloc(arows * nr * es)) || (st = b->tr.alloc((size_t)n_models * 4 * nr * b->N * es)))
        return st;
    if ((st = batch_alloc_flags(b, 0))) return st;
    const bool pair_ok = grtm::gemm_pair_ok(b->nk, nr, b->A0.p, b->A1.p, b->B0.p, b->B1.p);
    if (pair_ok && (precision == GRTM_PREC_BF16X3 || precision == GRTM_PREC_BF16X6)) {
        b->gemm_mode = 1;
        b->bsplit = precision == GRTM_PREC_BF16X3 ? grtm::kSplitBf16x3 : grtm::kSplitBf16x6;
    }
    if (precision == GRTM_PREC_FP16X3) {
        // the FP16 split on the pair layout only, its row scales from the
        // recursion's per-row peaks (kept on dense grids too)
        if (!pair_ok) return GRTM_E_UNSUPPORTED;
        b->bsplit = grtm::kSplitF16x3;
        if ((st = b->row_peak.alloc(sizeof(unsigned) * (size_t)n_models * b->nw))) return st;
    }
    size_t wsb = grtm::gemm_workspace_bytes(b->gemm_mode, (int)arows, nr, b->nk);
    if (b->gemm_mode != 2 && pair_ok) {
        b->pair = true;
        b->kc = grtm::gemm_pair_slice(b->gemm_mode, (int)arows, nr, b->nk, b->nk);
        const int kmin = std::max(256, b->kc / 2);
        wsb = std::max(wsb, grtm::gemm_pair_workspace_bytes((int)arows, nr, b->nk, kmin) * es / 4);
    }
    if (b->bsplit) {
        wsb = std::max(wsb, pair_sk_bytes(b->bsplit, (int)arows, nr, b->nk));
        if ((st = split_alloc(b->bsplit, b->nk, nr, b->S0, b->S1, b->ldk))) return st;
    }
    if ((st = b->ws.alloc(wsb))) return st;
    HIPCHK(hipMemcpy(b->dists.p, dists, sizeof(double) * nr, hipMemcpyHostToDevice));
    HIPCHK(grtm::launch_bessel_tab(precision == GRTM_PREC_FP64 ? 0 : 1, compat ? 0 : 1, b->nk, b->dk,
                                   b->dists.as<double>(), nr, b->B0.p, b->B1.p, nr, b->stream));
    if (b->bsplit) {
        HIPCHK(grtm::launch_split_table(b->bsplit, b->nk, nr, b->B0.as<float>(), nr, b->S0.p, b->ldk, b->stream));
        HIPCHK(grtm::launch_split_table(b->bsplit, b->nk, nr, b->B1.as<float>(), nr, b->S1.p, b->ldk, b->stream));
    }

    if (flags & GRTM_BATCH_AUTOTUNE) {
        if ((st = batch_autotune(b))) return st;
    }
    if ((st = batch_items(b, b->bs, 0, (int)b->groups.size()))) return st;
    // one eager pass of the first group (kernel attributes are set outside the capture)
    if ((st = batch_group(b, b->groups[0], b->stream))) return st;
    HIPCHK(hipStreamSynchronize(b->stream));
    if ((st = batch_capture(b))) return st;
    *out = b;
    b = nullptr;
    return GRTM_OK;
}

int grtm_batch_set_adaptive(grtm_batch *b, const double *kcut, const double *window, int stride, double tol)
{
    if (!b) return GRTM_E_INPUT;
    int st;
    if (kcut) {
        if (!window || stride < 1 || !(tol > 0)) return GRTM_E_INPUT;
        const size_t n = (size_t)b->nm * b->nw;
        for (int m = 0; m < b->nm; m++)  // validated before any batch state changes
            if (!(window[m] >= 0)) return GRTM_E_INPUT;
        for (size_t i = 0; i < n; i++)
            if (!(kcut[i] >= 0)) return GRTM_E_INPUT;
    }
    b->net = nullptr;  // host cuts replace learned grids
    b->pred_stale = false;
    if (b->resident && (st = batch_make_grouped(b))) return st;  // host-built items run grouped
    if (!kcut) {
        b->adaptive = false;
    } else {
        const size_t n = (size_t)b->nm * b->nw;
        b->nk_pred_h.assign(n, 0);
        b->nchk_h.assign(n, 0);
        b->max_chk.assign(b->nm, 0);
        for (int m = 0; m < b->nm; m++) {
            for (int iw = 0; iw < b->nw; iw++) {
                const size_t i = (size_t)m * b->nw + iw;
                int nkp, nc;
                row_cut(b->nk, kcut[i], b->dk, stride, window[m], nkp, nc);
                b->nk_pred_h[i] = nkp;
                b->nchk_h[i] = nc;
                b->max_chk[m] = std::max(b->max_chk[m], b->nchk_h[i]);
            }
        }
        if (!b->nk_pred.p) {
            if ((st = b->nk_pred.alloc(sizeof(int) * n)) || (st = b->nk_w.alloc(sizeof(int) * n)) ||
                (st = b->nchk.alloc(sizeof(int) * n)) || (st = b->row_peak.alloc(sizeof(unsigned) * n)) ||
                (st = b->tail_max.alloc(sizeof(unsigned) * n)) || (st = b->fb_count.alloc(sizeof(unsigned) * b->nm)) ||
                (st = b->fb_list.alloc(sizeof(unsigned) * 2 * n)) ||
                (st = b->fb_gcount.alloc(sizeof(unsigned) * b->groups.size())) ||
                (st = b->kend.alloc(sizeof(int) * b->groups.size() * batch_kend_per_group(b))))
                return st;
        }
        HIPCHK(hipMemcpy(b->nk_pred.p, b->nk_pred_h.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b->nchk.p, b->nchk_h.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        if (b->pair && b->gemm_mode == 1) {
            // the FP32 contraction of a learned grid runs stream-K
            int gmax = 0;
            for (const auto &gr : b->groups) gmax = std::max(gmax, gr.n);
            const size_t need = pair_sk_bytes(b->bsplit, 4 * b->nw * gmax, b->nr, b->nk);
            if (need > b->ws.n && (st = b->ws.alloc(need))) return st;
        }
        b->chk_stride = stride;
        b->chk_tol = tol;
        b->adaptive = true;
    }
    if ((st = batch_items(b, b->bs, 0, (int)b->groups.size()))) return st;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    if ((st = batch_group(b, b->groups[0], b->stream))) return st;  // eager pass outside the capture
    HIPCHK(hipStreamSynchronize(b->stream));
    return batch_capture(b);
}


int grtm_batch_set_learned(grtm_batch *b, grtm_mlp *net, double lmargin, const double *window, int stride, double tol)
{
    if (!b) return GRTM_E_INPUT;
    if (!net) return grtm_batch_set_adaptive(b, nullptr, nullptr, 1, 1.0);
    if (!b->pair || b->gemm_mode == 2) return GRTM_E_UNSUPPORTED;
    if (!window || stride < 1 || !(tol > 0) || !std::isfinite(lmargin)) return GRTM_E_INPUT;
    for (int m = 0; m < b->nm; m++) {
        if (!(window[m] >= 0)) return GRTM_E_INPUT;
        if ((int)b->layers_h[m].size() / 6 > grtm::kLearnSlots) return GRTM_E_UNSUPPORTED;
    }
    int st;
    // the grids are formed on the grouped layout first (its eager pass sizes
    // the resident one)
    if (b->resident && (st = batch_make_grouped(b))) return st;
    const size_t n = (size_t)b->nm * b->nw;
    if (!b->nk_pred.p) {
        if ((st = b->nk_pred.alloc(sizeof(int) * n)) || (st = b->nk_w.alloc(sizeof(int) * n)) ||
            (st = b->nchk.alloc(sizeof(int) * n)) || (st = b->row_peak.alloc(sizeof(unsigned) * n)) ||
            (st = b->tail_max.alloc(sizeof(unsigned) * n)) || (st = b->fb_count.alloc(sizeof(unsigned) * b->nm)) ||
            (st = b->fb_list.alloc(sizeof(unsigned) * 2 * n)) ||
            (st = b->fb_gcount.alloc(sizeof(unsigned) * b->groups.size())) ||
            (st = b->kend.alloc(sizeof(int) * b->groups.size() * batch_kend_per_group(b))))
            return st;
    }
    if (b->gemm_mode == 1) {
        int gmax = 0;
        for (const auto &gr : b->groups) gmax = std::max(gmax, gr.n);
        const size_t need = pair_sk_bytes(b->bsplit, 4 * b->nw * gmax, b->nr, b->nk);
        if (need > b->ws.n && (st = b->ws.alloc(need))) return st;
    }
    // predictor inputs of every model (the floor follows each model's FP64 rows)
    std::vector<grtm::LearnModel> lm(b->nm);
    for (int m = 0; m < b->nm; m++) {
        grtm::LearnModel &d = lm[m];
        d = grtm::LearnModel{};
        d.n = (int)b->layers_h[m].size() / 6;
        d.vs_min = 1e30;
        for (int j = 0; j < d.n; j++) {
            for (int c = 0; c < 6; c++) d.rows[j][c] = b->layers_h[m][(size_t)6 * j + c];
            d.vs_min = std::min(d.vs_min, d.rows[j][2]);
        }
        d.floor_rows = batch_mixed(b) ? b->rows64[m] : b->rows64_rule[m];
        d.src = b->src_h[m];
        d.rcv = b->rcv_h[m];
        d.window = window[m];
    }
    if ((st = b->lmodels.alloc(sizeof(grtm::LearnModel) * b->nm))) return st;
    HIPCHK(hipMemcpy(b->lmodels.p, lm.data(), sizeof(grtm::LearnModel) * b->nm, hipMemcpyHostToDevice));
    if ((st = batch_learned_slots(b))) return st;
    b->lg = grtm::LearnGrid{b->nw, b->nk, stride, 0, b->pp.dt, b->dk, b->dk * b->nk, lmargin};
    b->chk_stride = stride;
    b->chk_tol = tol;
    b->net = net;
    b->adaptive = true;
    b->nk_pred_h.assign(n, 0);
    b->nchk_h.assign(n, 0);
    b->max_chk.assign(b->nm, 0);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    // autotune: the items a persistent block takes per counter visit, timed
    // on the first group's recursion stage
    {
        const int cand[5] = {1, 2, 4, 8, 16};
        double best = 1e30;
        int bc = 1;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        for (int c : cand) {
            b->lchunk = c;
            for (int rep = 0; rep < 2 && !st; rep++) {  // the first run warms the kernels up
                HIPCHK(hipEventRecord(e0, b->stream));
                st = batch_cells(b, b->groups[0], b->stream);
                HIPCHK(hipEventRecord(e1, b->stream));
                HIPCHK(hipEventSynchronize(e1));
            }
            if (st) break;
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            b->tuned++;
            b->tune_ms += ms;
            if (ms < best) best = ms, bc = c;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b->lchunk = bc;
        if (st) {
            b->net = nullptr;
            b->adaptive = false;
            return st;
        }
    }
    // one eager pass of every group (the host's view of the grids; kernel
    // attributes set outside the capture), then the capture
    for (const grtm_batch::Group &g : b->groups)
        if ((st = batch_group(b, g, b->stream))) {
            b->net = nullptr;
            b->adaptive = false;
            return st;
        }
    if ((st = batch_pred_to_host(b))) return st;
    // FP32 batches: one group of every model with each row stored to the
    // extent this pass reached, when it fits the device (else the groups stay)
    if ((st = batch_make_resident(b))) {
        b->net = nullptr;
        b->adaptive = false;
        return st;
    }
    return batch_capture(b);
}

int grtm_batch_adaptive_stats(grtm_batch *b, double *out)
{
    if (!b || !out) return GRTM_E_INPUT;
    out[0] = out[1] = out[2] = out[3] = 0;
    if (!b->adaptive) return GRTM_OK;
    int st;
    if ((st = batch_pred_to_host(b))) return st;
    HIPCHK(hipStreamSynchronize(b->stream));
    const size_t n = (size_t)b->nm * b->nw;
    std::vector<int> nkw(n);
    std::vector<unsigned> fb(b->nm);
    HIPCHK(hipMemcpy(nkw.data(), b->nk_w.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fb.data(), b->fb_count.p, sizeof(unsigned) * b->nm, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) out[0] += b->nk_pred_h[i], out[1] += b->nchk_h[i], out[2] += nkw[i];
    for (int m = 0; m < b->nm; m++) out[3] += fb[m];
    return GRTM_OK;
}

int grtm_batch_plant_nonfinite(grtm_batch *b, const int *cells, int n)
{
    if (!b || n < 0 || n > GRTM_MAX_PLANT || (n && !cells)) return GRTM_E_INPUT;
    if (!batch_mixed(b)) return GRTM_E_UNSUPPORTED;
    const int nm = (int)b->L.size();
    for (int i = 0; i < n; i++) {
        const int m = cells[3 * i], iw = cells[3 * i + 1], ik = cells[3 * i + 2];
        if (m < 0 || m >= nm || iw < b->rows64[m] || iw >= b->nw || ik < 0 || ik >= b->nk) return GRTM_E_INPUT;
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    b->plant_h.assign(cells, cells + 3 * (size_t)n);
    int st;
    if ((st = batch_upload_plant(b))) return st;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return batch_capture(b);
}

int grtm_batch_execute(grtm_batch *b, void *stream)
{
    if (!b || !b->exec) return GRTM_E_INPUT;
    HIPCHK(hipGraphLaunch(b->exec, stream ? (hipStream_t)stream : b->stream));
    return GRTM_OK;
}

int grtm_batch_execute_eager(grtm_batch *b)
{
    if (!b) return GRTM_E_INPUT;
    return batch_all(b, b->stream);
}

int grtm_batch_sync(grtm_batch *b)
{
    if (!b) return GRTM_E_INPUT;
    HIPCHK(hipStreamSynchronize(b->stream));
    return GRTM_OK;
}

int grtm_batch_traces_out(grtm_batch *b, int m0, int m1, double *traces)
{
    if (!b || !traces || m0 < 0 || m1 > b->nm || m0 >= m1) return GRTM_E_INPUT;
    HIPCHK(hipStreamSynchronize(b->stream));
    const size_t per = (size_t)4 * b->nr * b->N, n = per * (m1 - m0);
    if (b->prec == GRTM_PREC_FP64) {
        HIPCHK(hipMemcpy(traces, b->tr.as<double>() + per * m0, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> f(n);
        HIPCHK(hipMemcpy(f.data(), b->tr.as<float>() + per * m0, n * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) traces[i] = f[i];
    }
    return GRTM_OK;
}

int grtm_batch_timing(grtm_batch *b, int reps, double *ms)
{
    if (!b || !ms || reps < 1) return GRTM_E_INPUT;
    int st = time_ms(b, reps, [&] { return grtm_batch_execute(b, nullptr); }, ms[0]);
    if (st) return st;
    // stage split of one eager pass: events around every group's stages
    const size_t ng = b->groups.size();
    const bool learned = b->adaptive && b->net;
    std::vector<hipEvent_t> ev(6 * ng, nullptr);
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    for (size_t i = 0; i < ng && !st; i++) {
        const grtm_batch::Group &g = b->groups[i];
        HIPCHK(hipEventRecord(ev[4 * i], b->stream));
        if (learned) b->ev_pred0 = ev[4 * ng + 2 * i], b->ev_pred1 = ev[4 * ng + 2 * i + 1];
        st = batch_cells(b, g, b->stream);
        b->ev_pred0 = b->ev_pred1 = nullptr;
        if (st) break;
        HIPCHK(hipEventRecord(ev[4 * i + 1], b->stream));
        if ((st = batch_gemm(b, g, b->stream))) break;
        HIPCHK(hipEventRecord(ev[4 * i + 2], b->stream));
        const size_t es = esize(b->prec);
        unsigned char *out = (unsigned char *)b->tr.p + (size_t)g.m0 * 4 * b->nr * b->N * es;
        HIPCHK(grtm::launch_post_ifft(batch_mixed(b) ? 1 : 0, b->pp, 1, 4 * b->nr * g.n, nullptr, b->C0.p, b->C1.p,
                                      b->nr, b->nr, out, nullptr, b->stream, (long)4 * b->nw * b->nr));
        HIPCHK(hipEventRecord(ev[4 * i + 3], b->stream));
    }
    if (!st) {
        HIPCHK(hipStreamSynchronize(b->stream));
        ms[1] = ms[2] = ms[3] = ms[4] = 0;
        for (size_t i = 0; i < ng; i++) {
            float a = 0, c = 0, d = 0, f = 0;
            HIPCHK(hipEventElapsedTime(&a, ev[4 * i], ev[4 * i + 1]));
            HIPCHK(hipEventElapsedTime(&c, ev[4 * i + 1], ev[4 * i + 2]));
            HIPCHK(hipEventElapsedTime(&d, ev[4 * i + 2], ev[4 * i + 3]));
            if (learned) HIPCHK(hipEventElapsedTime(&f, ev[4 * ng + 2 * i], ev[4 * ng + 2 * i + 1]));
            ms[1] += a, ms[2] += c, ms[3] += d, ms[4] += f;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return st;
}

int grtm_batch_config(grtm_batch *b, double *out)
{
    if (!b || !out) return GRTM_E_INPUT;
    out[0] = b->nb;
    out[1] = b->bs;
    out[2] = b->pair ? b->kc : 0;
    out[3] = (double)b->groups.size();
    out[4] = (double)b->nodes;
    out[5] = b->tuned;
    out[6] = b->tune_ms;
    out[7] = b->lchunk;
    out[8] = b->resident ? 1 : 0;
    out[9] = 0;
    if (b->resident) {
        unsigned over = 0;
        HIPCHK(hipStreamSynchronize(b->stream));
        HIPCHK(hipMemcpy(&over, b->ext_over.p, sizeof(unsigned), hipMemcpyDeviceToHost));
        out[9] = over;
    }
    return GRTM_OK;
}

int grtm_batch_work(grtm_batch *b, double *out)
{
    if (!b || !out) return GRTM_E_INPUT;
    {
        const int st = batch_pred_to_host(b);
        if (st) return st;
    }
    double cl = 0;
    for (int m = 0; m < b->nm; m++) cl += b->L[m];
    const double cells = (double)b->nw * b->nk;
    out[0] = cells * b->nm;
    out[1] = cells * cl;
    out[2] = 2.0 * 2.0 * (4.0 * b->nw) * b->nk * b->nr * b->nm;
    out[3] = 2.0 * b->nr * b->nm;
    out[4] = b->flops;
    // recursion flops of the cells the current grid solves (dense, or below
    // the predicted cuts), each cell counted with its upward sweep cut by
    // depth: estimated on every 8th row and every 16th wavenumber
    double fl = 0;
    std::vector<LayerW<double>> lw;
    for (int m = 0; m < b->nm; m++) {
        const grtm::Geo &g = b->md_h[m].g;
        lw.resize((size_t)b->nw * g.L);
        HIPCHK(hipMemcpy(lw.data(), b->lw64.as<LayerW<double>>() + b->md_h[m].lw_off, lw.size() * sizeof(LayerW<double>),
                         hipMemcpyDeviceToHost));
        std::vector<double> nbot(g.L, 0.0);
        for (int iw = 0; iw < b->nw; iw += 8) {
            const int nk = b->adaptive ? b->nk_pred_h[(size_t)m * b->nw + iw] : b->nk;
            for (int ik = 0; ik < nk; ik += 16)
                nbot[g.branch != 4 ? grtm::sweep_bottom(lw.data() + (size_t)iw * g.L, g.L, g.s,
                                                        g.branch == 3 ? g.zr : g.zs, b->dk * ik)
                                   : g.L - 1] += 8.0 * 16.0;
        }
        for (int j = 0; j < g.L; j++) {
            if (nbot[j] == 0) continue;
            double oc[2];
            if (grtm::opcount_cell_bot(0, g.L, g.s, g.rl, g.branch, g.up, j, oc, g.zr == 0.0)) return GRTM_E_INPUT;
            fl += nbot[j] * oc[0];
        }
    }
    out[5] = fl;
    return GRTM_OK;
}

void grtm_batch_destroy(grtm_batch *b) { delete b; }

}  // extern "C"
