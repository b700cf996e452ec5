// mlp.hip -- the learned wavenumber-grid predictor: a multilayer perceptron
// trained and evaluated on the GPU with BF16 operands and FP32 accumulation.
//
// Layers l = 0..D: Z_l = H_{l-1} W_l + b_l, H_l = bf16(relu(Z_l)) for hidden
// layers, the last layer is linear.  FP32 master weights, BF16 working
// copies, Adam.  Loss: masked mean squared error.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/grtm.h"
#include "kernels.h"

namespace {

#define HIPCHK(x)                                   \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) return GRTM_E_DEVICE; \
    } while (0)

typedef __bf16 bf16;

template <typename T> __device__ __forceinline__ float ld(const T *p, size_t i) { return float(p[i]); }

// C[M x N] = A[M x K] B[K x N] (+ bias[n]); A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn].  64x64 tile, 256 threads x 4x4, FP32 accumulation.
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_mm(int M, int N, int K, const TA *__restrict__ A, long sam, long sak,
                                            const TB *__restrict__ B, long sbk, long sbn, float *__restrict__ Cm,
                                            long ldc, const float *__restrict__ bias, int accumulate)
{
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ float As[BK][BM + 1];
    __shared__ float Bs[BK][BN + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += BK) {
        for (int e = threadIdx.x; e < BM * BK; e += 256) {
            const int mm = e / BK, kk = e % BK, gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < M && gk < K) ? ld(A, (size_t)gm * sam + (size_t)gk * sak) : 0.f;
        }
        for (int e = threadIdx.x; e < BN * BK; e += 256) {
            const int kk = e / BN, nn = e % BN, gk = k0 + kk, gn = n0 + nn;
            Bs[kk][nn] = (gk < K && gn < N) ? ld(B, (size_t)gk * sbk + (size_t)gn * sbn) : 0.f;
        }
        __syncthreads();
        for (int kk = 0; kk < BK; kk++) {
            float a[4], b[4];
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
            for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx * 4 + j];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            const int gm = m0 + ty * 4 + i, gn = n0 + tx * 4 + j;
            if (gm < M && gn < N) {
                float v = acc[i][j] + (bias ? bias[gn] : 0.f);
                if (accumulate) v += Cm[(size_t)gm * ldc + gn];
                Cm[(size_t)gm * ldc + gn] = v;
            }
        }
}

__global__ void k_act(int n, const float *__restrict__ Z, bf16 *__restrict__ H, int relu)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) H[t] = bf16(relu ? fmaxf(Z[t], 0.f) : Z[t]);
}

__global__ void k_relu_bwd(int n, const float *__restrict__ Z, float *__restrict__ dH)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && !(Z[t] > 0.f)) dH[t] = 0.f;
}

// Loss per output: the first nq outputs of every row use the pinball loss of
// quantile q, M max(q (T - Y), (q - 1)(T - Y)); the others the squared error
// M (Y - T)^2 (M weights the mask).  dY and per-block loss partials, / count.
__global__ void k_mse(int n, int n_out, int nq, float q, float inv_count, const float *__restrict__ Y,
                      const float *__restrict__ T, const float *__restrict__ Mk, float *__restrict__ dY,
                      float *__restrict__ lossp)
{
    __shared__ float red[256];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (t < n) {
        if (t % n_out < nq) {
            const float d = T[t] - Y[t];
            dY[t] = Mk[t] * (d > 0.f ? -q : 1.f - q) * inv_count;
            l = Mk[t] * fmaxf(q * d, (q - 1.f) * d) * inv_count;
        } else {
            const float d = (Y[t] - T[t]) * Mk[t];
            dY[t] = 2.f * d * inv_count;
            l = d * d * inv_count;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

__global__ void k_colsum(int B, int N, const float *__restrict__ dZ, float *__restrict__ db)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int b = 0; b < B; b++) s += dZ[(size_t)b * N + n];
    db[n] = s;
}

__global__ void k_adam(int n, float *__restrict__ w, bf16 *__restrict__ wb, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2, float eps,
                       float c1, float c2, int has_bf)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    m[t] = b1 * m[t] + (1.f - b1) * gt;
    v[t] = b2 * v[t] + (1.f - b2) * gt * gt;
    const float mh = m[t] / c1, vh = v[t] / c2;
    w[t] -= lr * mh / (sqrtf(vh) + eps);
    if (has_bf) wb[t] = bf16(w[t]);
}

__global__ void k_to_bf16(int n, const float *__restrict__ x, bf16 *__restrict__ y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = bf16(x[t]);
}

__global__ void k_gather(int B, int n_in, int n_out, const int *__restrict__ idx, const bf16 *__restrict__ X,
                         const float *__restrict__ T, const float *__restrict__ Mk, bf16 *__restrict__ Xb,
                         float *__restrict__ Tb, float *__restrict__ Mb)
{
    const int b = blockIdx.x;
    const int r = idx[b];
    for (int j = threadIdx.x; j < n_in; j += blockDim.x) Xb[(size_t)b * n_in + j] = X[(size_t)r * n_in + j];
    for (int j = threadIdx.x; j < n_out; j += blockDim.x) {
        Tb[(size_t)b * n_out + j] = T[(size_t)r * n_out + j];
        Mb[(size_t)b * n_out + j] = Mk[(size_t)r * n_out + j];
    }
}

// ---------------------------------------------------------------- inference
// The fused forward: one block of 8 waves takes 32 rows through every layer
// on v_mfma_f32_32x32x16_bf16 (BF16 operands, FP32 accumulation: the same
// arithmetic as the training forward), activations in LDS as BF16, weights
// read as 16-byte fragments of W^T from L2.  Hidden layers are 256 wide (each
// wave computes one 32-column tile, so a layer's serial chain per wave is 16
// products after one round of weight loads); the last layer's 32 padded
// columns are wave 0's.  MODE 0: rows of X in, outputs Y out.  MODE 1: the rows are
// (model, frequency) pairs; features are formed from the model (learned.py
// features()), and output 0 is decoded into the row's cut and residual-check
// count (learned.py decode_kcut, the low-frequency floor, row_cut).
constexpr int kMlpW = 256, kMlpKin = 192, kMlpNout = 32, kMlpRows = 32, kMlpMaxLayers = 8;
constexpr int kMlpThreads = 64 * kMlpW / 32;  // one wave per 32-column tile of a hidden layer
constexpr int kMlpLd = kMlpW + 8;  // LDS row stride (bf16): 16-byte row skew
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

struct MlpDev {
    const bf16 *wT[kMlpMaxLayers];
    const float *b[kMlpMaxLayers];
    int kp[kMlpMaxLayers], np[kMlpMaxLayers];
    int nl, n_in, n_out;
};

// Outputs of the network used by the grid: y[0] = log(kcut / s), then per
// pole cluster i < 8 y[1 + 2i] = centre / s, y[2 + 2i] = log(half width * 100).
constexpr int kLearnOut = 17;

__device__ __forceinline__ void learned_decode(const grtm::LearnModel &m, const grtm::LearnGrid &lg, int iw,
                                               const float *y, int &n, int &nc)
{
    const double w = 2.0 * M_PI * iw / (2.0 * lg.nw * lg.dt);
    const double dz = fmax(fabs(m.src - m.rcv), 0.5);
    const double s = fabs(w) / m.vs_min + 1.0 / dz;
    double kc = exp(double(y[0]) + lg.lmargin) * s;
    // rows solved in FP64 keep at least the analytic cut (models.analytic_kcut)
    if (iw < m.floor_rows) kc = fmax(kc, fmin(1.15 * w / m.vs_min + log(1e7) / dz, lg.kmax));
    kc = fmin(kc, lg.kmax);
    // the residual check covers the window past the cut and reaches every
    // predicted pole cluster that starts beyond it
    double reach = kc + m.window;
    for (int i = 0; i < 8; i++) {
        const double c = double(y[1 + 2 * i]) * s, hw = exp(double(y[2 + 2 * i])) / 100.0;
        if (c - hw > kc + m.window && c - hw < lg.kmax) reach = fmax(reach, c + hw);
    }
    // row_cut (host.hip): check cell j at n + stride j + stride / 2 < nk; a
    // cut row without a check cell stays dense
    n = (int)fmin((double)lg.nk, fmax(1.0, ceil(kc / lg.dk)));
    const int room = lg.nk - n;
    const int fit = room > lg.stride / 2 ? (room - lg.stride / 2 - 1) / lg.stride + 1 : 0;
    nc = max(0, min(fit, (int)ceil((reach - kc) / (lg.dk * lg.stride))));
    if (room > 0 && nc == 0) n = lg.nk;
}

template <int MODE>
__global__ __launch_bounds__(kMlpThreads) void k_mlp_infer(MlpDev net, int rows, const bf16 *__restrict__ X,
                                                   float *__restrict__ Y, const grtm::LearnModel *__restrict__ models,
                                                   grtm::LearnGrid lg, int *__restrict__ nk_pred,
                                                   int *__restrict__ nchk)
{
    constexpr int NF = 6 * grtm::kLearnSlots;  // layer features of a model
    __shared__ __attribute__((aligned(16))) bf16 act[2][kMlpRows][kMlpLd];
    __shared__ float mf[MODE ? kMlpRows + 1 : 1][MODE ? NF : 1];  // the block's models' layer features
    __shared__ float ysh[MODE ? kMlpRows : 1][MODE ? kLearnOut : 1];  // the rows' grid outputs
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r0 = blockIdx.x * kMlpRows;
    const int li = lane & 31, kq = (lane >> 5) * 8;
    // layer l's weight fragments of this wave (16-byte pieces of W^T: its 32
    // output columns of a hidden layer, wave 0 the last layer's 32 padded
    // ones) and the lane's bias.  The next layer's are requested before this
    // layer's products, so only layer 0's load latency is exposed, and it
    // overlaps the inputs.
    auto wload = [&](int l, v8bf(&b)[kMlpW / 16], float &bias) {
        const bool last = l == net.nl - 1;
        if (last && wv != 0) return;
        const int kp = l == 0 ? kMlpKin : kMlpW, n0 = last ? 0 : wv * 32;
        const bf16 *__restrict__ W = net.wT[l];
#pragma unroll
        for (int s = 0; s < kMlpW / 16; s++)
            if (s < kp / 16) b[s] = *reinterpret_cast<const v8bf *>(W + (size_t)(n0 + li) * kp + 16 * s + kq);
        bias = net.b[l][n0 + li];
    };
    v8bf wa[kMlpW / 16], wb[kMlpW / 16];
    float ba = 0.f, bb = 0.f;
    wload(0, wa, ba);
    int m_first = 0;
    if constexpr (MODE == 1) {
        // learned.py features(): (h / 10, vp / 8, vs / 4.5, rho / 3.3,
        // log10 qp / 3, log10 qs / 3) per layer slot, zero beyond the model's
        // layers, formed in FP64 and rounded to FP32 like the host's; once
        // per model of the block's rows
        m_first = r0 / lg.nw;
        const int m_last = min(rows - 1, r0 + kMlpRows - 1) / lg.nw;
        for (int e = tid; e < (m_last - m_first + 1) * NF; e += kMlpThreads) {
            const int mm = e / NF, k = e % NF, j = k / 6, c = k % 6;
            const grtm::LearnModel &md = models[m_first + mm];
            float v = 0.f;
            if (j < md.n) {
                const double x = md.rows[j][c];
                const double sc[6] = {10.0, 8.0, 4.5, 3.3, 3.0, 3.0};
                v = float(c >= 4 ? log10(x) / sc[c] : x / sc[c]);
            }
            mf[mm][k] = v;
        }
        __syncthreads();
    }
    // input rows, zero padded to kMlpKin
    for (int e = tid; e < kMlpRows * kMlpKin; e += kMlpThreads) {
        const int i = e / kMlpKin, k = e % kMlpKin, row = r0 + i;
        float v = 0.f;
        if (row < rows && k < net.n_in) {
            if constexpr (MODE == 0) {
                v = float(X[(size_t)row * net.n_in + k]);
            } else {
                const int m = row / lg.nw, iw = row - m * lg.nw;
                if (k < NF) v = mf[m - m_first][k];
                else if (k == NF) v = float(models[m].src / 40.0);
                else if (k == NF + 1) v = float(2.0 * M_PI * iw / (2.0 * lg.nw * lg.dt) / (2.0 * M_PI * 10.0));
            }
        }
        act[0][i][k] = bf16(v);
    }
    __syncthreads();
    int cur = 0;
    // one layer: this wave's products over the layer's K (kMlpKin in, kMlpW
    // hidden, kMlpNout out), then bias, ReLU and the BF16 activations of the
    // next layer (the last layer's outputs to Y or the grid decode)
    auto layer = [&](int l, const v8bf(&b)[kMlpW / 16], float bias) {
        const bool last = l == net.nl - 1;
        const int kp = l == 0 ? kMlpKin : kMlpW, n0 = last ? 0 : wv * 32;
        v16f acc;
        for (int r = 0; r < 16; r++) acc[r] = 0.f;
        if (!last || wv == 0) {
#pragma unroll
            for (int s = 0; s < kMlpW / 16; s++)
                if (s < kp / 16) {
                    const v8bf a = *reinterpret_cast<const v8bf *>(&act[cur][li][16 * s + kq]);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[s], acc, 0, 0, 0);
                }
            // accumulator r of lane l: row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
            for (int r = 0; r < 16; r++) {
                const int row = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int col = n0 + li;
                const float z = acc[r] + bias;
                if (!last) {
                    act[cur ^ 1][row][col] = bf16(fmaxf(z, 0.f));
                } else if (r0 + row < rows) {
                    if constexpr (MODE == 0) {
                        if (col < net.n_out) Y[(size_t)(r0 + row) * net.n_out + col] = z;
                    } else if (col < kLearnOut) {
                        ysh[row][col] = z;
                    }
                }
            }
        }
        __syncthreads();
        cur ^= 1;
    };
    for (int l = 0; l < net.nl; l += 2) {
        if (l + 1 < net.nl) wload(l + 1, wb, bb);
        layer(l, wa, ba);
        if (l + 1 < net.nl) {
            if (l + 2 < net.nl) wload(l + 2, wa, ba);
            layer(l + 1, wb, bb);
        }
    }
    if constexpr (MODE == 1) {
        // one thread per row: its cut and check cells from all 17 outputs
        if (tid < kMlpRows && r0 + tid < rows) {
            const int g = r0 + tid, m = g / lg.nw, iw = g - m * lg.nw;
            if (!lg.low_fixed || iw >= models[m].floor_rows) {
                int n, nc;
                learned_decode(models[m], lg, iw, ysh[tid], n, nc);
                nk_pred[g] = n;
                nchk[g] = nc;
            }
            if (Y) {
                // the pole clusters in km^-1 (centre, half width), as the decode reads them
                const double w = 2.0 * M_PI * iw / (2.0 * lg.nw * lg.dt);
                const double s = fabs(w) / models[m].vs_min + 1.0 / fmax(fabs(models[m].src - models[m].rcv), 0.5);
                for (int i = 0; i < 8; i++) {
                    Y[(size_t)g * 16 + 2 * i] = float(double(ysh[tid][1 + 2 * i]) * s);
                    Y[(size_t)g * 16 + 2 * i + 1] = float(exp(double(ysh[tid][2 + 2 * i])) / 100.0);
                }
            }
        }
    }
}

// W^T (bf16, padded) and padded biases of every layer from the FP32 master
// parameters.
__global__ void k_mlp_prep(int I, int O, int Kp, int Np, const float *__restrict__ w, const float *__restrict__ b,
                           bf16 *__restrict__ wT, float *__restrict__ bT)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < Kp * Np) {
        const int n = t / Kp, k = t % Kp;
        wT[t] = bf16(n < O && k < I ? w[(size_t)k * O + n] : 0.f);
    }
    if (t < Np) bT[t] = t < O ? b[t] : 0.f;
}

struct DB {
    void *p = nullptr;
    ~DB()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n)
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        return hipMalloc(&p, n ? n : 16) == hipSuccess ? GRTM_OK : GRTM_E_NOMEM;
    }
    template <typename T> T *as() const { return (T *)p; }
};

}  // namespace

struct grtm_mlp {
    int n_in = 0, width = 0, depth = 0, n_out = 0, bmax = 0;
    int nq = 0;     // leading outputs trained with the pinball loss
    float q = 0.5f; // its quantile
    std::vector<int> fin, fout;      // per layer
    std::vector<size_t> woff, boff;  // offsets into the flat parameter vector
    size_t nparam = 0;
    int step = 0;
    hipStream_t st = nullptr;
    DB w32, wbf, g, m, v;           // master params, bf16 weights, grads, Adam moments
    DB X, T, Mk;                    // resident training set
    int ndata = 0;
    DB Xb, Tb, Mb, idx, lossp;      // batch buffers
    std::vector<DB> Z, H, dH;       // per layer
    // inference copy for the fused matrix-core forward: per layer W^T as bf16
    // [Np][Kp] (K padded to 16, N to 32, zeros), biases [Np] in FP32
    DB wT, bT;
    bool fused = false;             // the shape fits the fused kernel
    ~grtm_mlp()
    {
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

int mlp_forward(grtm_mlp *n, int B, const bf16 *Xin)
{
    const bf16 *h = Xin;
    const int L = n->depth + 1;
    for (int l = 0; l < L; l++) {
        const int I = n->fin[l], O = n->fout[l];
        dim3 grid((O + 63) / 64, (B + 63) / 64);
        hipLaunchKernelGGL((k_mm<bf16, bf16>), grid, dim3(256), 0, n->st, B, O, I, h, (long)I, 1L,
                           n->wbf.as<bf16>() + n->woff[l], (long)O, 1L, n->Z[l].as<float>(), (long)O,
                           n->w32.as<float>() + n->boff[l], 0);
        if (l < L - 1) {
            const int cnt = B * O;
            hipLaunchKernelGGL(k_act, dim3((cnt + 255) / 256), dim3(256), 0, n->st, cnt, n->Z[l].as<float>(),
                               n->H[l].as<bf16>(), 1);
            h = n->H[l].as<bf16>();
        }
    }
    HIPCHK(hipGetLastError());
    return GRTM_OK;
}

// gradient of the masked MSE over a batch already in Xb/Tb/Mb; returns loss
int mlp_backward(grtm_mlp *n, int B, const bf16 *Xin, const float *Tb, const float *Mb, float *loss)
{
    int st;
    if ((st = mlp_forward(n, B, Xin))) return st;
    const int L = n->depth + 1;
    const int cnt = B * n->n_out;
    const int nb = (cnt + 255) / 256;
    float *dY = n->dH[L - 1].as<float>();
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, n->st, cnt, n->n_out, n->nq, n->q, 1.f / float(cnt),
                       n->Z[L - 1].as<float>(), Tb, Mb, dY, n->lossp.as<float>());
    for (int l = L - 1; l >= 0; l--) {
        const int I = n->fin[l], O = n->fout[l];
        float *dZ = n->dH[l].as<float>();
        const bf16 *hin = l == 0 ? Xin : n->H[l - 1].as<bf16>();
        float *gw = n->g.as<float>() + n->woff[l], *gb = n->g.as<float>() + n->boff[l];
        // dW = H_{l-1}^T dZ, db = colsum(dZ)
        dim3 gw_grid((O + 63) / 64, (I + 63) / 64);
        hipLaunchKernelGGL((k_mm<bf16, float>), gw_grid, dim3(256), 0, n->st, I, O, B, hin, 1L, (long)I, dZ, (long)O,
                           1L, gw, (long)O, nullptr, 0);
        hipLaunchKernelGGL(k_colsum, dim3((O + 255) / 256), dim3(256), 0, n->st, B, O, dZ, gb);
        if (l > 0) {
            // dH_{l-1} = dZ W^T, then through the ReLU
            float *dprev = n->dH[l - 1].as<float>();
            dim3 gd((I + 63) / 64, (B + 63) / 64);
            hipLaunchKernelGGL((k_mm<float, bf16>), gd, dim3(256), 0, n->st, B, I, O, dZ, (long)O, 1L,
                               n->wbf.as<bf16>() + n->woff[l], 1L, (long)O, dprev, (long)I, nullptr, 0);
            const int c2 = B * I;
            hipLaunchKernelGGL(k_relu_bwd, dim3((c2 + 255) / 256), dim3(256), 0, n->st, c2, n->Z[l - 1].as<float>(),
                               dprev);
        }
    }
    HIPCHK(hipGetLastError());
    if (loss) {
        std::vector<float> lp(nb);
        HIPCHK(hipMemcpyAsync(lp.data(), n->lossp.p, sizeof(float) * nb, hipMemcpyDeviceToHost, n->st));
        HIPCHK(hipStreamSynchronize(n->st));
        double s = 0;
        for (float x : lp) s += x;
        *loss = float(s);
    }
    return GRTM_OK;
}

int refresh_bf16(grtm_mlp *n)
{
    hipLaunchKernelGGL(k_to_bf16, dim3((n->nparam + 255) / 256), dim3(256), 0, n->st, (int)n->nparam,
                       n->w32.as<float>(), n->wbf.as<bf16>());
    HIPCHK(hipGetLastError());
    return GRTM_OK;
}

// Layout of the fused kernel's padded copy: K of layer 0 padded to kMlpKin,
// hidden layers kMlpW wide, the output padded to kMlpNout.
MlpDev mlp_dev(const grtm_mlp *n)
{
    MlpDev d{};
    const int L = n->depth + 1;
    size_t wo = 0, bo = 0;
    for (int l = 0; l < L; l++) {
        d.kp[l] = l == 0 ? kMlpKin : kMlpW;  // the fused kernel's compile-time shapes (zero padded)
        d.np[l] = l == L - 1 ? kMlpNout : kMlpW;
        d.wT[l] = n->wT.as<bf16>() + wo;
        d.b[l] = n->bT.as<float>() + bo;
        wo += (size_t)d.kp[l] * d.np[l];
        bo += d.np[l];
    }
    d.nl = L, d.n_in = n->fin[0], d.n_out = n->n_out;
    return d;
}

// The fused inference copy of the current parameters (after every change).
int refresh_infer(grtm_mlp *n)
{
    if (!n->fused) return GRTM_OK;
    const MlpDev d = mlp_dev(n);
    for (int l = 0; l < d.nl; l++) {
        const int cnt = d.kp[l] * d.np[l];
        hipLaunchKernelGGL(k_mlp_prep, dim3((cnt + 255) / 256), dim3(256), 0, n->st, n->fin[l], n->fout[l], d.kp[l],
                           d.np[l], n->w32.as<float>() + n->woff[l], n->w32.as<float>() + n->boff[l],
                           const_cast<bf16 *>(d.wT[l]), const_cast<float *>(d.b[l]));
    }
    HIPCHK(hipGetLastError());
    return GRTM_OK;
}

}  // namespace

extern "C" {

int grtm_mlp_create(int n_in, int width, int depth, int n_out, int max_batch, grtm_mlp **out)
{
    if (!out || n_in < 1 || width < 1 || depth < 1 || n_out < 1 || max_batch < 1) return GRTM_E_INPUT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return GRTM_E_DEVICE;
    grtm_mlp *n = new grtm_mlp;
    n->n_in = n_in, n->width = width, n->depth = depth, n->n_out = n_out, n->bmax = max_batch;
    const int L = depth + 1;
    size_t off = 0;
    for (int l = 0; l < L; l++) {
        const int I = l == 0 ? n_in : width, O = l == L - 1 ? n_out : width;
        n->fin.push_back(I);
        n->fout.push_back(O);
        n->woff.push_back(off);
        off += (size_t)I * O;
        n->boff.push_back(off);
        off += O;
    }
    n->nparam = off;
    int st = GRTM_OK;
    if (hipStreamCreate(&n->st) != hipSuccess) st = GRTM_E_DEVICE;
    if (!st) st = n->w32.alloc(sizeof(float) * off);
    if (!st) st = n->wbf.alloc(sizeof(bf16) * off);
    if (!st) st = n->g.alloc(sizeof(float) * off);
    if (!st) st = n->m.alloc(sizeof(float) * off);
    if (!st) st = n->v.alloc(sizeof(float) * off);
    if (!st) st = n->Xb.alloc(sizeof(bf16) * (size_t)max_batch * n_in);
    if (!st) st = n->Tb.alloc(sizeof(float) * (size_t)max_batch * n_out);
    if (!st) st = n->Mb.alloc(sizeof(float) * (size_t)max_batch * n_out);
    if (!st) st = n->idx.alloc(sizeof(int) * max_batch);
    if (!st) st = n->lossp.alloc(sizeof(float) * ((size_t)max_batch * n_out / 256 + 2));
    n->Z.resize(L), n->H.resize(L), n->dH.resize(L);
    for (int l = 0; l < L && !st; l++) {
        st = n->Z[l].alloc(sizeof(float) * (size_t)max_batch * n->fout[l]);
        if (!st) st = n->H[l].alloc(sizeof(bf16) * (size_t)max_batch * n->fout[l]);
        if (!st) st = n->dH[l].alloc(sizeof(float) * (size_t)max_batch * n->fout[l]);
    }
    if (!st && (hipMemset(n->m.p, 0, sizeof(float) * off) != hipSuccess ||
                hipMemset(n->v.p, 0, sizeof(float) * off) != hipSuccess))
        st = GRTM_E_DEVICE;
    // the fused matrix-core forward: inputs up to kMlpKin, hidden width kMlpW,
    // outputs up to kMlpNout, at most kMlpMaxLayers layers
    n->fused = n_in <= kMlpKin && width == kMlpW && n_out <= kMlpNout && L <= kMlpMaxLayers;
    if (!st && n->fused) {
        const MlpDev d = mlp_dev(n);
        size_t wn = 0, bn = 0;
        for (int l = 0; l < L; l++) wn += (size_t)d.kp[l] * d.np[l], bn += d.np[l];
        st = n->wT.alloc(sizeof(bf16) * wn);
        if (!st) st = n->bT.alloc(sizeof(float) * bn);
    }
    if (st) {
        delete n;
        return st;
    }
    *out = n;
    return GRTM_OK;
}

void grtm_mlp_destroy(grtm_mlp *n) { delete n; }

int grtm_mlp_set_loss(grtm_mlp *n, int n_quantile, float q)
{
    if (!n || n_quantile < 0 || n_quantile > n->n_out || !(q > 0.f && q < 1.f)) return GRTM_E_INPUT;
    n->nq = n_quantile;
    n->q = q;
    return GRTM_OK;
}

long grtm_mlp_nparams(grtm_mlp *n) { return n ? (long)n->nparam : 0; }

int grtm_mlp_set_params(grtm_mlp *n, const float *p)
{
    if (!n || !p) return GRTM_E_INPUT;
    HIPCHK(hipMemcpy(n->w32.p, p, sizeof(float) * n->nparam, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(n->m.p, 0, sizeof(float) * n->nparam));
    HIPCHK(hipMemset(n->v.p, 0, sizeof(float) * n->nparam));
    n->step = 0;
    int st = refresh_bf16(n);
    if (!st) st = refresh_infer(n);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(n->st));
    return GRTM_OK;
}

int grtm_mlp_get_params(grtm_mlp *n, float *p)
{
    if (!n || !p) return GRTM_E_INPUT;
    HIPCHK(hipStreamSynchronize(n->st));
    HIPCHK(hipMemcpy(p, n->w32.p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
    return GRTM_OK;
}

int grtm_mlp_set_data(grtm_mlp *n, const float *X, const float *T, const float *Mk, int ndata)
{
    if (!n || !X || !T || !Mk || ndata < 1) return GRTM_E_INPUT;
    int st;
    std::vector<bf16> xb((size_t)ndata * n->n_in);
    for (size_t i = 0; i < xb.size(); i++) xb[i] = bf16(X[i]);
    if ((st = n->X.alloc(sizeof(bf16) * xb.size())) || (st = n->T.alloc(sizeof(float) * (size_t)ndata * n->n_out)) ||
        (st = n->Mk.alloc(sizeof(float) * (size_t)ndata * n->n_out)))
        return st;
    HIPCHK(hipMemcpy(n->X.p, xb.data(), sizeof(bf16) * xb.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(n->T.p, T, sizeof(float) * (size_t)ndata * n->n_out, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(n->Mk.p, Mk, sizeof(float) * (size_t)ndata * n->n_out, hipMemcpyHostToDevice));
    n->ndata = ndata;
    return GRTM_OK;
}

// Adam steps on batches of the resident training set; batch rows chosen by a
// splitmix64 stream from `seed`.  losses[s] (optional) = batch loss of step s.
int grtm_mlp_train(grtm_mlp *n, int steps, int batch, float lr, unsigned long long seed, float *losses)
{
    if (!n || steps < 1 || batch < 1 || batch > n->bmax || n->ndata < 1 || !(lr > 0)) return GRTM_E_INPUT;
    std::vector<int> ib((size_t)steps * batch);
    unsigned long long s = seed;
    auto next = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        unsigned long long z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (auto &x : ib) x = int(next() % (unsigned long long)n->ndata);
    DB didx;  // every step's batch rows, uploaded once
    int st0 = didx.alloc(sizeof(int) * ib.size());
    if (st0) return st0;
    HIPCHK(hipMemcpy(didx.p, ib.data(), sizeof(int) * ib.size(), hipMemcpyHostToDevice));
    for (int it = 0; it < steps; it++) {
        hipLaunchKernelGGL(k_gather, dim3(batch), dim3(64), 0, n->st, batch, n->n_in, n->n_out,
                           didx.as<int>() + (size_t)it * batch,
                           n->X.as<bf16>(), n->T.as<float>(), n->Mk.as<float>(), n->Xb.as<bf16>(), n->Tb.as<float>(),
                           n->Mb.as<float>());
        float l = 0;
        int st = mlp_backward(n, batch, n->Xb.as<bf16>(), n->Tb.as<float>(), n->Mb.as<float>(), losses ? &l : nullptr);
        if (st) return st;
        if (losses) losses[it] = l;
        n->step++;
        const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
        hipLaunchKernelGGL(k_adam, dim3((n->nparam + 255) / 256), dim3(256), 0, n->st, (int)n->nparam,
                           n->w32.as<float>(), n->wbf.as<bf16>(), n->g.as<float>(), n->m.as<float>(),
                           n->v.as<float>(), lr, b1, b2, eps, c1, c2, 1);
        HIPCHK(hipGetLastError());
    }
    int st = refresh_infer(n);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(n->st));
    return GRTM_OK;
}

// Gradient of the masked MSE on the given samples (n <= max batch), for checks.
int grtm_mlp_grads(grtm_mlp *n, const float *X, const float *T, const float *Mk, int cnt, float *grads, float *loss)
{
    if (!n || !X || !T || !Mk || !grads || cnt < 1 || cnt > n->bmax) return GRTM_E_INPUT;
    std::vector<bf16> xb((size_t)cnt * n->n_in);
    for (size_t i = 0; i < xb.size(); i++) xb[i] = bf16(X[i]);
    HIPCHK(hipMemcpy(n->Xb.p, xb.data(), sizeof(bf16) * xb.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(n->Tb.p, T, sizeof(float) * (size_t)cnt * n->n_out, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(n->Mb.p, Mk, sizeof(float) * (size_t)cnt * n->n_out, hipMemcpyHostToDevice));
    int st = mlp_backward(n, cnt, n->Xb.as<bf16>(), n->Tb.as<float>(), n->Mb.as<float>(), loss);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(n->st));
    HIPCHK(hipMemcpy(grads, n->g.p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
    return GRTM_OK;
}

// fused: the matrix-core forward (k_mlp_infer); otherwise the training
// forward (k_mm).  Both compute BF16 products with FP32 accumulation.
static int mlp_predict(grtm_mlp *n, const float *X, int cnt, float *Y, bool fused)
{
    if (!n || !X || !Y || cnt < 1) return GRTM_E_INPUT;
    if (fused && !n->fused) return GRTM_E_UNSUPPORTED;
    for (int b0 = 0; b0 < cnt; b0 += n->bmax) {
        const int B = std::min(n->bmax, cnt - b0);
        std::vector<bf16> xb((size_t)B * n->n_in);
        for (size_t i = 0; i < xb.size(); i++) xb[i] = bf16(X[(size_t)b0 * n->n_in + i]);
        HIPCHK(hipMemcpy(n->Xb.p, xb.data(), sizeof(bf16) * xb.size(), hipMemcpyHostToDevice));
        const float *src = n->Z[n->depth].as<float>();
        if (fused) {
            hipLaunchKernelGGL((k_mlp_infer<0>), dim3((B + kMlpRows - 1) / kMlpRows), dim3(kMlpThreads), 0, n->st, mlp_dev(n),
                               B, n->Xb.as<bf16>(), n->Z[n->depth].as<float>(), nullptr, grtm::LearnGrid{}, nullptr,
                               nullptr);
            HIPCHK(hipGetLastError());
        } else {
            int st = mlp_forward(n, B, n->Xb.as<bf16>());
            if (st) return st;
        }
        HIPCHK(hipStreamSynchronize(n->st));
        HIPCHK(hipMemcpy(Y + (size_t)b0 * n->n_out, src, sizeof(float) * (size_t)B * n->n_out, hipMemcpyDeviceToHost));
    }
    return GRTM_OK;
}

int grtm_mlp_predict(grtm_mlp *n, const float *X, int cnt, float *Y)
{
    return mlp_predict(n, X, cnt, Y, n && n->fused);
}

int grtm_mlp_predict_train_forward(grtm_mlp *n, const float *X, int cnt, float *Y)
{
    return mlp_predict(n, X, cnt, Y, false);
}

}  // extern "C"

namespace grtm {
int learned_grid(grtm_mlp *net, const LearnModel *models, int n_models, const LearnGrid &lg, int *nk_pred,
                 int *nchk, hipStream_t st, float *poles)
{
    if (!net || !net->fused || net->fin[0] != 6 * kLearnSlots + 2 || net->n_out != kLearnOut || n_models < 1)
        return GRTM_E_UNSUPPORTED;
    const int rows = n_models * lg.nw;
    hipLaunchKernelGGL((k_mlp_infer<1>), dim3((rows + kMlpRows - 1) / kMlpRows), dim3(kMlpThreads), 0, st, mlp_dev(net),
                       rows,
                       nullptr, poles, models, lg, nk_pred, nchk);
    HIPCHK(hipGetLastError());
    return GRTM_OK;
}
}  // namespace grtm
