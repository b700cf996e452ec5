"""Learned wavenumber-grid predictor.

An MLP (182 -> 4 x 256 -> 17, BF16 weights, FP32 accumulation; kernels in
csrc/mlp.hip) maps (layered model, source depth, frequency) to the
truncation wavenumber of that frequency row and up to 8 pole-cluster centres
and widths of its k-kernel.  The prediction sets the per-row cut of the
adaptive grid (TracePlan.set_adaptive); the residual check and dense fallback
on the GPU guard it.

Inputs (182): 30 layer slots x (thickness/10, vp/8, vs/4.5, rho/3.3,
log10 qp/3, log10 qs/3), zero-padded, then source depth/40 and w/(2 pi 10).
Outputs (17): y0 = log(kcut / s), s = |w| / Vs_min + 1/dz (dz = |z_s - z_r|,
at least 0.5 km), trained with the pinball loss of a high quantile so it is a
per-row conservative bound on the cut; then 8 x (centre / s,
log(half width * 100)) for the pole clusters ordered by wavenumber, masked
when absent (squared error).
"""
import ctypes
import math
import os

import numpy as np

from .ffi import P, _lib, check

N_SLOTS = 30
N_IN = N_SLOTS * 6 + 2
N_OUT = 17
WIDTH = 256
DEPTH = 4
W_REF = 2 * math.pi * 10.0
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "predictor_weights.npz")


def _scale(layers, src_depth, rcv_depth, w):
    vs_min = float(np.min(layers[:, 2]))
    dz = max(abs(src_depth - rcv_depth), 0.5)
    return np.abs(w) / vs_min + 1.0 / dz


def features(layers, src_depth, w):
    """Input rows for frequencies w (array) of one model: [len(w), 182]."""
    layers = np.asarray(layers, dtype=np.float64)
    if layers.shape[0] > N_SLOTS:
        raise ValueError(f"at most {N_SLOTS} layers")
    f = np.zeros((len(w), N_IN), dtype=np.float32)
    for j, (h, vp, vs, rho, qp, qs) in enumerate(layers):
        f[:, 6 * j:6 * j + 6] = (h / 10.0, vp / 8.0, vs / 4.5, rho / 3.3, math.log10(qp) / 3.0, math.log10(qs) / 3.0)
    f[:, 180] = src_depth / 40.0
    f[:, 181] = np.asarray(w) / W_REF
    return f


def targets(labels, layers, src_depth, rcv_depth, w, dk):
    """Targets and mask [len(w), 17] from row labels (grtm_plan_row_labels)."""
    s = _scale(layers, src_depth, rcv_depth, w)
    t = np.zeros((len(w), N_OUT), dtype=np.float32)
    m = np.zeros((len(w), N_OUT), dtype=np.float32)
    kcut = np.maximum(labels[:, 1], 1.0) * dk
    t[:, 0] = np.log(kcut / s)
    m[:, 0] = 1.0
    for r in range(len(w)):
        pk = [(labels[r, 2 + 3 * i] * dk, labels[r, 3 + 3 * i] * dk) for i in range(8) if labels[r, 2 + 3 * i] >= 0]
        pk.sort()
        for i, (c, hw) in enumerate(pk):
            t[r, 1 + 2 * i] = c / s[r]
            t[r, 2 + 2 * i] = math.log(max(hw, 1e-6) * 100.0)
            m[r, 1 + 2 * i] = m[r, 2 + 2 * i] = 1.0
    return t, m


def decode_kcut(y, layers, src_depth, rcv_depth, w, lmargin=0.0):
    """Per-row truncation wavenumber from network outputs [len(w), 17];
    lmargin is the calibrated log-margin added to the cut output."""
    s = _scale(layers, src_depth, rcv_depth, w)
    return np.exp(y[:, 0] + lmargin) * s


class MLP:
    """GPU multilayer perceptron (libgrtm grtm_mlp_*)."""

    def __init__(self, n_in=N_IN, width=WIDTH, depth=DEPTH, n_out=N_OUT, max_batch=1024):
        self.ptr = ctypes.c_void_p()
        check(_lib().grtm_mlp_create(n_in, width, depth, n_out, max_batch, ctypes.byref(self.ptr)), "grtm_mlp_create")
        self.n_in, self.width, self.depth, self.n_out, self.max_batch = n_in, width, depth, n_out, max_batch
        self.nparams = int(_lib().grtm_mlp_nparams(self.ptr))

    def shapes(self):
        dims = [self.n_in] + [self.width] * self.depth + [self.n_out]
        return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]

    def init(self, seed=0):
        """He-normal weights, zero biases."""
        rng = np.random.default_rng(seed)
        parts = []
        for fi, fo in self.shapes():
            parts.append((rng.standard_normal((fi, fo)) * math.sqrt(2.0 / fi)).astype(np.float32).ravel())
            parts.append(np.zeros(fo, dtype=np.float32))
        self.set_params(np.concatenate(parts))

    def set_loss(self, n_quantile, q):
        check(_lib().grtm_mlp_set_loss(self.ptr, int(n_quantile), float(q)), "grtm_mlp_set_loss")

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        check(_lib().grtm_mlp_set_params(self.ptr, p.ctypes.data_as(P)), "grtm_mlp_set_params")

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        check(_lib().grtm_mlp_get_params(self.ptr, p.ctypes.data_as(P)), "grtm_mlp_get_params")
        return p

    def set_data(self, X, T, M):
        X, T, M = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, T, M))
        check(_lib().grtm_mlp_set_data(self.ptr, X.ctypes.data_as(P), T.ctypes.data_as(P), M.ctypes.data_as(P),
                                       X.shape[0]), "grtm_mlp_set_data")

    def train(self, steps, batch=512, lr=1e-3, seed=1):
        losses = np.zeros(steps, dtype=np.float32)
        check(_lib().grtm_mlp_train(self.ptr, int(steps), int(batch), float(lr), int(seed), losses.ctypes.data_as(P)),
              "grtm_mlp_train")
        return losses

    def grads(self, X, T, M):
        X, T, M = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, T, M))
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = ctypes.c_float()
        check(_lib().grtm_mlp_grads(self.ptr, X.ctypes.data_as(P), T.ctypes.data_as(P), M.ctypes.data_as(P),
                                    X.shape[0], g.ctypes.data_as(P), ctypes.byref(loss)), "grtm_mlp_grads")
        return g, loss.value

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], self.n_out), dtype=np.float32)
        check(_lib().grtm_mlp_predict(self.ptr, X.ctypes.data_as(P), X.shape[0], Y.ctypes.data_as(P)),
              "grtm_mlp_predict")
        return Y

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            _lib().grtm_mlp_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()


class Predictor:
    """Trained network + decoding: per-row truncation wavenumbers of a grid."""

    def __init__(self, path=WEIGHTS):
        data = np.load(path)
        self.mlp = MLP()
        self.mlp.set_params(data["params"])
        self.lmargin = float(data["lmargin"]) if "lmargin" in data else 0.0

    def kcut(self, layers, src_depth, rcv_depth, grid, floor_rows=None, low_fixed=True):
        """Per-row truncation wavenumbers of the grid.  Rows [0, floor_rows)
        (default: the plans' FP64 rows, fp64_rows) take the analytic cut, as
        the mixed-precision TracePlan's device grid does (low_fixed), or keep
        at least it, as batches and FP64 plans do (low_fixed=False)."""
        nw, dt = grid["nw"], grid["dt"]
        w = 2 * math.pi * np.arange(nw) / (2 * nw * dt)
        y = self.mlp.predict(features(layers, src_depth, w))
        kc = decode_kcut(y, layers, src_depth, rcv_depth, w, self.lmargin)
        # Rows solved in FP64 take (or keep at least) the analytic cut: they
        # are few and their cuts short, while a failed check sends such a row
        # to a dense FP64 completion, a latency-bound launch; taken without
        # the network, their grid lets their launch run beside it
        from grtm_cuda_amd import models
        if floor_rows is None:
            floor_rows = fp64_rows(layers, src_depth, rcv_depth, grid)
        low = np.arange(nw) < floor_rows
        ak = models.analytic_kcut(layers, src_depth, rcv_depth, grid)
        kc = np.where(low, ak if low_fixed else np.maximum(kc, ak), kc)
        return np.minimum(kc, grid["kmax"])

    def attach(self, plan, window=None, stride=8, tol=3e-6):
        """Give `plan` (TracePlan) the learned grid formed on the GPU at every
        execution by this predictor's network."""
        plan.set_learned(self, window=window, stride=stride, tol=tol)


def fp64_rows(layers, src_depth, rcv_depth, grid):
    """The mixed-precision plans' default FP64 rows: |w + i w_I| below
    2.5 Vs_min / max(|z_s - z_r|, 1e-3), at least 4 (host.hip model_tables)."""
    nw, dt = grid["nw"], grid["dt"]
    T = 2 * nw * dt
    oi = grid.get("omega_i", 0.0)
    oi = oi if oi > 0 else math.pi / T
    dz = max(abs(src_depth - rcv_depth), 1e-3)
    wc = 2.5 * float(np.min(np.asarray(layers)[:, 2])) / dz
    r = 0
    while r < nw and math.hypot(2 * math.pi * r / T, oi) < wc:
        r += 1
    return min(nw, max(r, 4))
