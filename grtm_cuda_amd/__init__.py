"""grtm_cuda_amd -- MI355X (gfx950) Green's functions for layered media.

Host-side mirror of the reference's interface (cangyeone/grtm_cuda):

* ``green(iname, oname_w, oname_t)`` -- the reference's only entry point
  (cusrc/popm/popm.h:159-161): reads input.txt, writes the spectrum and trace
  files, computed on the GPU.
* ``green_mem(iname)`` -- the same run returning arrays.
* ``TracePlan`` -- one model, many receiver distances, explosion and
  vertical-force Z/R traces from one recursion (the production path).
* ``BatchPlan`` -- many models (uneven layer counts, source depths) on one
  grid, replayed from one captured HIP graph.

All compute runs in libgrtm.so (HIP).  There is no CPU fallback.
"""
import ctypes

import numpy as np

from . import inputfile, models  # noqa: F401
from .ffi import (GRTM_OK, PREC_BF16X3, PREC_BF16X6, PREC_FP16X3, PREC_FP32, PREC_FP64, EXPORTS, GrtmError, Grid, Job, Model, D, P,  # noqa: F401
                   _lib, check)

__all__ = ["green", "green_mem", "TracePlan", "BatchPlan", "GrtmError", "PREC_FP64", "PREC_FP32", "PREC_BF16X3", "PREC_BF16X6", "PREC_FP16X3", "models",
           "inputfile"]


def green(iname, oname_w, oname_t, compat=True):
    """Run the drop-in path: read ``iname`` (input.txt format), write
    ``oname_w`` (N rows x 30 complex) and ``oname_t`` (N rows x 30 reals).
    Raises GrtmError on failure (the reference returns void and ignores
    errors; the library's ``green`` symbol keeps that signature)."""
    check(_lib().grtm_green(iname.encode(), oname_w.encode(), oname_t.encode(), int(bool(compat))), "green")


def green_mem(iname, compat=True, noalloc=False):
    """Returns (spec, traces): spec[30, nw] complex post-processed spectrum
    (rows 0..nw-1 of oname_w), traces[30, N] (columns of oname_t).
    noalloc: run through grtm_green_mem_noalloc (no device allocation; the
    library's fixed per-GPU scratch)."""
    lib = _lib()
    nw = ctypes.c_int(0)
    check(lib.grtm_green_mem(iname.encode(), int(bool(compat)), None, None, ctypes.byref(nw)), "green_mem")
    n = nw.value
    spec = np.zeros((30, n, 2))
    tr = np.zeros((30, 2 * n))
    if noalloc:
        check(lib.grtm_green_mem_noalloc(iname.encode(), int(bool(compat)), spec.ctypes.data_as(P),
                                         tr.ctypes.data_as(P), ctypes.byref(nw), None), "green_mem_noalloc")
    else:
        check(lib.grtm_green_mem(iname.encode(), int(bool(compat)), spec.ctypes.data_as(P), tr.ctypes.data_as(P),
                                 ctypes.byref(nw)), "green_mem")
    return spec[..., 0] + 1j * spec[..., 1], tr


class _Ctx:
    def __init__(self, device=0):
        self.ptr = ctypes.c_void_p()
        check(_lib().grtm_create(ctypes.byref(self.ptr), int(device)), "grtm_create")

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            _lib().grtm_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()


_CTX = {}


def _ctx(device=0):
    if device not in _CTX:
        _CTX[device] = _Ctx(device)
    return _CTX[device]


class TracePlan:
    """Explosion and vertical-force Z/R traces of one layered model at many
    distances.  ``layers`` rows: (thickness km, vp, vs, rho, qp, qs), last row
    the half-space.  ``grid``: dict with nw, dt, nk, kmax, omega_i, t0, taper."""

    _CREATE = "grtm_plan_traces"

    def __init__(self, layers, src_depth, rcv_depth, dists, grid, precision=PREC_FP64, compat=False, device=0):
        self._ctx = _ctx(device)
        self.src_depth, self.rcv_depth = float(src_depth), float(rcv_depth)
        self.layers = np.ascontiguousarray(layers, dtype=np.float64)
        self.dists = np.ascontiguousarray(dists, dtype=np.float64)
        self.grid = dict(grid)
        self.nw = int(grid["nw"])
        self.N = 2 * self.nw
        self.nk = int(grid["nk"])
        self.nr = len(self.dists)
        self.precision = precision
        g = Grid(self.nw, float(grid["dt"]), self.nk, float(grid["kmax"]), float(grid.get("omega_i", 0.0)),
                 float(grid.get("t0", 0.0)), float(grid.get("taper", 0.0)))
        j = Job(self.layers.shape[0], self.layers.ctypes.data_as(P), float(src_depth), float(rcv_depth), self.nr,
                self.dists.ctypes.data_as(P))
        self.ptr = ctypes.c_void_p()
        check(getattr(_lib(), self._CREATE)(self._ctx.ptr, ctypes.byref(j), ctypes.byref(g), int(precision),
                                            int(bool(compat)), *self._extra_args(), ctypes.byref(self.ptr)),
              self._CREATE)

    def _extra_args(self):
        return ()

    def execute(self):
        check(_lib().grtm_plan_execute(self.ptr, None), "grtm_plan_execute")

    def sync(self):
        check(_lib().grtm_plan_sync(self.ptr), "grtm_plan_sync")

    def traces(self):
        """Run and return traces[2 (explosion, vertical force), 2 (Z, R), nr, N]."""
        self.execute()
        out = np.zeros((2, 2, self.nr, self.N))
        check(_lib().grtm_plan_traces_out(self.ptr, out.ctypes.data_as(P)), "grtm_plan_traces_out")
        return out

    def kernels(self, q, w_lo, w_hi):
        """Unweighted k-kernels of source q for frequency rows [w_lo, w_hi):
        array [w_hi - w_lo, nk, 2 (Z, R)] complex (after an execute())."""
        out = np.zeros((w_hi - w_lo, self.nk, 4))
        check(_lib().grtm_plan_kernels_out(self.ptr, int(q), int(w_lo), int(w_hi), out.ctypes.data_as(P)),
              "grtm_plan_kernels_out")
        return out[..., 0::2] + 1j * out[..., 1::2]

    def timing(self, reps=5):
        """ms per execution: total, recursion stage, contraction, IFFT, Bessel
        tables, the main recursion launch alone (``main``), and a learned
        grid's formation on the device (``predictor``, part of the recursion
        stage)."""
        ms = np.zeros(7)
        check(_lib().grtm_plan_timing(self.ptr, int(reps), ms.ctypes.data_as(P)), "grtm_plan_timing")
        return dict(total=ms[0], cells=ms[1], ksum=ms[2], ifft=ms[3], bessel=ms[4], main=ms[5], predictor=ms[6])

    def work(self):
        w = np.zeros(11)
        check(_lib().grtm_plan_work(self.ptr, w.ctypes.data_as(P)), "grtm_plan_work")
        return dict(cells=w[0], cell_layers=w[1], ksum_flop=w[2], ifft_bytes=w[3], traces=w[4],
                    flops_per_cell=w[5], trans_per_cell=w[6], rows_fp64=int(w[7]), cells_refixed=int(w[8]),
                    main_cells=w[9], main_flops_per_cell=w[10])

    def set_rows64(self, rows64):
        """Mixed-precision plans: solve frequency rows [0, rows64) in FP64."""
        check(_lib().grtm_plan_set_rows64(self.ptr, int(rows64)), "grtm_plan_set_rows64")

    def set_ptam(self, n_ptam=10, tail=0):
        """Continue every k-sum beyond the dense grid by peak-trough averaging
        of its partial sums at the first n_ptam extrema (tail: wavenumbers
        solved beyond the grid, 0 = from the nearest receiver); n_ptam=0 removes it."""
        check(_lib().grtm_plan_set_ptam(self.ptr, int(n_ptam), int(tail)), "grtm_plan_set_ptam")

    def ptam_stats(self):
        o = np.zeros(4)
        check(_lib().grtm_plan_ptam_stats(self.ptr, o.ctypes.data_as(P)), "grtm_plan_ptam_stats")
        return dict(n_ptam=int(o[0]), tail=int(o[1]), extended=int(o[2]), sums=int(o[3]))

    def plant_nonfinite(self, cells):
        """Test hook: give the FP32 cells [(row, k), ...] non-finite operands at
        every execution, so that the FP64 re-solve of overflowed cells takes
        them (an empty list removes them)."""
        c = np.ascontiguousarray(np.asarray(cells, dtype=np.int32).reshape(-1, 2))
        check(_lib().grtm_plan_plant_nonfinite(self.ptr, c.ctypes.data_as(P), len(c)), "grtm_plan_plant_nonfinite")

    def set_flag64(self, branch_cells=-1, pole_widths=0.0):
        """Re-solve in FP64, at every execution, the FP32 cells within
        `branch_cells` wavenumber cells of every layer's branch points
        Re(w/alpha), Re(w/beta) and (learned plans) within `pole_widths` half
        widths of each pole cluster the network predicts; branch_cells < 0 and
        pole_widths = 0 turn the two parts off."""
        check(_lib().grtm_plan_set_flag64(self.ptr, int(branch_cells), float(pole_widths)), "grtm_plan_set_flag64")

    def flag_list(self):
        """(rows, ks) of the cells re-solved in FP64 in the last execution."""
        n = _lib().grtm_plan_flag_list(self.ptr, None, 0)
        if n < 0:
            check(n, "grtm_plan_flag_list")
        buf = np.zeros(max(n, 1), dtype=np.uint32)
        m = _lib().grtm_plan_flag_list(self.ptr, buf.ctypes.data_as(P), n)
        if m < 0:
            check(m, "grtm_plan_flag_list")
        c = buf[:min(n, m)].astype(np.int64)
        return c // self.nk, c % self.nk

    def set_adaptive(self, kcut, window=None, stride=8, tol=1e-5):
        """Use a truncated wavenumber grid: kcut[iw] (km^-1) per frequency row,
        residual check every `stride`-th wavenumber over `window` km^-1 beyond
        the cut (default: models.check_window of the plan's source and
        receiver depths, 4 / max(|z_s - z_r|, 0.5)), fallback to the dense row
        when the checked tail exceeds tol x the row peak.
        kcut=None restores the dense grid."""
        if kcut is None:
            check(_lib().grtm_plan_set_adaptive(self.ptr, None, 0.0, 1, 1.0), "grtm_plan_set_adaptive")
            return
        kc = np.ascontiguousarray(kcut, dtype=np.float64)
        if kc.shape != (self.nw,):
            raise ValueError("kcut must have one entry per frequency row")
        if window is None:
            window = models.check_window(self.src_depth, self.rcv_depth)
        check(_lib().grtm_plan_set_adaptive(self.ptr, kc.ctypes.data_as(P), float(window), int(stride), float(tol)),
              "grtm_plan_set_adaptive")

    def set_learned(self, predictor, window=None, stride=8, tol=3e-6):
        """Learned adaptive grid formed on the GPU at every execution: the
        network of `predictor` (learned.Predictor) evaluates this model's rows
        inside execute(), followed by the residual check and dense fallback of
        set_adaptive.  predictor=None returns to the dense grid."""
        if predictor is None:
            check(_lib().grtm_plan_set_learned(self.ptr, None, 0.0, 0.0, 1, 1.0), "grtm_plan_set_learned")
            self._predictor = None
            return
        if window is None:
            window = models.check_window(self.src_depth, self.rcv_depth)
        check(_lib().grtm_plan_set_learned(self.ptr, predictor.mlp.ptr, float(predictor.lmargin), float(window),
                                           int(stride), float(tol)), "grtm_plan_set_learned")
        self._predictor = predictor  # the plan uses its network: keep it alive

    def adaptive_rows(self):
        """(cut, checks) per frequency row of the current grid (a learned grid:
        as of the last execution)."""
        cut = np.zeros(self.nw, dtype=np.int32)
        chk = np.zeros(self.nw, dtype=np.int32)
        check(_lib().grtm_plan_adaptive_rows(self.ptr, cut.ctypes.data_as(P), chk.ctypes.data_as(P)),
              "grtm_plan_adaptive_rows")
        return cut, chk

    def row_labels(self, tau=1e-7):
        """Per-row training labels after a dense execute(): array [nw, 26]
        (peak, cut index, 8 x (peak index, half width, relative height))."""
        out = np.zeros((self.nw, 26), dtype=np.float32)
        check(_lib().grtm_plan_row_labels(self.ptr, float(tau), out.ctypes.data_as(P)), "grtm_plan_row_labels")
        return out

    def adaptive_stats(self):
        o = np.zeros(4)
        check(_lib().grtm_plan_adaptive_stats(self.ptr, o.ctypes.data_as(P)), "grtm_plan_adaptive_stats")
        return dict(cells_pred=o[0], cells_check=o[1], cells_final=o[2], rows_fallback=int(o[3]))

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            _lib().grtm_plan_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()


def selftest_cell(compat, src_type, k, w, lay, s_layer, zs, r_layer, zr, par0_rv=None, dd=False):
    """Host-compiled cell solve (test hook): returns 18 complex coefficients.
    dd: solve in double-double arithmetic."""
    lay = np.ascontiguousarray(lay, dtype=np.float64)
    out = np.zeros(36)
    p0 = None if par0_rv is None else np.ascontiguousarray(par0_rv, dtype=np.float64)
    fn = _lib().grtm_selftest_cell_dd if dd else _lib().grtm_selftest_cell
    check(fn(int(compat), int(src_type), float(k), float(w.real), float(w.imag),
                                    lay.shape[0], lay.ctypes.data_as(P), int(s_layer), float(zs), int(r_layer),
                                    float(zr), None if p0 is None else p0.ctypes.data_as(P),
                                    out.ctypes.data_as(P)), "grtm_selftest_cell")
    return out[0::2] + 1j * out[1::2]



# Displacement columns of a ColumnPlan (the reference's integrand columns,
# cusrc/popm/func/green.cu:1582-1586 single force, 1651-1662 tensor).
_T = ("tensor R n2", "tensor R n1", "tensor R n0 s2", "tensor R n0 s0", "tensor T n2", "tensor T n1",
      "tensor Z n2", "tensor Z n1", "tensor Z n0 s0", "tensor Z n0 s2")
_F = ("hforce R", "vforce R", "hforce T", "hforce Z", "vforce Z")
COLUMNS = _T + _F
N_COLUMNS = len(COLUMNS)
# With derivatives: the reference's physics-mode layout of both sources,
# tensor 0-29 then single force 0-14 (cusrc/popm/func/green.cu:1589-1600, 1665-1691).
STRAIN_COLUMNS = (_T + tuple("d/dz " + c for c in _T) + tuple("d/dr " + c for c in _T) + _F +
                  tuple("d/dz " + c for c in _F) + tuple("d/dr " + c for c in _F))


class ColumnPlan(TracePlan):
    """All displacement columns of both sources (tensor columns 0-9 and
    single-force columns 0-4 of green()'s output) of one layered model at many
    distances from one recursion; derivatives=True (physics mode) adds their
    d/dz and d/dr (STRAIN_COLUMNS).  Other arguments as TracePlan; traces()
    returns [len(columns), nr, N]."""

    _CREATE = "grtm_plan_columns"

    def __init__(self, *args, derivatives=False, **kw):
        self.derivatives = bool(derivatives)
        self.columns = STRAIN_COLUMNS if self.derivatives else COLUMNS
        super().__init__(*args, **kw)

    def _extra_args(self):
        return (int(self.derivatives),)

    def traces(self):
        self.execute()
        out = np.zeros((len(self.columns), self.nr, self.N))
        check(_lib().grtm_plan_traces_out(self.ptr, out.ctypes.data_as(P)), "grtm_plan_traces_out")
        return out

class BatchPlan:
    """Explosion and vertical-force Z/R traces of many layered models on one
    grid and one set of surface-receiver distances.  ``models``: list of
    (layers, src_depth) with layers rows (thickness km, vp, vs, rho, qp, qs).
    The launch sequence is captured once in a HIP graph; ``autotune`` times the
    recursion tile and contraction slice candidates on the first group."""

    def __init__(self, models, rcv_depth, dists, grid, precision=PREC_FP32, compat=False, autotune=True, device=0):
        self._ctx = _ctx(device)
        self._layers = [np.ascontiguousarray(l, dtype=np.float64) for l, _ in models]
        ms = (Model * len(models))()
        for i, (l, (_, zs)) in enumerate(zip(self._layers, models)):
            ms[i] = Model(l.shape[0], l.ctypes.data_as(P), float(zs), float(rcv_depth))
        self.dists = np.ascontiguousarray(dists, dtype=np.float64)
        self.src_depths = [float(zs) for _, zs in models]
        self.rcv_depths = [float(rcv_depth)] * len(models)
        self.nm, self.nr = len(models), len(self.dists)
        self.nw = int(grid["nw"])
        self.N = 2 * self.nw
        g = Grid(self.nw, float(grid["dt"]), int(grid["nk"]), float(grid["kmax"]), float(grid.get("omega_i", 0.0)),
                 float(grid.get("t0", 0.0)), float(grid.get("taper", 0.0)))
        self.ptr = ctypes.c_void_p()
        check(_lib().grtm_batch_create(self._ctx.ptr, ms, self.nm, self.dists.ctypes.data_as(P), self.nr,
                                       ctypes.byref(g), int(precision), int(bool(compat)), 1 if autotune else 0,
                                       ctypes.byref(self.ptr)), "grtm_batch_create")

    def execute(self):
        check(_lib().grtm_batch_execute(self.ptr, None), "grtm_batch_execute")

    def execute_eager(self):
        check(_lib().grtm_batch_execute_eager(self.ptr), "grtm_batch_execute_eager")

    def sync(self):
        check(_lib().grtm_batch_sync(self.ptr), "grtm_batch_sync")

    def set_adaptive(self, kcuts, windows, stride=8, tol=3e-6):
        """Adaptive wavenumber grids (as TracePlan.set_adaptive) for every model:
        kcuts [n_models, nw] km^-1, windows [n_models] km^-1; kcuts=None
        returns to the dense grids.  Recaptures the graph."""
        if kcuts is None:
            check(_lib().grtm_batch_set_adaptive(self.ptr, None, None, 1, 1.0), "grtm_batch_set_adaptive")
            return
        kc = np.ascontiguousarray(kcuts, dtype=np.float64)
        wn = np.ascontiguousarray(windows, dtype=np.float64)
        if kc.shape != (self.nm, self.nw) or wn.shape != (self.nm,):
            raise ValueError("kcuts must be [n_models, nw] and windows [n_models]")
        check(_lib().grtm_batch_set_adaptive(self.ptr, kc.ctypes.data_as(P), wn.ctypes.data_as(P), int(stride),
                                             float(tol)), "grtm_batch_set_adaptive")

    def set_learned(self, predictor, windows=None, stride=8, tol=3e-6):
        """Learned adaptive grids of every model formed on the GPU inside the
        captured graph (the network of `predictor` runs in every execution);
        windows[m] km^-1 (default models.check_window of each model's depths).
        predictor=None returns to the dense grids."""
        if predictor is None:
            check(_lib().grtm_batch_set_learned(self.ptr, None, 0.0, None, 1, 1.0), "grtm_batch_set_learned")
            self._predictor = None
            return
        if windows is None:
            windows = [models.check_window(zs, zr) for zs, zr in zip(self.src_depths, self.rcv_depths)]
        wn = np.ascontiguousarray(windows, dtype=np.float64)
        if wn.shape != (self.nm,):
            raise ValueError("windows must have one entry per model")
        check(_lib().grtm_batch_set_learned(self.ptr, predictor.mlp.ptr, float(predictor.lmargin),
                                            wn.ctypes.data_as(P), int(stride), float(tol)), "grtm_batch_set_learned")
        self._predictor = predictor  # the batch uses its network: keep it alive

    def plant_nonfinite(self, cells):
        """Test hook: non-finite FP32 operands in the cells [(model, row, k), ...]
        at every execution (TracePlan.plant_nonfinite); recaptures the graph."""
        c = np.ascontiguousarray(np.asarray(cells, dtype=np.int32).reshape(-1, 3))
        check(_lib().grtm_batch_plant_nonfinite(self.ptr, c.ctypes.data_as(P), len(c)), "grtm_batch_plant_nonfinite")

    def adaptive_stats(self):
        o = np.zeros(4)
        check(_lib().grtm_batch_adaptive_stats(self.ptr, o.ctypes.data_as(P)), "grtm_batch_adaptive_stats")
        return dict(cells_pred=o[0], cells_check=o[1], cells_final=o[2], rows_fallback=int(o[3]))

    def traces(self, m0=0, m1=None):
        """traces[m1 - m0, 2 (explosion, vertical force), 2 (Z, R), nr, N] of
        the last execution."""
        m1 = self.nm if m1 is None else m1
        out = np.zeros((m1 - m0, 2, 2, self.nr, self.N))
        check(_lib().grtm_batch_traces_out(self.ptr, int(m0), int(m1), out.ctypes.data_as(P)), "grtm_batch_traces_out")
        return out

    def timing(self, reps=3):
        """ms: graph replay (mean of reps) and the recursion / contraction /
        IFFT split of one eager pass, with the learned grids' formation
        (``predictor``, part of the recursion stage)."""
        ms = np.zeros(5)
        check(_lib().grtm_batch_timing(self.ptr, int(reps), ms.ctypes.data_as(P)), "grtm_batch_timing")
        return dict(total=float(ms[0]), cells=float(ms[1]), ksum=float(ms[2]), ifft=float(ms[3]),
                    predictor=float(ms[4]))

    def config(self):
        o = np.zeros(10)
        check(_lib().grtm_batch_config(self.ptr, o.ctypes.data_as(P)), "grtm_batch_config")
        return dict(models_per_group=int(o[0]), tile=int(o[1]), k_slice=int(o[2]), groups=int(o[3]),
                    graph_nodes=int(o[4]), tuned=int(o[5]), tune_ms=float(o[6]), item_chunk=int(o[7]),
                    resident=bool(o[8]), rows_cut=int(o[9]))

    def work(self):
        o = np.zeros(6)
        check(_lib().grtm_batch_work(self.ptr, o.ctypes.data_as(P)), "grtm_batch_work")
        return dict(cells=o[0], cell_layers=o[1], ksum_flop=o[2], traces=o[3], rec_flop=o[4], rec_flop_solved=o[5])

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            _lib().grtm_batch_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()


def selftest_gemm_cut(prec, A, B, nk_w):
    """The learned-grid contraction (test hook): C = A B with row r of A read up
    to column nk_w[r // 4]; prec 1 FP32 stream-K, 0 FP64 split-K, 3 / 6 the
    split-BF16 stream-K kernel with that many products."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    nk_w = np.ascontiguousarray(nk_w, dtype=np.int32)
    M, K = A.shape
    N = B.shape[1]
    if M % 4 or len(nk_w) != M // 4:
        raise ValueError("A needs four rows per entry of nk_w")
    C = np.zeros((M, N))
    check(_lib().grtm_selftest_gemm_cut(int(prec), M, N, K, A.ctypes.data_as(P), B.ctypes.data_as(P),
                                        nk_w.ctypes.data_as(P), C.ctypes.data_as(P)), "grtm_selftest_gemm_cut")
    return C


def selftest_gemm(mode, A, B):
    """C = A B on the GPU with a contraction kernel (test hook): mode -1 VALU
    FP64, 0 FP64 MFMA, 1 FP32 MFMA, 2 BF16x3 MFMA, 3 streamed FP32 MFMA pair
    kernel, 4 streamed FP64 MFMA pair kernel."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N))
    check(_lib().grtm_selftest_gemm(int(mode), M, N, K, A.ctypes.data_as(P), B.ctypes.data_as(P),
                                    C.ctypes.data_as(P)), "grtm_selftest_gemm")
    return C
