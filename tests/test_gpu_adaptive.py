"""Adaptive (truncated) wavenumber grid with residual check and dense fallback."""
import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import models

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


@pytest.fixture(scope="module")
def c1_dense():
    cfg = models.config("c1")
    print("gpu: c1 dense fp64 reference", flush=True)
    plan = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                       precision=g.PREC_FP64)
    tr = plan.traces()
    plan.sync()
    nw = cfg["grid"]["nw"]
    mags = np.zeros((nw, cfg["grid"]["nk"]))
    for q in (0, 1):
        k = plan.kernels(q, 0, nw)
        mags = np.maximum(mags, np.abs(k).max(axis=2))
    return cfg, tr, mags


def exact_cut(cfg, mags, tau):
    dk = cfg["grid"]["kmax"] / cfg["grid"]["nk"]
    pk = mags.max(axis=1, keepdims=True)
    last = np.array([np.nonzero(r)[0].max() + 1 for r in (mags > tau * pk)])
    return last * dk


@pytest.mark.parametrize("prec", [g.PREC_FP64, g.PREC_FP32])
def test_truncated_grid_matches_dense(c1_dense, prec):
    cfg, ref, mags = c1_dense
    kcut = exact_cut(cfg, mags, 1e-7)
    plan = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"], precision=prec)
    plan.set_adaptive(kcut, window=models.check_window(cfg["src_depth"], 0.0), stride=8, tol=1e-5)
    tr = plan.traces()
    st = plan.adaptive_stats()
    frac = st["cells_final"] / (cfg["grid"]["nw"] * cfg["grid"]["nk"])
    worst = max(rel_l2(tr[q], ref[q]).max() for q in (0, 1))
    print(f"truncated grid: {frac * 100:.1f}% cells, {st['rows_fallback']} fallback rows, rel-L2 {worst:.2e}")
    assert st["rows_fallback"] == 0 and frac < 0.4
    assert worst < (1e-5 if prec == g.PREC_FP64 else 1e-4)


def test_short_cut_falls_back_to_dense(c1_dense):
    """A cut placed far too low fails the residual check and the rows are
    completed on the dense grid: the result still matches."""
    cfg, ref, mags = c1_dense
    kcut = exact_cut(cfg, mags, 1e-7)
    bad = kcut.copy()
    bad[100:140] *= 0.5
    plan = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                       precision=g.PREC_FP32)
    plan.set_adaptive(bad, window=models.check_window(cfg["src_depth"], 0.0), stride=8, tol=1e-5)
    tr = plan.traces()
    st = plan.adaptive_stats()
    worst = max(rel_l2(tr[q], ref[q]).max() for q in (0, 1))
    print(f"short cut: {st['rows_fallback']} fallback rows, rel-L2 {worst:.2e}")
    assert st["rows_fallback"] >= 40
    assert worst < 1e-4


def test_analytic_cut(c1_dense):
    cfg, ref, _ = c1_dense
    kcut = models.analytic_kcut(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["grid"])
    plan = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                       precision=g.PREC_FP32)
    plan.set_adaptive(kcut, window=models.check_window(cfg["src_depth"], 0.0))
    tr = plan.traces()
    st = plan.adaptive_stats()
    worst = max(rel_l2(tr[q], ref[q]).max() for q in (0, 1))
    frac = st["cells_final"] / (cfg["grid"]["nw"] * cfg["grid"]["nk"])
    print(f"analytic cut: {frac * 100:.1f}% cells, {st['rows_fallback']} fallback rows, rel-L2 {worst:.2e}")
    assert worst < 1e-4


@pytest.mark.parametrize("prec,tol", [(g.PREC_FP64, 1e-9), (g.PREC_FP32, 1e-4)])
def test_cut_next_to_nk_keeps_the_row_dense(prec, tol):
    """A cut so close to the end of the grid that no residual-check cell fits
    below it (check cells sit at cut + stride j + stride / 2) keeps the row
    dense: for a source 30 m below the receivers the last wavenumbers still
    carry the trace, and an unchecked tail would be dropped."""
    layers = models.C1_LAYERS
    grid = models.dense_grid(128, 0.1, 1024, float(layers[:, 2].min()))
    dists = [5.0, 20.0]
    ref = g.TracePlan(layers, 0.03, 0.0, dists, grid, precision=g.PREC_FP64).traces()
    plan = g.TracePlan(layers, 0.03, 0.0, dists, grid, precision=prec)
    dk = grid["kmax"] / grid["nk"]
    print(f"gpu: cut next to nk, precision {prec}", flush=True)
    for room in (1, 4):
        plan.set_adaptive(np.full(grid["nw"], (grid["nk"] - room) * dk), window=1.0, stride=8, tol=3e-6)
        tr = plan.traces()
        st = plan.adaptive_stats()
        err = (np.linalg.norm(tr - ref, axis=-1) / np.linalg.norm(ref, axis=-1)).max()
        print(f"  room {room}: cells {st['cells_final']:.0f}, rel-L2 {err:.2e}")
        assert st["cells_final"] == grid["nw"] * grid["nk"], room
        assert err < tol, (room, err)
