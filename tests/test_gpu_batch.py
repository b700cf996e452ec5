"""Batched multi-model path: one captured HIP graph over groups of models
with uneven layer counts and source depths, against per-model TracePlans
(which are themselves pinned to the FP64 oracle)."""
import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import models

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


def small_set(n, seed=77, n_max=12):
    ms = models.batch_models(seed, n, n_min=4, n_max=n_max, src_lo=2.0, src_hi=30.0)
    vs_min = min(float(l[:, 2].min()) for l, _ in ms)
    return ms, models.dense_grid(128, 0.1, 1024, vs_min), np.linspace(5.0, 120.0, 8)


def per_model(ms, grid, dists, prec, compat=False):
    return np.stack([g.TracePlan(l, zs, 0.0, dists, grid, precision=prec, compat=compat).traces() for l, zs in ms])


@pytest.mark.parametrize("prec,tol", [(g.PREC_FP64, 1e-11), (g.PREC_FP32, 2e-6)])
def test_batch_matches_per_model_plans(prec, tol):
    ms, grid, dists = small_set(9)
    print(f"gpu: batch of {len(ms)} models, precision {prec}", flush=True)
    bp = g.BatchPlan(ms, 0.0, dists, grid, precision=prec)
    bp.execute()
    tr = bp.traces()
    ref = per_model(ms, grid, dists, prec)
    err = rel_l2(tr, ref).max()
    print("config", bp.config(), "max rel-L2 vs per-model plans", err)
    assert tr.shape == (len(ms), 2, 2, len(dists), 2 * grid["nw"])
    assert err < tol
    if prec == g.PREC_FP32:
        ref64 = per_model(ms, grid, dists, g.PREC_FP64)
        assert rel_l2(tr, ref64).max() < 1e-4


def test_batch_compat_mode():
    """Compat mode (the reference's quirks): the 3-layer stack with sources in
    several layers; depths a single compat plan rejects are left out."""
    _, grid, dists = small_set(1)
    grid = models.dense_grid(128, 0.1, 1024, float(models.C1_LAYERS[:, 2].min()))
    ms = []
    for zs in (1.0, 5.0, 10.0, 12.5, 20.0, 30.0):
        try:
            g.TracePlan(models.C1_LAYERS, zs, 0.0, dists, grid, compat=True)
            ms.append((models.C1_LAYERS, zs))
        except g.GrtmError:
            pass
    assert len(ms) >= 2
    print(f"gpu: batch compat, {len(ms)} source depths", flush=True)
    bp = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP64, compat=True)
    bp.execute()
    err = rel_l2(bp.traces(), per_model(ms, grid, dists, g.PREC_FP64, compat=True)).max()
    assert err < 1e-11, err


def test_batch_groups_and_graph_replays():
    """70 models make two groups (64 + 6).  Every graph replay, with eager
    launches and synchronisations between the replays, equals the eager
    result bit for bit."""
    ms = models.batch_models(91, 70, n_min=4, n_max=9, src_lo=3.0, src_hi=20.0)
    vs_min = min(float(l[:, 2].min()) for l, _ in ms)
    grid = models.dense_grid(64, 0.1, 512, vs_min)
    dists = np.linspace(5.0, 60.0, 4)
    print("gpu: batch of 70 models, graph replays", flush=True)
    bp = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP32)
    cfg = bp.config()
    assert cfg["groups"] >= 2 and cfg["graph_nodes"] > 0
    bp.execute_eager()
    eager = bp.traces()
    for rep in range(3):
        bp.execute()
        bp.sync()
        assert np.array_equal(bp.traces(), eager), rep
        bp.execute_eager()
        bp.sync()
    ref = per_model(ms[60:], grid, dists, g.PREC_FP32)
    assert rel_l2(eager[60:], ref).max() < 2e-6


def test_batch_autotune_choices():
    ms, grid, dists = small_set(5, seed=79)
    print("gpu: batch autotune", flush=True)
    tuned = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP32, autotune=True)
    plain = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP32, autotune=False)
    ct, cp = tuned.config(), plain.config()
    print("tuned", ct, "default", cp)
    assert ct["tile"] in (64, 128, 256) and ct["tuned"] >= 3
    assert cp["tuned"] == 0 and cp["tile"] == 256
    tuned.execute()
    plain.execute()
    assert rel_l2(tuned.traces(), plain.traces()).max() < 2e-6


def test_batch_rejects_bad_input():
    ms, grid, dists = small_set(2)
    with pytest.raises(g.GrtmError):
        g.BatchPlan(ms, 0.0, [-1.0], grid)
    with pytest.raises(g.GrtmError):
        g.BatchPlan(ms, 0.0, dists, dict(grid, nw=100))


def adaptive_inputs(ms, grid, scale=1.0):
    from grtm_cuda_amd import learned
    pred = learned.Predictor()
    kc = np.stack([pred.kcut(l, zs, 0.0, grid) * scale for l, zs in ms])
    win = np.array([models.check_window(zs, 0.0) for _, zs in ms])
    return kc, win


@pytest.mark.parametrize("prec,tol", [(g.PREC_FP64, 1e-11), (g.PREC_FP32, 2e-5)])
def test_batch_adaptive_matches_per_model_plans(prec, tol):
    """Learned cuts in a batch: the same cells, fallbacks and traces as one
    adaptive TracePlan per model.  In FP32 the batch's contraction splits its
    K range at other points than a single-model plan does, so the FP32
    partial sums round differently (about 1e-5 on far receivers)."""
    ms = models.batch_models(83, 6, n_min=4, n_max=14, src_lo=2.0, src_hi=30.0)
    vs_min = min(float(l[:, 2].min()) for l, _ in ms)
    grid = models.dense_grid(256, 0.05, 2048, vs_min)
    dists = np.linspace(5.0, 150.0, 8)
    kc, win = adaptive_inputs(ms, grid)
    print(f"gpu: adaptive batch of {len(ms)} models, precision {prec}", flush=True)
    bp = g.BatchPlan(ms, 0.0, dists, grid, precision=prec)
    bp.set_adaptive(kc, win, stride=8, tol=3e-6)
    bp.execute()
    tr = bp.traces()
    st = bp.adaptive_stats()
    ref, cells, fb = [], 0.0, 0
    for i, (l, zs) in enumerate(ms):
        p = g.TracePlan(l, zs, 0.0, dists, grid, precision=prec)
        p.set_adaptive(kc[i], window=win[i], stride=8, tol=3e-6)
        ref.append(p.traces())
        s = p.adaptive_stats()
        cells += s["cells_final"]
        fb += s["rows_fallback"]
    err = rel_l2(tr, np.stack(ref)).max()
    print("batch stats", st, "per-model cells", cells, "fallback rows", fb, "max rel-L2", err)
    assert st["cells_final"] == cells and st["rows_fallback"] == fb
    assert st["cells_final"] < 0.8 * len(ms) * grid["nw"] * grid["nk"]
    assert err < tol


def test_batch_adaptive_graph_replays():
    """Two groups with adaptive grids (one model's cut far too short, so rows
    fall back): graph replays equal the eager result bit for bit, and the
    traces stay within 1e-4 of the dense FP64 batch."""
    ms = models.batch_models(92, 70, n_min=4, n_max=9, src_lo=3.0, src_hi=20.0)
    vs_min = min(float(l[:, 2].min()) for l, _ in ms)
    grid = models.dense_grid(64, 0.1, 512, vs_min)
    dists = np.linspace(5.0, 60.0, 4)
    kc, win = adaptive_inputs(ms, grid)
    kc[65] *= 0.3
    print("gpu: adaptive batch of 70 models, graph replays", flush=True)
    bp = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP32)
    bp.set_adaptive(kc, win, stride=8, tol=3e-6)
    bp.execute_eager()
    eager = bp.traces()
    st = bp.adaptive_stats()
    assert st["rows_fallback"] > 0
    for rep in range(3):
        bp.execute()
        bp.sync()
        assert np.array_equal(bp.traces(), eager), rep
        bp.execute_eager()
        bp.sync()
    ref = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP64)
    ref.execute()
    assert rel_l2(eager, ref.traces()).max() < 1e-4
    bp.set_adaptive(None, None)
    bp.execute()
    assert rel_l2(bp.traces(), ref.traces()).max() < 1e-4
