"""The FP64 re-solve of FP32 cells whose result is not finite (the cell
flagging of the mixed-precision recursion): cells planted with non-finite
operands through the test hook must come out of the plan and the batch
re-solved, with traces equal to the unplanted run's and within the FP32 bar
of the FP64 path, and the plan's count of re-solved cells must include them."""
import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import learned, models

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


def _cells(rng, lo, hi, nk, n):
    """n distinct (row, k) cells with rows in [lo, hi) and k below nk."""
    flat = rng.choice((hi - lo) * nk, size=n, replace=False)
    return np.stack([lo + flat // nk, flat % nk], axis=1)


@pytest.mark.parametrize("learned_grid", [False, True], ids=["dense", "learned"])
def test_planted_cells_are_resolved_in_fp64(learned_grid):
    cfg = models.config("c1")
    mk = lambda p: g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                               precision=p)
    print(f"gpu: planted non-finite cells, C1 FP32 {'learned' if learned_grid else 'dense'}", flush=True)
    ref64 = mk(g.PREC_FP64).traces()
    plan = mk(g.PREC_FP32)
    pred = learned.Predictor() if learned_grid else None
    if pred:
        pred.attach(plan)
    base = plan.traces()
    base_fixed = plan.work()["cells_refixed"]
    rows64 = plan.work()["rows_fp64"]
    nw, nk = cfg["grid"]["nw"], cfg["grid"]["nk"]
    rng = np.random.default_rng(4)
    if pred:
        cut, _ = plan.adaptive_rows()
        cells = [(iw, int(rng.integers(0, cut[iw]))) for iw in rng.choice(np.arange(rows64, nw), 96, replace=False)]
        cells = np.array(cells)
    else:
        cells = _cells(rng, rows64, nw, nk, 96)
    plan.plant_nonfinite(cells)
    for rep in range(2):  # planted again at every execution
        tr = plan.traces()
        assert np.isfinite(tr).all()
        assert plan.work()["cells_refixed"] == base_fixed + len(cells)
        d = rel_l2(tr, base).max()
        e = rel_l2(tr, ref64).max()
        print(f"  execution {rep}: {len(cells)} planted, vs unplanted {d:.2e}, vs FP64 {e:.2e}")
        assert d < 1e-5 and e < 1e-4
    plan.plant_nonfinite([])
    tr = plan.traces()
    assert plan.work()["cells_refixed"] == base_fixed and rel_l2(tr, base).max() < 1e-6


def test_planted_cells_in_a_batch():
    ms = models.batch_models(77, 7, n_min=4, n_max=12, src_lo=2.0, src_hi=30.0)
    vs_min = min(float(l[:, 2].min()) for l, _ in ms)
    grid = models.dense_grid(128, 0.1, 1024, vs_min)
    dists = np.linspace(5.0, 120.0, 8)
    print(f"gpu: planted non-finite cells, batch of {len(ms)} models", flush=True)
    ref64 = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP64).traces()
    bp = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP32)
    base = bp.traces()
    rng = np.random.default_rng(9)
    cells = []
    for m in (0, 3, 6):
        rows64 = g.TracePlan(ms[m][0], ms[m][1], 0.0, dists, grid, precision=g.PREC_FP32).work()["rows_fp64"]
        cells += [(m, r, k) for r, k in _cells(rng, rows64, grid["nw"], grid["nk"], 40)]
    bp.plant_nonfinite(cells)
    for rep in range(2):
        bp.execute()
        tr = bp.traces()
        assert np.isfinite(tr).all()
        d = rel_l2(tr, base).max()
        e = rel_l2(tr, ref64).max()
        print(f"  replay {rep}: {len(cells)} planted, vs unplanted {d:.2e}, vs FP64 {e:.2e}")
        assert d < 1e-5 and e < 1e-4


def test_planted_cells_in_a_resident_learned_batch():
    """The learned FP32 batch in its resident layout (one group, rows stored to
    their extents, 64-bit flag list): planted cells below the cuts come out
    re-solved, the traces within 1e-5 of the unplanted batch."""
    from grtm_cuda_amd import learned
    ms = models.batch_models(77, 7, n_min=4, n_max=12, src_lo=2.0, src_hi=30.0)
    vs_min = min(float(l[:, 2].min()) for l, _ in ms)
    grid = models.dense_grid(128, 0.1, 1024, vs_min)
    dists = np.linspace(5.0, 120.0, 8)
    print(f"gpu: planted non-finite cells, resident learned batch of {len(ms)} models", flush=True)
    pred = learned.Predictor()
    bp = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP32)
    bp.set_learned(pred)
    cfg = bp.config()
    assert cfg["resident"] and cfg["groups"] == 1 and cfg["rows_cut"] == 0
    bp.execute()
    base = bp.traces()
    rng = np.random.default_rng(10)
    cells = []
    for m in (1, 4):
        rows64 = g.TracePlan(ms[m][0], ms[m][1], 0.0, dists, grid, precision=g.PREC_FP32).work()["rows_fp64"]
        cells += [(m, int(r), int(rng.integers(0, 8))) for r in rng.choice(np.arange(rows64, grid["nw"]), 30,
                                                                             replace=False)]
    bp.plant_nonfinite(cells)
    assert bp.config()["resident"]
    for rep in range(2):
        bp.execute()
        tr = bp.traces()
        assert np.isfinite(tr).all()
        d = rel_l2(tr, base).max()
        print(f"  replay {rep}: {len(cells)} planted, vs unplanted {d:.2e}")
        assert d < 1e-5
