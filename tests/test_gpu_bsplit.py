"""The split-precision contractions in plans and batches (GRTM_PREC_BF16X6:
three bf16 terms per FP32 operand, six products; GRTM_PREC_BF16X3: two
terms, three products; GRTM_PREC_FP16X3: two fp16 terms of each row-scaled
operand, three products), on the same FP32 recursion as
GRTM_PREC_FP32: traces against the plain FP64 plans, and against the FP32
plans they must track.
(The 3-layer model's traces of both forms against the oracle sit with the
other mixed-precision cases in test_gpu_traces.)"""
import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import learned, models

pytestmark = pytest.mark.gpu


def errs(a, b):
    d = a - b
    return (np.linalg.norm(d, axis=-1) / np.linalg.norm(b, axis=-1),
            np.abs(d).max(axis=-1) / np.abs(b).max(axis=-1))


def c2_plan(prec, learn):
    cfg = models.config("c2")
    p = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"], precision=prec)
    if learn:
        learned.Predictor().attach(p, window=models.check_window(cfg["src_depth"], cfg["rcv_depth"]),
                                   stride=8, tol=3e-6)
    return p


@pytest.fixture(scope="module")
def c2_ref():
    return c2_plan(g.PREC_FP64, False).traces()


@pytest.mark.parametrize("learn", [True, False])
def test_c2_bf16x6_traces(c2_ref, learn):
    """The 8-layer / 256-receiver model, learned adaptive and dense: BF16X6
    within 1e-4 rel-L2 and peak of FP64, and as close to it as the FP32 plan
    on the same grid, from which it differs (the split kernel ran): both sum
    16384 cancelling terms in FP32, so each is off by about 1e-5."""
    print(f"gpu: c2 bf16x6 {'learned' if learn else 'dense'}", flush=True)
    t6 = c2_plan(g.PREC_BF16X6, learn).traces()
    t32 = c2_plan(g.PREC_FP32, learn).traces()
    rl, pk = errs(t6, c2_ref)
    rl_32, _ = errs(t32, c2_ref)
    d32, _ = errs(t6, t32)
    print(f"  vs FP64: rel-L2 {rl.max():.2e} (FP32 plan {rl_32.max():.2e}), peak {pk.max():.2e}; "
          f"vs FP32 {d32.max():.2e}")
    assert rl.max() < 1e-4 and pk.max() < 1e-4
    assert rl.max() < 1.5 * rl_32.max() and np.median(rl) < 1.5 * np.median(rl_32)
    assert d32.max() > 0


def test_c4_sample_bf16x6_learned_batch():
    """Twelve models of the 1024-model batch, learned grids on the resident
    layout with the split contraction, against per-model FP64 plans."""
    cfg = models.config("c4")
    ms = [cfg["models"][i] for i in range(0, 1024, 86)][:12]
    grid, dists = cfg["grid"], cfg["dists"]
    print(f"gpu: c4 sample of {len(ms)} models, bf16x6 learned batch", flush=True)
    bp = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_BF16X6)
    bp.set_learned(learned.Predictor(), stride=8, tol=3e-6)
    bp.execute()
    tr = bp.traces()
    assert bp.config()["resident"] == 1
    worst = worst_pk = 0.0
    for m, (lay, zs) in enumerate(ms):
        ref = g.TracePlan(lay, zs, 0.0, dists, grid, precision=g.PREC_FP64).traces()
        rl, pk = errs(tr[m], ref)
        worst, worst_pk = max(worst, float(rl.max())), max(worst_pk, float(pk.max()))
    print(f"  vs FP64 plans: rel-L2 {worst:.2e}, peak {worst_pk:.2e}")
    assert worst < 1e-4 and worst_pk < 1e-4


@pytest.mark.parametrize("learn", [True, False])
def test_c2_fp16x3_traces(c2_ref, learn):
    """The scaled FP16 split (about 2^-22 per product) on the 8-layer /
    256-receiver model, learned adaptive and dense: within 1e-4 rel-L2 and
    peak of FP64 and as close to it as the FP32 plan, from which it differs
    (the split kernel ran)."""
    print(f"gpu: c2 fp16x3 {'learned' if learn else 'dense'}", flush=True)
    t16 = c2_plan(g.PREC_FP16X3, learn).traces()
    t32 = c2_plan(g.PREC_FP32, learn).traces()
    rl, pk = errs(t16, c2_ref)
    rl_32, _ = errs(t32, c2_ref)
    d32, _ = errs(t16, t32)
    print(f"  vs FP64: rel-L2 {rl.max():.2e} (FP32 plan {rl_32.max():.2e}), peak {pk.max():.2e}; "
          f"vs FP32 {d32.max():.2e}")
    assert rl.max() < 1e-4 and pk.max() < 1e-4
    assert rl.max() < 1.5 * rl_32.max() and np.median(rl) < 1.5 * np.median(rl_32)
    assert d32.max() > 0


def test_c4_sample_fp16x3_learned_batch():
    """The same twelve models of the 1024-model batch with the scaled FP16
    split, learned grids on the resident layout, against per-model FP64 plans."""
    cfg = models.config("c4")
    ms = [cfg["models"][i] for i in range(0, 1024, 86)][:12]
    grid, dists = cfg["grid"], cfg["dists"]
    print(f"gpu: c4 sample of {len(ms)} models, fp16x3 learned batch", flush=True)
    bp = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP16X3)
    bp.set_learned(learned.Predictor(), stride=8, tol=3e-6)
    bp.execute()
    tr = bp.traces()
    assert bp.config()["resident"] == 1
    worst = worst_pk = 0.0
    for m, (lay, zs) in enumerate(ms):
        ref = g.TracePlan(lay, zs, 0.0, dists, grid, precision=g.PREC_FP64).traces()
        rl, pk = errs(tr[m], ref)
        worst, worst_pk = max(worst, float(rl.max())), max(worst_pk, float(pk.max()))
    print(f"  vs FP64 plans: rel-L2 {worst:.2e}, peak {worst_pk:.2e}")
    assert worst < 1e-4 and worst_pk < 1e-4


def test_fp16x3_dense_batch_matches_plans():
    """A dense FP16X3 batch (row peaks kept by its recursion for the scales)
    against per-model FP16X3 plans on the same grid: the same arithmetic up to
    the order of the stream-K sums."""
    cfg = models.config("c4", n_models=6)
    grid = dict(cfg["grid"], nw=128, nk=1024)
    dists = cfg["dists"][:16]
    print("gpu: dense fp16x3 batch of 6 models vs plans", flush=True)
    bp = g.BatchPlan(cfg["models"], 0.0, dists, grid, precision=g.PREC_FP16X3)
    tr = bp.traces()
    for m, (lay, zs) in enumerate(cfg["models"]):
        ref = g.TracePlan(lay, zs, 0.0, dists, grid, precision=g.PREC_FP16X3).traces()
        rl, _ = errs(tr[m], ref)
        assert rl.max() < 2e-5, (m, rl.max())


def test_fp16x3_not_for_columns():
    """Column plans refuse the FP16 split (their operands are the column rows of
    four contractions, without per-row peaks)."""
    c1 = models.config("c1")
    grid = dict(c1["grid"], nw=64, nk=512)
    with pytest.raises(g.GrtmError):
        g.ColumnPlan(c1["layers"], c1["src_depth"], 0.0, c1["dists"][:4], grid, precision=g.PREC_FP16X3)
