"""Ablation: what the trained wavenumber-cut network gains over the method
without it.

Held-out layered models (seeds outside the training and calibration ranges)
run on the 1024 x 8192 grid with 64 receivers three ways: plain dense FP64
(the judge), the adaptive FP32 path with the analytic cut
(|w| * 1.15 / Vs_min + ln(1e6) / dz), and the adaptive FP32 path with the
network's cut.  Both adaptive paths keep the residual check and the dense
fallback, so both must stay within 1e-4 rel-L2 of the dense FP64 traces; the
network must solve fewer cells than the analytic cut.
"""
import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import learned, models
from grtm_cuda_amd.train_predictor import model_sample

pytestmark = pytest.mark.gpu

HELD_OUT = (990001, 990002, 990003)


def rel_l2(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


@pytest.fixture(scope="module")
def predictor():
    return learned.Predictor()


@pytest.mark.parametrize("seed", HELD_OUT)
def test_learned_cut_beats_analytic(predictor, seed):
    layers, src = model_sample(seed)
    grid = models.dense_grid(1024, 0.05, 8192, float(layers[:, 2].min()))
    dists = np.linspace(5.0, 300.0, 64)
    print(f"gpu: ablation seed {seed}, {len(layers)} layers, source {src:.1f} km", flush=True)
    mk = lambda p: g.TracePlan(layers, src, 0.0, dists, grid, precision=p)
    ref = mk(g.PREC_FP64).traces()
    dense_cells = grid["nw"] * grid["nk"]
    window = models.check_window(src, 0.0)
    out = {}
    for name, kcut in (("analytic", models.analytic_kcut(layers, src, 0.0, grid)),
                       ("learned", predictor.kcut(layers, src, 0.0, grid))):
        plan = mk(g.PREC_FP32)
        plan.set_adaptive(kcut, window=window, stride=8, tol=3e-6)
        tr = plan.traces()
        st = plan.adaptive_stats()
        ms = plan.timing(reps=3)["total"]
        err = max(rel_l2(tr[q], ref[q]).max() for q in (0, 1))
        out[name] = (st["cells_final"] / dense_cells, st["rows_fallback"], err, ms)
        print(f"  {name:8s}: {out[name][0] * 100:5.1f}% of dense cells, {st['rows_fallback']:4d} fallback rows, "
              f"rel-L2 {err:.2e}, {ms:.2f} ms", flush=True)
    for name in out:
        assert out[name][2] < 1e-4, name
    assert out["learned"][0] < 0.95 * out["analytic"][0]
    assert out["learned"][0] < 0.5
    assert out["learned"][1] <= 0.05 * grid["nw"]


def test_row_check_catches_every_trace_a_per_trace_check_would():
    """eval_grid on 3 held-out models: every learned trace within 1e-4 of the
    plain method, none of the traces whose 1/8-density tail check exceeds 1e-4
    left wrong, and no labelled pole cluster beyond the checked region (the
    1024-model run of the same evaluation is in the design notes; the
    evaluation's contraction runs on the host here)."""
    from grtm_cuda_amd import eval_grid
    print("gpu: eval_grid on 3 held-out models", flush=True)
    pred = learned.Predictor()
    dists = np.linspace(5.0, 300.0, 64)
    for i in range(3):
        layers, src = model_sample(eval_grid.HELDOUT_SEED + i)
        grid = models.dense_grid(1024, 0.05, 8192, float(layers[:, 2].min()))
        r = eval_grid.evaluate_model(layers, src, dists, grid, pred, eval_grid.numpy_contract,
                                     eval_grid.bessel_tables(grid, dists))
        bad = (r["res"] > 1e-4) & (r["err"] > 1e-4)
        print(f"  model {i}: worst error {r['err'].max():.2e}, traces the tail check fails "
              f"{int((r['res'] > 1e-4).sum())}, clusters beyond {r['beyond']}, cells {r['cells']:.3f}")
        assert r["err"].max() < 1e-4 and r["beyond"] == 0 and not bad.any()
