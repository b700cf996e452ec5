"""GPU parity of the multi-receiver trace path against the FP64 oracle."""
import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import models

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


@pytest.fixture(scope="module")
def c1_oracle(orc):
    cfg = models.config("c1")
    out = {}
    for compat in (False, True):
        for kind in (0, 1):
            tr, sp = orc.traces(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], kind, cfg["grid"],
                                compat=compat)
            out[(compat, kind)] = tr
    return cfg, out


@pytest.mark.parametrize("compat", [False, True])
def test_c1_fp64_traces_match_oracle(c1_oracle, compat):
    cfg, ref = c1_oracle
    print(f"gpu: c1 fp64 compat={compat}", flush=True)
    plan = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                       precision=g.PREC_FP64, compat=compat)
    tr = plan.traces()  # [q, c, r, N]
    for q in (0, 1):
        o = ref[(compat, q)]  # [r, c, N]
        rl = rel_l2(tr[q].transpose(1, 0, 2), o)
        # the residual comes from the lowest-frequency rows, where the FP64
        # oracle itself sits within 8e-9 of its binary128 build
        assert rl.max() < 1e-7, (q, rl)


def test_c1_kernels_match_oracle(c1_oracle, orc):
    cfg, _ = c1_oracle
    plan = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                       precision=g.PREC_FP64)
    plan.execute()
    plan.sync()
    for q in (0, 1):
        kg = plan.kernels(q, 100, 104)
        ko = orc.kernels(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], q, cfg["grid"], 100, 104)
        assert np.abs(kg - ko).max() <= 1e-11 * np.abs(ko).max()


@pytest.mark.parametrize("prec", [g.PREC_FP32, g.PREC_BF16X3, g.PREC_BF16X6])
def test_c1_mixed_precision_traces(c1_oracle, prec):
    """FP32 recursion (FP64 vertical wavenumbers, FP64 low-frequency rows,
    FP64 re-solve of non-finite cells) + FP32 contraction: every trace within
    1e-4 rel-L2 of the FP64 oracle."""
    cfg, ref = c1_oracle
    print(f"gpu: c1 mixed precision {prec}", flush=True)
    plan = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                       precision=prec)
    tr = plan.traces()
    w = plan.work()
    worst = 0.0
    for q in (0, 1):
        rl = rel_l2(tr[q].transpose(1, 0, 2), ref[(False, q)])
        worst = max(worst, rl.max())
    print("mixed worst rel-L2", worst, "fp64 rows", w["rows_fp64"], "re-solved cells", w["cells_refixed"])
    assert 4 <= w["rows_fp64"] < cfg["grid"]["nw"] // 8
    assert worst < 1e-4


def test_c1_learned_adaptive_traces_match_oracle(c1_oracle):
    """The learned adaptive FP32 path (the network's cuts formed on the GPU,
    residual check, dense fallback) against the FP64 oracle's dense traces:
    every trace within 1e-4 rel-L2 and peak error, on a fraction of the cells."""
    from grtm_cuda_amd import learned
    cfg, ref = c1_oracle
    print("gpu: c1 learned adaptive FP32", flush=True)
    plan = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                       precision=g.PREC_FP32)
    learned.Predictor().attach(plan)
    tr = plan.traces()
    st = plan.adaptive_stats()
    frac = st["cells_final"] / (cfg["grid"]["nw"] * cfg["grid"]["nk"])
    worst = worst_pk = 0.0
    for q in (0, 1):
        a, o = tr[q].transpose(1, 0, 2), ref[(False, q)]
        worst = max(worst, rel_l2(a, o).max())
        worst_pk = max(worst_pk, (np.abs(a - o).max(axis=-1) / np.abs(o).max(axis=-1)).max())
    print(f"  learned adaptive vs oracle: rel-L2 {worst:.2e}, peak {worst_pk:.2e}, cells {frac:.3f} of dense, "
          f"fallback rows {st['rows_fallback']}")
    assert worst < 1e-4 and worst_pk < 1e-4 and frac < 0.6


def test_c1_rows64_all_fp64_recursion(c1_oracle):
    """set_rows64(nw): every row's recursion in FP64, FP32 contraction; the
    error is then that of the FP32 contraction and FFT alone, and smaller than
    the mixed recursion's.  Out-of-range row counts are rejected."""
    cfg, ref = c1_oracle
    nw = cfg["grid"]["nw"]
    print("gpu: c1 rows64 = nw", flush=True)
    mk = lambda: g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                             precision=g.PREC_FP32)
    mixed, full = mk(), mk()
    full.set_rows64(nw)
    assert full.work()["rows_fp64"] == nw
    e = []
    for plan in (mixed, full):
        tr = plan.traces()
        e.append(max(rel_l2(tr[q].transpose(1, 0, 2), ref[(False, q)]).max() for q in (0, 1)))
    print("rel-L2 mixed", e[0], "FP64 recursion", e[1])
    assert e[1] < 1e-5 and e[1] <= e[0]
    for bad in (-1, nw + 1):
        with pytest.raises(g.GrtmError):
            full.set_rows64(bad)
    with pytest.raises(g.GrtmError):
        g.TracePlan(cfg["layers"], cfg["src_depth"], 0.0, cfg["dists"], cfg["grid"]).set_rows64(4)


def test_plan_rejects_bad_grid():
    cfg = models.config("c1")
    bad = dict(cfg["grid"], nw=500)
    with pytest.raises(g.GrtmError):
        g.TracePlan(cfg["layers"], 10.0, 0.0, cfg["dists"], bad)


@pytest.mark.parametrize("nr", [5, 8])
def test_ifft_receiver_groups(nr):
    """The inverse FFT stages and transforms the spectra of 2 adjacent
    receivers per workgroup when the receiver count is even (8), one
    otherwise (5); every trace equals the one of a one-receiver plan, for the
    Z/R traces and for the displacement and strain columns."""
    grid = models.dense_grid(128, 0.1, 1024, float(models.C1_LAYERS[:, 2].min()))
    dists = np.linspace(10.0, 80.0, nr)
    print(f"gpu: ifft receiver groups, {nr} receivers", flush=True)
    for make in (lambda d: g.TracePlan(models.C1_LAYERS, 10.0, 0.0, d, grid, precision=g.PREC_FP64),
                 lambda d: g.ColumnPlan(models.C1_LAYERS, 10.0, 0.0, d, grid, precision=g.PREC_FP64,
                                        derivatives=True)):
        tr = make(dists).traces()  # receivers on the next-to-last axis
        for i, d in enumerate(dists):
            one = make([d]).traces()
            a, b = tr[..., i, :], one[..., 0, :]
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            assert err < 1e-12, (nr, i, err)
