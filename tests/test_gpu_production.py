"""The production workloads at their full grids against the oracle:
- the 8-layer / 256-receiver model (Nw 2048, Nk 16384): sampled frequency
  rows of the FP64 and FP32 plans' k-kernels against oracle.kernels;
- a sample of the 1024-model batch (Nw 1024, Nk 8192, 64 receivers, the
  batch's own grid), with sources at the surface, on an interface and just
  below one added: the learned FP32 batch against per-model FP64 plans per
  trace (rel-L2 and peak), and those plans' kernels on sampled rows against
  oracle.kernels."""
import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import learned, models

pytestmark = pytest.mark.gpu

C2_ROWS = [(64, 67), (1000, 1003), (2045, 2048)]


def _row_rel(kg, ko):
    """Per frequency row and component: rel-L2 of the kernel over k."""
    return np.linalg.norm(kg - ko, axis=1) / np.linalg.norm(ko, axis=1)


@pytest.mark.parametrize("prec,tol", [(g.PREC_FP64, 1e-9), (g.PREC_FP32, 1e-4)])
def test_c2_sampled_rows_kernels_match_oracle(orc, prec, tol):
    cfg = models.config("c2")
    print(f"gpu: c2 sampled-row kernels, precision {prec}", flush=True)
    plan = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"], precision=prec)
    plan.execute()
    plan.sync()
    worst = 0.0
    for q in (0, 1):
        for lo, hi in C2_ROWS:
            kg = plan.kernels(q, lo, hi)
            ko = orc.kernels(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], q, cfg["grid"], lo, hi)
            e = _row_rel(kg, ko)
            worst = max(worst, float(e.max()))
            assert e.max() < tol, (q, lo, e)
    print(f"  worst per-row kernel rel-L2 {worst:.2e}")


def c4_sample():
    """12 models of the 1024-model batch on its grid, three of them with the
    source moved to the surface, onto the first interface, and 1 m below the
    second one."""
    cfg = models.config("c4")
    ms = [cfg["models"][i] for i in range(0, 1024, 86)][:12]
    ms[1] = (ms[1][0], 0.0)
    ms[4] = (ms[4][0], float(ms[4][0][0, 0]))
    ms[7] = (ms[7][0], float(ms[7][0][:2, 0].sum()) + 1e-3)
    return cfg, ms


def test_c4_sample_learned_batch_matches_fp64_plans(orc):
    cfg, ms = c4_sample()
    grid, dists = cfg["grid"], cfg["dists"]
    print(f"gpu: c4 sample of {len(ms)} models at Nw {grid['nw']}, Nk {grid['nk']}, {len(dists)} receivers",
          flush=True)
    bp = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP32)
    bp.set_learned(learned.Predictor(), stride=8, tol=3e-6)
    bp.execute()
    tr = bp.traces()
    st = bp.adaptive_stats()
    dense = len(ms) * grid["nw"] * grid["nk"]
    worst = worst_pk = worst_k = 0.0
    for m, (lay, zs) in enumerate(ms):
        p64 = g.TracePlan(lay, zs, 0.0, dists, grid, precision=g.PREC_FP64)
        ref = p64.traces()
        d = tr[m] - ref
        rl = np.linalg.norm(d, axis=-1) / np.linalg.norm(ref, axis=-1)
        pk = np.abs(d).max(axis=-1) / np.abs(ref).max(axis=-1)
        worst, worst_pk = max(worst, float(rl.max())), max(worst_pk, float(pk.max()))
        assert rl.max() < 1e-4 and pk.max() < 1e-4, (m, zs, rl.max(), pk.max())
        if m in (0, 1, 4, 7):
            for q in (0, 1):
                for lo in (64, 512, 1020):
                    kg = p64.kernels(q, lo, lo + 2)
                    ko = orc.kernels(lay, zs, 0.0, q, grid, lo, lo + 2)
                    e = _row_rel(kg, ko)
                    worst_k = max(worst_k, float(e.max()))
                    assert e.max() < 1e-8, (m, zs, q, lo, e)
        if m == 1:
            # Source and receivers at the surface: the kernels do not decay with
            # k, so a low row (|w| = 0.49 s^-1) weighs cells with k V / |w| near
            # 600, where an FP64 solve keeps about 16 - 4 log10(600) digits.
            # The binary128 oracle decides: both FP64 solves sit within 1e-5.
            kq = orc.kernels(lay, zs, 0.0, 1, grid, 8, 9, quad=True)
            e_gpu = _row_rel(p64.kernels(1, 8, 9), kq).max()
            e_orc = _row_rel(orc.kernels(lay, zs, 0.0, 1, grid, 8, 9), kq).max()
            print(f"  surface source, row 8 vs binary128: GPU FP64 {e_gpu:.1e}, oracle FP64 {e_orc:.1e}")
            assert e_gpu < 1e-5 and e_orc < 1e-5
    print(f"  learned FP32 batch vs FP64 plans: worst rel-L2 {worst:.2e}, peak {worst_pk:.2e}; "
          f"cells {st['cells_final'] / dense:.3f} of dense, fallback rows {st['rows_fallback']}; "
          f"FP64 kernels vs oracle {worst_k:.2e}")
    assert st["cells_final"] < 0.6 * dense
