"""FP64 cells next to branch points and predicted surface-wave poles
(grtm_plan_set_flag64): the listed cells of a mixed-precision plan must carry
the FP64 plan's kernel values (rounded to FP32), each cell once, and the
traces must stay within the FP32 bar of the FP64 path."""
import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import learned, models

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


def _plans(cfg, grid):
    mk = lambda p: g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], grid, precision=p)
    return mk(g.PREC_FP64), mk(g.PREC_FP32)


def _check_cells_are_fp64(p32, p64, base, rows, ks, nk):
    """Kernels of the flagged cells equal the FP64 plan's to FP32 rounding;
    base[row] holds the unflagged FP32 plan's kernels of the same rows."""
    worst, before = 0.0, []
    for iw in np.unique(rows)[:: max(1, len(np.unique(rows)) // 12)]:
        kk = ks[rows == iw]
        for q in (0, 1):
            a = p32.kernels(q, int(iw), int(iw) + 1)[0][kk]
            b = p64.kernels(q, int(iw), int(iw) + 1)[0][kk]
            scale = np.abs(b).max(axis=-1, keepdims=True) + 1e-300
            worst = max(worst, float((np.abs(a - b) / scale).max()))
            before.append(float((np.abs(base[(int(iw), q)][kk] - b) / scale).max()))
    return worst, max(before)


def test_branch_point_cells_are_resolved_in_fp64():
    cfg = models.config("c1")
    grid = cfg["grid"]
    p64, p32 = _plans(cfg, grid)
    print("gpu: FP64 branch-point cells, C1 FP32 dense", flush=True)
    ref = p64.traces()
    tr0 = p32.traces()
    n0 = p32.work()["cells_refixed"]
    assert n0 == 0
    rows64 = p32.work()["rows_fp64"]
    hb = 2
    # unflagged FP32 kernels of the rows checked below
    base = {}
    for iw in range(rows64, grid["nw"]):
        for q in (0, 1):
            base[(iw, q)] = p32.kernels(q, iw, iw + 1)[0]
    p32.set_flag64(branch_cells=hb)
    for rep in range(2):
        tr = p32.traces()
        rows, ks = p32.flag_list()
        n = len(rows)
        assert n > 0 and p32.work()["cells_refixed"] == n
        cells = rows * grid["nk"] + ks
        assert len(np.unique(cells)) == n, "a cell listed twice"
        assert rows.min() >= rows64 and ks.max() < grid["nk"]
        L = len(cfg["layers"])
        per_row = np.bincount(rows - rows64)
        assert per_row.max() <= 2 * L * (2 * hb + 1)
        worst, before = _check_cells_are_fp64(p32, p64, base, rows, ks, grid["nk"])
        e = rel_l2(tr, ref).max()
        d = rel_l2(tr, tr0).max()
        print(f"  execution {rep}: {n} cells in FP64 ({n / (grid['nw'] - rows64):.1f} per FP32 row); flagged "
              f"cells vs FP64 {worst:.2e} (unflagged FP32 {before:.2e}); traces vs FP64 {e:.2e} (unflagged "
              f"{rel_l2(tr0, ref).max():.2e}), vs unflagged {d:.2e}")
        assert worst < 2.0 ** -22
        assert e < 1e-4 and d < 1e-5
    p32.set_flag64()
    p32.traces()
    assert p32.work()["cells_refixed"] == n0


def test_pole_cells_of_a_learned_grid():
    cfg = models.config("c1")
    grid = cfg["grid"]
    p64, p32 = _plans(cfg, grid)
    print("gpu: FP64 branch-point and predicted-pole cells, C1 FP32 learned", flush=True)
    ref = p64.traces()
    pred = learned.Predictor()
    pred.attach(p32)
    tr0 = p32.traces()
    p32.set_flag64(branch_cells=1)
    p32.traces()
    n_branch = len(p32.flag_list()[0])
    p32.set_flag64(branch_cells=1, pole_widths=1.0)
    tr = p32.traces()
    rows, ks = p32.flag_list()
    cut, _ = p32.adaptive_rows()
    assert len(rows) > n_branch, "no cell near a predicted pole"
    assert (ks < cut[rows]).all(), "a flagged cell beyond its row's cut"
    assert len(np.unique(rows * grid["nk"] + ks)) == len(rows)
    base = {}
    for iw in np.unique(rows):
        for q in (0, 1):
            base[(int(iw), q)] = np.zeros((grid["nk"], 2), complex)  # error of the unflagged plan not needed here
    worst, _ = _check_cells_are_fp64(p32, p64, base, rows, ks, grid["nk"])
    e = rel_l2(tr, ref).max()
    d = rel_l2(tr, tr0).max()
    print(f"  {n_branch} branch-point cells, {len(rows)} with the pole clusters; flagged cells vs FP64 {worst:.2e}; "
          f"traces vs FP64 {e:.2e} (unflagged {rel_l2(tr0, ref).max():.2e}), vs unflagged {d:.2e}")
    assert worst < 2.0 ** -22
    assert e < 1e-4 and d < 1e-5


def test_flag64_refuses_what_it_cannot_do():
    cfg = models.config("c1")
    grid = dict(cfg["grid"], nw=64, nk=512)
    p64, p32 = _plans(cfg, grid)
    with pytest.raises(g.GrtmError):
        p64.set_flag64(branch_cells=1)  # nothing to re-solve in an FP64 plan
    with pytest.raises(g.GrtmError):
        p32.set_flag64(pole_widths=1.0)  # the clusters come from a network
    with pytest.raises(g.GrtmError):
        p32.set_flag64(branch_cells=65)
