"""Peak-trough averaging of the k-sum tails on the GPU (grtm_plan_set_ptam)
against the oracle's PTAM (the same method on the same grid) and against the
dense sum on a grid 32 times longer, for a source 0.5 km below surface
receivers, where the plain grid leaves up to 1e-2 of each trace out."""
import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import learned, models

pytestmark = pytest.mark.gpu

LAYERS = models.C1_LAYERS
ZS = 0.5
DISTS = np.array([5.0, 10.0, 30.0, 60.0])


def rel_l2(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


@pytest.fixture(scope="module")
def grids():
    g1 = models.dense_grid(64, 0.1, 512, float(LAYERS[:, 2].min()))
    dk = g1["kmax"] / g1["nk"]
    return g1, dict(g1, nk=512 * 32, kmax=dk * 512 * 32)


@pytest.fixture(scope="module")
def long_grid_traces(grids):
    """FP64 traces on the 32 times longer grid, [2 sources, 2 comps, nr, N]."""
    return g.TracePlan(LAYERS, ZS, 0.0, DISTS, grids[1], precision=g.PREC_FP64).traces()


@pytest.mark.parametrize("prec,tol", [(g.PREC_FP64, 1e-8), (g.PREC_FP32, 1e-4)])
def test_ptam_plan_matches_oracle(orc, grids, prec, tol):
    g1 = grids[0]
    plan = g.TracePlan(LAYERS, ZS, 0.0, DISTS, g1, precision=prec)
    plan.set_ptam(10)
    st = plan.ptam_stats()
    print(f"gpu: PTAM plan, precision {prec}, tail {st['tail']}", flush=True)
    tr = plan.traces()
    st = plan.ptam_stats()
    ref = np.stack([orc.traces(LAYERS, ZS, 0.0, DISTS, kind, g1, ptam=(10, st["tail"]))[0].transpose(1, 0, 2)
                    for kind in (0, 1)])
    err = rel_l2(tr, ref)
    print(f"  vs oracle PTAM: max rel-L2 {err.max():.2e}; extended sums {st['extended']} of {st['sums']}")
    assert err.max() < tol
    assert st["sums"] == 64 * 4 * len(DISTS) and st["extended"] < 0.05 * st["sums"]


@pytest.mark.parametrize("prec,tol", [(g.PREC_FP64, 1e-5), (g.PREC_FP32, 1e-4)])
def test_ptam_converges_to_the_long_grid(grids, long_grid_traces, prec, tol):
    g1 = grids[0]
    print(f"gpu: PTAM against the long grid, precision {prec}", flush=True)
    plain = g.TracePlan(LAYERS, ZS, 0.0, DISTS, g1, precision=prec)
    pt = g.TracePlan(LAYERS, ZS, 0.0, DISTS, g1, precision=prec)
    pt.set_ptam(10)
    e_plain = rel_l2(plain.traces(), long_grid_traces)  # [q, c, r]
    e_pt = rel_l2(pt.traces(), long_grid_traces)
    for ir, r in enumerate(DISTS):
        print(f"  r {r:5.1f} km: plain {e_plain[..., ir].max():.2e}, PTAM {e_pt[..., ir].max():.2e}")
    assert e_plain[..., :3].max() > 1e-3
    assert e_pt[..., :3].max() < tol
    assert e_pt[..., 3].max() < 0.1 * e_plain[..., 3].max()
    pt.set_ptam(0)  # removed: the plain grid again
    assert rel_l2(pt.traces(), plain.traces()).max() < 1e-12


def test_ptam_and_adaptive_grids_exclude_each_other(grids):
    g1 = grids[0]
    plan = g.TracePlan(LAYERS, ZS, 0.0, DISTS, g1, precision=g.PREC_FP32)
    plan.set_ptam(10)
    with pytest.raises(g.GrtmError):
        plan.set_learned(learned.Predictor())
    plan.set_ptam(0)
    plan.set_learned(learned.Predictor())
    with pytest.raises(g.GrtmError):
        plan.set_ptam(10)
