"""CPU checks of the host-side logic: the seeded model generators and
workloads, the predictor's feature / target / decode maps, the numpy MLP
reference's gradients, and the bench's counter-summary reader."""
import math
import os
import sys

import numpy as np
import pytest

from conftest import ROOT
from grtm_cuda_amd import learned, models


def test_splitmix64_known_answers():
    """Published splitmix64 outputs (seed 0; seed 1234567 as in the
    SplittableRandom test vectors)."""
    r = models.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert models.SplitMix64(1234567).next_u64() == 0x599ED017FB08FC85


def test_layered_model_distributions():
    for seed in range(20):
        m = models.layered_model(seed, 8)
        h, vp, vs, rho, qp, qs = m.T
        assert m.shape == (9, 6) and h[-1] == 0.0
        assert np.all((h[:-1] >= 1.0) & (h[:-1] <= 8.0))
        trend = 2.8 + (4.6 - 2.8) * np.arange(9) / 8
        assert np.all(np.abs(vs / trend - 1.0) <= 0.05 + 1e-12)
        assert np.all((vp / vs >= 1.70) & (vp / vs <= 1.80))
        assert np.allclose(rho, 1.74 * vp ** 0.25) and np.allclose(qs, 50 * vs ** 2) and np.allclose(qp, 2 * qs)
    assert np.array_equal(models.layered_model(7, 5), models.layered_model(7, 5))


def test_batch_models_and_c4_workload():
    ms = models.batch_models(models.C4_SEED, 1024)
    n = np.array([len(l) for l, _ in ms])
    z = np.array([zs for _, zs in ms])
    assert n.min() == 4 and n.max() == 30 and 15 < n.mean() < 19
    assert z.min() >= 0.0 and z.max() <= 40.0
    again = models.batch_models(models.C4_SEED, 4)
    assert all(np.array_equal(a[0], b[0]) and a[1] == b[1] for a, b in zip(ms[:4], again))
    cfg = models.config("c4", n_models=16)
    vs_min = min(float(l[:, 2].min()) for l, _ in cfg["models"])
    gr = cfg["grid"]
    assert gr["nw"] == 1024 and gr["nk"] == 8192 and len(cfg["dists"]) == 64
    assert math.isclose(gr["kmax"], 1.5 * (math.pi / gr["dt"]) / vs_min)
    assert math.isclose(gr["omega_i"], math.pi / (2 * gr["nw"] * gr["dt"]))


def test_learned_features_targets_decode():
    layers, zs = models.layered_model(3, 6), 12.5
    w = 2 * math.pi * np.arange(8) / 10.0
    f = learned.features(layers, zs, w)
    assert f.shape == (8, learned.N_IN)
    assert np.allclose(f[:, 6:12], (layers[1, 0] / 10, layers[1, 1] / 8, layers[1, 2] / 4.5, layers[1, 3] / 3.3,
                                    math.log10(layers[1, 4]) / 3, math.log10(layers[1, 5]) / 3))
    assert np.all(f[:, 6 * len(layers):180] == 0)
    assert np.allclose(f[:, 180], zs / 40) and np.allclose(f[:, 181], w / learned.W_REF)
    # a label's cut survives targets -> decode
    dk = 0.01
    lab = np.full((8, 26), -1.0, dtype=np.float32)
    lab[:, 0] = 1.0
    lab[:, 1] = np.arange(8) * 50 + 100
    lab[:, 2], lab[:, 3] = 40, 3  # one local maximum
    t, m = learned.targets(lab, layers, zs, 0.0, w, dk)
    assert np.all(m[:, 0] == 1) and np.all(m[:, 1:3] == 1) and np.all(m[:, 3:] == 0)
    kc = learned.decode_kcut(t, layers, zs, 0.0, w)
    assert np.allclose(kc, lab[:, 1] * dk, rtol=1e-6)
    assert np.allclose(learned.decode_kcut(t, layers, zs, 0.0, w, lmargin=0.1), kc * math.exp(0.1), rtol=1e-6)


def test_mlp_reference_gradients(monkeypatch):
    """The numpy reference's backward pass (masked MSE, and the pinball loss
    on the first output) against central differences, with the BF16
    rounding replaced by the identity so the forward pass is smooth."""
    sys.path.insert(0, ROOT)
    from oracle import mlp_ref
    monkeypatch.setattr(mlp_ref, "bf16", lambda x: np.asarray(x, dtype=np.float64))
    rng = np.random.default_rng(5)
    shapes = [(5, 7), (7, 6), (6, 3)]
    n = sum(fi * fo + fo for fi, fo in shapes)
    p = rng.standard_normal(n) * 0.5
    X = rng.standard_normal((9, 5))
    T = rng.standard_normal((9, 3))
    M = (rng.random((9, 3)) > 0.3).astype(np.float64)
    for nq, q in ((0, 0.5), (1, 0.9)):
        loss, g, _ = mlp_ref.loss_and_grads(p, shapes, X, T, M, nq=nq, q=q)
        eps = 1e-6
        for i in rng.choice(n, 40, replace=False):
            dp = np.zeros(n)
            dp[i] = eps
            lp = mlp_ref.loss_and_grads(p + dp, shapes, X, T, M, nq=nq, q=q)[0]
            lm = mlp_ref.loss_and_grads(p - dp, shapes, X, T, M, nq=nq, q=q)[0]
            assert abs((lp - lm) / (2 * eps) - g[i]) < 1e-6 * max(1.0, abs(g[i])), (nq, i)


def test_bench_reads_committed_counter_summary():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.pmc_traffic("k_cells_zr<float,float,false,1>")
    assert t is not None and 1e8 < t < 1e9
    assert bench.pmc_traffic("no_such_kernel") is None
    assert os.path.exists(bench.PMC_SUMMARY)
