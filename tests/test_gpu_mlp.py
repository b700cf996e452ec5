"""The predictor network's GPU kernels: gradients against the FP64 CPU
restatement (oracle/mlp_ref.py), and training actually reducing the loss."""
import numpy as np
import pytest

from grtm_cuda_amd import learned

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nq", [0, 1])
def test_gradients_match_cpu_restatement(nq):
    from oracle import mlp_ref
    mlp = learned.MLP(max_batch=64)
    mlp.init(seed=5)
    if nq:
        mlp.set_loss(nq, 0.9)
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (48, learned.N_IN)).astype(np.float32)
    T = rng.standard_normal((48, learned.N_OUT)).astype(np.float32)
    M = (rng.uniform(size=(48, learned.N_OUT)) > 0.3).astype(np.float32)
    print("gpu: mlp gradient check", flush=True)
    g, loss = mlp.grads(X, T, M)
    ref_loss, ref_g, _ = mlp_ref.loss_and_grads(mlp.get_params(), mlp.shapes(), X, T, M, nq=nq, q=0.9)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    o = 0
    for fi, fo in mlp.shapes():
        for n in (fi * fo, fo):
            a, b = g[o:o + n], ref_g[o:o + n]
            assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b) + 1e-12, (fi, fo, n)
            o += n
    # a planted error in the restatement's gradient (a dropped ReLU mask) is detected
    bad = ref_g.copy()
    bad[: learned.N_IN * learned.WIDTH] *= 1.5
    assert np.linalg.norm(g - bad) > 10 * np.linalg.norm(g - ref_g)


def test_training_reduces_loss():
    mlp = learned.MLP(max_batch=256)
    mlp.init(seed=1)
    rng = np.random.default_rng(4)
    X = rng.uniform(-1, 1, (4096, learned.N_IN)).astype(np.float32)
    W = rng.standard_normal((learned.N_IN, learned.N_OUT)) * 0.2
    T = np.tanh(X @ W).astype(np.float32)
    M = np.ones_like(T)
    mlp.set_data(X, T, M)
    losses = mlp.train(400, batch=256, lr=1e-3)
    assert losses[-20:].mean() < 0.1 * losses[:20].mean()
    Y = mlp.predict(X[:100])
    assert np.mean((Y - T[:100]) ** 2) < 0.1 * np.mean(T[:100] ** 2)
