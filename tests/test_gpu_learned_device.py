"""The learned grid formed on the GPU inside each execution
(grtm_plan_set_learned): the fused matrix-core forward against the training
forward and the FP64 restatement, the device-built grid against the host's
decode of the same network, and the traces against the plain FP64 dense path."""
import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import learned, models

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pred():
    return learned.Predictor()


def test_fused_forward_matches_training_forward_and_restatement(pred):
    from oracle import mlp_ref
    rng = np.random.default_rng(11)
    layers = models.layered_model(5, 12)
    w = 2 * np.pi * rng.uniform(0, 10, 300)
    X = learned.features(layers, 17.0, w)
    print("gpu: fused mlp forward", flush=True)
    a = pred.mlp.predict(X)  # fused matrix-core forward
    b = np.zeros_like(a)
    from grtm_cuda_amd.ffi import P, _lib, check
    Xc = np.ascontiguousarray(X, dtype=np.float32)
    check(_lib().grtm_mlp_predict_train_forward(pred.mlp.ptr, Xc.ctypes.data_as(P), Xc.shape[0], b.ctypes.data_as(P)),
          "train forward")
    # the same BF16 products, FP32 sums in another order
    assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max()
    _, _, y = mlp_ref.loss_and_grads(pred.mlp.get_params(), pred.mlp.shapes(), X, np.zeros_like(a),
                                     np.zeros_like(a))
    assert np.linalg.norm(a - y) <= 2e-3 * np.linalg.norm(y)
    # a planted error in the restatement (a dropped bias of the last layer) is seen
    p = pred.mlp.get_params().copy()
    p[-learned.N_OUT:] += 0.05
    _, _, y_bad = mlp_ref.loss_and_grads(p, pred.mlp.shapes(), X, np.zeros_like(a), np.zeros_like(a))
    assert np.linalg.norm(a - y_bad) > 10 * np.linalg.norm(a - y)


@pytest.fixture(scope="module")
def c2_plans(pred):
    cfg = models.config("c2")
    mk = lambda p: g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                               precision=p)
    print("gpu: c2 fp64 dense reference", flush=True)
    ref = mk(g.PREC_FP64).traces()
    print("gpu: c2 fp32 learned grid on the device", flush=True)
    dev = mk(g.PREC_FP32)
    pred.attach(dev, stride=8, tol=3e-6)
    tr_dev = dev.traces()
    print("gpu: c2 fp32 learned grid from the host", flush=True)
    host = mk(g.PREC_FP32)
    host.set_adaptive(pred.kcut(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["grid"]),
                      window=models.check_window(cfg["src_depth"], cfg["rcv_depth"]), stride=8, tol=3e-6)
    tr_host = host.traces()
    return cfg, ref, dev, tr_dev, host, tr_host


def test_device_grid_matches_host_decode(c2_plans):
    cfg, _, dev, _, host, _ = c2_plans
    cd, kd = dev.adaptive_rows()
    ch, kh = host.adaptive_rows()
    same = cd == ch
    print(f"rows with equal cuts: {same.mean() * 100:.2f}%, max |diff| {np.abs(cd - ch).max()}")
    # the two forwards sum in different orders: a cut may move by one cell
    assert np.abs(cd - ch).max() <= 1 and same.mean() > 0.95
    assert np.array_equal(kd[same], kh[same])
    # the floor follows the plan's FP64 rows, also in the host decode
    assert dev.work()["rows_fp64"] == learned.fp64_rows(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"],
                                                        cfg["grid"])


def test_device_learned_traces(c2_plans):
    cfg, ref, dev, tr_dev, host, tr_host = c2_plans
    rl = np.linalg.norm(tr_dev - ref, axis=-1) / np.linalg.norm(ref, axis=-1)
    pk = np.abs(tr_dev - ref).max(axis=-1) / np.abs(ref).max(axis=-1)
    d = np.linalg.norm(tr_dev - tr_host, axis=-1) / np.linalg.norm(tr_host, axis=-1)
    st = dev.adaptive_stats()
    print(f"device learned grid: {st['cells_final'] / (cfg['grid']['nw'] * cfg['grid']['nk']) * 100:.1f}% cells, "
          f"{st['rows_fallback']} fallback rows; vs FP64 dense rel-L2 {rl.max():.2e} peak {pk.max():.2e}; "
          f"vs host-built grid {d.max():.2e}")
    assert rl.max() < 1e-4 and pk.max() < 1e-4
    assert d.max() < 1e-5
    # every execution rebuilds the grid: a second run is identical
    again = dev.traces()
    assert np.array_equal(again, tr_dev)


def _row_cut(nk, kc, dk, stride, window):
    """The plans' cut and check count of one row (host.hip row_cut)."""
    n = int(min(nk, max(1.0, np.ceil(kc / dk))))
    room = nk - n
    fit = (room - stride // 2 - 1) // stride + 1 if room > stride // 2 else 0
    nc = max(0, min(fit, int(np.ceil(window / (dk * stride)))))
    if room > 0 and nc == 0:
        n = nk
    return n, nc


def test_fp64_rows_take_the_analytic_cut(pred):
    """Mixed-precision plans give the FP64 rows the analytic cut, without the
    network, so their launch runs beside it.  Model 152 of the 1024-model set
    (29 layers, source at 3.96 km) is one where the network's cut exceeds the
    analytic one on every FP64 row: there the rows' cuts and check counts are
    the host's row_cut of the analytic cut, the host decode with low_fixed
    agrees, and the traces stay within 1e-4 of the plain FP64 path (the
    residual check completes any row the shorter cut leaves short)."""
    cfg = models.config("c4")
    lay, zs = cfg["models"][152]
    grid, dists = cfg["grid"], cfg["dists"][:16]
    print("gpu: fp64 rows of model 152 on the analytic cut", flush=True)
    plan = g.TracePlan(lay, zs, 0.0, dists, grid, precision=g.PREC_FP32)
    window = models.check_window(zs, 0.0)
    pred.attach(plan, window=window, stride=8, tol=3e-6)
    tr = plan.traces()
    cut, chk = plan.adaptive_rows()
    r64 = plan.work()["rows_fp64"]
    assert r64 == learned.fp64_rows(lay, zs, 0.0, grid)
    ak = models.analytic_kcut(lay, zs, 0.0, grid)
    dk = grid["kmax"] / grid["nk"]
    want = np.array([_row_cut(grid["nk"], ak[iw], dk, 8, window) for iw in range(r64)])
    assert np.array_equal(cut[:r64], want[:, 0]) and np.array_equal(chk[:r64], want[:, 1])
    fixed = pred.kcut(lay, zs, 0.0, grid)
    floored = pred.kcut(lay, zs, 0.0, grid, low_fixed=False)
    assert np.array_equal(fixed[:r64], ak[:r64])
    # the network's cut is the larger on (nearly) every FP64 row here
    assert np.sum(floored[:r64] > ak[:r64]) >= r64 // 2
    ref = g.TracePlan(lay, zs, 0.0, dists, grid, precision=g.PREC_FP64).traces()
    rl = np.linalg.norm(tr - ref, axis=-1) / np.linalg.norm(ref, axis=-1)
    st = plan.adaptive_stats()
    print(f"  {r64} FP64 rows, {st['rows_fallback']} fallback rows, rel-L2 {rl.max():.2e}")
    assert rl.max() < 1e-4


def _batch_host_cuts(pred, ms, grid):
    kc = np.stack([pred.kcut(l, zs, 0.0, grid, low_fixed=False) for l, zs in ms])
    win = np.array([models.check_window(zs, 0.0) for _, zs in ms])
    return kc, win


@pytest.mark.parametrize("prec", [g.PREC_FP64, g.PREC_FP32])
def test_batch_device_grids_match_host_grids(pred, prec):
    """A batch's learned grids formed inside its graph against the same
    network's cuts decoded on the host: the same cells, fallbacks and traces."""
    ms = models.batch_models(83, 6, n_min=4, n_max=14, src_lo=2.0, src_hi=30.0)
    vs_min = min(float(l[:, 2].min()) for l, _ in ms)
    grid = models.dense_grid(256, 0.05, 2048, vs_min)
    dists = np.linspace(5.0, 150.0, 8)
    print(f"gpu: batch of {len(ms)} models, device learned grids, precision {prec}", flush=True)
    dev = g.BatchPlan(ms, 0.0, dists, grid, precision=prec)
    dev.set_learned(pred, stride=8, tol=3e-6)
    dev.execute()
    tr_dev = dev.traces()
    host = g.BatchPlan(ms, 0.0, dists, grid, precision=prec)
    kc, win = _batch_host_cuts(pred, ms, grid)
    host.set_adaptive(kc, win, stride=8, tol=3e-6)
    host.execute()
    tr_host = host.traces()
    sd, sh = dev.adaptive_stats(), host.adaptive_stats()
    print("device", sd, "host", sh)
    assert sd == sh
    d = np.linalg.norm(tr_dev - tr_host, axis=-1) / np.linalg.norm(tr_host, axis=-1)
    assert d.max() < 1e-6


def test_batch_device_grids_graph_replays(pred):
    """Two groups of device-formed learned grids: graph replays equal the
    eager pass bit for bit and stay within 1e-4 of the dense FP64 batch."""
    ms = models.batch_models(92, 70, n_min=4, n_max=9, src_lo=3.0, src_hi=20.0)
    vs_min = min(float(l[:, 2].min()) for l, _ in ms)
    grid = models.dense_grid(64, 0.1, 512, vs_min)
    dists = np.linspace(5.0, 60.0, 4)
    print("gpu: batch of 70 models, device learned grids, graph replays", flush=True)
    bp = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP32)
    bp.set_learned(pred, stride=8, tol=3e-6)
    bp.execute_eager()
    eager = bp.traces()
    for rep in range(3):
        bp.execute()
        bp.sync()
        assert np.array_equal(bp.traces(), eager), rep
        bp.execute_eager()
        bp.sync()
    ref = g.BatchPlan(ms, 0.0, dists, grid, precision=g.PREC_FP64)
    ref.execute()
    rl = np.linalg.norm(eager - ref.traces(), axis=-1) / np.linalg.norm(ref.traces(), axis=-1)
    assert rl.max() < 1e-4


def test_predicted_far_cluster_extends_the_check(pred):
    """The pole-cluster outputs steer the residual check: a network whose first
    cluster sits at 3 s (s = |w| / Vs_min + 1 / dz) with half width 0.05 km^-1,
    beyond most rows' cut plus window, must make the device grid check every
    such row up to that cluster (and no further elsewhere), and the traces stay
    within 1e-4 of FP64."""
    cfg = models.config("c1")
    grid = cfg["grid"]
    lay, zs, zr = cfg["layers"], cfg["src_depth"], cfg["rcv_depth"]
    far = learned.Predictor()
    p = far.mlp.get_params()
    fo = learned.N_OUT
    nW = learned.WIDTH * fo
    W = p[-(nW + fo):-fo].reshape(learned.WIDTH, fo)
    b = p[-fo:]
    W[:, 1:3] = 0.0
    b[1], b[2] = 3.0, np.log(0.05 * 100)
    p[-(nW + fo):-fo] = W.ravel()
    p[-fo:] = b
    far.mlp.set_params(p)
    mk = lambda: g.TracePlan(lay, zs, zr, cfg["dists"], grid, precision=g.PREC_FP32)
    print("gpu: c1 learned grid with a far predicted cluster", flush=True)
    base, ext = mk(), mk()
    pred.attach(base, stride=8, tol=3e-6)
    far.attach(ext, stride=8, tol=3e-6)
    tr = ext.traces()
    base.traces()
    cut_b, chk_b = base.adaptive_rows()
    cut_e, chk_e = ext.adaptive_rows()
    assert np.array_equal(cut_b, cut_e)  # the cut output is untouched
    nw, nk, dk, stride = grid["nw"], grid["nk"], grid["kmax"] / grid["nk"], 8
    w = 2 * np.pi * np.arange(nw) / (2 * nw * grid["dt"])
    s = np.abs(w) / lay[:, 2].min() + 1.0 / max(abs(zs - zr), 0.5)
    win = models.check_window(zs, zr)
    c, hw = 3.0 * s, 0.05
    # expected: the check reaches c + hw where the cluster starts beyond cut + window
    kc = cut_e * dk  # the cut in cells (the continuous cut lies within one cell below)
    hit = (c - hw > kc + win + dk) & (c - hw < grid["kmax"]) & (cut_e < nk)
    room = nk - cut_e
    fit = np.where(room > stride // 2, (room - stride // 2 - 1) // stride + 1, 0)
    want = np.minimum(fit, np.ceil((c + hw - kc) / (dk * stride)))
    print(f"  rows reaching a far cluster: {hit.sum()} of {nw}; check cells {chk_b.sum()} -> {chk_e.sum()}")
    assert hit.sum() > nw // 8
    assert (np.abs(chk_e[hit] - want[hit]) <= 1).all()
    assert (chk_e[hit] > chk_b[hit]).all() and np.array_equal(chk_e[~hit & (c + hw < kc)], chk_b[~hit & (c + hw < kc)])
    ref = g.TracePlan(lay, zs, zr, cfg["dists"], grid, precision=g.PREC_FP64).traces()
    rl = np.linalg.norm(tr - ref, axis=-1) / np.linalg.norm(ref, axis=-1)
    assert rl.max() < 1e-4
