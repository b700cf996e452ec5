"""Evaluate the learned wavenumber grid on held-out models against the plain
method.

For every model (the training generator's held-out seeds, 64 surface
receivers at 5-300 km, Nw 1024, Nk 8192, dt 0.05 s) it runs the plain FP64
dense plan and the learned FP32 plan (grid formed on the GPU, residual check,
dense fallback of the rows that fail it) and reports per trace
- the learned path's rel-L2 error against the plain method;
- the residual of a 1/8-density dense check of the tail beyond the learned
  cut: every 8th wavenumber past the cut, weighted 8 dk, contracted with the
  receivers' Bessel functions, relative to the full sum (spectral rel-L2);
  a trace whose residual exceeds 1e-4 is one a per-trace 1/8-density
  criterion would send to the dense grid;
and per frequency row the labelled local maxima of the kernel (pole
clusters, above 1e-6 of the row peak) that lie beyond the checked region
(cut plus the residual-check cells).

  python -m grtm_cuda_amd.eval_grid --models 1024 --report eval.json
"""
import argparse
import json
import time

import numpy as np
import scipy.special as sp

from . import PREC_FP32, PREC_FP64, TracePlan, learned, models
from .train_predictor import model_sample

HELDOUT_SEED = 900000
C5_TOL = 1e-4


def numpy_contract(K, w, J):
    """sum_k K[w, k] w[w, k] J[k, r] on the host: [nw, nr]."""
    return (K * w) @ J


def torch_contract(torch, dev):
    """The same contraction on the GPU through torch (the evaluation's own
    plumbing, not the library's path); J tables go in as numpy arrays."""
    cache = {}

    def f(K, w, J):
        key = id(J)
        if key not in cache:
            cache.clear()
            cache[key] = (J, torch.as_tensor(J, dtype=torch.complex128, device=dev))
        Jt = cache[key][1]
        Kt = torch.as_tensor(np.ascontiguousarray(K), dtype=torch.complex128, device=dev)
        wt = torch.as_tensor(np.ascontiguousarray(w), dtype=torch.float64, device=dev)
        return ((Kt * wt) @ Jt).cpu().numpy()
    return f


def bessel_tables(grid, dists):
    k = (grid["kmax"] / grid["nk"]) * np.arange(grid["nk"])
    return [f(np.outer(k, dists)).astype(np.complex128) for f in (sp.j0, sp.j1)]


def evaluate_model(layers, src, dists, grid, pred, contract, Jt):
    nw, nk = grid["nw"], grid["nk"]
    dk = grid["kmax"] / nk
    p64 = TracePlan(layers, src, 0.0, dists, grid, precision=PREC_FP64)
    t64 = p64.traces()
    lab = p64.row_labels(1e-6)
    pl = TracePlan(layers, src, 0.0, dists, grid, precision=PREC_FP32)
    pred.attach(pl, stride=8, tol=3e-6)
    tl = pl.traces()
    cut, chk = pl.adaptive_rows()
    st = pl.adaptive_stats()
    err = np.linalg.norm(tl - t64, axis=-1) / np.linalg.norm(t64, axis=-1)  # [q, c, r]
    j = np.arange(nk)[None, :]
    check = np.where((j >= cut[:, None]) & ((j - cut[:, None]) % 8 == 4), 8.0 * dk, 0.0)
    full = np.full((1, nk), dk)
    full[0, 0] = 0.5 * dk
    full = np.broadcast_to(full, (nw, nk))
    res = np.zeros_like(err)
    for q in (0, 1):
        K = p64.kernels(q, 0, nw)  # [nw, nk, (Z, R)]
        for c in (0, 1):
            # sum_k K[w, k] weight[w, k] J_c(k r): [nw, nr]
            G = contract(K[..., c], full, Jt[c])
            D = contract(K[..., c], check, Jt[c])
            res[q, c] = np.linalg.norm(D, axis=0) / np.linalg.norm(G, axis=0)
    # labelled local maxima beyond the checked region
    reach = cut + 8 * chk
    beyond = 0
    for i in range(8):
        idx, hw, h = lab[:, 2 + 3 * i], lab[:, 3 + 3 * i], lab[:, 4 + 3 * i]
        beyond += int(((idx >= 0) & (h > 1e-6) & (idx - hw > reach) & (cut < nk)).sum())
    return dict(err=err, res=res, beyond=beyond, cells=st["cells_final"] / (nw * nk),
                fallback=st["rows_fallback"])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--models", type=int, default=1024)
    ap.add_argument("--start", type=int, default=0, help="first held-out model")
    ap.add_argument("--report", default=None)
    args = ap.parse_args(argv)
    import torch
    contract = torch_contract(torch, torch.device("cuda"))
    pred = learned.Predictor()
    dists = np.linspace(5.0, 300.0, 64)
    tot = dict(traces=0, c5_fail=0, err_fail=0, c5_fail_err_fail=0, rows_beyond=0, rows=0, fallback=0)
    worst_err = worst_res = 0.0
    res_all = []
    cells = []
    t0 = time.time()
    last = t0
    for i in range(args.start, args.start + args.models):
        layers, src = model_sample(HELDOUT_SEED + i)
        g = models.dense_grid(1024, 0.05, 8192, float(layers[:, 2].min()))
        r = evaluate_model(layers, src, dists, g, pred, contract, bessel_tables(g, dists))
        c5 = r["res"] > C5_TOL
        bad = r["err"] > C5_TOL
        tot["traces"] += r["err"].size
        tot["c5_fail"] += int(c5.sum())
        tot["err_fail"] += int(bad.sum())
        tot["c5_fail_err_fail"] += int((c5 & bad).sum())
        tot["rows_beyond"] += r["beyond"]
        tot["rows"] += g["nw"]
        tot["fallback"] += r["fallback"]
        worst_err = max(worst_err, float(r["err"].max()))
        worst_res = max(worst_res, float(r["res"].max()))
        res_all.append(r["res"].ravel())
        cells.append(r["cells"])
        if time.time() - last > 30:
            last = time.time()
            print(f"  {i + 1 - args.start}/{args.models} models ({last - t0:.0f} s): {tot}", flush=True)
    out = dict(models=args.models, start=args.start, **tot, worst_trace_err=worst_err,
               worst_check_residual=worst_res,
               check_residual_p99=float(np.quantile(np.concatenate(res_all), 0.99)),
               mean_cells_frac=float(np.mean(cells)), seconds=time.time() - t0)
    print(json.dumps(out), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(out, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
