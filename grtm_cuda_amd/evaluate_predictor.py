"""Evaluate the learned wavenumber-cut predictor on held-out models.

Held-out models are drawn from the training distribution (3-29 layers over a
half-space, Vs rising 2.8 -> 4.6 km/s, sources U(0.5, 40) km) with seeds
outside the training and calibration ranges.  Each runs on the 1024 x 8192
grid (dt 0.05 s) with 64 surface receivers 5-300 km:

* plain FP64 dense (the reference),
* FP32 dense,
* FP32 adaptive with the analytic cut,
* FP32 adaptive with the learned cut (residual check + dense fallback).

Reports cells kept, rows that fell back, per-trace rel-L2 against the FP64
dense traces and GPU time of each path (HIP events, recursion + contraction
+ FFT).

  python -m grtm_cuda_amd.evaluate_predictor --models 1024 --report report.json
"""
import argparse
import json
import sys
import time

import numpy as np

from . import PREC_FP32, PREC_FP64, TracePlan, learned, models
from .train_predictor import model_sample

SEED0 = 990000


def rel_l2(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--models", type=int, default=1024)
    ap.add_argument("--chunk", type=int, default=32, help="models per progress line")
    ap.add_argument("--report", default=None)
    args = ap.parse_args()

    pred = learned.Predictor()
    dists = np.linspace(5.0, 300.0, 64)
    stats = {k: [] for k in ("cells_learned", "cells_analytic", "fb_learned", "fb_analytic", "err_learned",
                             "err_analytic", "err_dense32", "ms_learned", "ms_analytic", "ms_dense32",
                             "ms_dense64")}
    n_traces = 0
    t0 = time.time()
    for c0 in range(0, args.models, args.chunk):
        ms = [model_sample(SEED0 + i) for i in range(c0, min(args.models, c0 + args.chunk))]
        for l, zs in ms:
            grid = models.dense_grid(1024, 0.05, 8192, float(l[:, 2].min()))
            ref_plan = TracePlan(l, zs, 0.0, dists, grid, precision=PREC_FP64)
            ref = ref_plan.traces()
            stats["ms_dense64"].append(ref_plan.timing(reps=1)["total"])
            del ref_plan
            window = models.check_window(zs, 0.0)
            dense = grid["nw"] * grid["nk"]
            plan = TracePlan(l, zs, 0.0, dists, grid, precision=PREC_FP32)
            stats["err_dense32"].append(float(rel_l2(plan.traces(), ref).max()))
            stats["ms_dense32"].append(plan.timing(reps=1)["total"])
            for name, kcut in (("analytic", models.analytic_kcut(l, zs, 0.0, grid)),
                               ("learned", pred.kcut(l, zs, 0.0, grid))):
                plan.set_adaptive(kcut, window=window, stride=8, tol=3e-6)
                tr = plan.traces()
                st = plan.adaptive_stats()
                stats["cells_" + name].append(st["cells_final"] / dense)
                stats["fb_" + name].append(st["rows_fallback"] / grid["nw"])
                stats["err_" + name].append(float(rel_l2(tr, ref).max()))
                stats["ms_" + name].append(plan.timing(reps=1)["total"])
            n_traces += 2 * len(dists)
        done = min(args.models, c0 + args.chunk)
        print(f"evaluate_predictor: {done}/{args.models} models ({time.time() - t0:.0f} s), "
              f"learned cells {np.mean(stats['cells_learned']):.3f}, max rel-L2 {max(stats['err_learned']):.2e}",
              flush=True)

    a = {k: np.array(v) for k, v in stats.items()}
    rep = dict(
        models=args.models, receivers=len(dists), grid="Nw 1024 x Nk 8192, dt 0.05 s",
        cells_kept=dict(learned=float(a["cells_learned"].mean()), analytic=float(a["cells_analytic"].mean()),
                        dense=1.0),
        rows_fallback=dict(learned=float(a["fb_learned"].mean()), analytic=float(a["fb_analytic"].mean())),
        max_rel_l2=dict(learned=float(a["err_learned"].max()), analytic=float(a["err_analytic"].max()),
                        dense_fp32=float(a["err_dense32"].max())),
        p99_rel_l2_learned=float(np.quantile(a["err_learned"], 0.99)),
        models_over_1e4=dict(learned=int((a["err_learned"] > 1e-4).sum()),
                             analytic=int((a["err_analytic"] > 1e-4).sum()),
                             dense_fp32=int((a["err_dense32"] > 1e-4).sum())),
        gpu_ms_total=dict(learned=float(a["ms_learned"].sum()), analytic=float(a["ms_analytic"].sum()),
                          dense_fp32=float(a["ms_dense32"].sum()), dense_fp64=float(a["ms_dense64"].sum())),
        traces=n_traces,
    )
    rep["speedup_learned_vs_dense_fp64"] = rep["gpu_ms_total"]["dense_fp64"] / rep["gpu_ms_total"]["learned"]
    rep["speedup_learned_vs_dense_fp32"] = rep["gpu_ms_total"]["dense_fp32"] / rep["gpu_ms_total"]["learned"]
    rep["speedup_learned_vs_analytic"] = rep["gpu_ms_total"]["analytic"] / rep["gpu_ms_total"]["learned"]
    print(json.dumps(rep))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
