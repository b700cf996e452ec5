"""Train the wavenumber-grid predictor on the GPU.

Training data come from the plain method itself: dense FP64 runs of seeded
layered models (3-29 layers over a half-space, Vs rising 2.8 -> 4.6 km/s,
source depth U(0.5, 40) km, surface receivers, dt 0.05 s, Nw = 1024,
Nk = 8192), labelled per frequency row on the GPU (truncation index at
1e-6 of the row peak, pole clusters).  The network is trained with its own
GPU forward/backward kernels and saved to predictor_weights.npz.

  python -m grtm_cuda_amd.train_predictor --models 500 --rows 100
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import PREC_FP64, TracePlan, learned, models

# cut label: last wavenumber whose kernel exceeds this fraction of the row peak
LABEL_TAU = 1e-6
QUANTILE = 0.9


def model_sample(seed):
    rng = models.SplitMix64(seed)
    n_total = rng.randint(4, 30)
    layers = models.layered_model(seed * 7919 + 17, n_total - 1)
    src = rng.uniform(0.5, 40.0)
    return layers, src


def label_model(layers, src, nw=1024, nk=8192, dt=0.05):
    grid = models.dense_grid(nw, dt, nk, float(layers[:, 2].min()))
    plan = TracePlan(layers, src, 0.0, [10.0], grid, precision=PREC_FP64)
    plan.execute()
    plan.sync()
    lab = plan.row_labels(LABEL_TAU)
    dk = grid["kmax"] / nk
    w = 2 * math.pi * np.arange(nw) / (2 * nw * dt)
    return grid, lab, dk, w


def build_set(seeds, rows, rng, log_every=50):
    Xs, Ts, Ms, meta = [], [], [], []
    t0 = time.time()
    for i, sd in enumerate(seeds):
        layers, src = model_sample(sd)
        grid, lab, dk, w = label_model(layers, src)
        sel = np.arange(len(w)) if rows is None else np.sort(rng.choice(np.arange(1, len(w)), rows, replace=False))
        X = learned.features(layers, src, w[sel])
        T, M = learned.targets(lab[sel], layers, src, 0.0, w[sel], dk)
        M[:, 0] *= 10.0  # the truncation output drives the grid: weight it above the pole outputs
        Xs.append(X), Ts.append(T), Ms.append(M)
        meta.append(dict(seed=int(sd), layers=layers, src=src, grid=grid, labels=lab, dk=dk, rows=sel))
        if (i + 1) % log_every == 0:
            print(f"  labelled {i + 1}/{len(seeds)} models ({time.time() - t0:.1f} s)", flush=True)
    return np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms), meta


def calibrate(mlp, meta, coverage):
    """Log-margin added to the cut output so that a `coverage` fraction of the
    held-out rows is not undershot (split-conformal calibration)."""
    res = []
    for m in meta:
        grid, lab, dk = m["grid"], m["labels"], m["dk"]
        nw = grid["nw"]
        w = 2 * math.pi * np.arange(nw) / (2 * nw * grid["dt"])
        y = mlp.predict(learned.features(m["layers"], m["src"], w))
        kc = learned.decode_kcut(y, m["layers"], m["src"], 0.0, w, 0.0)
        exact = np.maximum(lab[:, 1], 1.0) * dk
        res.append(np.log(exact) - np.log(kc))
    return float(max(0.0, np.quantile(np.concatenate(res), coverage)))


def evaluate(mlp, meta, margin):
    """Cells kept (fraction of the dense grid) and undershoot statistics."""
    tot = kept_exact = kept_learn = kept_anal = sq = 0.0
    under = rows = 0
    for m in meta:
        grid, lab, dk = m["grid"], m["labels"], m["dk"]
        nw = grid["nw"]
        w = 2 * math.pi * np.arange(nw) / (2 * nw * grid["dt"])
        y = mlp.predict(learned.features(m["layers"], m["src"], w))
        kc = np.minimum(learned.decode_kcut(y, m["layers"], m["src"], 0.0, w, margin), grid["kmax"])
        ka = models.analytic_kcut(m["layers"], m["src"], 0.0, grid)
        exact = lab[:, 1] * dk
        tot += nw * grid["nk"]
        kept_exact += np.ceil(exact / dk).sum()
        kept_learn += np.minimum(np.ceil(kc / dk), grid["nk"]).sum()
        kept_anal += np.minimum(np.ceil(ka / dk), grid["nk"]).sum()
        under += int((kc < exact).sum())
        sq += float(((np.log(np.maximum(kc, 1e-9)) - np.log(np.maximum(exact, 1e-9))) ** 2).sum())
        rows += nw
    return dict(frac_exact=float(kept_exact / tot), frac_learned=float(kept_learn / tot),
                frac_analytic=float(kept_anal / tot), undershoot_rows=float(under / rows),
                kcut_log_rmse=float(np.sqrt(sq / rows)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--models", type=int, default=3125)
    ap.add_argument("--rows", type=int, default=16)
    ap.add_argument("--heldout", type=int, default=64)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--coverage", type=float, default=0.97,
                    help="fraction of calibration rows the margined cut must cover")
    ap.add_argument("--out", default=learned.WEIGHTS)
    ap.add_argument("--report", default=None)
    ap.add_argument("--dump", default=None, help="also save the labelled data set (npz)")
    ap.add_argument("--label-only", action="store_true")
    args = ap.parse_args(argv)
    rng = np.random.default_rng(7)
    print(f"train_predictor: labelling {args.models} training models", flush=True)
    X, T, M, _ = build_set(range(1000, 1000 + args.models), args.rows, rng)
    print(f"train_predictor: {X.shape[0]} samples; labelling {args.heldout} held-out models", flush=True)
    Xh, Th, Mh, held = build_set(range(900000, 900000 + args.heldout), None, rng, log_every=16)
    _, _, _, calib = build_set(range(800000, 800000 + args.heldout), None, rng, log_every=16)
    if args.dump:
        np.savez_compressed(args.dump, X=X, T=T, M=M, Xh=Xh, Th=Th, Mh=Mh)
    if args.label_only:
        return 0
    mlp = learned.MLP()
    mlp.init(seed=3)
    mlp.set_loss(1, QUANTILE)
    mlp.set_data(X, T, M)
    t0 = time.time()
    losses = []
    for phase, (steps, lr) in enumerate(((args.steps // 2, 1e-3), (args.steps // 3, 3e-4),
                                         (args.steps - args.steps // 2 - args.steps // 3, 1e-4))):
        l = mlp.train(steps, args.batch, lr, seed=11 + phase)
        losses.append(l)
        print(f"  phase {phase}: {steps} steps lr {lr:g}: loss {l[:50].mean():.4f} -> {l[-50:].mean():.4f}", flush=True)
    train_s = time.time() - t0
    losses = np.concatenate(losses)
    lmargin = calibrate(mlp, calib, args.coverage)
    print(f"train_predictor: calibrated log-margin {lmargin:.3f} for coverage {args.coverage}", flush=True)
    ev = evaluate(mlp, held, lmargin)
    ev0 = evaluate(mlp, held, 0.0)
    ev["undershoot_rows_uncalibrated"] = ev0["undershoot_rows"]
    ev["frac_learned_uncalibrated"] = ev0["frac_learned"]
    print("train_predictor: held-out", json.dumps(ev), flush=True)
    np.savez(args.out, params=mlp.get_params(), lmargin=np.float32(lmargin),
             meta=json.dumps(dict(models=args.models, rows=args.rows, steps=args.steps, batch=args.batch,
                                  label_tau=LABEL_TAU, quantile=QUANTILE)))
    report = dict(samples=int(X.shape[0]), train_seconds=train_s, first_loss=float(losses[:50].mean()),
                  final_loss=float(losses[-50:].mean()), heldout=ev)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    sys.exit(main())
