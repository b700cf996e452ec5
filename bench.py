"""Benchmark: Green's traces/s of the GPU path on the production workload,
learned adaptive-k mixed precision against the plain FP64 dense-k path.

Workload (one GPU): 8 layers over a half-space from a recorded seed,
explosion + vertical force at 12 km, 256 receivers 5-500 km, dt 0.05 s,
Nw = 2048 (N = 4096), Nk = 16384 uniform to 1.5 w_max / Vs_min.  One step =
one pass of recursion + k-sum contraction + inverse FFT producing
2 sources x 256 distances = 512 traces (Z and R each), inputs resident in HBM.

The default measures the learned adaptive path (MLP-predicted per-frequency
wavenumber cut, residual check and dense fallback, FP32 recursion with FP64
vertical wavenumbers and low-frequency rows, split-FP16 MFMA contraction: two
fp16 terms of each row-scaled FP32 operand, three products, FP32
accumulation); the
plain FP64 dense path runs in the same process as the denominator and as the
accuracy reference (max trace rel-L2 reported).

Prints one JSON line (rank 0).  The CPU baseline is the FP64 oracle on a
bounded sample of the same workload (a block of frequency rows), run on the
host cores and scaled to the full job.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0


def prec_code(g, name):
    return {"fp64": g.PREC_FP64, "fp32": g.PREC_FP32, "bf16x3": g.PREC_BF16X3, "bf16x6": g.PREC_BF16X6,
            "fp16x3": g.PREC_FP16X3}[name]


def dtype_name(g, prec):
    return {g.PREC_FP64: "f64", g.PREC_FP32: "f32", g.PREC_BF16X3: "f32+bf16x3", g.PREC_BF16X6: "f32+bf16x6",
            g.PREC_FP16X3: "f32+fp16x3"}[prec]


def ksum_peak(g, prec):
    """Dense matrix-core peak of the contraction's arithmetic (split-BF16: the
    bf16 rate over the products per term)."""
    return {g.PREC_FP64: FP64_PEAK_TFLOPS, g.PREC_FP32: FP32_PEAK_TFLOPS, g.PREC_BF16X3: BF16_PEAK_TFLOPS / 3,
            g.PREC_BF16X6: BF16_PEAK_TFLOPS / 6, g.PREC_FP16X3: BF16_PEAK_TFLOPS / 3}[prec]


def host_cpus():
    """The host cores this process may use (its CPU affinity, capped by
    OMP_NUM_THREADS where the machine sets a share), and what the host reports."""
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(aff, int(omp)) if omp.isdigit() and int(omp) > 0 else aff
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return threads, dict(nproc=os.cpu_count(), affinity=aff, omp_num_threads=omp or None, cpu_model=model)


def cpu_baseline(cfg, rows, rows_1core):
    """FP64 oracle (-O3, OpenMP over frequency rows) on frequency rows
    [nw/2, nw/2 + rows) of the workload, both sources, on every host core this
    process may use, and on one core over `rows_1core` rows; each scaled to
    the full job."""
    from oracle import oracle
    nw = cfg["grid"]["nw"]
    lo = nw // 2
    traces = 2 * len(cfg["dists"])
    threads, host = host_cpus()

    def run(nrows, nthreads):
        t0 = time.perf_counter()
        for kind in (0, 1):
            oracle.traces(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], kind, cfg["grid"],
                          w_range=(lo, lo + nrows), nthreads=nthreads)
        return time.perf_counter() - t0

    dt = run(rows, threads)
    dt1 = run(rows_1core, 1)
    return dict(value=traces / (dt * nw / rows), unit="traces/s", cores=threads, kind="port",
                sample=f"FP64 oracle, {rows} of {nw} frequency rows x {cfg['grid']['nk']} k x "
                       f"{len(cfg['dists'])} receivers, both sources, {dt:.1f} s on {threads} threads, "
                       f"scaled x{nw / rows:.0f}",
                value_1core=traces / (dt1 * nw / rows_1core),
                sample_1core=f"{rows_1core} rows, {dt1:.1f} s on 1 thread, scaled x{nw / rows_1core:.0f}",
                host=host)


def time_plan(plan, steps, warmup):
    for _ in range(warmup):
        plan.execute()
    plan.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute()
    plan.sync()
    return (time.perf_counter() - t0) / steps


def time_batch(plan, steps, warmup):
    for _ in range(warmup):
        plan.execute()
    plan.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute()
    plan.sync()
    return (time.perf_counter() - t0) / steps


def bench_batch(args, g, models):
    """Batched workload: many layered models in one captured graph, mixed
    precision against the plain FP64 batch."""
    cfg = models.config("c4", n_models=args.models)
    pname = args.precision or ("fp16x3" if args.path == "adaptive" else "fp32")
    prec = prec_code(g, pname)
    mk = lambda p: g.BatchPlan(cfg["models"], cfg["rcv_depth"], cfg["dists"], cfg["grid"], precision=p)
    nm = len(cfg["models"])
    print(f"bench: c4 plain FP64 batch of {nm} models", file=sys.stderr, flush=True)
    t_build = time.perf_counter()
    ref = mk(g.PREC_FP64)
    build_ref = time.perf_counter() - t_build
    t_ref = time_batch(ref, max(1, args.steps // 2), 1)
    print(f"bench: c4 {pname} batch, {args.path} grids", file=sys.stderr, flush=True)
    t_build = time.perf_counter()
    plan = mk(prec)
    build = time.perf_counter() - t_build
    extra = {}
    if args.path == "adaptive":
        # every model's learned grid is formed on the GPU inside the graph
        from grtm_cuda_amd import learned
        plan.set_learned(learned.Predictor(), stride=8, tol=3e-6)
    t_step = time_batch(plan, args.steps, args.warmup)
    if args.path == "adaptive":
        st = plan.adaptive_stats()
        dense_cells = nm * cfg["grid"]["nw"] * cfg["grid"]["nk"]
        extra.update(cells_frac=st["cells_final"] / dense_cells, check_frac=st["cells_check"] / dense_cells,
                     rows_fallback_frac=st["rows_fallback"] / (nm * cfg["grid"]["nw"]))
    worst = worst_pk = 0.0
    for m0 in range(0, nm, 64):
        a, b = plan.traces(m0, min(nm, m0 + 64)), ref.traces(m0, min(nm, m0 + 64))
        err = np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)
        worst = max(worst, float(err.max()))
        worst_pk = max(worst_pk, float((np.abs(a - b).max(axis=-1) / np.abs(b).max(axis=-1)).max()))
    work = plan.work()
    tm = plan.timing(reps=1)
    if args.path == "adaptive":
        extra.update(predictor_ms=round(tm["predictor"], 3), predictor_share=round(tm["predictor"] / tm["total"], 4),
                     predictor_in_step=True)
    peak = FP64_PEAK_TFLOPS if prec == g.PREC_FP64 else FP32_PEAK_TFLOPS
    # counted recursion flops of the cells actually solved, each with its upward
    # sweep as cut by depth (adaptive: the predicted cells, plus the check cells)
    frac = 1.0 + extra.get("check_frac", 0.0) / extra["cells_frac"] if extra else 1.0
    achieved = frac * work["rec_flop_solved"] / (tm["cells"] * 1e-3) / 1e12
    layers = [len(l) for l, _ in cfg["models"]]
    return {
        "metric": "Green's traces/s at <=1e-4 rel-L2 vs FP64; learned-adaptive vs plain dense-k",
        "value": work["traces"] / t_step,
        "unit": "traces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype_name(g, prec),
        "data": "synthetic (seeded layered models)",
        "config": {"workload": f"c4: {nm} models of {min(layers)}-{max(layers)} layers (mean {np.mean(layers):.1f}), "
                               f"sources 0-40 km, 64 receivers 5-300 km, Nw={cfg['grid']['nw']}, "
                               f"Nk={cfg['grid']['nk']}, dt={cfg['grid']['dt']}",
                   "path": f"batched {'learned adaptive' if args.path == 'adaptive' else 'dense'} {pname}, "
                           f"one captured HIP graph", "traces_per_step": work["traces"]},
        "plain_fp64_batch": {"value": work["traces"] / t_ref, "ms_per_step": 1e3 * t_ref},
        "speedup_vs_plain_fp64": t_ref / t_step,
        "max_rel_l2_vs_fp64": worst,
        "max_peak_err_vs_fp64": worst_pk,
        "batch": dict(plan.config(), build_s=round(build, 2), build_fp64_s=round(build_ref, 2)),
        "adaptive": extra or None,
        "kernel_ms": {"recursion": round(tm["cells"], 3), "ksum": round(tm["ksum"], 3), "ifft": round(tm["ifft"], 3)},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": None, "kernel": "k_cells_zr_batch", "flops": frac * work["rec_flop_solved"],
                     "cells": (extra["cells_frac"] + extra.get("check_frac", 0.0)) * work["cells"] if extra
                     else work["cells"]},
        "cpu_baseline": None,
    }


def bench_precision(args, g, models):
    """Precision study: the 3-layer model and geometry, every trace of each
    mixed precision against the plain FP64 path (itself within 1e-7 of the
    FP64 oracle in the tests); and the contraction kernels at the production
    shape (K = 16384 wavenumbers, 256 receivers, M = 4 x 2048 operand rows,
    i.e. four [2048 x 16384] x [16384 x 256] products)."""
    cfg = models.config("c1")
    ref = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                      precision=g.PREC_FP64).traces()
    errors = {}
    for pname, prec in (("fp32", g.PREC_FP32), ("bf16x3", g.PREC_BF16X3), ("bf16x6", g.PREC_BF16X6)):
        tr = g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                         precision=prec).traces()
        d = tr - ref
        rl = np.linalg.norm(d, axis=-1) / np.linalg.norm(ref, axis=-1)
        pk = np.abs(d).max(axis=-1) / np.abs(ref).max(axis=-1)
        errors[pname] = dict(max_rel_l2=float(rl.max()), median_rel_l2=float(np.median(rl)),
                             max_peak=float(pk.max()), traces=int(rl.size))
    c2 = models.config("c2")
    gemm = {}
    for pname in ("fp64", "fp32", "bf16x3", "bf16x6"):
        prec = prec_code(g, pname)
        peak = ksum_peak(g, prec)
        plan = g.TracePlan(c2["layers"], c2["src_depth"], c2["rcv_depth"], c2["dists"], c2["grid"], precision=prec)
        plan.execute()
        tm = plan.timing(reps=max(2, args.steps // 3))
        flop = plan.work()["ksum_flop"]
        gemm[pname] = dict(ms=tm["ksum"], tflops=flop / (tm["ksum"] * 1e-3) / 1e12,
                           frac=flop / (tm["ksum"] * 1e-3) / 1e12 / peak, peak=peak)
    fp32 = gemm["fp32"]
    return {
        "metric": "Green's traces/s at <=1e-4 rel-L2 vs FP64; learned-adaptive vs plain dense-k",
        "value": None, "unit": "traces/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 / f32+bf16x3 / f32+bf16x6", "data": "synthetic (3-layer model; seeded 8-layer model for the contraction)",
        "config": {"workload": "c3: 3 layers over a half-space, explosion at 10 km, 8 receivers 10-80 km, "
                               "Nw=512, Nk=4096, dt=0.1 (errors); contraction at Nw=2048, Nk=16384, 256 receivers",
                   "path": "precision study"},
        "errors_vs_fp64": errors,
        "contraction": gemm,
        "roofline": {"bound": "mfma", "achieved": fp32["tflops"], "peak": fp32["peak"], "unit": "TFLOP/s",
                     "frac": fp32["frac"], "traffic": None, "kernel": "k_gemm_f32s"},
        "cpu_baseline": None,
    }


def bench_columns(args, g, models):
    """Displacement columns: the 8-layer / 256-receiver workload with every
    tensor (0-9) and single-force (0-4) column of green()'s output, i.e.
    15 x 256 column traces per step; the learned adaptive grid (or --path
    dense) in mixed precision against the plain FP64 dense column plan
    (per-column rel-L2 and peak error)."""
    cfg = models.config("c2")
    dv = bool(args.derivatives)
    mk = lambda p: g.ColumnPlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                                precision=p, derivatives=dv)
    print("bench: columns plain FP64 dense", file=sys.stderr, flush=True)
    ref_plan = mk(g.PREC_FP64)
    t_ref = time_plan(ref_plan, max(3, args.steps // 2), 1)
    ref = ref_plan.traces()
    ref_plan.timing(reps=1)
    tm_ref = ref_plan.timing(reps=3)
    del ref_plan
    pname = args.precision or "fp32"
    prec = prec_code(g, pname)
    print(f"bench: columns {args.path} {pname}", file=sys.stderr, flush=True)
    plan = mk(prec)
    extra = None
    if args.path == "adaptive":
        from grtm_cuda_amd import learned
        kcut = learned.Predictor().kcut(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["grid"])
        plan.set_adaptive(kcut, window=models.check_window(cfg["src_depth"], cfg["rcv_depth"]), stride=8, tol=3e-6)
    t_step = time_plan(plan, args.steps, args.warmup)
    if args.path == "adaptive":
        st = plan.adaptive_stats()
        dense_cells = cfg["grid"]["nw"] * cfg["grid"]["nk"]
        extra = dict(cells_frac=st["cells_final"] / dense_cells, check_frac=st["cells_check"] / dense_cells,
                     rows_fallback=st["rows_fallback"])
    out = plan.traces()
    work = plan.work()
    plan.timing(reps=1)
    tm = plan.timing(reps=5)
    d = out - ref
    nrm = np.linalg.norm(ref, axis=-1)
    rl = np.linalg.norm(d, axis=-1) / np.maximum(nrm, 1e-6 * nrm.max(axis=-1, keepdims=True))
    pk = np.abs(d).max(axis=-1) / np.abs(ref).max(axis=-1)
    kernels = {"recursion": tm["cells"], "ksum": tm["ksum"], "ifft": tm["ifft"]}
    peak = FP64_PEAK_TFLOPS if prec == g.PREC_FP64 else FP32_PEAK_TFLOPS
    frac = extra["cells_frac"] if extra else 1.0  # kept share of the cells and of the contraction's K
    cells = work["main_cells"]  # the main recursion launch, timed alone
    rec = cells * work["main_flops_per_cell"] / (tm["main"] * 1e-3) / 1e12
    con = frac * work["ksum_flop"] / (tm["ksum"] * 1e-3) / 1e12
    if tm["cells"] >= tm["ksum"]:
        roof = dict(bound="valu", achieved=rec, peak=peak, unit="TFLOP/s", frac=rec / peak, traffic=None,
                    kernel="k_cells_mt", launch_ms=tm["main"], flops_per_cell=work["main_flops_per_cell"],
                    flops_per_cell_full_sweep=work["flops_per_cell"], cells=cells)
    else:
        roof = dict(bound="mfma", achieved=con, peak=peak, unit="TFLOP/s", frac=con / peak, traffic=None,
                    kernel=("k_gemm_f32sk" if extra else "k_gemm_f32t") if prec == g.PREC_FP32 else "k_gemm_f64s")
    ntr = work["traces"]
    return {
        "metric": ("displacement + strain" if dv else "displacement") +
                  "-column traces/s at <=1e-4 rel-L2 vs FP64 (plain dense-k FP64 alongside)",
        "value": ntr / t_step, "unit": "column traces/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype_name(g, prec),
        "data": "synthetic (seeded layered model)",
        "config": {"workload": f"mt: {len(cfg['layers']) - 1} layers + half-space, tensor and single-force sources "
                               f"at {cfg['src_depth']} km, {len(plan.columns)} columns"
                               f"{' (displacement, d/dz, d/dr)' if dv else ''} x {len(cfg['dists'])} receivers, "
                               f"Nw={cfg['grid']['nw']}, Nk={cfg['grid']['nk']}",
                   "path": f"{args.path} {pname}", "traces_per_step": ntr},
        "plain_fp64_dense": {"value": ntr / t_ref, "ms_per_step": 1e3 * t_ref,
                             "kernel_ms": {"recursion": round(tm_ref["cells"], 4), "ksum": round(tm_ref["ksum"], 4),
                                           "ifft": round(tm_ref["ifft"], 4)}},
        "speedup_vs_plain_fp64": t_ref / t_step,
        "max_rel_l2_vs_fp64": float(rl.max()),
        "max_peak_err_vs_fp64": float(pk.max()),
        "max_rel_l2_per_column": [float(v) for v in rl.max(axis=1)],
        "adaptive": extra,
        "kernel_ms": {k: round(v, 4) for k, v in kernels.items()},
        "contraction_tflops": con,
        "recursion_tflops": rec,
        "roofline": roof,
        "cpu_baseline": None,
    }


PMC_SUMMARY = os.path.join(ROOT, "profiles", "c2_adaptive_pmc_bytes.md")


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, each scaled
    by its reading for a plain copy of known size) from the committed counter
    summary of the default workload, or None."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "TRAFFIC_MB" and parts[1] == kernel:
                    return float(parts[2]) * 1e6
    except OSError:
        pass
    return None


def roofline(g, plan, prec, tm, work, counted=False):
    kernels = {"recursion": tm["cells"], "ksum": tm["ksum"], "ifft": tm["ifft"]}
    dom = max(kernels, key=kernels.get)
    peak = FP64_PEAK_TFLOPS if prec == g.PREC_FP64 else FP32_PEAK_TFLOPS
    if dom == "recursion":
        # the main recursion launch (FP32 rows of a mixed plan, all rows of an
        # FP64 plan) timed alone with HIP events on the plan's stream
        cells = work["main_cells"]
        # executed flops: each cell's count with its upward sweep as cut by depth
        achieved = cells * work["main_flops_per_cell"] / (tm["main"] * 1e-3) / 1e12
        traffic = pmc_traffic("k_cells_zr<float,float,false,1>") if counted else None
        return dict(bound="valu", achieved=achieved, peak=peak, unit="TFLOP/s", frac=achieved / peak, traffic=traffic,
                    traffic_source=os.path.relpath(PMC_SUMMARY, ROOT) if traffic else None,
                    kernel="k_cells_zr", launch_ms=tm["main"], flops_per_cell=work["main_flops_per_cell"],
                    flops_per_cell_full_sweep=work["flops_per_cell"],
                    transcendentals_per_cell=work["trans_per_cell"], cells=cells), kernels
    if dom == "ksum":
        achieved = work["ksum_flop"] / (tm["ksum"] * 1e-3) / 1e12
        kpeak = ksum_peak(g, prec)
        return dict(bound="mfma", achieved=achieved, peak=kpeak, unit="TFLOP/s", frac=achieved / kpeak,
                    traffic=None, kernel="k_gemm_mfma"), kernels
    achieved = work["ifft_bytes"] / (tm["ifft"] * 1e-3) / 1e9
    return dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                traffic=None, kernel="k_post_ifft"), kernels


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "mt"],
                    help="c2: one model, 256 receivers (the headline); c3: precision study; "
                         "c4: batch of --models models; mt: all displacement columns of the c2 model")
    ap.add_argument("--models", type=int, default=1024, help="c4: models in the batch")
    ap.add_argument("--path", default="adaptive", choices=["adaptive", "dense"],
                    help="adaptive: learned truncated grid, mixed precision (the headline); dense: plain grid")
    ap.add_argument("--precision", default=None, choices=["fp64", "fp32", "bf16x3", "bf16x6", "fp16x3"],
                    help="arithmetic of the measured path (default adaptive: FP32 recursion with the split-FP16 "
                         "contraction, fp16x3; dense: fp64)")
    ap.add_argument("--derivatives", action="store_true", help="mt: the d/dz and d/dr columns as well")
    ap.add_argument("--cpu-rows", type=int, default=512, help="CPU baseline sample: frequency rows on all cores")
    ap.add_argument("--cpu-rows-1core", type=int, default=48, help="CPU baseline sample: frequency rows on one core")
    ap.add_argument("--flag64", default=None, metavar="B,P",
                    help="c1/c2 mixed plans: also re-solve in FP64 the cells within B cells of every branch point "
                         "and (learned grid) within P half widths of each predicted pole cluster")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="c2: leave out the 1024-model batch figure")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one plain 256 MiB device copy (calibrates hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py: this path runs on one GPU; multi-GPU runs are not part of this build (use --gpus 1)")

    import grtm_cuda_amd as g
    from grtm_cuda_amd import models

    if args.workload == "c4":
        print(json.dumps(bench_batch(args, g, models)))
        return
    if args.workload == "c3":
        print(json.dumps(bench_precision(args, g, models)))
        return
    if args.workload == "mt":
        print(json.dumps(bench_columns(args, g, models)))
        return
    if args.copy_probe:
        from grtm_cuda_amd.ffi import _lib, check
        check(_lib().grtm_selftest_copy(64 << 20), "grtm_selftest_copy")
    cfg = models.config(args.workload)
    pname = args.precision or ("fp16x3" if args.path == "adaptive" else "fp64")
    prec = prec_code(g, pname)
    mk = lambda p: g.TracePlan(cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], cfg["grid"],
                               precision=p)

    # the plain FP64 dense-k path: the denominator and the accuracy reference
    print(f"bench: {args.workload} plain FP64 dense", file=sys.stderr, flush=True)
    ref_plan = mk(g.PREC_FP64)
    t_ref = time_plan(ref_plan, max(3, args.steps // 2), 1)
    ref = ref_plan.traces()
    work = ref_plan.work()
    traces_per_step = work["traces"]

    extra = {}
    if args.path == "dense" and prec == g.PREC_FP64:
        plan, t_step, out = ref_plan, t_ref, ref
    else:
        plan = mk(prec)
        if args.path == "adaptive":
            # the learned grid is formed on the GPU inside every execution:
            # network, per-row cuts, residual-check cells and item list
            from grtm_cuda_amd import learned
            learned.Predictor().attach(plan, window=models.check_window(cfg["src_depth"], cfg["rcv_depth"]),
                                       stride=8, tol=3e-6)
        if args.flag64:
            hb, pw = args.flag64.split(",")
            plan.set_flag64(int(hb), float(pw))
        print(f"bench: {args.path} {pname}", file=sys.stderr, flush=True)
        t_step = time_plan(plan, args.steps, args.warmup)
        out = plan.traces()
        if args.path == "adaptive":
            st = plan.adaptive_stats()
            dense_cells = cfg["grid"]["nw"] * cfg["grid"]["nk"]
            extra.update(cells_frac=st["cells_final"] / dense_cells, check_frac=st["cells_check"] / dense_cells,
                         rows_fallback=st["rows_fallback"])
        if args.flag64:
            extra.update(flag64=args.flag64, cells_fp64_flagged=int(plan.work()["cells_refixed"]))
    err = np.linalg.norm(out - ref, axis=-1) / np.linalg.norm(ref, axis=-1)
    peak = np.abs(out - ref).max(axis=-1) / np.abs(ref).max(axis=-1)
    value = traces_per_step / t_step
    plan.timing(reps=1)  # warm: the stage split is taken on the steady state
    tm = plan.timing(reps=5)
    if args.path == "adaptive":
        extra.update(predictor_ms=round(tm["predictor"], 4), predictor_share=round(tm["predictor"] / tm["total"], 4),
                     predictor_in_step=True)
    # the committed counter summary was taken on the default run (c2, learned
    # adaptive; the recursion kernel is the same for every contraction precision)
    counted = args.workload == "c2" and args.path == "adaptive" and prec != g.PREC_FP64
    roof, kernels = roofline(g, plan, prec, tm, plan.work(), counted)

    # the 1024-model batch end to end (learned grids formed inside its graph)
    c4 = None
    if args.workload == "c2" and args.path == "adaptive" and not args.no_c4:
        import copy
        a4 = copy.copy(args)
        a4.steps, a4.warmup, a4.precision = 3, 1, None
        print("bench: c4 batch", file=sys.stderr, flush=True)
        b = bench_batch(a4, g, models)
        c4 = {k: b[k] for k in ("value", "unit", "ms_per_step", "speedup_vs_plain_fp64", "max_rel_l2_vs_fp64",
                                "max_peak_err_vs_fp64")}
        c4.update(workload=b["config"]["workload"], path=b["config"]["path"], steps=a4.steps,
                  plain_fp64_batch=b["plain_fp64_batch"], adaptive=b["adaptive"], kernel_ms=b["kernel_ms"])

    cpu = None
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(cfg, args.cpu_rows, args.cpu_rows_1core)

    line = {
        "metric": "Green's traces/s at <=1e-4 rel-L2 vs FP64; learned-adaptive vs plain dense-k",
        "value": value,
        "unit": "traces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype_name(g, prec),
        "data": "synthetic (seeded layered model)",
        "config": {"workload": f"{args.workload}: {len(cfg['layers']) - 1} layers + half-space, explosion + "
                               f"vertical force at {cfg['src_depth']} km, {len(cfg['dists'])} receivers, "
                               f"Nw={cfg['grid']['nw']}, Nk={cfg['grid']['nk']}",
                   "path": f"{args.path} {pname}", "traces_per_step": traces_per_step},
        "plain_fp64_dense": {"value": traces_per_step / t_ref, "ms_per_step": 1e3 * t_ref},
        "speedup_vs_plain_fp64": t_ref / t_step,
        "max_rel_l2_vs_fp64": float(err.max()),
        "max_peak_err_vs_fp64": float(peak.max()),
        "adaptive": extra or None,
        "kernel_ms": {k: round(v, 4) for k, v in kernels.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "c4_batch": c4,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
