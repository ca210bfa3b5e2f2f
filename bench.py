#!/usr/bin/env python3
"""Throughput of the snapshot imager on one MI355X.

One step = one SKA-Low-like snapshot (512 stations log-radial within 35 km -> 130,816 rows,
64 channels 50-350 MHz -> 8,372,224 samples, 10% flags, 1024^2 image, cell 2e-5 rad) imaged
from raw device-resident visibilities (uvw, frequencies, XX/YY, flags, WEIGHT_SPECTRUM) to the
dirty image in HBM: FP64 rotation, planning, ES gridding of the Chebyshev moments, 2D inverse
FFT, image-plane w-correction, SIN splat.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--gridder auto|atomic|tiled]
                  [--no-cpu-baseline] [--cpu-samples N]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("snapshots/s and Mvis/s @1024², batched MFMA vs plain FP32, "
          "≤1e-4 err vs FP64 DFT")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
ES_SUPPORT = 8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic(cfg, S_eff, K, P):
    """Algorithmic bytes / flops of one snapshot (per-unit figures times units)."""
    R, C, N = cfg.num_rows, cfg.n_channels, cfg.image_size
    S = R * C
    G = cfg.grid_size
    return {
        "B_vis": 26 * S + 24 * R + 8 * C,
        "B_grid": 8 * P * G * G,
        "B_fft": 32 * P * G * G,
        "B_img": 8 * P * N * N + 8 * N * N,
        "F_grid": 4 * ES_SUPPORT * ES_SUPPORT * K * S_eff,
        "F_fft": 5 * P * G * G * math.log2(G * G),
    }


def cpu_baseline(snap, cfg, n_snapshots):
    """The oracle's FP32 restatement of the reference (OpenMP) on this host."""
    import numpy as np
    from oracle import oracle as O
    threads = O.lib().orc_num_threads()
    times = []
    for _ in range(n_snapshots):
        t0 = time.perf_counter()
        pre = O.preprocess_ref32(snap)
        O.fipipe_device_ref32(pre["bin"], pre["vin"], pre["vis_re"], pre["vis_im"],
                              cfg.image_size, cfg.cell)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": 1.0 / t, "unit": "snapshots/s", "cores": int(threads), "kind": "port",
            "sample": f"{n_snapshots} full snapshot(s) of the same workload "
                      f"({cfg.num_samples} samples, {cfg.image_size}^2), median "
                      f"{t:.2f} s each: FP32 preprocess + gridding + FP64 FFT + image pass",
            "mvis_per_s": cfg.num_samples / t / 1e6}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--gridder", default="auto", choices=["auto", "atomic", "tiled"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-samples", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(snapshots shard across GPUs with no collective, see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import svdimager_amd as S
    from svdimager_amd import synth

    cfg = synth.SKALOW
    log(f"bench: generating {cfg.name} ({cfg.num_samples} samples) ...")
    t0 = time.perf_counter()
    snap = synth.make_snapshot(cfg, device="cuda")
    log(f"bench: generated in {time.perf_counter() - t0:.1f} s")
    dsnap = S.DeviceSnapshot(snap, "cuda")
    w_ok = (snap["flag0"] == 0) | (snap["flag3"] == 0)
    S_eff = int(np.count_nonzero(w_ok))
    gridder = {"auto": S.GRIDDER_AUTO, "atomic": S.GRIDDER_ATOMIC, "tiled": S.GRIDDER_TILED}[args.gridder]

    ctx = S.Context(0)
    stream = torch.cuda.current_stream()
    image = torch.empty((cfg.image_size, cfg.image_size), dtype=torch.float32, device="cuda")
    for i in range(args.warmup):
        _, plan = S.image_snapshot(ctx, dsnap, cfg.image_size, cfg.cell, image=image, gridder=gridder)
        log(f"bench: warmup {i + 1}/{args.warmup} done ({plan['elapsed_ms']:.3f} ms device)")
    ctx.set_profiling(True)
    stage_sum = {k: 0.0 for k in ("preprocess", "plan", "grid", "fft", "image")}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, plan = S.image_snapshot(ctx, dsnap, cfg.image_size, cfg.cell, image=image, gridder=gridder)
        for k, v in ctx.stage_times().items():
            stage_sum[k] += v
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_profiling(False)
    log(f"bench: {args.steps} steps in {elapsed:.3f} s; plan {plan}")
    ms_per_step = elapsed * 1e3 / args.steps
    stages = {k: v / args.steps for k, v in stage_sum.items()}
    snaps_per_s = args.steps / elapsed
    K, P = plan["num_terms"], plan["num_planes"]
    alg = algorithmic(cfg, S_eff, K, P)

    # Dominant stage: the gridder.  One gridding launch per slab batch (one batch here).
    launches = max(1, math.ceil(plan["num_slabs"] / max(1, plan["batch_slabs"])))
    grid_ms = stages["grid"] / launches
    achieved_tflops = alg["F_grid"] / launches / (grid_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": achieved_tflops, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved_tflops / FP32_PEAK_TFLOPS, "traffic": None,
                "kernel": "grid", "kernel_ms": grid_ms,
                "algorithmic": "F_grid = 4 * 8^2 * K * S_unflagged per launch"}
    hbm_bytes = alg["B_vis"] + alg["B_grid"] + alg["B_fft"] + alg["B_img"]

    out = {
        "metric": METRIC, "value": snaps_per_s, "unit": "snapshots/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg.name, "stations": cfg.n_stations, "rows": cfg.num_rows,
                   "channels": cfg.n_channels, "samples": cfg.num_samples,
                   "image_size": cfg.image_size, "grid_size": cfg.grid_size, "cell": cfg.cell,
                   "num_terms": K, "num_slabs": plan["num_slabs"],
                   "gridder": args.gridder, "parallelism": "single-gpu"},
        "mvis_per_s": cfg.num_samples * snaps_per_s / 1e6,
        "stages_ms": stages,
        "path_hbm_bytes_algorithmic": hbm_bytes,
        "path_roofline_cap_snapshots_per_s": HBM_PEAK_GBS * 1e9 / hbm_bytes,
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        log("bench: timing the CPU baseline ...")
        host = {k: snap[k] for k in snap}
        out["cpu_baseline"] = cpu_baseline(host, cfg, args.cpu_samples)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
