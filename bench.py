#!/usr/bin/env python3
"""Throughput of the snapshot imager on one MI355X.

Workload: SKA-Low-like snapshots (512 stations log-radial within 35 km -> 130,816 rows, 64
channels 50-350 MHz -> 8,372,224 samples, 10% flags, 1024^2 image, cell 2e-5 rad) imaged from
raw device-resident visibilities (uvw, frequencies, XX/YY, flags, WEIGHT_SPECTRUM) to the dirty
image in HBM: FP64 rotation, planning, ES gridding of the Chebyshev moments, 2D inverse FFT,
image-plane w-correction, SIN splat.  One step = one batched call imaging `--batch` (8)
consecutive snapshots of the observation, 10 s apart, on concurrent lanes; `value` counts
snapshots.  Single-snapshot calls are timed too.  Prints one JSON line.

The headline gridder is the MFMA one (each 16 x 16 grid window contracted over its samples on
f16 hi/lo MFMA with FP32 accumulation); the plain FP32 tiled imager is timed in the same run on
the same snapshot.  The accuracy of both against an FP64 direct DFT is measured (untimed) at 64 seeded
pixels plus the peak pixel of the signed, pre-reprojection image.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--gridder mfma|tiled|atomic|auto]
                  [--batch B] [--no-compare] [--no-accuracy] [--no-cpu-baseline]
                  [--cpu-samples N]
  python bench.py --copy-calibration MIB     (plain copy, the PMC byte counters' reference)
  python bench.py --serve 50,100,200,400 [--serve-seconds T] [--serve-pool P]
                                             (streaming form: latency under Poisson arrivals)
  python bench.py --precision                (4096^2 form: error against the FP64 DFT)
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
FP32_PEAK_TFLOPS = 157.3          # MI355X FP32 matrix peak
FP16_PEAK_TFLOPS = 2500.0         # MI355X dense FP16 MFMA peak
ES_SUPPORT = 8
C_LIGHT = 299792458.0


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


def host_cpu():
    """CPU model and logical CPU count of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def cpu_baseline(snap, cfg, n_snapshots, small=None):
    """The oracle's FP32 restatement of the reference (OpenMP) on this host: `n_snapshots` runs
    of the workload's snapshot (median), and the small CPU-runnable workload in full."""
    import numpy as np
    from oracle import oracle as O
    threads = O.lib().orc_num_threads()

    def run(s, c):
        t0 = time.perf_counter()
        pre = O.preprocess_ref32(s)
        O.fipipe_device_ref32(pre["bin"], pre["vin"], pre["vis_re"], pre["vis_im"], c.image_size, c.cell)
        return time.perf_counter() - t0

    times = [run(snap, cfg) for _ in range(n_snapshots)]
    t = float(np.median(times))
    out = {"value": 1.0 / t, "unit": "snapshots/s", "cores": int(threads), "kind": "port",
           "sample": f"{n_snapshots} full snapshot(s) of the same workload "
                     f"({cfg.num_samples} samples, {cfg.image_size}^2), median "
                     f"{t:.2f} s each (runs {', '.join(f'{x:.2f}' for x in times)} s): FP32 "
                     f"preprocess + gridding + FP64 FFT + image pass",
           "mvis_per_s": cfg.num_samples / t / 1e6, **host_cpu()}
    if small is not None:
        s_snap, s_cfg = small
        ts = [run(s_snap, s_cfg) for _ in range(3)]
        out["small_workload"] = {"workload": s_cfg.name, "samples": s_cfg.num_samples,
                                 "image_size": s_cfg.image_size, "seconds_median": float(np.median(ts)),
                                 "snapshots_per_s": 1.0 / float(np.median(ts))}
    return out


def measured_traffic(kernel, workload):
    """HBM bytes per dispatch of `kernel` on `workload` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the reading of a plain copy of known
    size), or None when no such measurement is recorded."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    rec = table.get(workload, {}).get(kernel)
    if not rec:
        return None
    return {"hbm_bytes_per_dispatch": rec["hbm_bytes_per_dispatch"],
            "source": f"profiles/traffic.json ({rec.get('summary', '')})"}


def time_gridder(S, ctx, dsnap, cfg, image, gridder, steps, warmup, label, **opts):
    """Warm up, then time `steps` snapshots; returns (seconds, mean stage ms, last plan)."""
    import torch
    for i in range(warmup):
        _, plan = S.image_snapshot(ctx, dsnap, cfg.image_size, cfg.cell, image=image, gridder=gridder, **opts)
        log(f"bench[{label}]: warmup {i + 1}/{warmup} done ({plan['elapsed_ms']:.3f} ms device)")
    ctx.set_profiling(True)
    stage_sum = {k: 0.0 for k in S.STAGES}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        _, plan = S.image_snapshot(ctx, dsnap, cfg.image_size, cfg.cell, image=image, gridder=gridder, **opts)
        for k, v in ctx.stage_times().items():
            stage_sum[k] += v
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_profiling(False)
    log(f"bench[{label}]: {steps} steps in {elapsed:.3f} s")
    return elapsed, {k: v / steps for k, v in stage_sum.items()}, plan


def accuracy(S, ctx, snap, cfg, gridders, n_pixels=64, all_pixels=True):
    """Max / RMS error of the signed image against an FP64 direct DFT (torch, on the GPU).

    `gridders` maps a name to the imaging options (gridder, mfma_precision) of one imager.
    Samples whose ES footprint is not wholly inside the grid are flagged for this comparison
    (the gridder clips them, the DFT cannot), on both sides; S in the 1/S scaling still counts
    every sample.  Pixels: 64 seeded ones plus the peak of the first gridder's image."""
    import numpy as np
    import torch
    dev = "cuda"
    N, cell, G = cfg.image_size, cfg.cell, cfg.grid_size
    R, C = cfg.num_rows, cfg.n_channels
    dsnap = S.DeviceSnapshot(snap, dev)
    rot, _ = S.snapshot_geometry(ctx, dsnap)
    rot_t = torch.from_numpy(rot).to(dev)
    uvw = dsnap.tensors["uvw"]
    freq = dsnap.tensors["freq"]
    ruvw = uvw @ rot_t.T                                   # (R, 3) rotated, metres
    scale = freq / C_LIGHT                                 # (C,)
    b = (ruvw[None, :, :] * scale[:, None, None]).reshape(R * C, 3)   # idx = row + chan R
    p = b[:, :2] * (cell * G)
    lo = torch.ceil(p - 4.0)
    hi = torch.floor(p + 4.0)
    inside = ((lo >= -(G // 2)) & (hi <= (G - 1) // 2)).all(dim=1)
    t = dsnap.tensors
    f0 = t["flag0"] | (~inside).to(torch.uint8)
    f3 = t["flag3"] | (~inside).to(torch.uint8)
    acc_snap = dict(snap)
    acc_snap["flag0"] = f0
    acc_snap["flag3"] = f3
    asnap = S.DeviceSnapshot(acc_snap, dev)
    w0 = torch.where(f0 != 0, 0.0, t["weight0"].double())
    w3 = torch.where(f3 != 0, 0.0, t["weight3"].double())
    ws = w0 + w3
    keep = ws > 0
    vre = torch.where(keep, (t["vis0_re"].double() * w0 + t["vis3_re"].double() * w3) / ws, 0.0)
    vim = torch.where(keep, (t["vis0_im"].double() * w0 + t["vis3_im"].double() * w3) / ws, 0.0)
    det = abs(rot[0, 0] * rot[1, 1] - rot[0, 1] * rot[1, 0])

    images = {}
    plan = None
    for name, g in gridders.items():
        img, plan = S.image_snapshot(ctx, asnap, N, cell, output=S.OUTPUT_SIGNED, **g)
        images[name] = img.double()
    first = next(iter(images.values()))
    peak_idx = int(torch.argmax(first.abs()))
    rng = np.random.default_rng(20261203)
    pix = np.stack([rng.integers(0, N, n_pixels), rng.integers(0, N, n_pixels)], axis=1)
    pix = np.concatenate([pix, [[peak_idx % N, peak_idx // N]]], axis=0)   # (x, y)
    bk, vrk, vik = b[keep], vre[keep], vim[keep]
    dft = np.zeros(len(pix))
    for i0 in range(0, len(pix), 8):
        pp = torch.from_numpy(pix[i0:i0 + 8]).to(dev).double()
        t1 = cell * (pp[:, 1] - N // 2)
        t2 = cell * (pp[:, 0] - N // 2)
        ra = 1.0 - (t1 + rot[0, 2]) ** 2 - (t2 + rot[1, 2]) ** 2
        t3 = -rot[2, 2] + math.copysign(1.0, rot[2, 2]) * torch.sqrt(ra)
        ph = 2.0 * math.pi * (bk[None, :, 0] * t1[:, None] + bk[None, :, 1] * t2[:, None]
                              + bk[None, :, 2] * t3[:, None])
        val = (vrk[None, :] * torch.cos(ph) - vik[None, :] * torch.sin(ph)).sum(dim=1)
        dft[i0:i0 + 8] = (val / (det * R * C)).cpu().numpy()
    dft_peak = abs(dft[-1])
    out = {"pixels": int(len(pix)), "peak_pixel_xy": [int(pix[-1, 0]), int(pix[-1, 1])],
           "dft_peak": float(dft[-1]),
           "samples_flagged_outside_grid": int((~inside).sum()),
           "samples_kept": int(keep.sum())}
    for name, img in images.items():
        vals = img.cpu().numpy()[pix[:, 1], pix[:, 0]]
        d = np.abs(vals - dft) / dft_peak
        out[name] = {"max_rel_err": float(d.max()), "rms_rel_err": float(np.sqrt(np.mean(d * d)))}
    if all_pixels:
        # Every pixel against the FP64 reference of the same algorithm on the same plan: the
        # imagers' arithmetic error alone (the method's own error cancels).
        t0 = time.perf_counter()
        planes = ref64_grid(b, vre, vim, det, plan, np.float32(cell), G)
        ref = ref64_image(planes, rot, plan, N, np.float32(cell), R * C)
        del planes
        torch.cuda.synchronize()
        ref_peak = float(ref.abs().max())
        rv = ref.cpu().numpy()[pix[:, 1], pix[:, 0]]
        dr = np.abs(rv - dft) / dft_peak
        all_px = {"reference": "FP64 gridding (bench.ref64_grid / ref64_image) on the imagers' plan",
                  "seconds": time.perf_counter() - t0, "peak": ref_peak,
                  "reference_vs_dft": {"max_rel_err": float(dr.max()),
                                       "rms_rel_err": float(np.sqrt(np.mean(dr * dr)))}}
        for name, img in images.items():
            e = (img - ref).abs() / ref_peak
            all_px[name] = {"max_rel_err": float(e.max()), "rms_rel_err": float(torch.sqrt((e * e).mean()))}
        out["all_pixels_vs_fp64_gridder"] = all_px
    return out


PRECISIONS = {"default": 0, "split3": 1, "splitc": 2, "single": 3, "bf16": 4}
PRECISION_DTYPES = {
    "default": "f16 hi/lo MFMA (visibility term and both ES factors split), f32 accumulate, f64 geometry",
    "split3": "f16 hi/lo MFMA (visibility term and both ES factors split), f32 accumulate, f64 geometry",
    "splitc": "f16 MFMA (visibility term hi/lo, ES product one f16), f32 accumulate, f64 geometry",
    "single": "f16 MFMA (single f16 operands), f32 accumulate, f64 geometry",
    "bf16": "bf16 hi/lo MFMA (visibility term and ES product split), f32 accumulate, f64 geometry",
}


def imager_variants(S, headline=None):
    """Imaging options of every imager whose accuracy is measured: the headline one first, then
    each MFMA operand precision and the plain FP32 tile imager."""
    out = {}
    if headline is not None:
        out[headline[0]] = headline[1]
    for name in ("split3", "splitc", "single", "bf16"):
        key = f"mfma-{name}"
        if key not in out:
            out[key] = {"gridder": S.GRIDDER_MFMA, "mfma_precision": PRECISIONS[name]}
    out.setdefault("plain-fp32", {"gridder": S.GRIDDER_TILED})
    return out


# ---------------------------------------------------------------------------------------------
# FP64 reference of the gridding algorithm (not the oracle): the same plan (slabs, K), the same
# ES kernel, Chebyshev moments, inverse FFT and Jacobi-Anger recombination with the GL-28 grid
# correction, every step in FP64 with torch, so that an imager's arithmetic error can be measured
# on every pixel (the DFT measures the method's error too, and only at a few pixels).
ES_BETA = 15.3704324328                                   # SVDIntegratedImager.cu:1235
CONV_NORM = 2.4937047051153827                            # SVDIntegratedImager.cu:1512


def es64(x):
    """exp(beta (sqrt(1 - x^2) - 1)) for |x| <= 1, else 0 (SVDIntegratedImager.cu:884-888)."""
    import torch
    inside = x.abs() <= 1.0
    r = torch.sqrt(torch.clamp(1.0 - x * x, min=0.0))
    return torch.where(inside, torch.exp(ES_BETA * (r - 1.0)), torch.zeros_like(x))


def ref64_grid(b, vre, vim, det, plan, cell, G, chunk=1 << 18):
    """Planes (num_slabs K, G, G) complex128: every sample's ES footprint (cells ceil(p - 4) ..
    floor(p + 4) on each axis, clipped to the grid) times T_k(s) (Re, Im) v / det, with the
    checkerboard sign (-1)^(g1 + g2), at (g1 + G/2, g2 + G/2) of plane slab K + k
    (SVDIntegratedImager.cu:1218-1293).  b: (S, 3) rotated baselines in wavelengths, FP64;
    vre, vim: collapsed visibilities (S,), FP64, zero for dropped samples."""
    import torch
    dev = b.device
    K, nsl = int(plan["num_terms"]), int(plan["num_slabs"])
    r3_min, width, half = float(plan["r3_min"]), float(plan["slab_width"]), float(plan["slab_half_range"])
    cell = float(cell)
    gmin, gmax, off = -(G // 2), (G - 1) // 2, G // 2
    # (slab, cell) rows of 2 K values (Re and Im of every term): one index per footprint cell.
    grid = torch.zeros(nsl * G * G, 2 * K, dtype=torch.float64, device=dev)
    j = torch.arange(9, device=dev, dtype=torch.float64)
    keep = ((vre != 0) | (vim != 0)).nonzero().squeeze(1)
    for c0 in range(0, keep.numel(), chunk):
        sel = keep[c0:c0 + chunk]
        bb = b[sel]
        vr, vi = vre[sel] / det, vim[sel] / det
        p1, p2, r3 = bb[:, 0] * (cell * G), bb[:, 1] * (cell * G), bb[:, 2]
        slab = torch.clamp(torch.floor((r3 - r3_min) / width), 0, nsl - 1)
        centre = r3_min + (slab + 0.5) * width
        s = torch.clamp((r3 - centre) / half, -1.0, 1.0)
        g1 = torch.ceil(p1 - 4.0)[:, None] + j                 # (n, 9)
        g2 = torch.ceil(p2 - 4.0)[:, None] + j
        w1 = es64((g1 - p1[:, None]) / 4.0) * ((g1 >= gmin) & (g1 <= gmax))
        w2 = es64((g2 - p2[:, None]) / 4.0) * ((g2 >= gmin) & (g2 <= gmax))
        gi1 = torch.clamp(g1, gmin, gmax).long()
        gi2 = torch.clamp(g2, gmin, gmax).long()
        sign = 1.0 - 2.0 * ((gi1[:, :, None] + gi2[:, None, :]) & 1).double()
        w = (w1[:, :, None] * w2[:, None, :] * sign).reshape(-1, 81)
        rows = (slab.long()[:, None, None] * (G * G) + (gi1 + off)[:, :, None] * G
                + (gi2 + off)[:, None, :]).reshape(-1)
        t_prev = t_cur = None
        coef = torch.empty(sel.numel(), 2 * K, dtype=torch.float64, device=dev)
        for k in range(K):
            tk = torch.ones_like(s) if k == 0 else (s if k == 1 else 2.0 * s * t_cur - t_prev)
            t_prev, t_cur = t_cur, tk
            coef[:, 2 * k] = tk * vr
            coef[:, 2 * k + 1] = tk * vi
        vals = (w[:, :, None] * coef[:, None, :]).reshape(-1, 2 * K)
        nz = (w != 0).reshape(-1)
        grid.index_add_(0, rows[nz], vals[nz])
    planes = torch.view_as_complex(grid.reshape(nsl, G * G, K, 2))
    return planes.permute(0, 2, 1).reshape(nsl * K, G, G).contiguous()


def bessel_j64(alpha, K, terms=48):
    """J_0 .. J_{K-1} of each alpha (|alpha| <= 12) by the power series, FP64: (K,) + alpha.shape."""
    import torch
    if float(alpha.abs().max()) > 12.0:
        raise ValueError("bessel_j64: |alpha| above 12 (the series would lose digits)")
    h = alpha / 2.0
    h2 = h * h
    pk = torch.ones_like(alpha)                            # (alpha / 2)^k / k!
    out = []
    for k in range(K):
        term = pk.clone()
        acc = term.clone()
        for m in range(1, terms):
            term = term * (-h2) / (m * (m + k))
            acc = acc + term
        out.append(acc)
        pk = pk * h / (k + 1)
    return torch.stack(out)


def conv_corr64(N, G, dev):
    """Grid correction c(t), t = 0 .. N/2, from the 28-point Gauss-Legendre rule (positive
    half) of the ES kernel (SVDIntegratedImager.cu:1204-1216), FP64."""
    import numpy as np
    import torch
    x, w = np.polynomial.legendre.leggauss(28)
    pos = x > 0
    x, w = torch.from_numpy(x[pos]).to(dev), torch.from_numpy(w[pos]).to(dev)
    t = torch.arange(N // 2 + 1, device=dev, dtype=torch.float64) / G
    k = es64(x)
    corr = (k[None, :] * w[None, :] * torch.cos(math.pi * t[:, None] * ES_SUPPORT * x[None, :])).sum(1)
    return corr * ES_SUPPORT / CONV_NORM


def ref64_image(planes, rot, plan, N, cell, n_samples):
    """Signed, grid-corrected image (N, N) of FP64 planes (SVDIntegratedImager.cu:1295-1445,
    before |.|): unnormalised inverse FFT, then per pixel sum over slabs and terms of
    Re[(k ? 2 : 1) i^k J_k(2 pi t3 half) e^{2 pi i t3 centre} M], times (-1)^(x' + y'), divided
    by c(|x'|) c(|y'|) norm^2 and by S.  rot: FP64 rotation rows (numpy 3 x 3)."""
    import torch
    dev = planes.device
    G = planes.shape[-1]
    if G % 2:
        raise ValueError("ref64_image: even grid sizes only")
    K, nsl = int(plan["num_terms"]), int(plan["num_slabs"])
    r3_min, width, half = float(plan["r3_min"]), float(plan["slab_width"]), float(plan["slab_half_range"])
    cell = float(cell)
    F = torch.fft.ifft2(planes, norm="forward")
    r = torch.arange(N, device=dev) - N // 2               # x' = x - N/2, y' = y - N/2
    M = F[:, G // 2 + r[:, None], G // 2 + r[None, :]]     # (P, y, x)
    t1 = (cell * r.double())[:, None]                      # rows (y)
    t2 = (cell * r.double())[None, :]                      # columns (x)
    ra = 1.0 - (t1 + rot[0, 2]) ** 2 - (t2 + rot[1, 2]) ** 2
    valid = ra >= 0
    t3 = -rot[2, 2] + math.copysign(1.0, rot[2, 2]) * torch.sqrt(torch.clamp(ra, min=0.0))
    J = bessel_j64(2.0 * math.pi * t3 * half, K)
    val = torch.zeros(N, N, dtype=torch.float64, device=dev)
    for ls in range(nsl):
        centre = r3_min + (ls + 0.5) * width
        e = torch.exp(1j * (2.0 * math.pi * t3 * centre))
        for k in range(K):
            coef = (2.0 if k else 1.0) * (1j ** k) * J[k] * e
            val += (coef * M[ls * K + k]).real
    sign = 1.0 - 2.0 * ((r[:, None] + r[None, :]) & 1).double()
    conv = conv_corr64(N, G, dev)
    val = val * sign / (conv[r.abs()][:, None] * conv[r.abs()][None, :] * CONV_NORM ** 2) / n_samples
    return torch.where(valid, val, torch.zeros_like(val))


def serve_main(args):
    """Streaming form: requests for snapshots of the batching form (hour angles over +-4 h, a
    fraction U(0, 0.6) of rows removed, images of 512^2, 1024^2 or 2048^2) arrive as a Poisson
    process; the service admits them against a device memory budget, batches them by image
    size and reports latency from arrival to image-ready."""
    import numpy as np
    import torch
    import svdimager_amd as S
    from svdimager_amd import serve, synth

    rates = [float(r) for r in args.serve.split(",") if r.strip()]
    cfg = synth.SKALOW_BATCHING
    n = args.serve_pool
    pool = []
    t0 = time.perf_counter()
    for i in range(n):
        s = synth.make_batching_snapshot(i, n, cfg, device="cuda")
        pool.append((S.DeviceSnapshot(s, "cuda"), int(s["image_size"]), cfg.cell))
        log(f"serve: snapshot {i + 1}/{n}: {s['num_rows'] * s['num_channels']} samples, "
            f"{s['image_size']}^2")
        del s
    log(f"serve: pool generated in {time.perf_counter() - t0:.1f} s")
    budget = int(args.serve_budget_gib * (1 << 30))
    results = {}
    workspace = {}
    # The MFMA imager through the library's own request loop (svdi_service_*) and through the
    # Python loop of serve.py, and the plain FP32 imager through the Python loop.
    for name, gridder, native in (("mfma", S.GRIDDER_MFMA, True), ("mfma-python-loop", S.GRIDDER_MFMA, False),
                                  ("plain-fp32", S.GRIDDER_TILED, False)):
        ctx = S.Context(0)
        opts = dict(gridder=gridder, output=S.OUTPUT_DIRTY)
        make = serve.NativeService if native else serve.SnapshotService
        svc = make(ctx, budget, max_batch=8, max_inflight=2, **opts)
        workspace[name] = svc.reserve([serve.Request(i, d, N, c, 0.0) for i, (d, N, c) in enumerate(pool)])
        log(f"serve[{name}]: workspace {workspace[name] / 2**30:.2f} GiB after the reserve pass")
        serve.run_poisson(svc, pool, rate=min(rates), duration=0.3, seed=12345)     # warm-up
        if native:
            svc.close()
        rows = []
        for rate in rates:
            svc = make(ctx, budget, max_batch=8, max_inflight=2, **opts)
            out = serve.run_poisson(svc, pool, rate=rate, duration=args.serve_seconds,
                                    seed=int(rate), drain_limit=120.0)
            row = {k: out[k] for k in ("offered_rate", "count", "snapshots_per_s", "mvis_per_s",
                                       "latency_ms", "queue_ms", "mean_batch")}
            if native:
                row["request_loop"] = "native (svdi_service)"
                svc.close()
                row["workspace_bytes"] = ctx.workspace_bytes()
            else:
                row["request_loop"] = "python (serve.SnapshotService)"
                row["peak_admitted_bytes"] = svc.peak_held
                row["workspace_bytes"] = svc.workspace
            rows.append(row)
            lat = out["latency_ms"]
            log(f"serve[{name}] {rate:g} req/s: {out['count']} requests, p50 {lat['p50']:.2f} ms, "
                f"p99 {lat['p99']:.2f} ms, {out['snapshots_per_s']:.0f} snapshots/s, "
                f"mean batch {out['mean_batch']:.2f}")
        results[name] = rows
        torch.cuda.synchronize()
        ctx.close()
    sizes = [N for _, N, _ in pool]
    samples = [d.num_samples for d, _, _ in pool]
    print(json.dumps({
        "metric": "p50/p99 latency per snapshot, Poisson arrivals (streaming form)",
        "unit": "ms", "higher_is_better": False, "n_gpus": 1, "data": "synthetic",
        "config": {"workload": "skalow-512st-64ch batching form (hour angles over +-4 h, "
                               "U(0, 0.6) of rows removed, images 512/1024/2048)",
                   "pool": n, "image_sizes": {str(k): sizes.count(k) for k in sorted(set(sizes))},
                   "samples_min": min(samples), "samples_max": max(samples),
                   "arrival_window_s": args.serve_seconds, "budget_bytes": budget,
                   "max_batch": 8, "max_inflight_batches": 2, "output": "dirty"},
        "workspace_bytes": workspace, "results": results}), flush=True)


def precision_main(args):
    """The 4096^2 precision form: the SKA-Low array with 128 channels (16,744,448 samples),
    cell 2e-5, G = 6144.  Error against the FP64 DFT as in the default run, and the single-call
    time of each imager."""
    import torch
    import svdimager_amd as S
    from svdimager_amd import synth

    cfg = synth.SKALOW_4K
    log(f"precision: generating {cfg.name} ({cfg.num_samples} samples) ...")
    snap = synth.make_snapshot(cfg, device="cuda")
    dsnap = S.DeviceSnapshot(snap, "cuda")
    ctx = S.Context(0)
    gridders = imager_variants(S)
    times, stages = {}, {}
    plan = None
    for name, g in gridders.items():
        image = torch.empty((cfg.image_size, cfg.image_size), dtype=torch.float32, device="cuda")
        for _ in range(2):
            _, plan = S.image_snapshot(ctx, dsnap, cfg.image_size, cfg.cell, image=image, **g)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            S.image_snapshot(ctx, dsnap, cfg.image_size, cfg.cell, image=image, **g)
        torch.cuda.synchronize()
        times[name] = (time.perf_counter() - t0) / 3 * 1e3
        ctx.set_profiling(True)
        S.image_snapshot(ctx, dsnap, cfg.image_size, cfg.cell, image=image, **g)
        stages[name] = ctx.stage_times()
        ctx.set_profiling(False)
        log(f"precision[{name}]: {times[name]:.2f} ms per snapshot, stages {stages[name]}")
    acc = None
    if not args.no_accuracy:
        log("precision: accuracy against the FP64 DFT ...")
        acc = accuracy(S, ctx, snap, cfg, gridders)
    print(json.dumps({
        "metric": "max / RMS error vs FP64 DFT (64 seeded pixels + peak), 4096^2 precision form",
        "n_gpus": 1, "data": "synthetic",
        "config": {"workload": cfg.name, "samples": cfg.num_samples, "image_size": cfg.image_size,
                   "grid_size": cfg.grid_size, "cell": cfg.cell, "num_terms": plan["num_terms"],
                   "num_slabs": plan["num_slabs"]},
        "ms_per_snapshot": times, "stages_ms": stages, "accuracy_vs_fp64_dft": acc}), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # The kernels reach steady clocks after about 20 calls (k_win_grid 489 -> 420 us over the
    # first 20), so the default warm-up covers them.
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--gridder", default="mfma", choices=["mfma", "tiled", "atomic", "auto"])
    ap.add_argument("--mfma-precision", default="default", choices=sorted(PRECISIONS),
                    help="operand precision of the MFMA gridder (default: the library's choice)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip timing the plain FP32 tiled imager")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the MFMA gridder's default work shape instead of running "
                         "svdi_autotune on the workload first")
    ap.add_argument("--cpu-samples", type=int, default=3,
                    help="snapshots of the workload timed by the CPU baseline (median)")
    ap.add_argument("--batch", type=int, default=8,
                    help="snapshots per batched call (consecutive 10 s snapshots of the same "
                         "observation); 0 = time single-snapshot calls only")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (three times) and exit: the "
                         "reference reading for the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--serve", default="", metavar="RATES",
                    help="only run the streaming form and exit: Poisson arrivals at each of the "
                         "comma-separated RATES (requests/s, e.g. 50,100,200,400) of snapshots "
                         "drawn from the batching form, through the MFMA and the plain FP32 "
                         "imagers; prints p50/p99 latency and throughput per rate")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window per rate")
    ap.add_argument("--serve-pool", type=int, default=16,
                    help="distinct device-resident snapshots the requests are drawn from")
    ap.add_argument("--serve-budget-gib", type=float, default=64.0,
                    help="device memory budget of the service (workspace + admitted requests)")
    ap.add_argument("--precision", action="store_true",
                    help="only measure the 4096^2 precision form (512 stations, 128 channels, "
                         "16.7 M samples): error of the MFMA and plain FP32 imagers against the "
                         "FP64 DFT at 64 seeded pixels plus the peak, and their single-call times")
    args = ap.parse_args()
    if args.precision:
        precision_main(args)
        return
    if args.serve:
        if args.gpus != 1:
            print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported",
                  file=sys.stderr)
            sys.exit(2)
        serve_main(args)
        return
    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 * n, "dispatches": 3}), flush=True)
        return
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
    gmap = {"auto": S.GRIDDER_AUTO, "atomic": S.GRIDDER_ATOMIC, "tiled": S.GRIDDER_TILED,
            "mfma": S.GRIDDER_MFMA}
    gridder = gmap[args.gridder]

    ctx = S.Context(0)
    image = torch.empty((cfg.image_size, cfg.image_size), dtype=torch.float32, device="cuda")
    gopts = {"mfma_precision": PRECISIONS[args.mfma_precision]}
    tuning = None
    if args.gridder == "mfma" and not args.no_autotune:
        # The project's autotuner picks the window gridder's work shape for this size on the
        # first snapshot (untimed); every call below, single and batched, uses it.
        log("bench: autotuning the MFMA gridder's work shape ...")
        tuning = S.autotune(ctx, dsnap, cfg.image_size, cfg.cell, repeats=3, **gopts)
        log(f"bench: tuning {tuning}")
    elapsed, stages, plan = time_gridder(S, ctx, dsnap, cfg, image, gridder, args.steps,
                                         args.warmup, args.gridder, **gopts)
    log(f"bench: plan {plan}")
    single_snaps_per_s = args.steps / elapsed
    ms_per_step = elapsed * 1e3 / args.steps
    snaps_per_s = single_snaps_per_s
    batch = max(0, args.batch)
    batched = None
    mfma_headline = args.gridder in ("mfma", "auto")
    if batch > 0:
        # The snapshot above plus batch - 1 followers, 10 s apart in hour angle.
        dsnaps = [dsnap]
        for k in range(1, batch):
            s_k = synth.make_snapshot(cfg, ha_hours=cfg.ha_hours + k * 10.0 / 3600.0, device="cuda")
            dsnaps.append(S.DeviceSnapshot(s_k, "cuda"))
            del s_k
        log(f"bench: {batch} snapshots resident")
        images = [torch.empty_like(image) for _ in range(batch)]
        for i in range(args.warmup):
            S.image_snapshot_batch(ctx, dsnaps, cfg.image_size, cfg.cell, images=images, gridder=gridder, **gopts)
            log(f"bench[batch {batch}]: warmup {i + 1}/{args.warmup} done")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        # Pipelined calls: each returns once its work is enqueued, so the next call's planning
        # overlaps the previous call's last snapshots; the synchronize below waits for all.
        for _ in range(args.steps):
            S.image_snapshot_batch(ctx, dsnaps, cfg.image_size, cfg.cell, images=images, gridder=gridder,
                                   wait=False, **gopts)
        torch.cuda.synchronize()
        el_b = time.perf_counter() - t0
        log(f"bench[batch {batch}]: {args.steps} steps in {el_b:.3f} s")
        batched = {"snapshots_per_s": args.steps * batch / el_b, "ms_per_step": el_b * 1e3 / args.steps}
        snaps_per_s = batched["snapshots_per_s"]
        ms_per_step = batched["ms_per_step"]
        if mfma_headline:
            # Re-imaging the same geometry with its plans known (svdi_image_snapshots_planned):
            # the batch in one group (one window-contraction launch over all its snapshots),
            # eager calls, then the same call captured once as a graph and replayed.  Not the
            # headline: a new snapshot needs its plan first.
            _, plans = S.image_snapshot_batch(ctx, dsnaps, cfg.image_size, cfg.cell, images=images,
                                              gridder=gridder, **gopts)
            for _ in range(max(1, args.warmup)):
                S.image_snapshot_batch(ctx, dsnaps, cfg.image_size, cfg.cell, images=images, plans=plans,
                                       gridder=gridder, **gopts)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                S.image_snapshot_batch(ctx, dsnaps, cfg.image_size, cfg.cell, images=images, plans=plans,
                                       gridder=gridder, **gopts)
            torch.cuda.synchronize()
            el_p = time.perf_counter() - t0
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                S.image_snapshot_batch(ctx, dsnaps, cfg.image_size, cfg.cell, images=images, plans=plans,
                                       gridder=gridder, **gopts)
            graph.replay()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                graph.replay()
            torch.cuda.synchronize()
            el_g = time.perf_counter() - t0
            del graph
            batched["planned"] = {"snapshots_per_s": args.steps * batch / el_p,
                                  "graph_replay_snapshots_per_s": args.steps * batch / el_g,
                                  "note": "same snapshots re-imaged on known plans: one group, one "
                                          "window-contraction launch per batch"}
            log(f"bench[batch {batch}]: planned {batched['planned']}")
        del dsnaps, images
    K, P = plan["num_terms"], plan["num_planes"]
    alg = algorithmic(cfg, S_eff, K, P)

    # Dominant kernel: the gridder's accumulation kernel, timed alone with HIP events on the
    # library's stream.  One launch per slab batch (one batch here).
    launches = max(1, math.ceil(plan["num_slabs"] / max(1, plan["batch_slabs"])))
    kernel_ms = stages["grid_accumulate"] / launches
    achieved_tflops = alg["F_grid"] / launches / (kernel_ms * 1e-3) / 1e12
    mfma = args.gridder == "mfma"
    peak = FP16_PEAK_TFLOPS if mfma else FP32_PEAK_TFLOPS
    kernel = {"mfma": "k_win_grid", "tiled": "k_tile_grid",
              "auto": "k_tile_grid", "atomic": "k_grid_atomic"}[args.gridder]
    roofline = {"bound": "mfma", "achieved": achieved_tflops, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved_tflops / peak, "traffic": None,
                "kernel": kernel, "kernel_ms": kernel_ms,
                "algorithmic": "F_grid = 4 * 8^2 * K * S_unflagged per launch",
                "peak_kind": "dense FP16 MFMA" if mfma else "FP32 matrix"}
    traffic = measured_traffic(kernel, cfg.name)
    if traffic is not None:
        roofline["traffic"] = traffic["hbm_bytes_per_dispatch"]
        roofline["traffic_source"] = traffic["source"]
    hbm_bytes = alg["B_vis"] + alg["B_grid"] + alg["B_fft"] + alg["B_img"]

    out = {
        "metric": METRIC, "value": snaps_per_s, "unit": "snapshots/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": PRECISION_DTYPES[args.mfma_precision] if mfma else "f32, f64 geometry",
        "data": "synthetic",
        "config": {"workload": cfg.name, "stations": cfg.n_stations, "rows": cfg.num_rows,
                   "channels": cfg.n_channels, "samples": cfg.num_samples,
                   "image_size": cfg.image_size, "grid_size": cfg.grid_size, "cell": cfg.cell,
                   "num_terms": K, "num_slabs": plan["num_slabs"],
                   "gridder": args.gridder, "mfma_precision": args.mfma_precision, "batch": batch,
                   "batch_calls": "pipelined (svdi_image_snapshot_batch_async)" if batch > 0 else None,
                   "parallelism": "single-gpu"},
        "mvis_per_s": cfg.num_samples * snaps_per_s / 1e6,
        "single_snapshot": {"snapshots_per_s": single_snaps_per_s,
                            "ms_per_snapshot": elapsed * 1e3 / args.steps},
        "batched": batched,
        "stages_ms": stages,
        "tuning": tuning,
        "records": plan["records"],
        "path_hbm_bytes_algorithmic": hbm_bytes,
        "path_roofline_cap_snapshots_per_s": HBM_PEAK_GBS * 1e9 / hbm_bytes,
        "roofline": roofline,
    }
    if not args.no_compare and mfma:
        # The other operand precisions of the MFMA gridder, single calls (their accuracy is in
        # accuracy_vs_fp64_dft).
        modes = {}
        for name in ("split3", "splitc", "single", "bf16"):
            el3, st3, _ = time_gridder(S, ctx, dsnap, cfg, image, S.GRIDDER_MFMA, args.steps, min(args.warmup, 2),
                                       f"mfma-{name}", mfma_precision=PRECISIONS[name])
            modes[name] = {"snapshots_per_s": args.steps / el3, "k_win_grid_ms": st3["grid_accumulate"],
                           "grid_stage_ms": st3["grid"], "dtype": PRECISION_DTYPES[name]}
        out["mfma_precision_modes"] = modes
    if not args.no_compare and args.gridder not in ("tiled", "atomic"):
        # The plain FP32 imagers: the FP32 tile gridder, and the reference's own schedule (one
        # thread per sample, a device atomic per cell and term, SVDIntegratedImager.cu:1276-1291).
        # plain_fp32 is the faster of the two.
        plain = {}
        for name, g, steps in (("tiled", S.GRIDDER_TILED, args.steps),
                               ("atomic", S.GRIDDER_ATOMIC, max(1, min(args.steps, 5)))):
            el2, st2, _ = time_gridder(S, ctx, dsnap, cfg, image, g, steps, min(args.warmup, 2),
                                       f"plain-fp32-{name}")
            plain[name] = {"snapshots_per_s": steps / el2, "ms_per_step": el2 * 1e3 / steps,
                           "mvis_per_s": cfg.num_samples * steps / el2 / 1e6, "stages_ms": st2}
        best = max(plain, key=lambda k: plain[k]["snapshots_per_s"])
        out["plain_fp32"] = dict(plain[best], gridder=best, both=plain)
        out["speedup_vs_plain_fp32"] = snaps_per_s / plain[best]["snapshots_per_s"]
        out["speedup_single_vs_plain_fp32"] = single_snaps_per_s / plain[best]["snapshots_per_s"]
    if not args.no_accuracy:
        log("bench: accuracy against the FP64 DFT ...")
        gl = imager_variants(S, (args.gridder, {"gridder": gridder, **gopts}))
        out["accuracy_vs_fp64_dft"] = accuracy(S, ctx, snap, cfg, gl)
    # The host-buffer boundary (FIpipe on host Bin / Visreal / Visimag, archive semantics): the
    # PCIe-inclusive single-call time, the rate a caller of the C++ FIpipe symbol sees.
    pre = S.preprocess_measurement_set_gpu(ctx, dsnap)
    host_in = {k: pre[k].cpu().numpy() for k in ("bin", "vin", "vis_re", "vis_im")}
    del pre
    S.fipipe_host(ctx, host_in["vis_re"], host_in["vis_im"], host_in["bin"], host_in["vin"],
                  cfg.image_size, cfg.cell)
    t0 = time.perf_counter()
    for _ in range(3):
        S.fipipe_host(ctx, host_in["vis_re"], host_in["vis_im"], host_in["bin"], host_in["vin"],
                      cfg.image_size, cfg.cell)
    fip_ms = (time.perf_counter() - t0) / 3 * 1e3
    h2d = sum(host_in[k].nbytes for k in host_in)
    out["fipipe_host"] = {"ms_per_call": fip_ms, "h2d_bytes": int(h2d),
                          "d2h_bytes": 4 * cfg.image_size * cfg.image_size,
                          "note": "PCIe-inclusive: host Bin / Visreal / Visimag / Vin in, dirty image out"}
    log(f"bench: FIpipe on host buffers {fip_ms:.2f} ms per call ({h2d / 1e6:.0f} MB in)")
    if not args.no_cpu_baseline:
        log("bench: timing the CPU baseline ...")
        host = {k: snap[k] for k in snap}
        small = (synth.make_snapshot(synth.SMALL), synth.SMALL)
        out["cpu_baseline"] = cpu_baseline(host, cfg, args.cpu_samples, small)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
