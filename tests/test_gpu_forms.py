"""The batching and streaming forms at their workload (SKA-Low array, 512 stations x 64
channels, hour angles over +-4 h, a fraction U(0, 0.6) of rows removed, images of 512^2, 1024^2
or 2048^2):

* batching form: 16 such snapshots of mixed image sizes imaged in one call
  (svdi_image_snapshots), each image against the same snapshot imaged alone, and the MFMA
  gridder against the FP32 tile gridder on the 2048^2 members;
* streaming form: the request service (svdimager_amd/serve.py) fed by Poisson arrivals at 200
  requests/s drawn from the same snapshots, every served image against its single call.

Tolerances: the MFMA gridder's FP32 sums follow the record order that the binning's LDS atomics
set, so repeated calls agree to rounding (1e-5 of peak over the image, 1e-6 over the central
half, signed images); the two gridders differ by their own rounding (measured 3.4e-6 / 4.1e-7 at
these sizes).  Dirty images: DIRTY_CENTRAL over the central half.
"""
import numpy as np
import pytest
import torch

from conftest import DIRTY_CENTRAL, announce, assert_parity
import svdimager_amd as S
from svdimager_amd import serve, synth

pytestmark = pytest.mark.gpu

COUNT = 16


@pytest.fixture(scope="module")
def pool():
    cfg = synth.SKALOW_BATCHING
    out = []
    for i in range(COUNT):
        s = synth.make_batching_snapshot(i, COUNT, cfg, device="cuda")
        out.append((S.DeviceSnapshot(s, "cuda"), int(s["image_size"]), cfg.cell))
        del s
    sizes = sorted({N for _, N, _ in out})
    print(f"pool: sizes {sizes}, samples {[d.num_samples for d, _, _ in out]}", flush=True)
    return out


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    torch.cuda.synchronize()
    c.close()


def test_mixed_size_batch_matches_single_calls(ctx, pool):
    dsnaps = [d for d, _, _ in pool]
    sizes = [N for _, N, _ in pool]
    cells = [c for _, _, c in pool]
    assert set(sizes) == set(synth.BATCHING_SIZES)
    announce(f"batching form: {COUNT} snapshots, sizes {sizes}, one call")
    imgs, plans = S.image_snapshot_batch(ctx, dsnaps, sizes, cells, output=S.OUTPUT_SIGNED,
                                         gridder=S.GRIDDER_MFMA)
    torch.cuda.synchronize()
    tiled_checked = 0
    for i, (d, N, cell) in enumerate(pool):
        announce(f"batching form: snapshot {i} alone, N={N}, {d.num_samples} samples")
        ref, plan = S.image_snapshot(ctx, d, N, cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_MFMA)
        assert imgs[i].shape == (N, N)
        assert plans[i]["num_terms"] == plan["num_terms"] and plans[i]["num_slabs"] == plan["num_slabs"]
        assert plans[i]["grid_size"] == (3 * N + 1) // 2
        ref_np = ref.cpu().numpy()
        assert_parity(imgs[i].cpu().numpy(), ref_np, f"batch[{i}] vs single ({N}^2)", whole=1e-5,
                      central=1e-6)
        if N == 2048 and tiled_checked < 2:
            announce(f"batching form: snapshot {i}, FP32 tiles, N={N}")
            tiles, _ = S.image_snapshot(ctx, d, N, cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_TILED)
            assert_parity(ref_np, tiles.cpu().numpy(), f"MFMA vs tiles [{i}] (slabs {plan['num_slabs']}, "
                          f"K {plan['num_terms']})", whole=1e-5, central=1e-6)
            tiled_checked += 1
    assert tiled_checked == 2


def test_streaming_form_serves_the_single_call_images(ctx, pool):
    singles = []
    for d, N, cell in pool:
        img, _ = S.image_snapshot(ctx, d, N, cell, gridder=S.GRIDDER_MFMA, output=S.OUTPUT_DIRTY)
        singles.append(img.cpu().numpy())
    announce("streaming form: 200 req/s for 0.25 s")
    svc = serve.SnapshotService(ctx, budget_bytes=1 << 40, max_batch=8, max_inflight=2,
                                keep_images=True, gridder=S.GRIDDER_MFMA, output=S.OUTPUT_DIRTY)
    ws = svc.reserve([serve.Request(i, d, N, cell, 0.0) for i, (d, N, cell) in enumerate(pool)])
    svc.budget = ws + (8 << 30)
    out = serve.run_poisson(svc, pool, rate=200.0, duration=0.25, seed=7, drain_limit=60.0)
    torch.cuda.synchronize()
    assert out["count"] >= 20
    lat = out["latency_ms"]
    print(f"streaming form: {out['count']} requests, p50 {lat['p50']:.2f} ms, p99 {lat['p99']:.2f} ms, "
          f"{out['snapshots_per_s']:.0f} snapshots/s", flush=True)
    index = {id(d): i for i, (d, _, _) in enumerate(pool)}
    for r in out["requests"]:
        i = index[id(r.snapshot)]
        assert r.image.shape == (r.image_size, r.image_size)
        assert_parity(r.image.cpu().numpy(), singles[i], f"request {r.rid} (pool entry {i})",
                      central=DIRTY_CENTRAL)
