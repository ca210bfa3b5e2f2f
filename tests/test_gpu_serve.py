"""GPU tests of the streaming service (svdimager_amd/serve.py) on the batching form at small
size: snapshots of the 64-station array at hour angles over +-4 h with a random fraction of rows
removed and image sizes 64 and 128, arriving as a Poisson process.  Every served image equals
the same snapshot imaged alone (to the FP32 summation-order tolerance), under a loose budget and
under one that admits about two requests at a time."""
import dataclasses

import numpy as np
import pytest
import torch

from conftest import DIRTY_CENTRAL, announce, assert_parity
import svdimager_amd as S
from svdimager_amd import serve, synth

pytestmark = pytest.mark.gpu

SIZES = (64, 128)


@pytest.fixture(scope="module")
def pool():
    cfg = dataclasses.replace(synth.SMALL, flag_fraction=0.0)
    out = []
    for i in range(6):
        s = synth.make_batching_snapshot(i, 6, cfg, sizes=SIZES)
        out.append((S.DeviceSnapshot(s, "cuda"), s["image_size"], cfg.cell))
    return out


@pytest.fixture(scope="module")
def ctx():
    # A context of its own: the workspace of the session context depends on the tests before.
    c = S.Context(0)
    yield c
    torch.cuda.synchronize()
    c.close()


@pytest.fixture(scope="module")
def singles(ctx, pool):
    refs = []
    for d, N, cell in pool:
        img, _ = S.image_snapshot(ctx, d, N, cell, gridder=S.GRIDDER_MFMA, output=S.OUTPUT_DIRTY)
        refs.append(img.cpu().numpy())
    torch.cuda.synchronize()
    return refs


def check_served(out, pool, singles):
    index = {id(d): i for i, (d, _, _) in enumerate(pool)}
    for r in out["requests"]:
        i = index[id(r.snapshot)]
        assert r.image.shape == (r.image_size, r.image_size)
        assert_parity(r.image.cpu().numpy(), singles[i], f"request {r.rid} (pool entry {i})",
                      central=DIRTY_CENTRAL)


def test_served_images_match_single_calls(ctx, pool, singles):
    announce("service: loose budget, 400 req/s for 0.25 s")
    svc = serve.SnapshotService(ctx, budget_bytes=1 << 40, max_batch=4, max_inflight=2,
                                keep_images=True, gridder=S.GRIDDER_MFMA, output=S.OUTPUT_DIRTY)
    ws = svc.reserve([serve.Request(i, d, N, cell, 0.0) for i, (d, N, cell) in enumerate(pool)])
    svc.budget = ws + (1 << 30)
    out = serve.run_poisson(svc, pool, rate=400.0, duration=0.25, seed=1, drain_limit=30.0)
    torch.cuda.synchronize()
    assert out["count"] > 20
    assert {r.image_size for r in out["requests"]} == set(SIZES)
    assert svc.workspace + svc.peak_held <= svc.budget
    check_served(out, pool, singles)


def test_tight_budget_admits_two_at_a_time(ctx, pool, singles):
    announce("service: budget for two requests")
    svc = serve.SnapshotService(ctx, budget_bytes=1 << 40, max_batch=4, max_inflight=2,
                                keep_images=True, gridder=S.GRIDDER_MFMA, output=S.OUTPUT_DIRTY)
    ws = svc.reserve([serve.Request(i, d, N, cell, 0.0) for i, (d, N, cell) in enumerate(pool)])
    costs = [serve.request_bytes(serve.Request(0, d, N, cell, 0.0)) for d, N, cell in pool]
    biggest, smallest = max(costs), min(costs)
    svc.budget = ws + 2 * biggest
    out = serve.run_poisson(svc, pool, rate=1000.0, duration=0.05, seed=2, drain_limit=30.0)
    torch.cuda.synchronize()
    assert out["count"] > 10
    assert svc.workspace == ws                   # no growth after the reserve pass
    assert svc.peak_held <= 2 * biggest
    assert max(svc.batch_sizes) <= (2 * biggest) // smallest
    check_served(out, pool, singles)


def test_native_service_images_match_single_calls(ctx, pool, singles):
    """The library's own request loop (svdi_service_*, serve.NativeService): the same Poisson
    stream, every served image against its single call, batches of one key at most max_batch,
    latencies positive and ordered (queue <= latency)."""
    announce("native service: loose budget, 400 req/s for 0.25 s")
    svc = serve.NativeService(ctx, budget_bytes=1 << 40, max_batch=4, max_inflight=2, keep_images=True,
                              gridder=S.GRIDDER_MFMA, output=S.OUTPUT_DIRTY)
    ws = svc.reserve([serve.Request(i, d, N, cell, 0.0) for i, (d, N, cell) in enumerate(pool)])
    svc.budget = ws + (1 << 30)
    try:
        out = serve.run_poisson(svc, pool, rate=400.0, duration=0.25, seed=3, drain_limit=30.0)
        torch.cuda.synchronize()
        assert out["count"] > 20
        assert {r.image_size for r in out["requests"]} == set(SIZES)
        assert max(svc.batch_sizes) <= 4 and sum(svc.batch_sizes) == out["count"]
        for r in out["requests"]:
            assert r.arrival <= r.launched <= r.done
        print(f"native service: {out['count']} requests, p50 {out['latency_ms']['p50']:.2f} ms, "
              f"p99 {out['latency_ms']['p99']:.2f} ms, mean batch {out['mean_batch']:.2f}", flush=True)
        check_served(out, pool, singles)
    finally:
        svc.close()


def test_native_service_tight_budget(ctx, pool, singles):
    """The native loop under a budget of two of the largest requests beside the workspace: the
    stats never show more held than the budget leaves, and every image is right."""
    announce("native service: budget for two requests")
    svc = serve.NativeService(ctx, budget_bytes=1 << 40, max_batch=4, max_inflight=2, keep_images=True,
                              gridder=S.GRIDDER_MFMA, output=S.OUTPUT_DIRTY)
    ws = svc.reserve([serve.Request(i, d, N, cell, 0.0) for i, (d, N, cell) in enumerate(pool)])
    costs = [serve.request_bytes(serve.Request(0, d, N, cell, 0.0)) for d, N, cell in pool]
    svc.budget = ws + 2 * max(costs)
    reqs = [serve.Request(i, *pool[i % len(pool)], arrival=svc.clock()) for i in range(12)]
    peak = 0
    try:
        for r in reqs:
            svc.submit(r)
            peak = max(peak, svc.stats()["held_bytes"])
        while svc.busy():
            svc.poll(wait_ms=50)
            peak = max(peak, svc.stats()["held_bytes"])
        torch.cuda.synchronize()
        assert peak <= 2 * max(costs)
        check_served({"requests": reqs}, pool, singles)
    finally:
        svc.close()
    with pytest.raises(S.SvdiError):
        tiny = serve.NativeService(ctx, budget_bytes=ws + 1, keep_images=True)
        try:
            tiny.submit(serve.Request(0, *pool[0], arrival=tiny.clock()))
        finally:
            tiny.close()
