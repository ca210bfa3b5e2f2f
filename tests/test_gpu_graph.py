"""svdi_image_snapshot_planned under HIP graph capture: a call on a plan that an earlier call
reported for the same geometry waits on nothing and records no memset, so torch's graph capture
records it; the graph is then replayed with new visibilities written in place into the same
device buffers (the geometry fixed), with the program's own eager work and synchronisations
between replays, and every replay is checked against an eager full call on the same data.

Tolerance: the MFMA gridder's FP32 sums follow the record order that the binning's LDS atomics
set, so two calls agree to rounding (1e-5 of peak over the image, 1e-6 over the central half).
"""
import numpy as np
import pytest
import torch

from conftest import announce, assert_parity, load_small_snapshot
import svdimager_amd as S
from svdimager_amd import _lib

pytestmark = pytest.mark.gpu

VIS = ("vis0_re", "vis0_im", "vis3_re", "vis3_im")


def test_planned_call_matches_the_full_call():
    snap = load_small_snapshot()
    N, cell = snap["image_size"], snap["cell"]
    with S.Context(0) as ctx:
        d = S.DeviceSnapshot(snap, "cuda")
        announce("planned call, eager")
        ref, plan = S.image_snapshot(ctx, d, N, cell, output=S.OUTPUT_SIGNED)
        img = torch.empty((N, N), dtype=torch.float32, device="cuda")
        S.image_snapshot(ctx, d, N, cell, image=img, plan=plan, output=S.OUTPUT_SIGNED)
        torch.cuda.synchronize()
        assert_parity(img.cpu().numpy(), ref.cpu().numpy(), "planned vs full", whole=1e-5, central=1e-6)
        bad = dict(plan, grid_size=plan["grid_size"] + 1)
        with pytest.raises(_lib.SvdiError):
            S.image_snapshot(ctx, d, N, cell, image=img, plan=bad)


def test_graph_replays_with_new_visibilities():
    snap = load_small_snapshot()
    N, cell = snap["image_size"], snap["cell"]
    with S.Context(0) as ctx:
        d = S.DeviceSnapshot(snap, "cuda")
        _, plan = S.image_snapshot(ctx, d, N, cell, output=S.OUTPUT_SIGNED)
        img = torch.zeros((N, N), dtype=torch.float32, device="cuda")
        # One eager planned call sizes the workspace and sets the kernels' LDS limits.
        S.image_snapshot(ctx, d, N, cell, image=img, plan=plan, output=S.OUTPUT_SIGNED)
        torch.cuda.synchronize()
        announce("graph capture of a planned call")
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            S.image_snapshot(ctx, d, N, cell, image=img, plan=plan, output=S.OUTPUT_SIGNED)
        torch.cuda.synchronize()
        base = {k: d.tensors[k].clone() for k in VIS}
        gen = torch.Generator(device="cuda").manual_seed(11)
        for rep, scale in enumerate((1.0, -0.5, 2.0)):
            for k in VIS:
                noise = torch.randn(base[k].shape, device="cuda", generator=gen)
                d.tensors[k].copy_(base[k] * scale + 0.05 * rep * noise)
            torch.cuda.synchronize()
            announce(f"graph replay {rep}")
            img.fill_(float("nan"))
            graph.replay()
            torch.cuda.synchronize()
            got = img.cpu().numpy()
            eager, _ = S.image_snapshot(ctx, d, N, cell, output=S.OUTPUT_SIGNED)
            torch.cuda.synchronize()
            assert np.isfinite(got).all()
            assert_parity(got, eager.cpu().numpy(), f"replay {rep} vs eager", whole=1e-5, central=1e-6)
        del graph
        torch.cuda.synchronize()


def test_batch_graph_replays_mixed_sizes():
    """svdi_image_snapshots_planned: three snapshots of mixed image sizes on known plans, captured
    as one graph (the lanes fork from the capturing stream and it joins them), replayed with new
    visibilities; every image of every replay is checked against an eager full call."""
    snap = load_small_snapshot()
    cell = snap["cell"]
    sizes = [128, 64, 96]                      # G = 192, 96 and 145 (an odd grid: whole plane)
    cells = [cell] * len(sizes)
    with S.Context(0) as ctx:
        ds = [S.DeviceSnapshot(snap, "cuda") for _ in sizes]
        announce("planned batch, eager")
        full, plans = S.image_snapshot_batch(ctx, ds, sizes, cells, output=S.OUTPUT_SIGNED)
        imgs = [torch.zeros((N, N), dtype=torch.float32, device="cuda") for N in sizes]
        # One eager planned batch sizes the lanes' workspaces for exactly this batch.
        S.image_snapshot_batch(ctx, ds, sizes, cells, images=imgs, plans=plans, output=S.OUTPUT_SIGNED)
        torch.cuda.synchronize()
        for i, N in enumerate(sizes):
            assert_parity(imgs[i].cpu().numpy(), full[i].cpu().numpy(), f"planned batch [{N}] vs full",
                          whole=1e-5, central=1e-6)
        announce("graph capture of a planned batch")
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            S.image_snapshot_batch(ctx, ds, sizes, cells, images=imgs, plans=plans, output=S.OUTPUT_SIGNED)
        torch.cuda.synchronize()
        base = [{k: d.tensors[k].clone() for k in VIS} for d in ds]
        gen = torch.Generator(device="cuda").manual_seed(12)
        for rep, scale in enumerate((1.0, 0.5, -1.5)):
            for i, d in enumerate(ds):
                for k in VIS:
                    noise = torch.randn(base[i][k].shape, device="cuda", generator=gen)
                    d.tensors[k].copy_(base[i][k] * (scale + 0.1 * i) + 0.05 * rep * noise)
            torch.cuda.synchronize()
            announce(f"batch graph replay {rep}")
            for im in imgs:
                im.fill_(float("nan"))
            graph.replay()
            torch.cuda.synchronize()
            for i, (d, N) in enumerate(zip(ds, sizes)):
                got = imgs[i].cpu().numpy()
                eager, _ = S.image_snapshot(ctx, d, N, cell, output=S.OUTPUT_SIGNED)
                torch.cuda.synchronize()
                assert np.isfinite(got).all()
                assert_parity(got, eager.cpu().numpy(), f"replay {rep} image {i} ({N}^2) vs eager",
                              whole=1e-5, central=1e-6)
        del graph
        torch.cuda.synchronize()


def test_grouped_planned_batch_matches_single_calls():
    """svdi_image_snapshots_planned's grouped path: 18 snapshots (two groups, of 16 and 2) of three
    image sizes, each group binned on its lanes and contracted by one window-kernel launch over
    all its snapshots' windows; every image against the same snapshot imaged alone."""
    snap = load_small_snapshot()
    cell = snap["cell"]
    sizes = [64, 96, 128] * 6
    with S.Context(0) as ctx:
        base = S.DeviceSnapshot(snap, "cuda")
        ds = []
        gen = torch.Generator(device="cuda").manual_seed(13)
        for i, _ in enumerate(sizes):
            d = S.DeviceSnapshot(snap, "cuda")
            for k in VIS:
                d.tensors[k].copy_(base.tensors[k] * (1.0 + 0.05 * i)
                                   + 0.01 * torch.randn(base.tensors[k].shape, device="cuda", generator=gen))
            ds.append(d)
        _, plans = S.image_snapshot_batch(ctx, ds, sizes, [cell] * len(sizes), output=S.OUTPUT_SIGNED)
        announce("grouped planned batch: 18 snapshots")
        imgs, _ = S.image_snapshot_batch(ctx, ds, sizes, [cell] * len(sizes), plans=plans, output=S.OUTPUT_SIGNED)
        torch.cuda.synchronize()
        for i, (d, N) in enumerate(zip(ds, sizes)):
            ref, _ = S.image_snapshot(ctx, d, N, cell, output=S.OUTPUT_SIGNED)
            torch.cuda.synchronize()
            assert_parity(imgs[i].cpu().numpy(), ref.cpu().numpy(), f"grouped [{i}] ({N}^2) vs single",
                          whole=1e-5, central=1e-6)
