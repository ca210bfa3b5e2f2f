"""Ordering between the library and torch's default (null) stream, in both directions: snapshot
inputs written by torch kernels just before a call are read complete, and torch work after an
asynchronous batched call sees the call's images.  The context's streams are blocking streams,
so neither needs synchronisation by the caller."""
import numpy as np
import pytest
import torch

from conftest import announce, assert_parity
import svdimager_amd as S
from svdimager_amd import synth

pytestmark = pytest.mark.gpu


def test_inputs_written_by_torch_just_before_the_call(gpu_ctx, small_snap):
    announce("flags written by a torch kernel queued behind a large matmul, then imaged")
    N, cell = small_snap["image_size"], small_snap["cell"]
    ref, _ = S.image_snapshot(gpu_ctx, S.DeviceSnapshot(small_snap, "cuda"), N, cell,
                              output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_MFMA)
    ref = ref.cpu().numpy()
    for rep in range(3):
        d = S.DeviceSnapshot(small_snap, "cuda")
        want = d.tensors["flag0"].clone()
        d.tensors["flag0"].fill_(1)                 # every sample flagged in XX ...
        d.tensors["flag3"].fill_(1)                 # ... and in YY, until the copies below
        torch.cuda.synchronize()
        a = torch.randn(4096, 4096, device="cuda")
        for _ in range(4):
            a = a @ a / 64.0                        # keeps the default stream busy
        d.tensors["flag0"].copy_(want)
        d.tensors["flag3"].copy_(want)
        img, _ = S.image_snapshot(gpu_ctx, d, N, cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_MFMA)
        got = img.cpu().numpy()
        assert np.abs(got).max() > 0, rep
        assert_parity(got, ref, f"torch-written flags, repeat {rep}")


def test_torch_sees_async_batch_images(gpu_ctx):
    announce("async batched call, then an event and reads on the default stream")
    cfg = synth.SMALL
    snaps = [S.DeviceSnapshot(synth.make_snapshot(cfg, ha_hours=h, seed_offset=i), "cuda")
             for i, h in enumerate((-1.0, 0.0, 1.0, 2.0, -2.0))]
    refs, _ = S.image_snapshot_batch(gpu_ctx, snaps, cfg.image_size, cfg.cell, gridder=S.GRIDDER_MFMA,
                                     output=S.OUTPUT_SIGNED)
    refs = [r.cpu().numpy() for r in refs]
    imgs = [torch.full((cfg.image_size, cfg.image_size), float("nan"), device="cuda") for _ in snaps]
    torch.cuda.synchronize()
    S.image_snapshot_batch(gpu_ctx, snaps, cfg.image_size, cfg.cell, images=imgs, wait=False,
                           gridder=S.GRIDDER_MFMA, output=S.OUTPUT_SIGNED)
    ev = torch.cuda.Event()
    ev.record()
    ev.synchronize()
    sums = torch.stack([i.abs().sum() for i in imgs])     # default-stream kernels on the images
    assert torch.isfinite(sums).all()
    for i, (img, ref) in enumerate(zip(imgs, refs)):
        assert_parity(img.cpu().numpy(), ref, f"async batch image {i}")
