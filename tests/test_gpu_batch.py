"""GPU tests of the batched snapshot call: every image of a batch equals the image of the same
snapshot imaged alone (to the FP32 summation-order tolerance), for batches longer than the
number of concurrent lanes (lane workspaces reused across rounds), with both tiled gridders and
both outputs."""
import numpy as np
import pytest
import torch

from conftest import DIRTY_CENTRAL, announce, assert_parity
import svdimager_amd as S
from svdimager_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def snaps():
    # Five snapshots of the small array at different hour angles (distinct uvw, same shapes).
    out = []
    for i, ha in enumerate((-1.0, -0.4, 0.0, 0.5, 1.2)):
        s = synth.make_snapshot(synth.SMALL, ha_hours=ha, seed_offset=i)
        out.append(S.DeviceSnapshot(s, "cuda"))
    return out


@pytest.mark.parametrize("gridder", [S.GRIDDER_MFMA, S.GRIDDER_TILED])
@pytest.mark.parametrize("output", [S.OUTPUT_SIGNED, S.OUTPUT_DIRTY])
def test_batch_matches_single(gpu_ctx, snaps, gridder, output):
    cfg = synth.SMALL
    N, cell = cfg.image_size, cfg.cell
    announce(f"batch of {len(snaps)} gridder={gridder} output={output}")
    imgs, plans = S.image_snapshot_batch(gpu_ctx, snaps, N, cell, gridder=gridder, output=output,
                                         num_terms=cfg.num_terms)
    torch.cuda.synchronize()
    assert len(imgs) == len(snaps) == len(plans)
    for i, d in enumerate(snaps):
        ref, plan = S.image_snapshot(gpu_ctx, d, N, cell, gridder=gridder, output=output,
                                     num_terms=cfg.num_terms)
        assert plans[i]["num_terms"] == plan["num_terms"]
        assert plans[i]["num_slabs"] == plan["num_slabs"]
        a, b = imgs[i].cpu().numpy(), ref.cpu().numpy()
        assert np.abs(b).max() > 0
        assert_parity(a, b, f"batch image {i}",
                      central=DIRTY_CENTRAL if output == S.OUTPUT_DIRTY else 1e-5)


def test_batch_of_one_and_reuse(gpu_ctx, snaps):
    """A batch of one, then a longer batch on the same context (lanes already created)."""
    cfg = synth.SMALL
    N, cell = cfg.image_size, cfg.cell
    announce("batch of one, then of three")
    one, _ = S.image_snapshot_batch(gpu_ctx, snaps[:1], N, cell, gridder=S.GRIDDER_MFMA,
                                    output=S.OUTPUT_SIGNED)
    three, _ = S.image_snapshot_batch(gpu_ctx, snaps[:3], N, cell, gridder=S.GRIDDER_MFMA,
                                      output=S.OUTPUT_SIGNED)
    assert_parity(one[0].cpu().numpy(), three[0].cpu().numpy(), "first image, two batches")


def test_pipelined_batches_match_single(gpu_ctx, snaps):
    """Two batched calls enqueued back to back without waiting (svdi_image_snapshot_batch_async,
    the second over the snapshots in reverse order into other images), then one synchronize:
    every image equals the snapshot imaged alone."""
    cfg = synth.SMALL
    N, cell = cfg.image_size, cfg.cell
    announce(f"two pipelined batches of {len(snaps)}")
    first, _ = S.image_snapshot_batch(gpu_ctx, snaps, N, cell, gridder=S.GRIDDER_MFMA,
                                      output=S.OUTPUT_SIGNED, num_terms=cfg.num_terms, wait=False)
    rev = list(reversed(snaps))
    second, plans = S.image_snapshot_batch(gpu_ctx, rev, N, cell, gridder=S.GRIDDER_MFMA,
                                           output=S.OUTPUT_SIGNED, num_terms=cfg.num_terms, wait=False)
    torch.cuda.synchronize()
    assert len(plans) == len(snaps) and all(p["num_terms"] == cfg.num_terms for p in plans)
    for i, d in enumerate(snaps):
        ref, _ = S.image_snapshot(gpu_ctx, d, N, cell, gridder=S.GRIDDER_MFMA, output=S.OUTPUT_SIGNED,
                                  num_terms=cfg.num_terms)
        r = ref.cpu().numpy()
        assert_parity(first[i].cpu().numpy(), r, f"pipelined batch 1, snapshot {i}")
        assert_parity(second[len(snaps) - 1 - i].cpu().numpy(), r, f"pipelined batch 2, snapshot {i}")
