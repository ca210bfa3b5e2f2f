"""GPU cross-checks at the full SKA-Low shape (512 stations, 64 channels, 8,372,224 samples,
1024^2, G = 1536): the MFMA window gridder against the FP32 tile gridder, both already pinned
to the oracle at small sizes.  These sizes reach code paths the small cases cannot:

* hour angle 0: one slab, 4-row record classes, windows split over several workgroup pieces
  whose partial windows are summed in piece order;
* hour angle 4 h: two slabs and more Chebyshev terms (two or more eight-term passes), where
  only the 8-row class histogram fits in LDS;
* 2048^2 at hour angle -4 h and 4096^2: bins split into LDS bands, and the G = 3072 and general
  long FFTs.

The two gridders differ in summation order and in the MFMA's split-f16 operands (~22
significant bits per product), so the signed images agree to the gridder tolerance of the
small cases: 1e-5 of peak over the central half, 1e-4 over the image.
"""
import numpy as np
import pytest
import torch

from conftest import announce, assert_parity
import svdimager_amd as S
from svdimager_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ha", [0.0, 4.0])
def test_skalow_mfma_vs_tiled(gpu_ctx, ha):
    cfg = synth.SKALOW
    announce(f"SKA-Low full size, HA {ha} h: MFMA vs tiled FP32")
    snap = synth.make_snapshot(cfg, ha_hours=ha, device="cuda")
    d = S.DeviceSnapshot(snap, "cuda")
    N, cell = cfg.image_size, cfg.cell
    img_t, plan_t = S.image_snapshot(gpu_ctx, d, N, cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_TILED)
    img_m, plan_m = S.image_snapshot(gpu_ctx, d, N, cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_MFMA)
    torch.cuda.synchronize()
    assert plan_m["num_terms"] == plan_t["num_terms"] and plan_m["num_slabs"] == plan_t["num_slabs"]
    print(f"HA {ha}: slabs {plan_m['num_slabs']} K {plan_m['num_terms']} records {plan_m['records']}")
    if ha > 0:
        assert plan_m["num_slabs"] >= 2 and plan_m["num_terms"] > 8
    a, b = img_m.cpu().numpy(), img_t.cpu().numpy()
    assert np.isfinite(a).all() and np.abs(b).max() > 0
    assert_parity(a, b, f"MFMA vs tiled, HA {ha} h")


def test_skalow_4k_mfma_vs_tiled(gpu_ctx):
    """The 4096^2 form (512 stations, 128 channels, 16,744,448 samples, G = 6144): no row-class
    histogram fits in LDS there, so the window binning splits its 294,912 bins into eight LDS
    bands.  MFMA against the FP32 tile gridder, signed images."""
    cfg = synth.SKALOW_4K
    announce("SKA-Low 4096^2 (128 channels): MFMA vs tiled FP32")
    snap = synth.make_snapshot(cfg, device="cuda")
    d = S.DeviceSnapshot(snap, "cuda")
    del snap
    N, cell = cfg.image_size, cfg.cell
    img_t, plan_t = S.image_snapshot(gpu_ctx, d, N, cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_TILED)
    img_m, plan_m = S.image_snapshot(gpu_ctx, d, N, cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_MFMA)
    torch.cuda.synchronize()
    assert plan_m["grid_size"] == 6144 and plan_m["num_terms"] == plan_t["num_terms"]
    print(f"4k: slabs {plan_m['num_slabs']} K {plan_m['num_terms']} records {plan_m['records']}")
    a, b = img_m.cpu().numpy(), img_t.cpu().numpy()
    assert np.isfinite(a).all() and np.abs(b).max() > 0
    assert_parity(a, b, "MFMA vs tiled, 4096^2")


def test_skalow_2048_far_from_transit_mfma_vs_tiled(gpu_ctx):
    """A 2048^2 snapshot of the batching form at hour angle -4 h (G = 3072, four slabs, K = 16:
    64 planes; window bins split into eight LDS bands; the G = 3072 transforms on 64 planes).
    MFMA against the FP32 tile gridder, signed images."""
    cfg = synth.SKALOW_BATCHING
    announce("SKA-Low batching form, snapshot 0 of 16 (2048^2, HA -4 h): MFMA vs tiled FP32")
    snap = synth.make_batching_snapshot(0, 16, cfg, device="cuda")
    N = snap["image_size"]
    d = S.DeviceSnapshot(snap, "cuda")
    del snap
    img_t, plan_t = S.image_snapshot(gpu_ctx, d, N, cfg.cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_TILED)
    img_m, plan_m = S.image_snapshot(gpu_ctx, d, N, cfg.cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_MFMA)
    torch.cuda.synchronize()
    print(f"2048: slabs {plan_m['num_slabs']} K {plan_m['num_terms']} records {plan_m['records']}")
    assert N == 2048 and plan_m["grid_size"] == 3072
    assert plan_m["num_slabs"] == plan_t["num_slabs"] >= 3 and plan_m["num_terms"] == plan_t["num_terms"] == 16
    a, b = img_m.cpu().numpy(), img_t.cpu().numpy()
    assert np.isfinite(a).all() and np.abs(b).max() > 0
    assert_parity(a, b, "MFMA vs tiled, 2048^2 at HA -4 h")


def test_skalow_preprocessed_mfma_vs_tiled(gpu_ctx):
    """The SKA-Low snapshot through the reference's two-step form (svdi_preprocess, then
    svdi_image_preprocessed on its FP32 Bin / Vin / visibilities): 8,372,224 samples make 512
    binning blocks of 16,384 samples, more than the CUs, so the window gridder's count pass keeps
    16-bit histograms two bins to an LDS word and 16-bit per-block counts in memory.  Its image
    against the FP32 tile gridder's (independent binning) on the same inputs."""
    cfg = synth.SKALOW
    announce("SKA-Low full size, preprocessed input: MFMA vs tiled FP32")
    snap = synth.make_snapshot(cfg, device="cuda")
    d = S.DeviceSnapshot(snap, "cuda")
    pre = S.preprocess_measurement_set_gpu(gpu_ctx, d)
    N, cell = cfg.image_size, cfg.cell
    args = (pre["bin"], pre["vin"], pre["vis_re"], pre["vis_im"], N, cell)
    img_t, plan_t = S.FIpipeDevice(gpu_ctx, *args, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_TILED)
    img_m, plan_m = S.FIpipeDevice(gpu_ctx, *args, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_MFMA)
    torch.cuda.synchronize()
    assert plan_m["num_terms"] == plan_t["num_terms"] and plan_m["num_slabs"] == plan_t["num_slabs"]
    assert plan_m["gridder"] == S.GRIDDER_MFMA and plan_m["records"] > cfg.num_samples
    a, b = img_m.cpu().numpy(), img_t.cpu().numpy()
    assert np.isfinite(a).all() and np.abs(b).max() > 0
    assert_parity(a, b, "preprocessed MFMA vs tiled")
