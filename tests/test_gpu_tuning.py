"""The MFMA gridder's work shape (svdi_tuning) and the autotuner (svdi_autotune) on the SKA-Low
snapshot (512 stations x 64 channels, 1024^2, G = 1536):

* any valid shape gives the default shape's image to the gridder's rounding: the shape decides
  which windows one wavefront contracts, how larger ones are cut into workgroup pieces and the
  record class height, i.e. the order of the FP32 sums, not the sums;
* shapes outside the documented ranges are refused;
* the autotuner's pick is stored on the context for its grid size, is used by single and batched
  calls, and is no slower than the defaults by the autotuner's own measure (the defaults are
  one of the candidates it keeps).

Tolerances as for two summation orders of the same gridder (the batching form's): 1e-5 of peak
over the image, 1e-6 over the central half, signed images.
"""
import numpy as np
import pytest
import torch

from conftest import announce, assert_parity
import svdimager_amd as S
from svdimager_amd import synth

pytestmark = pytest.mark.gpu

CFG = synth.SKALOW
G = (3 * CFG.image_size + 1) // 2


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    torch.cuda.synchronize()
    c.close()


@pytest.fixture(scope="module")
def snap():
    s = synth.make_snapshot(CFG, device="cuda")
    return S.DeviceSnapshot(s, "cuda")


def _image(ctx, d):
    img, plan = S.image_snapshot(ctx, d, CFG.image_size, CFG.cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_MFMA)
    torch.cuda.synchronize()
    return img.cpu().numpy(), plan


def test_invalid_shapes_are_refused(ctx):
    for kw in (dict(coop_min=1), dict(coop_min=1 << 20), dict(coop_cap=100), dict(class_log2=1),
               dict(class_log2=4)):
        with pytest.raises(S.SvdiError) as e:
            ctx.set_tuning(G, **kw)
        assert e.value.status == 1, kw
    assert ctx.get_tuning(G)["coop_min"] == 0


@pytest.mark.parametrize("shape", [(32, 4096, 3), (128, 8192, 2), (4096, 49152, 2), (2048, 16384, 3)])
def test_every_shape_gives_the_default_image(ctx, snap, shape):
    ctx.clear_tuning(G)
    ref, plan_ref = _image(ctx, snap)
    cmin, ccap, cls = shape
    announce(f"tuning: coop_min {cmin} coop_cap {ccap} class_log2 {cls}")
    ctx.set_tuning(G, coop_min=cmin, coop_cap=ccap, class_log2=cls)
    assert ctx.get_tuning(G)["coop_cap"] == ccap
    img, plan = _image(ctx, snap)
    ctx.clear_tuning(G)
    assert plan["num_terms"] == plan_ref["num_terms"] and plan["records"] == plan_ref["records"]
    assert np.isfinite(img).all()
    assert_parity(img, ref, f"shape {shape} vs defaults", whole=1e-5, central=1e-6)


def test_autotune_pick_is_kept_and_used(ctx, snap):
    ctx.clear_tuning(G)
    ref, _ = _image(ctx, snap)
    announce("autotune: SKA-Low 1024^2, 1 repeat per shape")
    best = S.autotune(ctx, snap, CFG.image_size, CFG.cell, repeats=1)
    print(f"autotune: {best}", flush=True)
    assert best["grid_ms"] > 0 and best["default_ms"] > 0
    assert best["grid_ms"] <= best["default_ms"]
    kept = ctx.get_tuning(G)
    for k in ("coop_min", "coop_cap", "class_log2"):
        assert kept[k] == best[k]
    img, _ = _image(ctx, snap)
    assert_parity(img, ref, "autotuned vs defaults", whole=1e-5, central=1e-6)
    # Batched calls hand the context's shapes to their lanes.
    imgs, _ = S.image_snapshot_batch(ctx, [snap, snap], CFG.image_size, CFG.cell, output=S.OUTPUT_SIGNED,
                                     gridder=S.GRIDDER_MFMA)
    torch.cuda.synchronize()
    for i in imgs:
        assert_parity(i.cpu().numpy(), img, "batched with the tuned shape", whole=1e-5, central=1e-6)
    ctx.clear_tuning(G)
