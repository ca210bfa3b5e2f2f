"""The MFMA gridder's operand precisions (svdi_options.mfma_precision) at the SKA-Low workload
(8,372,224 samples, 1024^2): each mode's signed image against the default split3 mode.

Measured on MI355X against an FP64 gridding of the same plan over all pixels: split3 2.1e-6 of
peak (max), bf16 1.9e-6, the single-f16 ES product modes splitc / single 1.1e-5 (RMS 2.7e-6;
the grid correction amplifies their rounding about 85x at the image edge).  Bounds here:
bf16 within 6e-6 of split3 over the image and 1e-6 over the central half; splitc and single
within 3e-5 and 5e-6.  An unknown precision value is an argument error.
"""
import numpy as np
import pytest

from conftest import announce, assert_parity
import svdimager_amd as S
from svdimager_amd import _lib, synth

pytestmark = pytest.mark.gpu

BOUNDS = {S.PRECISION_BF16: (6e-6, 1e-6), S.PRECISION_SPLITC: (3e-5, 5e-6),
          S.PRECISION_SINGLE: (3e-5, 5e-6)}


@pytest.fixture(scope="module")
def skalow():
    snap = synth.make_snapshot(synth.SKALOW, device="cuda")
    return S.DeviceSnapshot(snap, "cuda"), synth.SKALOW


def test_precision_modes_against_split3(gpu_ctx, skalow):
    d, cfg = skalow
    announce("precision modes: split3 reference")
    ref, plan = S.image_snapshot(gpu_ctx, d, cfg.image_size, cfg.cell, output=S.OUTPUT_SIGNED,
                                 gridder=S.GRIDDER_MFMA, mfma_precision=S.PRECISION_SPLIT3)
    ref = ref.cpu().numpy()
    dflt, _ = S.image_snapshot(gpu_ctx, d, cfg.image_size, cfg.cell, output=S.OUTPUT_SIGNED,
                               gridder=S.GRIDDER_MFMA)
    assert_parity(dflt.cpu().numpy(), ref, "default vs split3", whole=1e-5, central=1e-6)
    for prec, (whole, central) in BOUNDS.items():
        announce(f"precision mode {prec}")
        img, p = S.image_snapshot(gpu_ctx, d, cfg.image_size, cfg.cell, output=S.OUTPUT_SIGNED,
                                  gridder=S.GRIDDER_MFMA, mfma_precision=prec)
        img = img.cpu().numpy()
        assert np.isfinite(img).all()
        assert p["num_terms"] == plan["num_terms"]
        assert_parity(img, ref, f"mode {prec} vs split3", whole=whole, central=central)


def test_unknown_precision_is_an_argument_error(gpu_ctx, skalow):
    d, cfg = skalow
    with pytest.raises(_lib.SvdiError) as e:
        S.image_snapshot(gpu_ctx, d, cfg.image_size, cfg.cell, mfma_precision=7)
    assert _lib.STATUS_NAMES[e.value.status] == "INVALID_ARGUMENT"
