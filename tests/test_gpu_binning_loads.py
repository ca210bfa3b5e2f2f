"""The window binning's two ways of reading a snapshot: four rows per load (rows a multiple of 4
and the arrays aligned) and one row per load (otherwise).  Both must bin the same samples: the
MFMA images of the same snapshot read either way agree to FP32 summation order, and a snapshot
whose row count is not a multiple of 4 matches the tile gridder, whose binning is separate code.
"""
import numpy as np
import pytest
import torch

from conftest import announce, assert_parity
import svdimager_amd as S

pytestmark = pytest.mark.gpu


def _misaligned(dsnap):
    """The same snapshot with every per-sample array stored one element past an aligned address,
    so that no 4-sample load is aligned (the one-row path)."""
    out = S.DeviceSnapshot.__new__(S.DeviceSnapshot)
    out.num_rows, out.num_channels, out.array_center_x = dsnap.num_rows, dsnap.num_channels, dsnap.array_center_x
    out.tensors = {}
    for k, t in dsnap.tensors.items():
        if t is None or k in ("uvw", "freq"):
            out.tensors[k] = t
            continue
        buf = torch.empty(t.numel() + 1, dtype=t.dtype, device=t.device)
        view = buf[1:]
        view.copy_(t)
        assert view.data_ptr() % 4 != 0 or t.element_size() == 4
        out.tensors[k] = view
    return out


def test_four_row_and_one_row_loads_agree(gpu_ctx, small_snap):
    announce("binning loads: aligned (four rows) vs misaligned (one row), small snapshot")
    assert small_snap["num_rows"] % 4 == 0
    N, cell = small_snap["image_size"], small_snap["cell"]
    d = S.DeviceSnapshot(small_snap, "cuda")
    m = _misaligned(d)
    for output in (S.OUTPUT_SIGNED, S.OUTPUT_DIRTY):
        a, pa = S.image_snapshot(gpu_ctx, d, N, cell, output=output, gridder=S.GRIDDER_MFMA)
        b, pb = S.image_snapshot(gpu_ctx, m, N, cell, output=output, gridder=S.GRIDDER_MFMA)
        for k in ("elapsed_ms", "enqueued_early"):
            pa.pop(k)
            pb.pop(k)
        assert pa == pb
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.isfinite(a).all() and np.abs(a).max() > 0
        assert_parity(a, b, f"four-row vs one-row loads, output {output}", whole=1e-5, central=1e-6)


def test_rows_not_a_multiple_of_four(gpu_ctx, small_snap):
    announce("binning loads: 2015 rows (one-row path) vs the tile gridder")
    snap = dict(small_snap)
    R, C = small_snap["num_rows"] - 1, small_snap["num_channels"]
    snap["num_rows"] = R
    snap["uvw"] = np.ascontiguousarray(np.asarray(small_snap["uvw"]).reshape(-1, 3)[:R])
    for k in ("vis0_re", "vis0_im", "vis3_re", "vis3_im", "flag0", "flag3", "weight0", "weight3"):
        if small_snap.get(k) is not None:
            v = np.asarray(small_snap[k]).reshape(C, R + 1)
            snap[k] = np.ascontiguousarray(v[:, :R]).reshape(-1)
    N, cell = small_snap["image_size"], small_snap["cell"]
    d = S.DeviceSnapshot(snap, "cuda")
    ref, _ = S.image_snapshot(gpu_ctx, d, N, cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_TILED)
    img, _ = S.image_snapshot(gpu_ctx, d, N, cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_MFMA)
    ref, img = ref.cpu().numpy(), img.cpu().numpy()
    assert np.isfinite(img).all() and np.abs(ref).max() > 0
    assert_parity(img, ref, "2015 rows, MFMA vs tiles")
