"""svdi_image_snapshot_noalloc, the entry point that allocates no device memory: its scratch is
carved from the library's fixed device array.  Checked through the C-ABI against
svdi_image_snapshot on the same context, on two calls in a row, the second with a smaller
snapshot and image (the array keeps the first call's contents), then a request beyond the
array's size, which must fail with a clear error and leave the context usable.

Both calls run the MFMA window gridder, whose FP32 sums follow the record order that the
binning's LDS atomics set, so the images agree to rounding, not bit for bit.
"""
import numpy as np
import pytest

from conftest import DIRTY_CENTRAL, announce, assert_parity, load_small_snapshot
import svdimager_amd as S
from svdimager_amd import _lib

pytestmark = pytest.mark.gpu


def _first_rows(snap, rows):
    """The snapshot's first `rows` baselines, every channel (sample index row + chan R)."""
    R, C = snap["num_rows"], snap["num_channels"]
    out = dict(snap)
    out["uvw"] = np.ascontiguousarray(snap["uvw"][:rows])
    for k in ("vis0_re", "vis0_im", "vis3_re", "vis3_im", "flag0", "flag3", "weight0", "weight3"):
        out[k] = np.ascontiguousarray(snap[k].reshape(C, R)[:, :rows].reshape(-1))
    out["num_rows"] = rows
    return out


def _compare(ctx, snap, N, cell):
    d = S.DeviceSnapshot(snap, "cuda")
    for output in (S.OUTPUT_SIGNED, S.OUTPUT_DIRTY):
        announce(f"noalloc rows={snap['num_rows']} N={N} output={output}")
        ref, plan_r = S.image_snapshot(ctx, d, N, cell, output=output, max_batch_slabs=1)
        img, plan_n = S.image_snapshot(ctx, d, N, cell, output=output, noalloc=True)
        assert plan_n["gridder"] == plan_r["gridder"] == S.GRIDDER_MFMA
        assert plan_n["num_terms"] == plan_r["num_terms"]
        assert plan_n["num_slabs"] == plan_r["num_slabs"]
        assert np.isfinite(img.cpu().numpy()).all()
        assert_parity(img.cpu().numpy(), ref.cpu().numpy(), f"noalloc vs allocating ({output})",
                      whole=1e-5, central=1e-6 if output == S.OUTPUT_SIGNED else DIRTY_CENTRAL)


def test_noalloc_matches_the_allocating_call():
    snap = load_small_snapshot()
    cell = snap["cell"]
    with S.Context(0) as ctx:
        _compare(ctx, snap, snap["image_size"], cell)          # 2016 rows, 128^2
        small = _first_rows(snap, 700)
        _compare(ctx, small, 64, cell)                          # 700 rows, 64^2


def test_noalloc_rejects_a_call_beyond_its_scratch():
    snap = load_small_snapshot()
    with S.Context(0) as ctx:
        d = S.DeviceSnapshot(snap, "cuda")
        announce("noalloc oversize request (4096^2 planes exceed the scratch)")
        with pytest.raises(_lib.SvdiError) as e:
            S.image_snapshot(ctx, d, 4096, snap["cell"], noalloc=True)
        assert _lib.STATUS_NAMES[e.value.status] == "INVALID_ARGUMENT"
        assert "scratch" in str(e.value)
        _compare(ctx, _first_rows(snap, 500), 48, snap["cell"])
