"""GPU parity of the FP32 tile gridder and the MFMA window gridder against the reference-style
atomic gridder and the oracle.

The gridders grid the same sample model.  The FP32 ones differ only in summation order; the
MFMA one contracts each 16 x 16 window's samples on f16 hi/lo operands with FP32 accumulation
(~22 significant bits per product).  The signed images agree to 1e-5 of peak over the central
half and 1e-4 over the image (the grid correction amplifies rounding up to ~85x at the image
edge).  Cases: the small snapshot (snapshot and preprocessed inputs), several slabs,
exact-integer positions (nine-cell footprints), one hot position that forces split work items,
footprints clipped at the grid edge, every supported K range (the MFMA gridder takes terms eight
at a time, so K = 12, 16, 24, 32 exercise its multi-pass path).
"""
import numpy as np
import pytest
import torch

from conftest import DIRTY_CENTRAL, announce, assert_parity
from oracle import oracle as O
import svdimager_amd as S

pytestmark = pytest.mark.gpu

GRIDDERS = {
    "tiled": dict(gridder=S.GRIDDER_TILED),
    "mfma": dict(gridder=S.GRIDDER_MFMA),
}
gridders = pytest.mark.parametrize("gname", list(GRIDDERS))


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _dev(arrs):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in arrs.items()}


def _image(ctx, t, N, cell, **kw):
    img, plan = S.FIpipeDevice(ctx, t["bin"], t["vin"], t["vis_re"], t["vis_im"], N, cell, **kw)
    return img.cpu().numpy(), plan


@pytest.fixture(scope="module")
def small_pre(small_snap):
    return O.preprocess_ref32(small_snap)


@gridders
def test_snapshot_tiled_vs_atomic(gpu_ctx, small_snap, gname):
    announce(f"snapshot {gname} vs atomic")
    d = S.DeviceSnapshot(small_snap, "cuda")
    N, cell = small_snap["image_size"], small_snap["cell"]
    ref, _ = S.image_snapshot(gpu_ctx, d, N, cell, output=S.OUTPUT_SIGNED, num_terms=4,
                              gridder=S.GRIDDER_ATOMIC)
    img, plan = S.image_snapshot(gpu_ctx, d, N, cell, output=S.OUTPUT_SIGNED, num_terms=4,
                                 **GRIDDERS[gname])
    assert plan["records"] > 0
    assert_parity(img.cpu().numpy(), ref.cpu().numpy(), f"{gname} vs atomic records {plan['records']}")


@gridders
@pytest.mark.parametrize("K", [1, 4, 8, 12, 16, 24, 32])
def test_preprocessed_tiled_vs_oracle(gpu_ctx, small_snap, small_pre, K, gname):
    announce(f"preprocessed {gname} K={K}")
    N, cell = small_snap["image_size"], small_snap["cell"]
    ref = O.fipipe_device_ref32(small_pre["bin"], small_pre["vin"], small_pre["vis_re"],
                                small_pre["vis_im"], N, cell, num_terms=K)
    img, plan = _image(gpu_ctx, _dev(small_pre), N, cell, output=S.OUTPUT_SIGNED, num_terms=K,
                       **GRIDDERS[gname])
    assert_parity(img, ref["signed"], f"K={K} {gname} vs oracle")


@gridders
def test_tiled_multislab_batches(gpu_ctx, small_snap, small_pre, gname):
    N, cell = small_snap["image_size"], small_snap["cell"]
    t = _dev(small_pre)
    for batch in (0, 2):
        announce(f"{gname} multislab batch={batch}")
        ref = O.fipipe_device_ref32(small_pre["bin"], small_pre["vin"], small_pre["vis_re"],
                                    small_pre["vis_im"], N, cell, target_alpha=0.005,
                                    max_batch_slabs=(None if batch == 0 else batch))
        for output, key in ((S.OUTPUT_SIGNED, "signed"), (S.OUTPUT_DIRTY, "dirty")):
            img, plan = _image(gpu_ctx, t, N, cell, output=output, target_alpha=0.005,
                               max_batch_slabs=batch, **GRIDDERS[gname])
            assert plan["num_slabs"] == ref["plan"]["num_slabs"] > 2
            assert_parity(img, ref[key], f"batch={batch} {key} slabs {plan['num_slabs']}",
                          central=DIRTY_CENTRAL if key == "dirty" else 1e-5)


def _synthetic_bin(n, rng, scale, spread):
    """Samples with random Bin in a given spread (cells), a fixed rotation, random vis."""
    p = rng.uniform(-spread, spread, size=(n, 2)) / scale
    r3 = rng.uniform(-5.0, 5.0, size=n)
    b = np.concatenate([p, r3[:, None]], axis=1).astype(np.float32)
    ang = 0.3
    vin = np.array([[np.cos(ang), -np.sin(ang), 0.02], [np.sin(ang), np.cos(ang), 0.03],
                    [-0.02, -0.03, 0.9993]], np.float32)
    return {"bin": b, "vin": vin, "vis_re": rng.normal(size=n).astype(np.float32),
            "vis_im": rng.normal(size=n).astype(np.float32)}


def test_integer_positions_nine_cell_footprints(gpu_ctx):
    """Positions Bin * (cell G) that are exact integers in FP32, where the footprint has nine
    cells per axis (SVDIntegratedImager.cu:1261-1267 includes both ends)."""
    announce("integer positions")
    rng = np.random.default_rng(11)
    N, cell = 128, 1e-3
    G = (3 * N + 1) // 2
    scale = np.float32(cell) * np.float32(G)
    a = _synthetic_bin(3000, rng, float(scale), 60.0)
    ints = rng.integers(-60, 60, size=(4000, 2)).astype(np.float32)
    cand = (ints / scale).astype(np.float32)
    exact = np.all(cand * scale == ints, axis=1)
    cand = cand[exact][:1500]
    assert cand.shape[0] > 200
    a["bin"][:cand.shape[0], :2] = cand
    ref = O.fipipe_device_ref32(a["bin"], a["vin"], a["vis_re"], a["vis_im"], N, cell, num_terms=8)
    t = _dev(a)
    cases = dict(GRIDDERS, atomic=dict(gridder=S.GRIDDER_ATOMIC))
    for gname, kw in cases.items():
        img, _ = _image(gpu_ctx, t, N, cell, output=S.OUTPUT_SIGNED, num_terms=8, **kw)
        assert_parity(img, ref["signed"], f"gridder {gname}:")


@gridders
def test_hot_position_splits_items(gpu_ctx, gname):
    """Thousands of samples at nearly one position exceed a work item and exercise the split
    (zeroed tile + atomic write-back) path; edge samples exercise grid clipping."""
    announce(f"hot position + grid edge, {gname}")
    rng = np.random.default_rng(12)
    N, cell = 128, 1e-4
    G = (3 * N + 1) // 2
    scale = np.float32(cell) * np.float32(G)
    a = _synthetic_bin(20000, rng, float(scale), 40.0)
    a["bin"][:6000, :2] = (rng.normal(0.0, 0.3, size=(6000, 2)) / scale).astype(np.float32)
    edge = np.array([G // 2 - 2.5, -(G // 2) + 1.7, G // 2 + 3.2, G // 2 - 6.0], np.float32)
    for i, e in enumerate(edge):
        a["bin"][6000 + i, 0] = e / scale
        a["bin"][6010 + i, 1] = e / scale
    ref = O.fipipe_device_ref32(a["bin"], a["vin"], a["vis_re"], a["vis_im"], N, cell, num_terms=8)
    img, plan = _image(gpu_ctx, _dev(a), N, cell, output=S.OUTPUT_SIGNED, num_terms=8,
                       **GRIDDERS[gname])
    assert_parity(img, ref["signed"], f"hot/edge {gname}: records {plan['records']}")


@gridders
def test_archive_tiled_vs_oracle(gpu_ctx, small_snap, small_pre, gname):
    announce(f"archive semantics {gname}")
    N, cell = small_snap["image_size"], small_snap["cell"]
    ref = O.fipipe_archive_ref32(small_pre["bin"], small_pre["vin"], small_pre["vis_re"],
                                 small_pre["vis_im"], N, cell)
    img, plan = _image(gpu_ctx, _dev(small_pre), N, cell, output=S.OUTPUT_SIGNED,
                       semantics=S.SEMANTICS_ARCHIVE, **GRIDDERS[gname])
    assert plan["num_terms"] == ref["plan"]["num_terms"]
    assert_parity(img, ref["signed"], f"archive {gname}")
