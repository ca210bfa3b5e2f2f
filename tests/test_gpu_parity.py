"""GPU parity of the HIP imager against the oracle and the FP64 DFT (small snapshot).

Tolerances (relative to the peak |DFT| of the signed, pre-reprojection image):
* vs FP64 DFT: max <= 1.1e-4 over the whole image (the ES kernel's own aliasing error at the
  image edge is 9.6e-5 for support 8 at 1.5x oversampling), <= 2e-5 over the central half;
* vs the oracle's FP32 restatement of the reference on the same Bin / Vin / vis: <= 1e-5 over
  the central half, <= 1e-4 over the image (the grid correction amplifies FP32 rounding ~85x
  at the edge)
  (FP32 summation order, FP64 vs FP32 Bessel / w-phase in the image pass).
"""
import numpy as np
import pytest
import torch

from conftest import ARCHIVE_DIRTY, DIRTY_CENTRAL, announce, assert_parity
from oracle import oracle as O
import svdimager_amd as S

pytestmark = pytest.mark.gpu




def _peak(d):
    return float(np.nanmax(np.abs(d)))


def _rel(a, b, peak):
    return float(np.nanmax(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)))) / peak


def _central(a, N):
    q = N // 4
    return a[q:N - q, q:N - q]


@pytest.fixture(scope="module")
def dsnap(small_snap):
    return S.DeviceSnapshot(small_snap, "cuda")


@pytest.fixture(scope="module")
def oracle_pre(small_snap):
    return O.preprocess_ref32(small_snap)


def test_preprocess_matches_fp64_rotation(gpu_ctx, small_snap, small_dft, dsnap):
    announce("preprocess small snapshot")
    out = S.preprocess_measurement_set_gpu(gpu_ctx, dsnap)
    vin = out["vin"].cpu().numpy()
    assert np.abs(vin - small_dft["vin64"]).max() < 1e-6
    p64 = O.preprocess_f64(small_snap)
    b = out["bin"].cpu().numpy().astype(np.float64)
    assert np.abs(b - p64["bin"]).max() <= 1e-6 * np.abs(p64["bin"]).max()
    vre, vim = O.collapse_f64(small_snap)
    assert np.abs(out["vis_re"].cpu().numpy() - vre).max() < 1e-6
    assert np.abs(out["vis_im"].cpu().numpy() - vim).max() < 1e-6
    R = vin.astype(np.float64)
    assert np.abs(R @ R.T - np.eye(3)).max() < 1e-6
    assert np.linalg.det(R) > 0.999


@pytest.mark.parametrize("gridder", [S.GRIDDER_ATOMIC, S.GRIDDER_TILED, S.GRIDDER_MFMA])
def test_snapshot_signed_image_vs_dft(gpu_ctx, small_snap, small_dft, dsnap, gridder):
    announce(f"image_snapshot signed small K=4 gridder={gridder}")
    N, cell = small_snap["image_size"], small_snap["cell"]
    img, plan = S.image_snapshot(gpu_ctx, dsnap, N, cell, output=S.OUTPUT_SIGNED,
                                 num_terms=small_snap["num_terms"], gridder=gridder)
    img = img.cpu().numpy()
    dft = small_dft["dft"]
    peak = _peak(dft)
    err = _rel(img, dft, peak)
    err_c = _rel(_central(img, N), _central(dft, N), peak)
    rms = float(np.sqrt(np.mean((img - dft) ** 2))) / peak
    print(f"signed vs DFT: max {err:.3e} central {err_c:.3e} rms {rms:.3e} plan {plan}")
    assert plan["num_terms"] == small_snap["num_terms"]
    # The gridding method itself (ES support 8 at 1.5x oversampling, evaluated in FP64 by the
    # oracle) is 9.6e-5 of peak off the DFT at the image edge; FP32 sums then move edge pixels
    # by a few 1e-6 (the grid correction amplifies them ~85x there).
    assert err <= 1.1e-4
    assert err_c <= 2e-5


def test_fipipedevice_matches_oracle_ref32(gpu_ctx, small_snap, oracle_pre):
    announce("FIpipeDevice small vs oracle ref32 (bucket K)")
    N, cell = small_snap["image_size"], small_snap["cell"]
    t = {k: torch.from_numpy(np.ascontiguousarray(oracle_pre[k])).cuda()
         for k in ("bin", "vin", "vis_re", "vis_im")}
    ref = O.fipipe_device_ref32(oracle_pre["bin"], oracle_pre["vin"], oracle_pre["vis_re"],
                                oracle_pre["vis_im"], N, cell)
    for output, key in ((S.OUTPUT_SIGNED, "signed"), (S.OUTPUT_DIRTY, "dirty")):
        img, plan = S.FIpipeDevice(gpu_ctx, t["bin"], t["vin"], t["vis_re"], t["vis_im"], N, cell,
                                   output=output, gridder=S.GRIDDER_ATOMIC)
        img = img.cpu().numpy()
        assert plan["num_terms"] == ref["plan"]["num_terms"]
        assert plan["num_slabs"] == ref["plan"]["num_slabs"]
        assert abs(plan["r3_min"] - ref["plan"]["r3_min"]) == 0.0
        assert abs(plan["r3_max"] - ref["plan"]["r3_max"]) == 0.0
        assert plan["valid_pixels"] == ref["plan"]["valid_pixels"]
        assert_parity(img, ref[key], key, central=DIRTY_CENTRAL if key == "dirty" else 1e-5)


def test_default_options_take_the_mfma_gridder(gpu_ctx, small_snap, small_dft, dsnap, oracle_pre):
    """With default options (gridder AUTO) the snapshot path, FIpipeDevice and the archive FIpipe
    path (svdi_fipipe_host, which the C++ FIpipe symbol calls with default options) run the MFMA
    window gridder and meet their parity bounds."""
    announce("default options -> MFMA gridder")
    N, cell = small_snap["image_size"], small_snap["cell"]
    img, plan = S.image_snapshot(gpu_ctx, dsnap, N, cell, output=S.OUTPUT_SIGNED,
                                 num_terms=small_snap["num_terms"])
    assert plan["gridder"] == S.GRIDDER_MFMA
    dft = small_dft["dft"]
    peak = _peak(dft)
    img = img.cpu().numpy()
    assert _rel(img, dft, peak) <= 1.1e-4
    assert _rel(_central(img, N), _central(dft, N), peak) <= 2e-5
    t = {k: torch.from_numpy(np.ascontiguousarray(oracle_pre[k])).cuda()
         for k in ("bin", "vin", "vis_re", "vis_im")}
    ref = O.fipipe_device_ref32(oracle_pre["bin"], oracle_pre["vin"], oracle_pre["vis_re"],
                                oracle_pre["vis_im"], N, cell)
    for output, key in ((S.OUTPUT_SIGNED, "signed"), (S.OUTPUT_DIRTY, "dirty")):
        img, plan = S.FIpipeDevice(gpu_ctx, t["bin"], t["vin"], t["vis_re"], t["vis_im"], N, cell,
                                   output=output)
        assert plan["gridder"] == S.GRIDDER_MFMA
        assert_parity(img.cpu().numpy(), ref[key], f"default {key}",
                      central=DIRTY_CENTRAL if key == "dirty" else 1e-5)
    aref = O.fipipe_archive_ref32(oracle_pre["bin"], oracle_pre["vin"], oracle_pre["vis_re"],
                                  oracle_pre["vis_im"], N, cell)
    aimg, plan = S.fipipe_host(gpu_ctx, oracle_pre["vis_re"], oracle_pre["vis_im"], oracle_pre["bin"],
                               oracle_pre["vin"], N, cell, output=S.OUTPUT_SIGNED)
    assert plan["gridder"] == S.GRIDDER_MFMA
    assert_parity(aimg, aref["signed"], "default archive signed")


def test_fipipe_symbol_matches_archive_oracle(small_snap, oracle_pre):
    announce("FIpipe C++ symbol small vs oracle archive")
    N, cell = small_snap["image_size"], small_snap["cell"]
    S_ = oracle_pre["vis_re"].size
    dirty = np.zeros(N * N, np.float32)
    rc = S.FIpipe(oracle_pre["vis_re"].copy(), oracle_pre["vis_im"].copy(),
                  np.ascontiguousarray(oracle_pre["bin"]).reshape(-1).copy(),
                  np.ascontiguousarray(oracle_pre["vin"]).reshape(-1).copy(), dirty, S_, N, cell)
    assert rc == 0
    ref = O.fipipe_archive_ref32(oracle_pre["bin"], oracle_pre["vin"], oracle_pre["vis_re"],
                                 oracle_pre["vis_im"], N, cell)
    peak = _peak(ref["dirty"])
    d2 = dirty.reshape(N, N)
    err = _rel(d2, ref["dirty"], peak)
    err_c = _rel(_central(d2, N), _central(ref["dirty"], N), peak)
    print(f"archive dirty max rel {err:.3e} central {err_c:.3e}, K {ref['plan']['num_terms']}")
    assert err <= ARCHIVE_DIRTY and err_c <= ARCHIVE_DIRTY
    assert abs(float(dirty.sum()) - float(ref["dirty"].sum())) <= 1e-4 * float(np.abs(ref["dirty"]).sum())
    with S.Context(0) as ctx:
        ctx_img, plan = S.fipipe_host(ctx, oracle_pre["vis_re"], oracle_pre["vis_im"],
                                      oracle_pre["bin"], oracle_pre["vin"], N, cell,
                                      output=S.OUTPUT_SIGNED)
    assert plan["num_terms"] == ref["plan"]["num_terms"]
    assert_parity(ctx_img, ref["signed"], "archive signed")


def test_multislab_and_batching_vs_oracle(gpu_ctx, small_snap, oracle_pre):
    """Force several r3 slabs (small alpha target) and one slab per batch."""
    N, cell = small_snap["image_size"], small_snap["cell"]
    t = {k: torch.from_numpy(np.ascontiguousarray(oracle_pre[k])).cuda()
         for k in ("bin", "vin", "vis_re", "vis_im")}
    for batch in (0, 1):
        announce(f"multislab target_alpha=0.01 batch={batch}")
        ref = O.fipipe_device_ref32(oracle_pre["bin"], oracle_pre["vin"], oracle_pre["vis_re"],
                                    oracle_pre["vis_im"], N, cell, target_alpha=0.01,
                                    max_batch_slabs=(None if batch == 0 else 1))
        assert ref["plan"]["num_slabs"] > 1
        for output, key in ((S.OUTPUT_SIGNED, "signed"), (S.OUTPUT_DIRTY, "dirty")):
            img, plan = S.FIpipeDevice(gpu_ctx, t["bin"], t["vin"], t["vis_re"], t["vis_im"], N,
                                       cell, output=output, target_alpha=0.01,
                                       max_batch_slabs=batch, gridder=S.GRIDDER_ATOMIC)
            assert plan["num_slabs"] == ref["plan"]["num_slabs"]
            assert plan["batch_slabs"] == ref["plan"]["batch_slabs"]
            assert_parity(img.cpu().numpy(), ref[key], f"batch={batch} {key}",
                          central=DIRTY_CENTRAL if key == "dirty" else 1e-5)


@pytest.mark.parametrize("gridder", [S.GRIDDER_ATOMIC, S.GRIDDER_TILED, S.GRIDDER_MFMA])
@pytest.mark.parametrize("N", [6, 10, 14, 33, 90, 170])
def test_odd_grid_sizes_vs_oracle(gpu_ctx, small_snap, oracle_pre, N, gridder):
    """Image sizes whose grid G = ceil(1.5 N) is odd or whose N is odd; G = 9 .. 255, so the
    MFMA gridder's last 16-cell window and the tiled gridder's last tile are partial."""
    announce(f"odd sizes N={N} gridder={gridder}")
    cell = small_snap["cell"] * 128 / N
    t = {k: torch.from_numpy(np.ascontiguousarray(oracle_pre[k])).cuda()
         for k in ("bin", "vin", "vis_re", "vis_im")}
    ref = O.fipipe_device_ref32(oracle_pre["bin"], oracle_pre["vin"], oracle_pre["vis_re"],
                                oracle_pre["vis_im"], N, cell, num_terms=4)
    img, _ = S.FIpipeDevice(gpu_ctx, t["bin"], t["vin"], t["vis_re"], t["vis_im"], N, cell,
                            output=S.OUTPUT_SIGNED, num_terms=4, gridder=gridder)
    assert_parity(img.cpu().numpy(), ref["signed"], f"N={N} G={(3 * N + 1) // 2} gridder={gridder}")


@pytest.mark.parametrize("gridder", [S.GRIDDER_TILED, S.GRIDDER_MFMA])
@pytest.mark.parametrize("N", [1024, 2048, 4096])
def test_large_grid_vs_oracle(gpu_ctx, small_snap, oracle_pre, gridder, N):
    """N = 1024 (G = 1536: the wavefront transform of 24 x 64 points), N = 2048 (G = 3072:
    one radix-2 step over two such transforms, by two wavefronts per row or column; bins split
    into two LDS bands; a window layout of 192 x 192 windows) and N = 4096 (G = 6144: the general
    Stockham kernels' longest transform; eight LDS bands)."""
    announce(f"large grid N={N} gridder={gridder}")
    cell = small_snap["cell"] * 128 / N
    t = {k: torch.from_numpy(np.ascontiguousarray(oracle_pre[k])).cuda()
         for k in ("bin", "vin", "vis_re", "vis_im")}
    ref = O.fipipe_device_ref32(oracle_pre["bin"], oracle_pre["vin"], oracle_pre["vis_re"],
                                oracle_pre["vis_im"], N, cell, num_terms=4)
    img, _ = S.FIpipeDevice(gpu_ctx, t["bin"], t["vin"], t["vis_re"], t["vis_im"], N, cell,
                            output=S.OUTPUT_SIGNED, num_terms=4, gridder=gridder)
    assert_parity(img.cpu().numpy(), ref["signed"], f"N={N} G={(3 * N + 1) // 2}")


def test_zero_weight_samples(gpu_ctx, small_snap):
    """A sample whose unflagged polarisation has zero weight collapses to V = 0
    (SVDIntegratedImager.cu:588-617, the else branch), NaN visibilities included.  The MFMA
    window binning keeps such samples as records with V = 0 (its count pass reads flags, not
    weights); the tile gridder skips them.  Both must give the image of the same snapshot with
    those samples flagged in both polarisations."""
    announce("zero-weight samples with NaN visibilities")
    n = small_snap["num_rows"] * small_snap["num_channels"]
    pick = np.random.default_rng(11).choice(n, n // 5, replace=False)
    zw, ff = dict(small_snap), dict(small_snap)
    for k in ("flag0", "flag3", "weight0", "vis0_re", "vis0_im", "vis3_re", "vis3_im"):
        zw[k] = np.array(small_snap[k], copy=True)
        ff[k] = np.array(small_snap[k], copy=True)
    zw["flag0"][pick] = 0
    zw["flag3"][pick] = 1
    zw["weight0"][pick] = 0.0
    for k in ("vis0_re", "vis0_im", "vis3_re", "vis3_im"):
        zw[k][pick] = np.nan
    ff["flag0"][pick] = 1
    ff["flag3"][pick] = 1
    N, cell = small_snap["image_size"], small_snap["cell"]
    d_zw, d_ff = S.DeviceSnapshot(zw, "cuda"), S.DeviceSnapshot(ff, "cuda")
    ref, _ = S.image_snapshot(gpu_ctx, d_ff, N, cell, output=S.OUTPUT_SIGNED, gridder=S.GRIDDER_TILED)
    ref = ref.cpu().numpy()
    assert np.isfinite(ref).all() and np.abs(ref).max() > 0
    for g in (S.GRIDDER_MFMA, S.GRIDDER_TILED):
        img, _ = S.image_snapshot(gpu_ctx, d_zw, N, cell, output=S.OUTPUT_SIGNED, gridder=g)
        a = img.cpu().numpy()
        assert np.isfinite(a).all(), g
        assert_parity(a, ref, f"zero weights, gridder {g}")


def test_invalid_arguments(gpu_ctx, small_snap, dsnap):
    announce("invalid arguments")
    with pytest.raises(S.SvdiError):
        S.image_snapshot(gpu_ctx, dsnap, 0, 1e-4)
    with pytest.raises(S.SvdiError):
        S.image_snapshot(gpu_ctx, dsnap, 64, -1.0)
    with pytest.raises(S.SvdiError):
        S.image_snapshot(gpu_ctx, dsnap, 64, 1e-4, num_terms=33)


# Where the sampled Chebyshev error of K = 8 terms crosses the reference's 1e-2 target (FP64;
# the oracle pins the reference's FP32 rule within 1e-5 of it, tests/test_oracle_pins.py).
K8_THRESHOLD = 4.0946


@pytest.mark.parametrize("side", [-1, 1])
def test_bucket_choice_next_to_the_k8_threshold(gpu_ctx, small_snap, oracle_pre, side):
    """The slab alpha driven 0.1% below / above the K = 8 threshold (r3 scaled, one slab): the
    device picks K = 8 / 16 as the reference's FP32 bucket rule does on the same alpha."""
    N, cell = small_snap["image_size"], small_snap["cell"]
    b = np.ascontiguousarray(oracle_pre["bin"], np.float32).copy()
    stats = O.cheb_stats(b, oracle_pre["vin"], N, cell)
    alpha0 = 2.0 * np.pi * float(stats[2]) * 0.5 * (float(stats[1]) - float(stats[0]))
    target = K8_THRESHOLD * (1.0 + 1e-3 * side)
    b[:, 2] *= np.float32(target / alpha0)
    announce(f"bucket choice at alpha {target:.5f}")
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda()
         for k, v in (("bin", b), ("vin", oracle_pre["vin"]), ("vis_re", oracle_pre["vis_re"]),
                      ("vis_im", oracle_pre["vis_im"]))}
    _, plan = S.FIpipeDevice(gpu_ctx, t["bin"], t["vin"], t["vis_re"], t["vis_im"], N, cell,
                             output=S.OUTPUT_SIGNED)
    assert plan["num_slabs"] == 1
    assert abs(plan["slab_alpha_max"] / target - 1.0) < 1e-4
    expected = 8 if side < 0 else 16
    assert plan["num_terms"] == expected
    assert O.select_bucket(O.bucket_errors_ref32(plan["slab_alpha_max"])) == expected
