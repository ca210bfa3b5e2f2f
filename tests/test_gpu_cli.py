"""The archive CLI end to end on the GPU (archive/pipe_interface.cpp:139-186): the small
snapshot prepared into FITS files by the port of SVD_MFS.py, then sharedlibrary_gpu
Visreal.fits Visimag.fits B.fits V.fits Image_Size Number_of_Rows Cell_Size Output.fits, and
the output FITS read back and compared with the oracle's archive restatement (FIpipe semantics)
on the same FP32 inputs.

Bound: the dirty image within ARCHIVE_DIRTY of its peak (the reference's degree-trig SIN
reprojection is ill-conditioned near 90 degrees: host and device libm move splat targets by
about 1e-3 pixel), total flux within 1e-4.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ARCHIVE_DIRTY, ROOT, announce, assert_parity
from oracle import oracle as O
from svdimager_amd import prepare, read_fits_image

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "svdimager_amd", "bin", "sharedlibrary_gpu")


def test_sharedlibrary_gpu_end_to_end(tmp_path, small_snap):
    assert os.path.exists(CLI), "the CLI is built with the library"
    cols = prepare.measurement_set_columns(small_snap)
    out = prepare.svd_mfs(*cols, out_dir=str(tmp_path), index=0)
    p = out["paths"]
    N = small_snap["image_size"]
    cell = float(np.float32(small_snap["cell"]))
    S = out["Visreal"].size
    dirty_path = str(tmp_path / "dirty.fits")
    announce(f"sharedlibrary_gpu CLI small snapshot N={N} S={S}")
    r = subprocess.run([CLI, p["Visreal"], p["Visimag"], p["Bin"], p["Vin"], str(N), str(S),
                        repr(cell), dirty_path], capture_output=True, text=True, timeout=180)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    assert "Chebyshev range check" in r.stdout and "Time elapsed" in r.stdout
    dirty = read_fits_image(dirty_path).reshape(N, N)
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    ref = O.fipipe_archive_ref32(f32(out["Bin"]), f32(out["Vin"]), f32(out["Visreal"]),
                                 f32(out["Visimag"]), N, cell)
    peak = float(np.abs(ref["dirty"]).max())
    err = float(np.abs(dirty - ref["dirty"]).max()) / peak
    print(f"CLI dirty vs oracle archive: max {err:.3e}, K {ref['plan']['num_terms']}")
    assert err <= ARCHIVE_DIRTY
    assert abs(float(dirty.sum()) - float(ref["dirty"].sum())) <= 1e-4 * float(np.abs(ref["dirty"]).sum())
    assert np.all(dirty >= 0.0)


INTEGRATED_CLI = os.path.join(ROOT, "svdimager_amd", "bin", "svd_integrated_imager_gpu")


def test_svd_integrated_imager_gpu_end_to_end(tmp_path, small_snap):
    """The integrated CLI (SVDIntegratedImager.cu:1763-1854) on a column directory of two TIME
    snapshots: the small snapshot at TIME 100 and a copy with half its visibilities at TIME 50.
    Snapshot 1 must be the small snapshot's dirty image as the library's device path makes it
    (the same kernels; the order in which records reach a bin and float atomics reach a pixel
    varies from run to run, so the bound is the one between two runs of the device path: 1e-5 of
    peak over the image, 1e-6 over its central half), snapshot 0 half of it (the dirty image is
    linear in |v|), and the stdout / stderr lines the reference's."""
    import torch
    import svdimager_amd as S
    from svdimager_amd import prepare
    assert os.path.exists(INTEGRATED_CLI), "the CLI is built with the library"
    half = dict(small_snap)
    for k in ("vis0_re", "vis0_im", "vis3_re", "vis3_im"):
        half[k] = (0.5 * small_snap[k]).astype(np.float32)
    ms = prepare.save_measurement_set_columns(str(tmp_path / "ms"), [small_snap, half], [100.0, 50.0])
    N = small_snap["image_size"]
    cell = float(np.float32(small_snap["cell"]))
    images = {}
    for idx in (1, 0):
        path = str(tmp_path / f"snap{idx}.fits")
        announce(f"svd_integrated_imager_gpu small snapshot index {idx} N={N}")
        r = subprocess.run([INTEGRATED_CLI, ms, str(N), repr(cell), str(idx), path], capture_output=True, text=True,
                           timeout=180)
        print(r.stderr[-1500:], r.stdout[-1500:])
        assert r.returncode == 0, r.stderr
        R, C = small_snap["num_rows"], small_snap["num_channels"]
        assert "Found 2 TIME snapshot(s)" in r.stderr
        assert f"Selected snapshot {idx} time {100 if idx else 50} rows {R} samples {R * C}" in r.stderr
        assert f"Wrote image FITS: {path}" in r.stderr and "Total pipeline time:" in r.stdout
        images[idx] = read_fits_image(path).reshape(N, N)
    ctx = S.Context(0)
    try:
        announce("device-path reference image of the small snapshot")
        ref, _ = S.image_snapshot(ctx, S.DeviceSnapshot(small_snap, "cuda"), N, cell, output=S.OUTPUT_DIRTY)
        ref = ref.cpu().numpy()
        announce("host-snapshot entry point on the read-back snapshot")
        host, _ = S.image_host_snapshot(ctx, S.read_measurement_set_snapshot(ms, 1), N, cell, output=S.OUTPUT_DIRTY)
        torch.cuda.synchronize()
    finally:
        ctx.close()
    assert float(np.abs(ref).max()) > 0
    for name, img in (("CLI", images[1]), ("host entry", host), ("CLI half / 0.5", images[0] / 0.5)):
        assert_parity(img, ref, f"{name} vs device path", whole=1e-5, central=1e-6)


@pytest.mark.parametrize("weights", ["both", "none", "xx_only"])
def test_host_snapshot_entry_matches_device_path(small_snap, weights):
    """svdi_image_host_snapshot (host arrays copied in by the library) against
    svdi_image_snapshot on the same arrays already on the device, signed image: with both weight
    arrays, without WEIGHT_SPECTRUM (weights 1), and with only polarisation 0's weights."""
    import svdimager_amd as S
    snap = dict(small_snap)
    if weights == "none":
        snap["weight0"] = snap["weight3"] = None
    elif weights == "xx_only":
        snap["weight3"] = None
    N = small_snap["image_size"]
    cell = float(np.float32(small_snap["cell"]))
    ctx = S.Context(0)
    try:
        announce(f"host-snapshot entry, weights {weights}")
        host, hplan = S.image_host_snapshot(ctx, snap, N, cell, output=S.OUTPUT_SIGNED)
        dev, dplan = S.image_snapshot(ctx, S.DeviceSnapshot(snap, "cuda"), N, cell, output=S.OUTPUT_SIGNED)
        dev = dev.cpu().numpy()
    finally:
        ctx.close()
    assert hplan["num_terms"] == dplan["num_terms"] and hplan["num_slabs"] == dplan["num_slabs"]
    assert_parity(host, dev, f"host entry vs device path ({weights})", whole=1e-5, central=1e-6)
