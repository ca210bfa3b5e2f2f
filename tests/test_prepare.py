"""The port of the archive's preparation step (svdimager_amd/prepare.py, archive/SVD_MFS.py:
33-112) against the libraries the script itself calls: scikit-learn's PCA (mean, components and
transform) and the oracle's FP64 polarisation collapse; and its FITS files read back by the
library's FITS reader (the one the archive CLI uses)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
from svdimager_amd import prepare, read_fits_image


@pytest.fixture(scope="module")
def columns(small_snap):
    return prepare.measurement_set_columns(small_snap)


def test_pca_matches_scikit_learn(columns):
    sk = pytest.importorskip("sklearn.decomposition")
    vis, uvw, flag, freq, wgt = columns
    out = prepare.svd_mfs(vis, uvw, flag, freq, wgt, out_dir=None)
    D = prepare.fold_baselines(uvw, freq)
    pca = sk.PCA(3).fit(D)
    scale = np.abs(D).max()
    assert np.abs(out["Vin"] - pca.components_).max() <= 1e-10
    assert np.abs(out["Min"][0] - pca.mean_).max() <= 1e-12 * scale
    B1 = pca.transform(D)
    M = out["Min"]
    expect = B1 - np.stack([np.dot(-M, pca.components_[i]) for i in range(3)], axis=1).reshape(1, 3)
    assert np.abs(out["Bin"] - expect).max() <= 1e-9 * scale
    # Bin is the uncentred rotation D V^T, rows of V orthonormal.
    assert np.abs(out["Bin"] - D @ out["Vin"].T).max() <= 1e-9 * scale
    assert np.abs(out["Vin"] @ out["Vin"].T - np.eye(3)).max() <= 1e-12


def test_collapse_matches_the_oracle(small_snap, columns):
    vis, uvw, flag, freq, wgt = columns
    out = prepare.svd_mfs(vis, uvw, flag, freq, wgt, out_dir=None)
    re, im = O.collapse_f64(small_snap)
    assert np.abs(out["Visreal"] - re).max() <= 1e-6
    assert np.abs(out["Visimag"] - im).max() <= 1e-6
    # No WEIGHT_SPECTRUM: unit weights, flags still applied.
    out1 = prepare.svd_mfs(vis, uvw, flag, freq, None, out_dir=None)
    both = flag[:, :, 0] & flag[:, :, 3]
    assert np.all(out1["Visreal"][both.ravel(order="F")] == 0.0)


def test_fits_files_read_back(tmp_path, columns):
    vis, uvw, flag, freq, wgt = columns
    out = prepare.svd_mfs(vis, uvw, flag, freq, wgt, out_dir=str(tmp_path), index=3)
    S = vis.shape[0] * vis.shape[1]
    for name, shape in (("Bin", (S, 3)), ("Min", (1, 3)), ("Vin", (3, 3)), ("Visreal", (S,)),
                        ("Visimag", (S,))):
        p = os.path.join(str(tmp_path), f"{name}3.fits")
        assert os.path.getsize(p) % 2880 == 0
        back = read_fits_image(p)
        assert back.size == int(np.prod(shape)), name
        assert np.array_equal(back.reshape(-1), out[name].reshape(-1).astype(np.float32)), name
