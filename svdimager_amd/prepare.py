"""The archive's input preparation (archive/SVD_MFS.py:33-112), ported: one snapshot's
MeasurementSet columns -> the FITS inputs of the archive CLI sharedlibrary_gpu and of FIpipe.

    out = svd_mfs(vis_data, uvw, flag, chan_freq, weight_spectrum, out_dir=".", index=0)

writes Bin{index}.fits, Min{index}.fits, Vin{index}.fits, Visreal{index}.fits and
Visimag{index}.fits as the script does (FP64 primary-HDU images, overwriting):

* MFS fold: D = (u, v, w) nu / c per (row, channel), sample index row + chan R
  (SVD_MFS.py:33-47, the reference's Fortran-order ravel);
* PCA of D (SVD_MFS.py:49-72): mean M, components V = the eigenvectors of the centred scatter by
  decreasing variance, each row's sign set so that its largest-magnitude entry is positive (the
  convention of scikit-learn's PCA, svd_flip with u_based_decision=False, in the 1.7 series the
  script runs under here; earlier releases chose signs from U, which the script does not pin);
* Bin = PCA transform + M V^T = D V^T, the rotated uncentred baselines (SVD_MFS.py:95-106);
* Visreal / Visimag: the XX / YY collapse with WEIGHT_SPECTRUM and flags, 0 where no weight
  remains (SVD_MFS.py:74-93).

The arithmetic is FP64 numpy on the host, as in the script: this is the data preparation ahead of
the imager, not the imaging path.  casacore is not needed; the caller passes the columns.
"""
from __future__ import annotations

import os
from typing import Dict, Optional

import numpy as np

C_LIGHT = 299792458.0          # SVD_MFS.py:7


def fold_baselines(uvw: np.ndarray, chan_freq: np.ndarray) -> np.ndarray:
    """(R, 3) metres x (C,) Hz -> (R C, 3) wavelengths, sample index row + chan R."""
    nu = np.ravel(np.asarray(chan_freq, np.float64))
    uvw = np.asarray(uvw, np.float64)
    # (C, R, 3): channel-major, so that the flattened index is row + chan R.
    return (uvw[None, :, :] * (nu / C_LIGHT)[:, None, None]).reshape(-1, 3)


def pca_basis(D: np.ndarray):
    """Mean and principal axes (rows, by decreasing variance) of the rows of D, with each row's
    largest-magnitude entry made positive."""
    mean = D.mean(axis=0)
    X = D - mean
    scatter = X.T @ X
    evals, evecs = np.linalg.eigh(scatter)                 # ascending
    V = evecs[:, ::-1].T.copy()                            # rows = axes, decreasing variance
    for i in range(3):
        j = int(np.argmax(np.abs(V[i])))
        if V[i, j] < 0:
            V[i] = -V[i]
    return mean, V, evals[::-1]


def collapse(vis_data: np.ndarray, flag: np.ndarray, weight_spectrum: Optional[np.ndarray]):
    """Weighted XX / YY average per (row, channel), flattened as row + chan R."""
    vis_data = np.asarray(vis_data)
    if weight_spectrum is None:
        weight_spectrum = np.ones(vis_data.shape, np.float64)
    w0 = np.asarray(weight_spectrum[:, :, 0], np.float64) * (~np.asarray(flag[:, :, 0], bool))
    w3 = np.asarray(weight_spectrum[:, :, 3], np.float64) * (~np.asarray(flag[:, :, 3], bool))
    ws = w0 + w3
    safe = np.where(ws > 0, ws, 1.0)
    acc = vis_data[:, :, 0].astype(np.complex128) * w0 + vis_data[:, :, 3].astype(np.complex128) * w3
    re = np.where(ws > 0, acc.real / safe, 0.0)
    im = np.where(ws > 0, acc.imag / safe, 0.0)
    return re.ravel(order="F"), im.ravel(order="F")


def write_fits_f64(path: str, data: np.ndarray) -> None:
    """Primary-HDU FITS image of FP64 values (BITPIX -64), NAXIS1 = the last numpy axis,
    replacing an existing file."""
    a = np.ascontiguousarray(data, dtype=">f8")
    cards = ["SIMPLE  =                    T", "BITPIX  =                  -64",
             f"NAXIS   = {a.ndim:20d}"]
    for k, n in enumerate(reversed(a.shape)):
        cards.append(f"NAXIS{k + 1:<3d}= {n:20d}")
    cards += ["EXTEND  =                    T", "END"]
    header = "".join(c.ljust(80) for c in cards).encode("ascii")
    header += b" " * (-len(header) % 2880)
    body = a.tobytes()
    body += b"\0" * (-len(body) % 2880)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(header)
        f.write(body)
    os.replace(tmp, path)


def svd_mfs(vis_data: np.ndarray, uvw: np.ndarray, flag: np.ndarray, chan_freq: np.ndarray,
            weight_spectrum: Optional[np.ndarray] = None, out_dir: Optional[str] = ".",
            index: int = 0) -> Dict:
    """SVD_MFS.py on one snapshot's columns: DATA (R, C, npol) complex, UVW (R, 3), FLAG
    (R, C, npol) bool, CHAN_FREQ (C,) or (1, C), WEIGHT_SPECTRUM (R, C, npol) or None.
    Writes the five FITS files into out_dir (None: write nothing) and returns the arrays."""
    D = fold_baselines(uvw, chan_freq)
    mean, V, _ = pca_basis(D)
    B = (D - mean) @ V.T                                   # the PCA transform
    Min = mean[None, :]                                    # gro = 1 group: shape (1, 3)
    Bin = B + (Min @ V.T)                                  # SVD_MFS.py:100-106: = D V^T
    Visreal, Visimag = collapse(vis_data, flag, weight_spectrum)
    out = {"Bin": Bin, "Min": Min, "Vin": V, "Visreal": Visreal, "Visimag": Visimag}
    if out_dir is not None:
        paths = {}
        for name in ("Bin", "Min", "Vin", "Visreal", "Visimag"):
            p = os.path.join(out_dir, f"{name}{index}.fits")
            write_fits_f64(p, out[name])
            paths[name] = p
        out["paths"] = paths
    return out


def measurement_set_columns(snap: Dict):
    """A snapshot dict (svdi_snapshot fields, sample index row + chan R) as the MeasurementSet
    columns SVD_MFS.py reads: DATA and FLAG and WEIGHT_SPECTRUM with four polarisations (XX in
    0, YY in 3), UVW and CHAN_FREQ."""
    R, C = int(snap["num_rows"]), int(snap["num_channels"])
    rc = lambda a: np.asarray(a).reshape(C, R).T           # (R, C) from row + chan R
    vis = np.zeros((R, C, 4), np.complex64)
    vis[:, :, 0] = rc(snap["vis0_re"]) + 1j * rc(snap["vis0_im"])
    vis[:, :, 3] = rc(snap["vis3_re"]) + 1j * rc(snap["vis3_im"])
    flag = np.zeros((R, C, 4), bool)
    flag[:, :, 0] = rc(snap["flag0"]) != 0
    flag[:, :, 3] = rc(snap["flag3"]) != 0
    wgt = np.ones((R, C, 4), np.float32)
    if snap.get("weight0") is not None:
        wgt[:, :, 0] = rc(snap["weight0"])
        wgt[:, :, 3] = rc(snap["weight3"])
    return vis, np.asarray(snap["uvw"], np.float64), flag, np.asarray(snap["freq"], np.float64), wgt


def save_measurement_set_columns(out_dir: str, snaps, times) -> str:
    """Snapshot dicts (svdi_snapshot fields, one channel list shared by all) saved as the column
    directory that svd_integrated_imager_gpu and svdi_ms_read_snapshot read: the rows of
    snapshot i carry TIME times[i]; DATA / FLAG / WEIGHT_SPECTRUM as (rows, channels, 4) with XX
    in polarisation 0 and YY in 3, as python-casacore's getcol returns them; CHAN_FREQ one
    spectral window; ARRAY_CENTER (x, 0, 0) from the first snapshot.  Returns out_dir."""
    cols = [measurement_set_columns(s) for s in snaps]
    freq = cols[0][3]
    if any(not np.array_equal(c[3], freq) for c in cols):
        raise ValueError("the snapshots must share one channel list")
    os.makedirs(os.path.join(out_dir, "SPECTRAL_WINDOW"), exist_ok=True)
    os.makedirs(os.path.join(out_dir, "OBSERVATION"), exist_ok=True)
    time = np.concatenate([np.full(c[1].shape[0], float(t)) for c, t in zip(cols, times)])
    np.save(os.path.join(out_dir, "TIME.npy"), time)
    np.save(os.path.join(out_dir, "UVW.npy"), np.concatenate([c[1] for c in cols]))
    np.save(os.path.join(out_dir, "DATA.npy"), np.concatenate([c[0] for c in cols]))
    np.save(os.path.join(out_dir, "FLAG.npy"), np.concatenate([c[2] for c in cols]))
    np.save(os.path.join(out_dir, "WEIGHT_SPECTRUM.npy"), np.concatenate([c[4] for c in cols]))
    np.save(os.path.join(out_dir, "SPECTRAL_WINDOW", "CHAN_FREQ.npy"), freq.reshape(1, -1))
    np.save(os.path.join(out_dir, "OBSERVATION", "ARRAY_CENTER.npy"),
            np.array([[float(snaps[0]["array_center_x"]), 0.0, 0.0]]))
    return out_dir
