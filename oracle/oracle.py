"""ORACLE — test infrastructure only (see oracle/ref_imager.cpp).

ctypes front end of ``liboracle.so`` plus the reference's host-side orchestration restated in
Python: ``fipipe_device_ref32`` follows FIpipeDevice (SVDIntegratedImager.cu:1498-1761) and
``fipipe_archive_ref32`` follows FIpipe (archive/FIPkernels.cu:679-911).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
CONV_CORR_NORM = 2.4937047051153827      # SVDIntegratedImager.cu:1512
CHEB_BUCKETS = (8, 16, 24, 32)           # SVDIntegratedImager.cu:1577
DEFAULT_TARGET_ALPHA = 10.0              # SVDIntegratedImager.cu:37

_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
sz = ctypes.c_size_t


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_gl28": (None, [f64p, f64p, f64p]),
            "orc_conv_corr_norm_factor": (ctypes.c_double, []),
            "orc_es_f32": (ctypes.c_float, [ctypes.c_float]),
            "orc_cheb_seq_f32": (None, [ctypes.c_float, ctypes.c_int, f32p]),
            "orc_bessel_seq_f32": (None, [ctypes.c_float, ctypes.c_int, f32p]),
            "orc_jacobi3": (None, [f64p, f64p, f64p]),
            "orc_preprocess_ref32": (None, [sz, sz, f32p, f32p, f32p, f32p, f32p, f32p, u8p, u8p,
                                            f32p, f32p, ctypes.c_float, f32p, f32p, f32p, f32p,
                                            f32p]),
            "orc_preprocess_f64": (None, [sz, sz, f64p, f64p, ctypes.c_double, f64p, f64p, f64p]),
            "orc_cheb_stats": (None, [f32p, sz, f32p, sz, ctypes.c_float, f32p]),
            "orc_slab_config": (None, [ctypes.c_float] * 4 + [f32p]),
            "orc_bucket_errors_ref32": (None, [ctypes.c_float, f32p]),
            "orc_select_bucket": (ctypes.c_int, [f32p]),
            "orc_archive_select_terms": (ctypes.c_int, [ctypes.c_float, f32p]),
            "orc_conv_corr": (None, [sz, sz, ctypes.c_int, f32p]),
            "orc_grid_svdi": (None, [f32p, f32p, f32p, sz, f32p, ctypes.c_float, sz,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     f32p]),
            "orc_grid_archive": (None, [f32p, f32p, f32p, sz, ctypes.c_float, ctypes.c_float, sz,
                                        ctypes.c_float, ctypes.c_float, ctypes.c_int, f32p]),
            "orc_ifft2": (None, [f32p, sz, sz]),
            "orc_ifft1_f64": (None, [f64p, sz]),
            "orc_reproject": (None, [sz, ctypes.c_float, f32p, ctypes.c_int, f32p]),
            "orc_splat": (None, [sz, f32p, f32p, f32p]),
            "orc_image_svdi": (None, [f32p, sz, sz, ctypes.c_float, f32p, f32p, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      f32p, ctypes.c_void_p]),
            "orc_image_archive": (None, [f32p, sz, sz, ctypes.c_float, f32p, f32p,
                                         ctypes.c_float, sz, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_int, f32p, ctypes.c_void_p]),
            "orc_dft64": (None, [f64p, f64p, f64p, ctypes.c_void_p, sz, f64p, sz,
                                 ctypes.c_double, i32p, sz, f64p]),
            "orc_num_threads": (ctypes.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ------------------------------------------------------------------ tables and helpers
def gl28():
    n, w, k = (np.zeros(14) for _ in range(3))
    lib().orc_gl28(n, w, k)
    return n, w, k


def norm_factor() -> float:
    return lib().orc_conv_corr_norm_factor()


def es_f32(x: float) -> float:
    return lib().orc_es_f32(x)


def cheb_seq(s: float, n: int) -> np.ndarray:
    out = np.zeros(max(n, 2), np.float32)
    lib().orc_cheb_seq_f32(s, n, out)
    return out[:n]


def bessel_seq_f32(alpha: float, max_order: int) -> np.ndarray:
    out = np.zeros(max_order + 1, np.float32)
    lib().orc_bessel_seq_f32(alpha, max_order, out)
    return out


def jacobi3(a):
    ev, vec = np.zeros(3), np.zeros(9)
    lib().orc_jacobi3(_c(a, np.float64).reshape(9), ev, vec)
    return ev, vec.reshape(3, 3).T      # columns are eigenvectors


def conv_corr(N: int, G: int, variant: int = 0) -> np.ndarray:
    out = np.zeros(N // 2 + 1, np.float32)
    lib().orc_conv_corr(N, G, variant, out)
    return out


def grid_size(N: int) -> int:
    """G = ceil(3N/2) (SVDIntegratedImager.cu:1510)."""
    return (3 * N + 1) // 2


# ------------------------------------------------------------------ preprocess
def preprocess_ref32(snap: Dict) -> Dict:
    R, C = int(snap["num_rows"]), int(snap["num_channels"])
    S = R * C
    bin_ = np.zeros(S * 3, np.float32)
    vin = np.zeros(9, np.float32)
    vre = np.zeros(S, np.float32)
    vim = np.zeros(S, np.float32)
    ev = np.zeros(3, np.float32)
    lib().orc_preprocess_ref32(
        R, C, _c(snap["uvw"], np.float32).reshape(-1), _c(snap["freq"], np.float32),
        _c(snap["vis0_re"], np.float32), _c(snap["vis0_im"], np.float32),
        _c(snap["vis3_re"], np.float32), _c(snap["vis3_im"], np.float32),
        _c(snap["flag0"], np.uint8), _c(snap["flag3"], np.uint8),
        _c(snap["weight0"], np.float32), _c(snap["weight3"], np.float32),
        float(snap["array_center_x"]), bin_, vin, vre, vim, ev)
    return {"bin": bin_.reshape(S, 3), "vin": vin.reshape(3, 3), "vis_re": vre, "vis_im": vim,
            "evals": ev}


def preprocess_f64(snap: Dict) -> Dict:
    R, C = int(snap["num_rows"]), int(snap["num_channels"])
    S = R * C
    bin_ = np.zeros(S * 3, np.float64)
    vin = np.zeros(9, np.float64)
    ev = np.zeros(3, np.float64)
    lib().orc_preprocess_f64(R, C, _c(snap["uvw"], np.float64).reshape(-1),
                             _c(snap["freq"], np.float64), float(snap["array_center_x"]),
                             bin_, vin, ev)
    return {"bin": bin_.reshape(S, 3), "vin": vin.reshape(3, 3), "evals": ev}


def collapse_f64(snap: Dict):
    """Flag / WEIGHT_SPECTRUM pol collapse (SVDIntegratedImager.cu:588-617) in FP64."""
    w0 = np.where(snap["flag0"] != 0, 0.0, snap["weight0"].astype(np.float64))
    w3 = np.where(snap["flag3"] != 0, 0.0, snap["weight3"].astype(np.float64))
    ws = w0 + w3
    safe = np.where(ws > 0, ws, 1.0)
    re = np.where(ws > 0, (snap["vis0_re"] * w0 + snap["vis3_re"] * w3) / safe, 0.0)
    im = np.where(ws > 0, (snap["vis0_im"] * w0 + snap["vis3_im"] * w3) / safe, 0.0)
    return re, im


# ------------------------------------------------------------------ planning
def cheb_stats(bin_, vin, N: int, cell: float) -> np.ndarray:
    out = np.zeros(4, np.float32)
    b = _c(bin_, np.float32).reshape(-1)
    lib().orc_cheb_stats(b, b.size // 3, _c(vin, np.float32).reshape(9), N, cell, out)
    return out


def slab_config(r3min, r3max, max_abs_t3, target_alpha: Optional[float] = None) -> Dict:
    out = np.zeros(8, np.float32)
    lib().orc_slab_config(r3min, r3max, max_abs_t3,
                          0.0 if target_alpha is None else target_alpha, out)
    keys = ["alpha_max", "target_alpha", "num_slabs", "r3_range", "slab_width", "slab_half_range",
            "inv_slab_half_range", "slab_alpha_max"]
    d = {k: float(v) for k, v in zip(keys, out)}
    d["num_slabs"] = int(d["num_slabs"])
    return d


def bucket_errors_ref32(alpha: float) -> np.ndarray:
    out = np.zeros(4, np.float32)
    lib().orc_bucket_errors_ref32(alpha, out)
    return out


def select_bucket(errs) -> int:
    return lib().orc_select_bucket(_c(errs, np.float32))


def archive_select_terms(alpha_max: float):
    errs = np.zeros(32, np.float32)
    k = lib().orc_archive_select_terms(alpha_max, errs)
    return k, errs


# ------------------------------------------------------------------ kernels
def ifft2(planes: np.ndarray) -> np.ndarray:
    """planes: (P, G, G) complex64 -> unnormalised inverse 2D DFT, computed in FP64."""
    P, G, _ = planes.shape
    buf = np.ascontiguousarray(planes.astype(np.complex64)).view(np.float32).reshape(-1).copy()
    lib().orc_ifft2(buf, G, P)
    return buf.view(np.complex64).reshape(P, G, G)


def ifft1_f64(x: np.ndarray) -> np.ndarray:
    buf = np.ascontiguousarray(x.astype(np.complex128)).view(np.float64).reshape(-1).copy()
    lib().orc_ifft1_f64(buf, x.size)
    return buf.view(np.complex128)


def reproject(N: int, cell: float, vin, variant: int = 0) -> np.ndarray:
    out = np.zeros(N * N * 2, np.float32)
    lib().orc_reproject(N, cell, _c(vin, np.float32).reshape(9), variant, out)
    return out.reshape(N, N, 2)


def splat(N: int, mapping: np.ndarray, values: np.ndarray) -> np.ndarray:
    dirty = np.zeros(N * N, np.float32)
    lib().orc_splat(N, _c(mapping, np.float32).reshape(-1), _c(values, np.float32).reshape(-1),
                    dirty)
    return dirty.reshape(N, N)


def grid_svdi(bin_, vre, vim, vin, N, cell, r3_min, slab_width, inv_half, K, slab_start,
              nslabs, total_slabs) -> np.ndarray:
    G = grid_size(N)
    b = _c(bin_, np.float32).reshape(-1)
    S = b.size // 3
    grid = np.zeros(nslabs * K * G * G * 2, np.float32)
    scale = np.float32(cell) * np.float32(G)
    lib().orc_grid_svdi(b, _c(vre, np.float32), _c(vim, np.float32), S,
                        _c(vin, np.float32).reshape(9), float(scale), G, r3_min, slab_width,
                        inv_half, K, slab_start, nslabs, total_slabs, grid)
    return grid.view(np.complex64).reshape(nslabs * K, G, G)


def grid_archive(bin_, vre, vim, vin, N, cell, r3_centre, inv_half, K) -> np.ndarray:
    """Archive gridding (FIPk.cu:383-452): one slab centred at r3_centre, vis / |det| first.
    Returns (K, G, G) complex64 planes."""
    G = grid_size(N)
    b = _c(bin_, np.float32).reshape(-1)
    S = b.size // 3
    vf = _c(vin, np.float32).reshape(-1)
    inten = np.float32(abs(vf[0] * vf[4] - vf[1] * vf[3]))
    grid = np.zeros(K * G * G * 2, np.float32)
    lib().orc_grid_archive(b, _c(vre, np.float32), _c(vim, np.float32), S, float(inten),
                           float(np.float32(cell) * np.float32(G)), G, float(r3_centre),
                           float(inv_half), int(K), grid)
    return grid.view(np.complex64).reshape(K, G, G)


def dft64(bin64, vre, vim, vin64, N: int, cell: float, pixels=None, keep=None) -> np.ndarray:
    """FP64 direct DFT on the rotated frame (signed image, / (|det| S)).

    ``pixels``: (n, 2) int array of (x, y); None for the whole N x N image (returned (N, N)).
    ``keep``: optional uint8 mask of samples to include (S still counts every sample).
    """
    b = _c(bin64, np.float64).reshape(-1)
    S = b.size // 3
    full = pixels is None
    if full:
        yy, xx = np.mgrid[0:N, 0:N]
        pixels = np.stack([xx.reshape(-1), yy.reshape(-1)], axis=1)
    px = _c(pixels, np.int32).reshape(-1)
    out = np.zeros(px.size // 2, np.float64)
    keep_arr = None if keep is None else _c(keep, np.uint8)
    lib().orc_dft64(b, _c(vre, np.float64), _c(vim, np.float64),
                    None if keep_arr is None else keep_arr.ctypes.data, S,
                    _c(vin64, np.float64).reshape(9), N, float(cell), px, px.size // 2, out)
    return out.reshape(N, N) if full else out


# ------------------------------------------------------------------ pipelines
def fipipe_device_ref32(bin_, vin, vre, vim, N: int, cell: float,
                        target_alpha: Optional[float] = None,
                        max_batch_slabs: Optional[int] = None,
                        num_terms: Optional[int] = None) -> Dict:
    """FIpipeDevice (SVDIntegratedImager.cu:1498-1761) restated on the CPU.

    ``max_batch_slabs``: None = every slab in one batch (the HBM-budget rule of
    choose_slab_batch_size always admits them at the sizes tested), else the
    SVDIMAGER_MAX_BATCH_SLABS value.  ``num_terms`` overrides the bucket choice.
    Returns dirty (N, N), signed (N, N; 0 where root_arg < 0) and the plan.
    """
    b = _c(bin_, np.float32).reshape(-1, 3)
    vin = _c(vin, np.float32).reshape(3, 3)
    S = b.shape[0]
    G = grid_size(N)
    stats = cheb_stats(b, vin, N, cell)
    sc = slab_config(stats[0], stats[1], stats[2], target_alpha)
    errs = bucket_errors_ref32(sc["slab_alpha_max"])
    K = select_bucket(errs) if num_terms is None else int(num_terms)
    ns = sc["num_slabs"]
    bs = ns if max_batch_slabs is None else min(ns, int(max_batch_slabs))
    conv = conv_corr(N, G, 0)
    dirty = np.zeros(N * N, np.float32)
    signed = np.zeros(N * N, np.float32)
    invS = np.float32(1.0) / np.float32(S)
    for start in range(0, ns, bs):
        nb = min(bs, ns - start)
        grid = grid_svdi(b, vre, vim, vin, N, cell, float(stats[0]), sc["slab_width"],
                         sc["inv_slab_half_range"], K, start, nb, ns)
        planes = ifft2(grid)
        sig = np.zeros(N * N, np.float32)
        lib().orc_image_svdi(_c(planes.view(np.float32).reshape(-1), np.float32), G, N,
                             float(cell), _c(vin, np.float32).reshape(9), conv,
                             float(np.float32(CONV_CORR_NORM)), float(invS), float(stats[0]),
                             sc["slab_width"], sc["slab_half_range"], K, start, nb, dirty,
                             sig.ctypes.data)
        signed += sig
    plan = dict(sc)
    plan.update(r3_min=float(stats[0]), r3_max=float(stats[1]), max_abs_t3=float(stats[2]),
                valid_pixels=int(stats[3]), bucket_errors=errs, num_terms=K, batch_slabs=bs,
                grid_size=G)
    return {"dirty": dirty.reshape(N, N), "signed": signed.reshape(N, N), "plan": plan}


def fipipe_archive_ref32(bin_, vin, vre, vim, N: int, cell: float,
                         num_terms: Optional[int] = None) -> Dict:
    """FIpipe (archive/FIPkernels.cu:679-911) restated on the CPU (single slab)."""
    b = _c(bin_, np.float32).reshape(-1, 3)
    vin = _c(vin, np.float32).reshape(3, 3)
    S = b.shape[0]
    G = grid_size(N)
    stats = cheb_stats(b, vin, N, cell)                    # FIPk.cu:320-358 (same values)
    r3min, r3max = np.float32(stats[0]), np.float32(stats[1])
    centre = np.float32(0.5) * (r3max + r3min)
    half = max(np.float32(0.5) * (r3max - r3min), np.float32(1e-6))
    inv_half = np.float32(1.0) / half
    alpha_max = np.float32(2.0 * math.pi * float(stats[2]) * float(half))
    if num_terms is None:
        K, errs = archive_select_terms(float(alpha_max))
    else:
        K, errs = int(num_terms), None
    vf = vin.reshape(-1)
    inten = np.float32(abs(vf[0] * vf[4] - vf[1] * vf[3]))
    uv_scale = np.float32(cell) * np.float32(G)
    grid = np.zeros(K * G * G * 2, np.float32)
    lib().orc_grid_archive(_c(b.reshape(-1), np.float32), _c(vre, np.float32),
                           _c(vim, np.float32), S, float(inten), float(uv_scale), G,
                           float(centre), float(inv_half), K, grid)
    planes = ifft2(grid.view(np.complex64).reshape(K, G, G))
    conv = conv_corr(N, G, 1)
    dirty = np.zeros(N * N, np.float32)
    signed = np.zeros(N * N, np.float32)
    lib().orc_image_archive(_c(planes.view(np.float32).reshape(-1), np.float32), G, N,
                            float(cell), _c(vin, np.float32).reshape(9), conv,
                            float(np.float32(CONV_CORR_NORM)), S, float(centre), float(half), K,
                            dirty, signed.ctypes.data)
    plan = {"r3_min": float(r3min), "r3_max": float(r3max), "r3_centre": float(centre),
            "r3_half_range": float(half), "max_abs_t3": float(stats[2]),
            "valid_pixels": int(stats[3]), "alpha_max": float(alpha_max), "num_terms": K,
            "term_errors": errs, "grid_size": G}
    return {"dirty": dirty.reshape(N, N), "signed": signed.reshape(N, N), "plan": plan}


def ms_read_snapshot(column_dir: str, snapshot_index: int) -> Dict:
    """One TIME snapshot of a MeasurementSet whose main-table columns were saved as .npy files
    (TIME, UVW, DATA, FLAG, optional WEIGHT_SPECTRUM; SPECTRAL_WINDOW/CHAN_FREQ; optional
    OBSERVATION/ARRAY_CENTER; the arrays as python-casacore's getcol returns them), restating
    group_rows_by_time, read_measurement_set_snapshot and read_measurement_set_rows
    (SVDIntegratedImager.cu:180-378) with detect_row_axis_layout (:99-134).  uvw and the
    frequencies are kept in FP64 (the reference narrows them to float).  Raises ValueError with
    the reference's message where it has one."""
    p = lambda *a: os.path.join(column_dir, *a)
    time = np.load(p("TIME.npy"))
    times = sorted(set(time.tolist()))                      # std::map<double, rows>, ascending
    if snapshot_index >= len(times):
        raise ValueError("Snapshot_Index is outside available TIME snapshots")
    t = times[snapshot_index]
    rows = [r for r in range(time.shape[0]) if time[r] == t]
    uvw = np.load(p("UVW.npy"))
    data = np.load(p("DATA.npy"))
    flag = np.load(p("FLAG.npy"))
    weight = np.load(p("WEIGHT_SPECTRUM.npy")) if os.path.exists(p("WEIGHT_SPECTRUM.npy")) else None
    freq = [float(f) for spw in np.atleast_2d(np.load(p("SPECTRAL_WINDOW", "CHAN_FREQ.npy"))) for f in spw]
    if not freq:
        raise ValueError("No frequencies found in SPECTRAL_WINDOW")
    centre = 0.0
    if os.path.exists(p("OBSERVATION", "ARRAY_CENTER.npy")):
        ac = np.atleast_2d(np.load(p("OBSERVATION", "ARRAY_CENTER.npy")))
        if ac.shape[0] > 0 and ac.shape[1] > 0:
            centre = float(ac[0][0])
    # Row shape in casacore's IPosition order: the reverse of the numpy row axes.
    d0, d1 = data.shape[2], data.shape[1]
    if d1 in (4, 2, 1):
        chan_axis, nchan, npol = 0, d0, d1
    elif d0 in (4, 2, 1):
        chan_axis, nchan, npol = 1, d1, d0
    elif d0 >= d1:
        chan_axis, nchan, npol = 0, d0, d1
    else:
        chan_axis, nchan, npol = 1, d1, d0
    if npol < 4:
        raise ValueError("DATA has fewer than 4 polarisations")

    def at(arr, row, chan, pol):
        ip = [0, 0]
        ip[chan_axis] = chan
        ip[1 - chan_axis] = pol
        return arr[row, ip[1], ip[0]]                      # numpy axes: (row, IPosition[1], IPosition[0])

    R, C = len(rows), len(freq)
    out = {k: np.zeros(R * C, np.float32) for k in ("vis0_re", "vis0_im", "vis3_re", "vis3_im")}
    out["flag0"] = np.zeros(R * C, np.uint8)
    out["flag3"] = np.zeros(R * C, np.uint8)
    out["weight0"] = None if weight is None else np.zeros(R * C, np.float32)
    out["weight3"] = None if weight is None else np.zeros(R * C, np.float32)
    out["uvw"] = np.zeros((R, 3), np.float64)
    for lr, row in enumerate(rows):
        out["uvw"][lr] = uvw[row]
        for c in range(C):
            i, cr = lr + c * R, c % nchan
            v0, v3 = np.complex64(at(data, row, cr, 0)), np.complex64(at(data, row, cr, 3))
            out["vis0_re"][i], out["vis0_im"][i] = v0.real, v0.imag
            out["vis3_re"][i], out["vis3_im"][i] = v3.real, v3.imag
            out["flag0"][i] = 1 if at(flag, row, cr, 0) else 0
            out["flag3"][i] = 1 if at(flag, row, cr, 3) else 0
            if weight is not None:
                out["weight0"][i] = at(weight, row, cr, 0)
                out["weight3"][i] = at(weight, row, cr, 3)
    out.update(num_rows=R, num_channels=C, freq_hz=np.array(freq, np.float64), array_center_x=centre,
               time=t, index=snapshot_index, total_snapshots=len(times))
    return out
