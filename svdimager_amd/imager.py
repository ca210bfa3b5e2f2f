"""Python front end of libsvdimager, mirroring the reference's host API.

Reference entry points and their counterparts here:

* ``preprocess_measurement_set_gpu`` (SVDIntegratedImager.cu:635-853) -> ``preprocess_measurement_set_gpu``
* ``FIpipeDevice`` (SVDIntegratedImager.cu:1498-1761)               -> ``FIpipeDevice``
* preprocess + FIpipeDevice as ``main`` chains them (:1822-1828)   -> ``image_snapshot``
* ``FIpipe`` (archive/FIPkernels.cu:679-911, pipe_interface.cpp:8)  -> ``FIpipe`` (the exported C++ symbol)
* ``write_fits_2d`` (SVDIntegratedImager.cu:153-176)                -> ``write_fits_2d``

Device buffers are torch tensors on a ROCm device; torch is used for memory and streams only.
Every call goes to the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib
from ._lib import Context, PlanInfo, Snapshot, check, default_options, lib


def _torch():
    import torch
    return torch


def _ptr(t) -> int:
    return t.data_ptr()


class DeviceSnapshot:
    """A snapshot (the HostMeasurementSetData fields) resident in device memory."""

    FIELDS = ("uvw", "freq", "vis0_re", "vis0_im", "vis3_re", "vis3_im", "flag0", "flag3",
              "weight0", "weight3")

    def __init__(self, snap: Dict, device: str = "cuda"):
        torch = _torch()
        dt = {"uvw": torch.float64, "freq": torch.float64, "flag0": torch.uint8,
              "flag3": torch.uint8}
        self.num_rows = int(snap["num_rows"])
        self.num_channels = int(snap["num_channels"])
        self.array_center_x = float(snap["array_center_x"])
        self.tensors = {}
        for k in self.FIELDS:
            if snap.get(k) is None:
                self.tensors[k] = None
                continue
            v = snap[k]
            if isinstance(v, np.ndarray):
                v = torch.from_numpy(np.ascontiguousarray(v))
            self.tensors[k] = v.to(device=device, dtype=dt.get(k, torch.float32)).contiguous()

    @property
    def num_samples(self) -> int:
        return self.num_rows * self.num_channels

    @property
    def nbytes(self) -> int:
        """Device bytes of the snapshot's arrays."""
        return sum(t.numel() * t.element_size() for t in self.tensors.values() if t is not None)

    def struct(self) -> Snapshot:
        t = self.tensors
        p = lambda k: (None if t[k] is None else _ptr(t[k]))
        return Snapshot(self.num_rows, self.num_channels, p("uvw"), p("freq"), p("vis0_re"),
                        p("vis0_im"), p("vis3_re"), p("vis3_im"), p("flag0"), p("flag3"),
                        p("weight0"), p("weight3"), self.array_center_x)


def _bind_stream(ctx: Context, device) -> None:
    torch = _torch()
    ctx.set_stream(torch.cuda.current_stream(device).cuda_stream)


def preprocess_measurement_set_gpu(ctx: Context, dsnap: DeviceSnapshot) -> Dict:
    """Bin (S, 3), Vin (3, 3), vis_re (S,), vis_im (S,) as device tensors."""
    torch = _torch()
    dev = dsnap.tensors["uvw"].device
    S = dsnap.num_samples
    out = {"bin": torch.empty((S, 3), dtype=torch.float32, device=dev),
           "vin": torch.empty((3, 3), dtype=torch.float32, device=dev),
           "vis_re": torch.empty(S, dtype=torch.float32, device=dev),
           "vis_im": torch.empty(S, dtype=torch.float32, device=dev)}
    _bind_stream(ctx, dev)
    s = dsnap.struct()
    check(lib().svdi_preprocess(ctx.handle, ctypes.byref(s), _ptr(out["bin"]), _ptr(out["vin"]),
                                _ptr(out["vis_re"]), _ptr(out["vis_im"])))
    return out


def snapshot_geometry(ctx: Context, dsnap: DeviceSnapshot) -> Tuple[np.ndarray, np.ndarray]:
    """The FP64 rotation (3, 3; rows = oriented principal axes) and ascending eigenvalues that
    the snapshot imager uses."""
    rot = np.zeros(9, np.float64)
    evals = np.zeros(3, np.float64)
    _bind_stream(ctx, dsnap.tensors["uvw"].device)
    s = dsnap.struct()
    check(lib().svdi_snapshot_geometry(ctx.handle, ctypes.byref(s), rot.ctypes.data, evals.ctypes.data))
    return rot.reshape(3, 3), evals


def FIpipeDevice(ctx: Context, bin_, vin, vis_re, vis_im, image_size: int, cell_size: float,
                 image=None, **options) -> Tuple[object, Dict]:
    """Image preprocessed device buffers; returns (image tensor (N, N), plan dict)."""
    torch = _torch()
    dev = bin_.device
    N = int(image_size)
    if image is None:
        image = torch.empty((N, N), dtype=torch.float32, device=dev)
    opts = default_options(**options)
    info = PlanInfo()
    _bind_stream(ctx, dev)
    check(lib().svdi_image_preprocessed(ctx.handle, _ptr(bin_), _ptr(vin), _ptr(vis_re),
                                        _ptr(vis_im), bin_.shape[0], _ptr(image), N,
                                        float(cell_size), ctypes.byref(opts),
                                        ctypes.byref(info)))
    return image, info.as_dict()


def image_snapshot(ctx: Context, dsnap: DeviceSnapshot, image_size: int, cell_size: float,
                   image=None, noalloc: bool = False, plan: Optional[Dict] = None,
                   **options) -> Tuple[object, Dict]:
    """Preprocess + image one device-resident snapshot; returns (image, plan dict).
    noalloc=True calls svdi_image_snapshot_noalloc (no device allocation inside the call; the
    image tensor, when not given, is still allocated here by torch).  plan=<a plan dict an
    earlier call on the same geometry returned> calls svdi_image_snapshot_planned: no planning,
    no host wait, capturable in a CUDA/HIP graph; it returns (image, plan) with the image
    complete when torch's current stream reaches this point."""
    torch = _torch()
    dev = dsnap.tensors["uvw"].device
    N = int(image_size)
    if image is None:
        image = torch.empty((N, N), dtype=torch.float32, device=dev)
    opts = default_options(**options)
    info = PlanInfo()
    _bind_stream(ctx, dev)
    s = dsnap.struct()
    if plan is not None:
        pin = PlanInfo.from_dict(plan)
        check(lib().svdi_image_snapshot_planned(ctx.handle, ctypes.byref(s), _ptr(image), N,
                                                float(cell_size), ctypes.byref(opts), ctypes.byref(pin)))
        return image, plan
    fn = lib().svdi_image_snapshot_noalloc if noalloc else lib().svdi_image_snapshot
    check(fn(ctx.handle, ctypes.byref(s), _ptr(image), N, float(cell_size), ctypes.byref(opts),
             ctypes.byref(info)))
    return image, info.as_dict()


def autotune(ctx: Context, dsnap: DeviceSnapshot, image_size: int, cell_size: float, repeats: int = 2,
             **options) -> Dict:
    """svdi_autotune: time the MFMA gridder's candidate work shapes on this snapshot and size,
    keep the fastest on the context for this grid size, and return it (with its grid-stage time
    and the defaults' time, ms)."""
    from ._lib import Tuning
    dev = dsnap.tensors["uvw"].device
    opts = default_options(**options)
    best = Tuning()
    _bind_stream(ctx, dev)
    s = dsnap.struct()
    check(lib().svdi_autotune(ctx.handle, ctypes.byref(s), int(image_size), float(cell_size), ctypes.byref(opts),
                              int(repeats), ctypes.byref(best)))
    return best.as_dict()


def image_snapshot_batch(ctx: Context, dsnaps: List[DeviceSnapshot], image_size, cell_size,
                         images=None, wait: bool = True, plans: Optional[List[Dict]] = None,
                         **options) -> Tuple[List, List[Dict]]:
    """Image several device-resident snapshots in one call (concurrent lanes); returns (images,
    plan dicts).  image_size / cell_size: one value for all, or one per snapshot (mixed sizes:
    svdi_image_snapshots).  `images`: optional list of (N_i, N_i) float32 device tensors.
    wait=False returns once the work is enqueued: the images are complete when the current
    stream reaches this point, and the snapshots must stay alive until then.  plans=<the plan
    dicts an earlier call returned for the same snapshots and sizes> calls
    svdi_image_snapshots_planned: no planning and no host wait, so the batch can be captured in
    one CUDA/HIP graph; it returns (images, plans) at once."""
    torch = _torch()
    n = len(dsnaps)
    if n == 0:
        return [], []
    dev = dsnaps[0].tensors["uvw"].device
    mixed = isinstance(image_size, (list, tuple)) or isinstance(cell_size, (list, tuple))
    sizes = [int(x) for x in image_size] if isinstance(image_size, (list, tuple)) else [int(image_size)] * n
    cells = [float(x) for x in cell_size] if isinstance(cell_size, (list, tuple)) else [float(cell_size)] * n
    if len(sizes) != n or len(cells) != n:
        raise ValueError("one image size and cell per snapshot")
    if images is None:
        images = [torch.empty((N, N), dtype=torch.float32, device=dev) for N in sizes]
    if len(images) != n:
        raise ValueError("one image per snapshot")
    opts = default_options(**options)
    infos = (PlanInfo * n)()
    structs = (Snapshot * n)(*[d.struct() for d in dsnaps])
    ptrs = (ctypes.c_void_p * n)(*[_ptr(im) for im in images])
    _bind_stream(ctx, dev)
    if plans is not None:
        if len(plans) != n:
            raise ValueError("one plan per snapshot")
        pins = (PlanInfo * n)(*[PlanInfo.from_dict(p) for p in plans])
        csz = (ctypes.c_size_t * n)(*sizes)
        ccell = (ctypes.c_float * n)(*cells)
        check(lib().svdi_image_snapshots_planned(ctx.handle, structs, n, ptrs, csz, ccell, ctypes.byref(opts), pins))
        return images, list(plans)
    if mixed:
        csz = (ctypes.c_size_t * n)(*sizes)
        ccell = (ctypes.c_float * n)(*cells)
        check(lib().svdi_image_snapshots(ctx.handle, structs, n, ptrs, csz, ccell, ctypes.byref(opts),
                                         infos, 1 if wait else 0))
    else:
        fn = lib().svdi_image_snapshot_batch if wait else lib().svdi_image_snapshot_batch_async
        check(fn(ctx.handle, structs, n, ptrs, sizes[0], cells[0], ctypes.byref(opts), infos))
    return images, [infos[i].as_dict() for i in range(n)]


def fft2_inverse(ctx: Context, planes, row_lo: int = 0, rows: Optional[int] = None,
                 col_lo: int = 0, cols: Optional[int] = None):
    """The library's unnormalised 2D inverse FFT of (P, G, G) complex64 device planes, window
    rows [row_lo, row_lo + rows) x columns [col_lo, col_lo + cols); returns (P, rows, cols)."""
    torch = _torch()
    P, G, G2 = planes.shape
    if G != G2 or planes.dtype != torch.complex64 or not planes.is_contiguous():
        raise ValueError("planes: contiguous (P, G, G) complex64")
    rows = G - row_lo if rows is None else int(rows)
    cols = G - col_lo if cols is None else int(cols)
    out = torch.empty((P, rows, cols), dtype=torch.complex64, device=planes.device)
    _bind_stream(ctx, planes.device)
    check(lib().svdi_fft2_inverse(ctx.handle, _ptr(planes), P, G, int(row_lo), rows, int(col_lo), cols,
                                  _ptr(out)))
    return out


def fipipe_host(ctx: Context, vis_re: np.ndarray, vis_im: np.ndarray, bin_: np.ndarray,
                vin: np.ndarray, image_size: int, cell_size: float, **options):
    """FIpipe semantics on host numpy arrays through a given context; returns (image, plan)."""
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    vr, vi, b, v = f(vis_re), f(vis_im), f(bin_).reshape(-1), f(vin).reshape(-1)
    N = int(image_size)
    out = np.zeros(N * N, np.float32)
    opts = default_options(**({"semantics": _lib.SEMANTICS_ARCHIVE} | options))
    info = PlanInfo()
    check(lib().svdi_fipipe_host(ctx.handle, vr.ctypes.data, vi.ctypes.data, b.ctypes.data,
                                 v.ctypes.data, out.ctypes.data, vr.size, N, float(cell_size),
                                 ctypes.byref(opts), ctypes.byref(info)))
    return out.reshape(N, N), info.as_dict()


def FIpipe(Visreal: np.ndarray, Visimag: np.ndarray, Bin: np.ndarray, Vin: np.ndarray,
           dirty_image: np.ndarray, num_baselines: int, image_size: int, cell_size: float) -> int:
    """Call the exported C++ FIpipe symbol (archive/pipe_interface.cpp:8) on host arrays."""
    arrs = [Visreal, Visimag, Bin, Vin, dirty_image]
    for a in arrs:
        if a.dtype != np.float32 or not a.flags["C_CONTIGUOUS"]:
            raise TypeError("FIpipe takes C-contiguous float32 arrays")
    fn = getattr(lib(), "_Z6FIpipePfS_S_S_S_mmf")
    return fn(*(a.ctypes.data for a in arrs), int(num_baselines), int(image_size),
              float(cell_size))


def write_fits_2d(filename: str, values: np.ndarray) -> None:
    v = np.ascontiguousarray(values, dtype=np.float32)
    n1 = v.shape[-1]
    n2 = v.size // n1
    rc = lib().svdi_fits_write(filename.encode(), v.ctypes.data, n1, n2)
    check(rc)


def read_fits_image(filename: str) -> np.ndarray:
    """Primary-HDU image as float32, shape (NAXIS2, NAXIS1) (1-D images: (NAXIS1,))."""
    data = ctypes.c_void_p()
    naxes = (ctypes.c_long * 2)()
    check(lib().svdi_fits_read(filename.encode(), ctypes.byref(data), naxes))
    n1, n2 = naxes[0], naxes[1]
    try:
        arr = np.ctypeslib.as_array(ctypes.cast(data, ctypes.POINTER(ctypes.c_float)),
                                    shape=(n1 * n2,)).copy()
    finally:
        lib().svdi_free(data)
    return arr if n2 == 1 else arr.reshape(n2, n1)


def read_measurement_set_snapshot(column_dir: str, snapshot_index: int) -> Dict:
    """read_measurement_set_snapshot (SVDIntegratedImager.cu:359-378) on a MeasurementSet whose
    main-table columns were saved as .npy files (svdi_ms_read_snapshot): TIME.npy, UVW.npy,
    DATA.npy, FLAG.npy, optional WEIGHT_SPECTRUM.npy, SPECTRAL_WINDOW/CHAN_FREQ.npy and optional
    OBSERVATION/ARRAY_CENTER.npy, each as python-casacore's getcol returns it.  Returns the
    snapshot dict DeviceSnapshot takes (uvw, freq, vis0_re ... weight3, num_rows, num_channels,
    array_center_x) plus index, time and total_snapshots."""
    ms = _lib.MsSnapshot()
    check(lib().svdi_ms_read_snapshot(column_dir.encode(), int(snapshot_index), ctypes.byref(ms)))
    try:
        d = ms.data
        R, C = int(d.num_rows), int(d.num_channels)
        S = R * C

        def take(ptr, ctype, n, dtype):
            if not ptr:
                return None
            a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(n,))
            return a.astype(dtype, copy=True)
        out = {"uvw": take(d.uvw, ctypes.c_double, 3 * R, np.float64).reshape(R, 3),
               "freq": take(d.freq_hz, ctypes.c_double, C, np.float64)}
        for k in ("vis0_re", "vis0_im", "vis3_re", "vis3_im", "weight0", "weight3"):
            out[k] = take(getattr(d, k), ctypes.c_float, S, np.float32)
        for k in ("flag0", "flag3"):
            out[k] = take(getattr(d, k), ctypes.c_uint8, S, np.uint8)
        out.update(num_rows=R, num_channels=C, array_center_x=float(d.array_center_x), index=int(ms.index),
                   time=float(ms.time), total_snapshots=int(ms.total_snapshots))
        return out
    finally:
        lib().svdi_ms_free_snapshot(ctypes.byref(ms))


def image_host_snapshot(ctx: Context, snap: Dict, image_size: int, cell_size: float, **options):
    """svdi_image_host_snapshot: a snapshot dict of host numpy arrays (read_measurement_set_snapshot,
    synth.make_snapshot) imaged through the context; returns (image, plan)."""
    keep = []

    def ptr(k, dtype):
        if snap.get(k) is None:
            return None
        a = np.ascontiguousarray(snap[k], dtype=dtype)
        keep.append(a)
        return a.ctypes.data
    s = Snapshot()
    s.num_rows, s.num_channels = int(snap["num_rows"]), int(snap["num_channels"])
    s.uvw, s.freq_hz = ptr("uvw", np.float64), ptr("freq", np.float64)
    for k in ("vis0_re", "vis0_im", "vis3_re", "vis3_im", "weight0", "weight3"):
        setattr(s, k, ptr(k, np.float32))
    s.flag0, s.flag3 = ptr("flag0", np.uint8), ptr("flag3", np.uint8)
    s.array_center_x = float(snap["array_center_x"])
    N = int(image_size)
    out = np.zeros(N * N, np.float32)
    opts = default_options(**options)
    info = PlanInfo()
    check(lib().svdi_image_host_snapshot(ctx.handle, ctypes.byref(s), out.ctypes.data, N, float(cell_size),
                                         ctypes.byref(opts), ctypes.byref(info)))
    return out.reshape(N, N), info.as_dict()
