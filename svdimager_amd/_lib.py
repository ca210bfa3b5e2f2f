"""ctypes binding of libsvdimager (the C-ABI declared in include/svdimager.h).

There is no CPU fallback: importing a function that needs the library raises when
libsvdimager.so is missing, and every call that fails raises SvdiError with the library's
message.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsvdimager.so")

SVDI_OK = 0
GRIDDER_AUTO, GRIDDER_ATOMIC, GRIDDER_TILED, GRIDDER_MFMA = 0, 1, 2, 3
PRECISION_DEFAULT, PRECISION_SPLIT3, PRECISION_SPLITC, PRECISION_SINGLE, PRECISION_BF16 = 0, 1, 2, 3, 4
OUTPUT_DIRTY, OUTPUT_SIGNED = 0, 1
SEMANTICS_INTEGRATED, SEMANTICS_ARCHIVE = 0, 1

STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "HIP", 3: "OUT_OF_MEMORY", 4: "UNSUPPORTED",
                5: "IO", 6: "INTERNAL"}

# Every symbol include/svdimager.h declares.
EXPORTED_SYMBOLS = [
    "svdi_version", "svdi_last_error", "svdi_default_options", "svdi_create", "svdi_destroy",
    "svdi_set_stream", "svdi_trim", "svdi_preprocess", "svdi_image_preprocessed",
    "svdi_image_snapshot", "svdi_fipipe_host", "svdi_fits_write", "svdi_fits_read", "svdi_free",
    "svdi_set_profiling", "svdi_last_stage_times", "svdi_snapshot_geometry",
    "svdi_image_snapshot_batch", "svdi_image_snapshot_batch_async", "svdi_workspace_bytes",
    "svdi_image_snapshot_noalloc", "svdi_image_snapshots", "svdi_fft2_inverse",
    "svdi_image_snapshot_planned", "svdi_set_tuning", "svdi_get_tuning", "svdi_autotune",
    "svdi_image_snapshots_planned", "svdi_service_create", "svdi_service_destroy", "svdi_service_submit",
    "svdi_service_collect", "svdi_service_stats", "svdi_ms_read_snapshot", "svdi_ms_free_snapshot",
    "svdi_image_host_snapshot",
    "_Z6FIpipePfS_S_S_S_mmf",
]
STAGES = ("preprocess", "plan", "grid", "fft", "image", "grid_accumulate")


class SvdiError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"svdimager: {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Snapshot(ctypes.Structure):
    _fields_ = [
        ("num_rows", ctypes.c_size_t), ("num_channels", ctypes.c_size_t),
        ("uvw", ctypes.c_void_p), ("freq_hz", ctypes.c_void_p),
        ("vis0_re", ctypes.c_void_p), ("vis0_im", ctypes.c_void_p),
        ("vis3_re", ctypes.c_void_p), ("vis3_im", ctypes.c_void_p),
        ("flag0", ctypes.c_void_p), ("flag3", ctypes.c_void_p),
        ("weight0", ctypes.c_void_p), ("weight3", ctypes.c_void_p),
        ("array_center_x", ctypes.c_double),
    ]


class MsSnapshot(ctypes.Structure):
    _fields_ = [("data", Snapshot), ("index", ctypes.c_size_t), ("total_snapshots", ctypes.c_size_t),
                ("time", ctypes.c_double), ("internal", ctypes.c_void_p)]


class Options(ctypes.Structure):
    _fields_ = [
        ("gridder", ctypes.c_int), ("output", ctypes.c_int), ("semantics", ctypes.c_int),
        ("num_terms", ctypes.c_int), ("target_alpha", ctypes.c_float),
        ("max_batch_slabs", ctypes.c_size_t), ("verbose", ctypes.c_int),
        ("mfma_precision", ctypes.c_int),
    ]


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("r3_min", ctypes.c_float), ("r3_max", ctypes.c_float), ("max_abs_t3", ctypes.c_float),
        ("valid_pixels", ctypes.c_uint32), ("alpha_max", ctypes.c_float),
        ("target_alpha", ctypes.c_float), ("num_slabs", ctypes.c_int),
        ("slab_width", ctypes.c_float), ("slab_half_range", ctypes.c_float),
        ("slab_alpha_max", ctypes.c_float), ("num_terms", ctypes.c_int),
        ("term_error", ctypes.c_float), ("bucket_errors", ctypes.c_float * 4),
        ("batch_slabs", ctypes.c_size_t), ("grid_size", ctypes.c_size_t),
        ("num_planes", ctypes.c_size_t), ("records", ctypes.c_size_t),
        ("elapsed_ms", ctypes.c_float), ("gridder", ctypes.c_int),
        ("r3_centre", ctypes.c_float), ("inv_slab_half_range", ctypes.c_float),
        ("vis_det", ctypes.c_float), ("enqueued_early", ctypes.c_int),
    ]

    @classmethod
    def from_dict(cls, d):
        p = cls()
        for name, _ in cls._fields_:
            if name == "bucket_errors":
                for i, v in enumerate(d[name]):
                    p.bucket_errors[i] = v
            else:
                setattr(p, name, d[name])
        return p

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name == "bucket_errors" else v
        return d


class Tuning(ctypes.Structure):
    """svdi_tuning: the MFMA window gridder's work shape (0 fields: the library's defaults)."""
    _fields_ = [
        ("coop_min", ctypes.c_uint32), ("coop_cap", ctypes.c_uint32), ("class_log2", ctypes.c_int32),
        ("grid_ms", ctypes.c_float), ("default_ms", ctypes.c_float),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServiceConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_size_t), ("max_batch", ctypes.c_uint32),
                ("max_inflight", ctypes.c_uint32), ("options", Options)]


class ServiceResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("batch", ctypes.c_uint32),
                ("call", ctypes.c_uint64), ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]


class ServiceStats(ctypes.Structure):
    _fields_ = [("outstanding", ctypes.c_size_t), ("waiting", ctypes.c_size_t), ("admitted", ctypes.c_size_t),
                ("inflight_calls", ctypes.c_size_t), ("held_bytes", ctypes.c_size_t),
                ("workspace_bytes", ctypes.c_size_t)]


_lib = None


def lib():
    """Load libsvdimager.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run svdimager_amd.build.build() first")
        # torch's wheel carries its own HIP runtime, which its libraries load by file name
        # (libamdhip64.so), while this library needs it by soname (libamdhip64.so.7).  Loading
        # torch first makes this library bind to that same runtime; loaded the other way round,
        # the process would hold two runtimes and torch would find no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, f = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float
        sigs = {
            "svdi_version": (ctypes.c_int, []),
            "svdi_last_error": (ctypes.c_char_p, []),
            "svdi_default_options": (None, [ctypes.POINTER(Options)]),
            "svdi_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int]),
            "svdi_destroy": (ctypes.c_int, [vp]),
            "svdi_set_stream": (ctypes.c_int, [vp, vp]),
            "svdi_trim": (ctypes.c_int, [vp]),
            "svdi_workspace_bytes": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_size_t)]),
            "svdi_preprocess": (ctypes.c_int, [vp, ctypes.POINTER(Snapshot), vp, vp, vp, vp]),
            "svdi_snapshot_geometry": (ctypes.c_int, [vp, ctypes.POINTER(Snapshot), vp, vp]),
            "svdi_image_preprocessed": (ctypes.c_int, [vp, vp, vp, vp, vp, sz, vp, sz, f,
                                                       ctypes.POINTER(Options),
                                                       ctypes.POINTER(PlanInfo)]),
            "svdi_image_snapshot": (ctypes.c_int, [vp, ctypes.POINTER(Snapshot), vp, sz, f,
                                                   ctypes.POINTER(Options),
                                                   ctypes.POINTER(PlanInfo)]),
            "svdi_image_snapshot_noalloc": (ctypes.c_int, [vp, ctypes.POINTER(Snapshot), vp, sz, f,
                                                           ctypes.POINTER(Options),
                                                           ctypes.POINTER(PlanInfo)]),
            "svdi_image_snapshot_batch": (ctypes.c_int, [vp, ctypes.POINTER(Snapshot), sz,
                                                         ctypes.POINTER(ctypes.c_void_p), sz, f,
                                                         ctypes.POINTER(Options),
                                                         ctypes.POINTER(PlanInfo)]),
            "svdi_image_snapshot_batch_async": (ctypes.c_int, [vp, ctypes.POINTER(Snapshot), sz,
                                                               ctypes.POINTER(ctypes.c_void_p), sz, f,
                                                               ctypes.POINTER(Options),
                                                               ctypes.POINTER(PlanInfo)]),
            "svdi_image_snapshots": (ctypes.c_int, [vp, ctypes.POINTER(Snapshot), sz,
                                                    ctypes.POINTER(ctypes.c_void_p),
                                                    ctypes.POINTER(ctypes.c_size_t),
                                                    ctypes.POINTER(ctypes.c_float),
                                                    ctypes.POINTER(Options),
                                                    ctypes.POINTER(PlanInfo), ctypes.c_int]),
            "svdi_fft2_inverse": (ctypes.c_int, [vp, vp, sz, sz, sz, sz, sz, sz, vp]),
            "svdi_image_snapshot_planned": (ctypes.c_int, [vp, ctypes.POINTER(Snapshot), vp, sz, f,
                                                           ctypes.POINTER(Options),
                                                           ctypes.POINTER(PlanInfo)]),
            "svdi_fipipe_host": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, sz, sz, f,
                                                ctypes.POINTER(Options), ctypes.POINTER(PlanInfo)]),
            "svdi_fits_write": (ctypes.c_int, [ctypes.c_char_p, vp, ctypes.c_long, ctypes.c_long]),
            "svdi_fits_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p),
                                              ctypes.POINTER(ctypes.c_long)]),
            "svdi_free": (None, [vp]),
            "svdi_ms_read_snapshot": (ctypes.c_int, [ctypes.c_char_p, sz, ctypes.POINTER(MsSnapshot)]),
            "svdi_ms_free_snapshot": (None, [ctypes.POINTER(MsSnapshot)]),
            "svdi_image_host_snapshot": (ctypes.c_int, [vp, ctypes.POINTER(Snapshot), vp, sz, f,
                                                        ctypes.POINTER(Options), ctypes.POINTER(PlanInfo)]),
            "svdi_set_profiling": (ctypes.c_int, [vp, ctypes.c_int]),
            "svdi_last_stage_times": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_float),
                                                     ctypes.c_int]),
            "svdi_image_snapshots_planned": (ctypes.c_int, [vp, ctypes.POINTER(Snapshot), sz,
                                                            ctypes.POINTER(ctypes.c_void_p),
                                                            ctypes.POINTER(ctypes.c_size_t),
                                                            ctypes.POINTER(ctypes.c_float),
                                                            ctypes.POINTER(Options), ctypes.POINTER(PlanInfo)]),
            "svdi_service_create": (ctypes.c_int, [vp, ctypes.POINTER(ServiceConfig), ctypes.POINTER(vp)]),
            "svdi_service_destroy": (ctypes.c_int, [vp]),
            "svdi_service_submit": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.POINTER(Snapshot), sz, vp, sz, f]),
            "svdi_service_collect": (ctypes.c_int, [vp, ctypes.POINTER(ServiceResult), sz,
                                                    ctypes.POINTER(ctypes.c_size_t), ctypes.c_int]),
            "svdi_service_stats": (ctypes.c_int, [vp, ctypes.POINTER(ServiceStats)]),
            "svdi_set_tuning": (ctypes.c_int, [vp, sz, ctypes.POINTER(Tuning)]),
            "svdi_get_tuning": (ctypes.c_int, [vp, sz, ctypes.POINTER(Tuning)]),
            "svdi_autotune": (ctypes.c_int, [vp, ctypes.POINTER(Snapshot), sz, f, ctypes.POINTER(Options),
                                             ctypes.c_int, ctypes.POINTER(Tuning)]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        fip = getattr(L, "_Z6FIpipePfS_S_S_S_mmf")
        fip.restype = ctypes.c_int
        fip.argtypes = [vp, vp, vp, vp, vp, sz, sz, f]
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != SVDI_OK:
        raise SvdiError(status, lib().svdi_last_error().decode(errors="replace"))


def default_options(**kw) -> Options:
    o = Options()
    lib().svdi_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


class Context:
    """One libsvdimager context (device, stream, workspace)."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        check(lib().svdi_create(ctypes.byref(self._h), device))

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr: Optional[int]) -> None:
        check(lib().svdi_set_stream(self._h, stream_ptr))

    def trim(self) -> None:
        check(lib().svdi_trim(self._h))

    def workspace_bytes(self) -> int:
        """Device bytes the context's workspace holds (batch lanes included)."""
        n = ctypes.c_size_t()
        check(lib().svdi_workspace_bytes(self._h, ctypes.byref(n)))
        return int(n.value)

    def set_profiling(self, enable: bool) -> None:
        check(lib().svdi_set_profiling(self._h, 1 if enable else 0))

    def stage_times(self) -> dict:
        """Device milliseconds per stage of the last call (profiling must be enabled)."""
        arr = (ctypes.c_float * len(STAGES))()
        check(lib().svdi_last_stage_times(self._h, arr, len(STAGES)))
        return {name: float(arr[i]) for i, name in enumerate(STAGES)}

    def set_tuning(self, grid_size: int, coop_min: int = 0, coop_cap: int = 0, class_log2: int = 0) -> None:
        """The MFMA gridder's work shape for grids of grid_size (0: every size without its own)."""
        t = Tuning(coop_min=coop_min, coop_cap=coop_cap, class_log2=class_log2)
        check(lib().svdi_set_tuning(self._h, grid_size, ctypes.byref(t)))

    def clear_tuning(self, grid_size: int) -> None:
        check(lib().svdi_set_tuning(self._h, grid_size, None))

    def get_tuning(self, grid_size: int) -> dict:
        t = Tuning()
        check(lib().svdi_get_tuning(self._h, grid_size, ctypes.byref(t)))
        return t.as_dict()

    def close(self) -> None:
        if self._h:
            check(lib().svdi_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
