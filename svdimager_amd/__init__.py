"""svdimager_amd — MI355X snapshot imager behind the egbdfX/SVDimager host API.

The compute path is libsvdimager.so (hand-written HIP kernels for gfx950, C-ABI in
include/svdimager.h); this package builds it, binds it with ctypes, and generates synthetic
snapshots.  Importing the package does not load the library; the first imaging call does, and
raises if the library has not been built.
"""
from ._lib import (GRIDDER_ATOMIC, GRIDDER_AUTO, GRIDDER_MFMA, GRIDDER_TILED, OUTPUT_DIRTY, OUTPUT_SIGNED,
                   PRECISION_BF16, PRECISION_DEFAULT, PRECISION_SINGLE, PRECISION_SPLIT3, PRECISION_SPLITC,
                   SEMANTICS_ARCHIVE, SEMANTICS_INTEGRATED, STAGES, Context, SvdiError)
from .imager import (DeviceSnapshot, FIpipe, FIpipeDevice, autotune, fft2_inverse, fipipe_host, image_host_snapshot,
                     image_snapshot, image_snapshot_batch,
                     preprocess_measurement_set_gpu, read_fits_image, read_measurement_set_snapshot,
                     snapshot_geometry, write_fits_2d)

__all__ = [
    "Context", "SvdiError", "DeviceSnapshot", "FIpipe", "FIpipeDevice", "autotune", "fft2_inverse", "fipipe_host",
    "image_host_snapshot", "image_snapshot", "image_snapshot_batch", "preprocess_measurement_set_gpu",
    "read_fits_image", "read_measurement_set_snapshot", "snapshot_geometry", "write_fits_2d",
    "GRIDDER_ATOMIC", "GRIDDER_AUTO", "GRIDDER_MFMA", "GRIDDER_TILED", "OUTPUT_DIRTY", "OUTPUT_SIGNED",
    "PRECISION_BF16", "PRECISION_DEFAULT", "PRECISION_SINGLE", "PRECISION_SPLIT3", "PRECISION_SPLITC",
    "SEMANTICS_ARCHIVE", "SEMANTICS_INTEGRATED", "STAGES",
]
