import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


def load_small_snapshot():
    z = np.load(os.path.join(GOLDEN, "small_snapshot.npz"))
    snap = {k: z[k] for k in z.files}
    for k in ("num_rows", "num_channels", "image_size", "num_terms"):
        snap[k] = int(snap[k])
    for k in ("array_center_x", "cell"):
        snap[k] = float(snap[k])
    return snap


@pytest.fixture(scope="session")
def small_snap():
    return load_small_snapshot()


@pytest.fixture(scope="session")
def small_dft():
    z = np.load(os.path.join(GOLDEN, "small_dft64.npz"))
    return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    from svdimager_amd import Context
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    ctx = Context(0)
    yield ctx
    ctx.close()


def parity_errors(img, ref):
    """(max over the image, max over the central half), relative to the peak |ref|.

    The grid correction divides by c(t)^2, which falls to ~0.01 at the image edge, so FP32
    rounding that differs between two summation orders is amplified ~85x at the edge and ~3x
    inside the central half."""
    a = np.asarray(img, np.float64)
    b = np.asarray(ref, np.float64)
    peak = max(float(np.abs(b).max()), 1e-30)
    d = np.abs(a - b)
    n0, n1 = b.shape
    q0, q1 = n0 // 4, n1 // 4
    return float(d.max()) / peak, float(d[q0:n0 - q0, q1:n1 - q1].max()) / peak


# Dirty images (|.| then bilinear splat): the splat targets come from FP32 hypot / division
# on the device vs glibc on the host, a few ulp apart, which moves each splat weight by ~1e-5.
DIRTY_CENTRAL = 5e-5

# Archive dirty image (after |.| and the degree-trig SIN splat) against the oracle.  The
# archive's p2p takes cos of an FP32 angle in degrees near 90 degrees: one ulp of that angle
# (7.6e-6 degree) moves a splat target by about 1e-3 pixel at this cell size, host and device
# libm may round it differently, and a bilinear splat moved by d pixels changes a pixel by up to
# about 4 d of the values splatted onto it.  Measured on MI355X: 3.0e-3 over the image, 2.7e-3
# over the central half, i.e. uniform (the signed image before the splat agrees to 1e-5).
ARCHIVE_DIRTY = 5e-3


def assert_parity(img, ref, what="", whole=1e-4, central=1e-5):
    e, c = parity_errors(img, ref)
    print(f"{what}: max {e:.2e} central {c:.2e}", flush=True)
    assert e <= whole and c <= central, (what, e, c)


def announce(msg: str) -> None:
    """Name the case on stdout before a GPU step, flushed, so a fault names its case."""
    print(f"[case] {msg}", flush=True)
