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


def announce(msg: str) -> None:
    """Name the case on stdout before a GPU step, flushed, so a fault names its case."""
    print(f"[case] {msg}", flush=True)
