"""CPU checks of bench.py's FP64 reference of the gridding algorithm (ref64_grid / ref64_image),
which measures the imagers' arithmetic error on every pixel of the large workloads.  It is a
separate FP64 restatement (torch), pinned here by closed forms and independent results:
* one sample's grid is its ES footprint times T_k(s) v / det with the checkerboard sign;
* bessel_j64 against scipy's jv; conv_corr64 at t = 0 is 1 (the normalisation);
* the whole small-snapshot image against the committed FP64 direct DFT (the method's own error:
  about 1e-4 of peak at the image edge, under 2e-6 inside the central half) and against the
  oracle's FP32 restatement of the reference on the same plan.
"""
import math

import numpy as np
import pytest
import torch
from scipy import special

import bench
from oracle import oracle as O


def test_single_sample_grid_is_the_es_footprint():
    G, cell = 64, 1e-3
    b = torch.tensor([[3.3 / (cell * G), -7.6 / (cell * G), 12.5]], dtype=torch.float64)
    vre, vim = torch.tensor([0.8], dtype=torch.float64), torch.tensor([-0.3], dtype=torch.float64)
    plan = {"num_terms": 3, "num_slabs": 2, "r3_min": 10.0, "slab_width": 2.0, "slab_half_range": 1.0}
    det = 0.9
    g = bench.ref64_grid(b, vre, vim, det, plan, cell, G).numpy()
    # r3 = 12.5 lies in slab 1 (centre 13): s = -0.5; T = 1, -0.5, -0.5.
    T = [1.0, -0.5, 2 * 0.25 - 1]
    p1, p2 = 3.3, -7.6
    es = lambda x: math.exp(bench.ES_BETA * (math.sqrt(1 - x * x) - 1)) if abs(x) <= 1 else 0.0
    expect = np.zeros((6, G, G), np.complex128)
    for g1 in range(math.ceil(p1 - 4), math.floor(p1 + 4) + 1):
        for g2 in range(math.ceil(p2 - 4), math.floor(p2 + 4) + 1):
            w = es((g1 - p1) / 4) * es((g2 - p2) / 4) * (-1) ** ((g1 + g2) & 1)
            for k in range(3):
                expect[3 + k, g1 + G // 2, g2 + G // 2] = w * T[k] * complex(0.8, -0.3) / det
    assert np.abs(g - expect).max() <= 1e-15
    assert np.abs(g[:3]).max() == 0.0


def test_bessel_and_grid_correction():
    a = torch.linspace(-11.5, 11.5, 97, dtype=torch.float64)
    J = bench.bessel_j64(a, 12).numpy()
    ref = np.stack([special.jv(k, a.numpy()) for k in range(12)])
    assert np.abs(J - ref).max() <= 1e-12
    c = bench.conv_corr64(128, 192, "cpu").numpy()
    assert abs(c[0] - 1.0) <= 1e-8
    assert np.all(np.diff(c) < 0) and c[-1] > 0


@pytest.fixture(scope="module")
def small_ref64(small_snap, small_dft):
    snap = small_snap
    R, C, N, cell = snap["num_rows"], snap["num_channels"], snap["image_size"], snap["cell"]
    rot = small_dft["vin64"]
    uvw = torch.from_numpy(snap["uvw"])
    freq = torch.from_numpy(snap["freq"])
    b = ((uvw @ torch.from_numpy(rot).T)[None, :, :] * (freq / bench.C_LIGHT)[:, None, None]).reshape(R * C, 3)
    vre, vim = O.collapse_f64(snap)
    pre = O.preprocess_ref32(snap)
    ref32 = O.fipipe_device_ref32(pre["bin"], pre["vin"], pre["vis_re"], pre["vis_im"], N, cell,
                                  num_terms=snap["num_terms"])
    plan = ref32["plan"]
    det = abs(rot[0, 0] * rot[1, 1] - rot[0, 1] * rot[1, 0])
    G = (3 * N + 1) // 2
    planes = bench.ref64_grid(b, torch.from_numpy(vre), torch.from_numpy(vim), det, plan,
                              np.float32(cell), G)
    img = bench.ref64_image(planes, rot, plan, N, np.float32(cell), R * C).numpy()
    return img, ref32


def test_small_snapshot_against_the_dft_and_the_oracle(small_snap, small_dft, small_ref64):
    img, ref32 = small_ref64
    N = small_snap["image_size"]
    dft = small_dft["dft"]
    peak = np.abs(dft).max()
    q = N // 4
    err = np.abs(img - dft).max() / peak
    err_c = np.abs(img[q:N - q, q:N - q] - dft[q:N - q, q:N - q]).max() / peak
    print(f"ref64 vs DFT: max {err:.3e} central {err_c:.3e}")
    assert err <= 1.1e-4 and err_c <= 2e-5
    e32 = np.abs(img - ref32["signed"]).max() / peak
    e32_c = np.abs(img[q:N - q, q:N - q] - ref32["signed"][q:N - q, q:N - q]).max() / peak
    print(f"ref64 vs oracle ref32: max {e32:.3e} central {e32_c:.3e}")
    # The oracle's FP32 preprocess (FP32 Bin and Vin) puts its own image 1.2e-5 of peak off the
    # DFT in the central half, where this FP64 restatement is within 1e-6 of it.
    assert e32 <= 1e-4 and e32_c <= 2e-5
    assert err_c <= 2e-6
