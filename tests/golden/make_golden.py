"""Write the golden fixtures for the small (64-station, 16-channel, 128^2) snapshot.

The reference ships no tests, fixtures or golden vectors, so these are produced here: the
seeded synthetic snapshot itself, the FP64 rotation from the oracle's FP64 preprocess, and the
FP64 direct-sum DFT image of the collapsed visibilities in the rotated frame (the signed,
pre-reprojection image every imager output is compared against).

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O           # noqa: E402
from svdimager_amd import synth          # noqa: E402


def main() -> None:
    cfg = synth.SMALL
    snap = synth.make_snapshot(cfg)
    p64 = O.preprocess_f64(snap)
    vre, vim = O.collapse_f64(snap)
    N, cell = cfg.image_size, cfg.cell
    dft = O.dft64(p64["bin"], vre, vim, p64["vin"], N, cell)
    rng = np.random.default_rng(cfg.seed)
    pix = np.stack([rng.integers(0, N, 64), rng.integers(0, N, 64)], axis=1).astype(np.int32)
    np.savez_compressed(
        os.path.join(HERE, "small_snapshot.npz"),
        uvw=snap["uvw"], freq=snap["freq"], vis0_re=snap["vis0_re"], vis0_im=snap["vis0_im"],
        vis3_re=snap["vis3_re"], vis3_im=snap["vis3_im"], flag0=snap["flag0"],
        flag3=snap["flag3"], weight0=snap["weight0"], weight3=snap["weight3"],
        num_rows=snap["num_rows"], num_channels=snap["num_channels"],
        array_center_x=snap["array_center_x"], image_size=N, cell=cell,
        num_terms=cfg.num_terms)
    np.savez_compressed(
        os.path.join(HERE, "small_dft64.npz"),
        vin64=p64["vin"], evals64=p64["evals"], dft=dft, pixels=pix,
        dft_pixels=dft[pix[:, 1], pix[:, 0]])
    print("peak |DFT| =", np.nanmax(np.abs(dft)))


if __name__ == "__main__":
    main()
