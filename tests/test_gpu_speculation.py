"""Snapshot calls enqueued before their plan is read back (svdi_image_snapshot on a context that
imaged the same size, cell and options before with one slab): the guess (that K, one slab) is
kept when the plan agrees and the work is enqueued again when it does not.  Every call's image
and plan must equal those of the same snapshot imaged on a fresh context, where nothing is
guessed.  The snapshots: the small array at two hour angles (K = 8, one slab), with 250 m
station heights (alpha 6.0: K = 16, one slab), and with 200 m heights at 1.5 h (alpha 13.4: two
slabs).
"""
import dataclasses

import numpy as np
import pytest

from conftest import announce, assert_parity
import svdimager_amd as S
from svdimager_amd import synth

pytestmark = pytest.mark.gpu

IGNORED = ("elapsed_ms", "enqueued_early")


def _snap(height_sigma, ha):
    cfg = dataclasses.replace(synth.SMALL, height_sigma=height_sigma)
    return S.DeviceSnapshot(synth.make_snapshot(cfg, ha_hours=ha), "cuda")


def _image(ctx, d, output):
    img, plan = S.image_snapshot(ctx, d, synth.SMALL.image_size, synth.SMALL.cell, output=output)
    return img.cpu().numpy(), plan


def _fresh(d, output):
    ctx = S.Context(0)
    try:
        return _image(ctx, d, output)
    finally:
        ctx.close()


@pytest.mark.parametrize("output", [S.OUTPUT_SIGNED, S.OUTPUT_DIRTY])
def test_guessed_calls_match_fresh_contexts(output):
    a = _snap(5.0, 0.0)
    a2 = _snap(5.0, 1.5)
    b = _snap(250.0, 0.0)
    c = _snap(200.0, 1.5)
    # (snapshot, expected enqueued_early, expected K, expected slabs)
    calls = [(a, 0, 8, 1), (a2, 1, 8, 1), (b, 2, 16, 1), (b, 1, 16, 1), (c, 2, 16, 2), (a, 0, 8, 1)]
    ctx = S.Context(0)
    try:
        for i, (d, early, K, slabs) in enumerate(calls):
            announce(f"guessed call {i} output {output}: expect K {K}, {slabs} slab(s), early {early}")
            img, plan = _image(ctx, d, output)
            ref, ref_plan = _fresh(d, output)
            assert plan["enqueued_early"] == early, (i, plan["enqueued_early"])
            assert ref_plan["enqueued_early"] == 0
            assert plan["num_terms"] == K and plan["num_slabs"] == slabs, (i, plan["num_terms"], plan["num_slabs"])
            assert plan["gridder"] == S.GRIDDER_MFMA
            p = {k: v for k, v in plan.items() if k not in IGNORED}
            q = {k: v for k, v in ref_plan.items() if k not in IGNORED}
            assert p == q, i
            assert np.isfinite(img).all() and np.abs(ref).max() > 0
            assert_parity(img, ref, f"call {i}", whole=1e-5, central=1e-6)
    finally:
        ctx.close()


def test_guessed_batched_calls_match_fresh_contexts():
    """A batched call on a context with a record (K = 8, one slab): every lane's snapshot is
    enqueued on that guess; the K = 16 and the two-slab snapshots miss and are enqueued again."""
    snaps = [_snap(5.0, 0.0), _snap(5.0, 1.5), _snap(250.0, 0.0), _snap(200.0, 1.5)]
    want_early = [1, 1, 2, 2]
    N, cell = synth.SMALL.image_size, synth.SMALL.cell
    ctx = S.Context(0)
    try:
        _image(ctx, snaps[0], S.OUTPUT_SIGNED)             # the record: K = 8, one slab
        announce("guessed batched call of 4 snapshots (2 hits, 2 misses)")
        imgs, plans = S.image_snapshot_batch(ctx, snaps, N, cell, output=S.OUTPUT_SIGNED)
        for i, d in enumerate(snaps):
            ref, ref_plan = _fresh(d, S.OUTPUT_SIGNED)
            assert plans[i]["enqueued_early"] == want_early[i], (i, plans[i]["enqueued_early"])
            assert plans[i]["num_terms"] == ref_plan["num_terms"] and plans[i]["num_slabs"] == ref_plan["num_slabs"]
            img = imgs[i].cpu().numpy()
            assert np.isfinite(img).all()
            assert_parity(img, ref, f"batched {i}", whole=1e-5, central=1e-6)
    finally:
        ctx.close()
