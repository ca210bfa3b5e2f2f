"""The window gridder's ES polynomials (csrc/es_poly.hpp) against the kernel itself.

The header's coefficients are parsed as the device reads them and evaluated in float32 in the
device's order of operations; every cell of the footprint, over a dense grid of offsets, must
lie within 2e-7 of the peak of exp(beta (sqrt(1 - x^2) - 1)) evaluated in float64
(SVDIntegratedImager.cu:884-888, beta of SVDIntegratedImager.cu:1235).  The header must also be
what the generator writes, so that the table and its fit cannot drift apart.
"""
import importlib.util
import math
import os
import re

import numpy as np

from conftest import ROOT

CSRC = os.path.join(ROOT, "svdimager_amd", "csrc")
BETA = 15.3704324328


def _generator():
    spec = importlib.util.spec_from_file_location("gen_es_poly", os.path.join(CSRC, "gen_es_poly.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _parse_header():
    text = open(os.path.join(CSRC, "es_poly.hpp")).read()

    def table(name):
        body = re.search(name + r"\[4\]\[\d+\] = \{(.*?)\n\};", text, re.S).group(1)
        rows = re.findall(r"\{([^}]*)\}", body)
        return np.array([[np.float32(float(v.strip().rstrip("f"))) for v in r.split(",")] for r in rows], dtype=np.float32)

    edge = float(re.search(r"kEsPolyEdge = ([0-9.eE+-]+)f;", text).group(1))
    return table("kEsPolyEven"), table("kEsPolyOdd"), np.float32(edge), text


def _es(x):
    x = np.asarray(x, dtype=np.float64)
    inside = np.abs(x) <= 1.0
    return np.where(inside, np.exp(BETA * (np.sqrt(np.where(inside, 1.0 - x * x, 0.0)) - 1.0)), 0.0)


def test_header_matches_generator():
    gen = _generator()
    even, odd = gen.coefficients()
    assert gen.header(even, odd) == _parse_header()[3]


def test_cells_within_2e7_of_peak():
    gen = _generator()
    even, odd, edge, _ = _parse_header()
    assert even.shape == (4, 5) and odd.shape == (4, 4)
    f = np.linspace(0.0, 1.0, 200001)
    got = gen.evaluate(even, odd, f).astype(np.float64) / 128.0
    for j in range(8):
        want = _es((f + j) / 4.0 - 1.0)
        err = np.abs(got[j] - want).max()
        assert err < 2e-7, (j, err)
    # Interior quarters are far below the float32 rounding of the peak.
    mid = gen.evaluate(even, odd, f)[3].astype(np.float64) / 128.0
    assert np.abs(mid - _es((f + 3) / 4.0 - 1.0)).max() < 1.5e-7
    # The ninth cell (x = 1 exactly when f = 0) and the kernel's edge value.
    assert math.isclose(float(edge), math.exp(-BETA) * 128.0, rel_tol=1e-6)


def test_mirror_pairs_and_peak():
    gen = _generator()
    even, odd, _, _ = _parse_header()
    f = np.linspace(0.0, 1.0, 1001)
    v = gen.evaluate(even, odd, f)
    # ES is even: cell j at f is cell 7 - j at 1 - f.
    np.testing.assert_allclose(v[::-1, ::-1], v, rtol=0, atol=2e-7 * 128)
    # Peak: x = 0 is cell 4 at f = 0 (and cell 3 at f = 1).
    assert abs(float(gen.evaluate(even, odd, np.array([0.0]))[4, 0]) - 128.0) < 128.0 * 2e-7
