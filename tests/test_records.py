"""The window gridder's 16-byte records (csrc/records.hpp), compiled for the CPU.

Round trips against the values that went in: positions to 2^-19 of a cell with the footprint's
first cell and fraction consistent (an integer position has fraction 0, its ninth cell), s to
2^-24 (2^-23 at s = 1, -1 exact), and the visibility pair to 2^-23 of its larger component, over
magnitudes from 1e-30 to 1e30 and both signs.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "svdimager_amd", "csrc")

PROBE = r"""
#include "records.hpp"
extern "C" {
void enc(long n, const int* b1, const float* q1, const int* b2, const float* q2, const float* s,
         const float* vr, const float* vi, unsigned* w) {
    for (long i = 0; i < n; ++i) {
        const svdi::RecordWords r = svdi::rec_encode(b1[i], q1[i], b2[i], q2[i], s[i], vr[i], vi[i]);
        for (int k = 0; k < 4; ++k) w[4 * i + k] = r.w[k];
    }
}
void dec(long n, const unsigned* w, float* p1, float* p2, int* c1, unsigned* f1, float* s, float* vr, float* vi) {
    for (long i = 0; i < n; ++i) {
        const unsigned* r = w + 4 * i;
        p1[i] = svdi::rec_pos(r[2]);
        p2[i] = svdi::rec_pos(r[3]);
        c1[i] = svdi::rec_cell(r[2]);
        f1[i] = svdi::rec_frac_bits(r[2]);
        s[i] = svdi::rec_s(r[0], r[1], r[2]);
        vr[i] = svdi::rec_v(r[0], r[3], 0);
        vi[i] = svdi::rec_v(r[1], r[3], 0);
    }
}
}
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("records")
    src = d / "probe.cpp"
    src.write_text(PROBE)
    so = d / "probe.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", CSRC, str(src), "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    u32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.enc.argtypes = [ctypes.c_long, i32, f32, i32, f32, f32, f32, f32, u32]
    L.dec.argtypes = [ctypes.c_long, u32, f32, f32, i32, u32, f32, f32, f32]
    return L


def roundtrip(L, b1, q1, b2, q2, s, vr, vi):
    n = len(q1)
    a = [np.ascontiguousarray(x, dtype=t) for x, t in ((b1, np.int32), (q1, np.float32), (b2, np.int32),
                                                       (q2, np.float32), (s, np.float32), (vr, np.float32),
                                                       (vi, np.float32))]
    w = np.zeros(4 * n, np.uint32)
    L.enc(n, *a, w)
    out = [np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.int32), np.zeros(n, np.uint32),
           np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)]
    L.dec(n, w, *out)
    return a, out


def test_positions_cells_and_fractions(lib):
    rng = np.random.default_rng(5)
    n = 200000
    b1 = rng.integers(-8, 15, n)
    b2 = rng.integers(-15, 16, n)
    q1 = rng.uniform(-5.0, 5.0, n).astype(np.float32)
    q2 = rng.uniform(-5.0, 5.0, n).astype(np.float32)
    q1[:1000] = np.round(q1[:1000])                           # integer positions
    a, (p1, p2, c1, f1, *_rest) = roundtrip(lib, b1, q1, b2, q2, np.zeros(n), np.ones(n), np.zeros(n))
    want1 = b1 + q1.astype(np.float64)
    want2 = b2 + q2.astype(np.float64)
    assert np.abs(p1 - want1).max() <= 2.0 ** -19
    assert np.abs(p2 - want2).max() <= 2.0 ** -19
    # First footprint cell and fraction: p = cell + 4 - F, 0 <= F < 1.
    F = f1 / 2.0 ** 18
    assert (F >= 0).all() and (F < 1).all()
    np.testing.assert_allclose(c1 + 4 - F, want1, atol=2.0 ** -19, rtol=0)
    # Integer positions: fraction 0 (the ninth footprint cell), cell = p - 4.
    ints = np.arange(1000)
    assert (f1[ints] == 0).all() and (c1[ints] == (want1[ints] - 4)).all()
    # Elsewhere the cell is ceil(p - 4), up to positions within a rounding step of an integer.
    far = np.abs(want1 - np.round(want1)) > 2.0 ** -18
    assert (c1[far] == np.ceil(want1[far] - 4)).all()


def test_s_precision_and_ends(lib):
    s = np.concatenate([[-1.0, 1.0, 0.0, 0.5, -0.5], np.random.default_rng(6).uniform(-1, 1, 100000)]).astype(np.float32)
    n = len(s)
    z = np.zeros(n)
    a, out = roundtrip(lib, z, z, z, z, s, np.ones(n), z)
    got = out[4]
    assert got[0] == -1.0 and abs(got[1] - 1.0) <= 2.0 ** -23 and got[2] == 0.0
    top = s.astype(np.float64) > 1.0 - 2.0 ** -24
    assert np.abs(got - s.astype(np.float64))[~top].max() <= 2.0 ** -24
    assert np.abs(got - s.astype(np.float64)).max() <= 2.0 ** -23


@pytest.mark.parametrize("scale", [1e-30, 1e-6, 1.0, 3.7e3, 1e30])
def test_visibility_pairs(lib, scale):
    rng = np.random.default_rng(int(np.log10(scale) + 40))
    n = 100000
    vr = (rng.standard_normal(n) * scale).astype(np.float32)
    vi = (rng.standard_normal(n) * scale * rng.choice([1.0, 1e-3, 1e3], n)).astype(np.float32)
    vr[:10] = 0.0
    vi[:5] = 0.0
    z = np.zeros(n)
    a, out = roundtrip(lib, z, z, z, z, z, vr, vi)
    gr, gi = out[5].astype(np.float64), out[6].astype(np.float64)
    m = np.maximum(np.abs(vr), np.abs(vi)).astype(np.float64)
    assert (gr[:5] == 0).all() and (gi[:5] == 0).all()
    ok = m > 0
    assert (np.abs(gr - vr)[ok] <= 2.0 ** -23 * m[ok]).all()
    assert (np.abs(gi - vi)[ok] <= 2.0 ** -23 * m[ok]).all()
    assert (np.sign(gr) == np.sign(vr))[np.abs(vr) > 2.0 ** -23 * m].all()


def test_bits_match_the_fp64_rules(lib):
    """Cell and fraction bits equal the FP64 rule k = ceil(q - 4), F = rint((k + 4 - q) 2^18) (a
    carry at F = 2^18), and the mantissas rint(v 2^(23 - E)), bit for bit, including exact ties
    of the fraction rounding, tiny q and |q| up to 4100 cells."""
    rng = np.random.default_rng(11)
    n = 400000
    q1 = np.concatenate([rng.uniform(-4100, 4100, n // 4), rng.uniform(2.9, 4.1, n // 4),
                         (rng.integers(-2 ** 20, 2 ** 20, n // 4) + 0.5) / 2.0 ** 18,
                         10.0 ** rng.uniform(-30, 0, n // 4) * rng.choice([-1, 1], n // 4)]).astype(np.float32)
    q2 = rng.permutation(q1)
    b1 = rng.integers(-8, 8, n)
    b2 = rng.integers(-8, 8, n)
    vr = (rng.standard_normal(n) * 10.0 ** rng.uniform(-20, 20, n)).astype(np.float32)
    vi = (rng.standard_normal(n) * 10.0 ** rng.uniform(-20, 20, n)).astype(np.float32)
    a = [np.ascontiguousarray(x, dtype=t) for x, t in ((b1, np.int32), (q1, np.float32), (b2, np.int32),
                                                       (q2, np.float32), (np.zeros(n), np.float32),
                                                       (vr, np.float32), (vi, np.float32))]
    w = np.zeros(4 * n, np.uint32)
    lib.enc(n, *a, w)
    w = w.reshape(n, 4)

    def pos_bits(base, q):
        qd = q.astype(np.float64)
        k = np.ceil(qd - 4.0)
        cell = base + k.astype(np.int64)
        frac = np.rint((k + 4.0 - qd) * 2.0 ** 18).astype(np.int64)
        carry = frac >= 2 ** 18
        frac[carry] -= 2 ** 18
        cell[carry] -= 1
        cell = np.clip(cell, -32, 31)
        return ((cell.astype(np.int64) & 0x3f) << 26) | (frac << 8)

    assert (w[:, 2] & 0xffffff00 == pos_bits(b1, q1)).all()
    assert (w[:, 3] & 0xffffff00 == pos_bits(b2, q2)).all()
    m = np.maximum(np.abs(vr), np.abs(vi)).astype(np.float64)
    e = np.frexp(m)[1]
    for v, word in ((vr, w[:, 0]), (vi, w[:, 1])):
        mant = np.clip(np.rint(np.ldexp(v.astype(np.float64), 23 - e)), -2 ** 23, 2 ** 23 - 1).astype(np.int64)
        assert ((word >> 8).astype(np.int64) == (mant & 0xffffff)).all()
    assert ((w[:, 3] & 0xff).astype(np.int64) == (e & 0xff)).all()
