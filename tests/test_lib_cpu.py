"""CPU checks of libsvdimager: it loads, exports every declared symbol, its host-only entry
points (FITS I/O) work, device entry points fail cleanly without a GPU, and the scalar device
math (svdi_math.hpp, compiled here for the host) matches scipy / numpy / the oracle."""
import ctypes
import os
import re
import struct
import subprocess

import numpy as np
import pytest
from scipy import special

from conftest import ROOT
from oracle import oracle as O
from svdimager_amd import _lib

HEADER = os.path.join(ROOT, "include", "svdimager.h")
CSRC = os.path.join(ROOT, "svdimager_amd", "csrc")


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(_lib.LIB_PATH):
        from svdimager_amd import build
        build.build()
    return _lib.lib()


def test_exports_every_declared_symbol(L):
    text = open(HEADER).read()
    declared = set(re.findall(r"\b(svdi_[a-z_0-9]+)\s*\(", text))
    assert declared == set(s for s in _lib.EXPORTED_SYMBOLS if s.startswith("svdi_"))
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines())
    for sym in _lib.EXPORTED_SYMBOLS:
        assert sym in exported, sym
    assert L.svdi_version() == 1


def test_fipipe_has_the_reference_cxx_linkage(L):
    # int FIpipe(float*, float*, float*, float*, float*, size_t, size_t, float) with C++
    # linkage (archive/pipe_interface.cpp:8) mangles to this name on x86-64.
    assert hasattr(L, "_Z6FIpipePfS_S_S_S_mmf")


def test_create_without_gpu_returns_an_error(L):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = L.svdi_create(ctypes.byref(h), 0)
    assert rc != 0 and not h
    assert L.svdi_last_error()


def test_null_context_is_an_error(L):
    """Entry points that take a context reject a null one with a status and a message (no
    device call is made)."""
    n = ctypes.c_size_t(7)
    for rc in (L.svdi_workspace_bytes(None, ctypes.byref(n)), L.svdi_trim(None),
               L.svdi_set_profiling(None, 1), L.svdi_destroy(None)):
        if rc == 0:                      # svdi_destroy(NULL) may be a no-op, as free(NULL)
            continue
        assert L.svdi_last_error()
    assert L.svdi_workspace_bytes(None, ctypes.byref(n)) != 0 and n.value == 7
    assert L.svdi_trim(None) != 0
    from svdimager_amd._lib import Tuning
    t = Tuning()
    assert L.svdi_set_tuning(None, 0, ctypes.byref(t)) == 1 and L.svdi_last_error()
    assert L.svdi_get_tuning(None, 0, ctypes.byref(t)) == 1
    assert L.svdi_autotune(None, None, 1024, 1e-5, None, 1, ctypes.byref(t)) == 1
    from svdimager_amd._lib import ServiceConfig, ServiceStats
    h = ctypes.c_void_p()
    cfg = ServiceConfig(budget_bytes=1 << 30, max_batch=4, max_inflight=2)
    assert L.svdi_service_create(None, ctypes.byref(cfg), ctypes.byref(h)) == 1 and L.svdi_last_error()
    assert L.svdi_service_destroy(None) == 1
    assert L.svdi_service_stats(None, ctypes.byref(ServiceStats())) == 1


def test_fits_round_trip(tmp_path):
    from svdimager_amd import read_fits_image, write_fits_2d
    img = np.random.default_rng(1).normal(size=(12, 7)).astype(np.float32)
    p = str(tmp_path / "a.fits")
    write_fits_2d(p, img)
    raw = open(p, "rb").read()
    assert len(raw) % 2880 == 0
    header = raw[:2880].decode("ascii")
    cards = [header[i:i + 80] for i in range(0, 2880, 80)]
    assert cards[0].startswith("SIMPLE  =                    T")
    assert cards[1].startswith("BITPIX  =                  -32")
    assert cards[3].startswith("NAXIS1  =                    7")
    assert cards[4].startswith("NAXIS2  =                   12")
    data = np.frombuffer(raw[2880:2880 + img.size * 4], dtype=">f4").reshape(12, 7)
    assert np.array_equal(data, img)
    assert np.array_equal(read_fits_image(p), img)
    write_fits_2d(p, img[:3])                 # overwrite, as cfitsio's "!" prefix does
    assert read_fits_image(p).shape == (3, 7)


def _fits_bytes(cards, payload):
    h = "".join(c.ljust(80) for c in cards + ["END"])
    h = h.ljust((len(h) + 2879) // 2880 * 2880)
    pad = (len(payload) + 2879) // 2880 * 2880 - len(payload)
    return h.encode("ascii") + payload + b"\0" * pad


def test_fits_reads_fp64_bin_and_scaled_int(tmp_path):
    # A float64 S x 3 array as astropy writes Bin (archive/SVD_MFS.py:108), and a 1-D
    # 16-bit image with BSCALE / BZERO.
    b = np.random.default_rng(2).normal(size=(5, 3))
    p = tmp_path / "bin.fits"
    p.write_bytes(_fits_bytes(["SIMPLE  =                    T", "BITPIX  =                  -64",
                               "NAXIS   =                    2", "NAXIS1  =                    3",
                               "NAXIS2  =                    5"], b.astype(">f8").tobytes()))
    from svdimager_amd import read_fits_image
    assert np.allclose(read_fits_image(str(p)), b.astype(np.float32))
    q = tmp_path / "v.fits"
    vals = np.array([1, -2, 300, 7], ">i2")
    q.write_bytes(_fits_bytes(["SIMPLE  =                    T", "BITPIX  =                   16",
                               "NAXIS   =                    1", "NAXIS1  =                    4",
                               "BSCALE  =                  0.5", "BZERO   =                 10.0"],
                              vals.tobytes()))
    assert np.allclose(read_fits_image(str(q)), 10.0 + 0.5 * vals.astype(np.float32))


def test_cli_usage_exit_code():
    cli = os.path.join(ROOT, "svdimager_amd", "bin", "sharedlibrary_gpu")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 2 and "Usage" in r.stderr
    r = subprocess.run([cli, "/nonexistent.fits", "a", "b", "c", "16", "1", "1e-3", "o.fits"],
                       capture_output=True, text=True)
    assert r.returncode == 1


def _cli_inputs(tmp_path, S):
    from svdimager_amd import write_fits_2d
    rng = np.random.default_rng(5)
    paths = []
    for name, arr in (("re", rng.normal(size=S)), ("im", rng.normal(size=S)),
                      ("b", rng.normal(size=(S, 3))), ("v", np.eye(3))):
        p = str(tmp_path / f"{name}.fits")
        write_fits_2d(p, np.asarray(arr, np.float32).reshape(-1, 3) if name in ("b", "v")
                      else np.asarray(arr, np.float32).reshape(1, -1))
        paths.append(p)
    return paths


def test_cli_rejects_rows_beyond_the_inputs(tmp_path):
    """Number_of_Rows larger than the FITS inputs hold is an error (exit 1, no output), checked
    before any device work."""
    cli = os.path.join(ROOT, "svdimager_amd", "bin", "sharedlibrary_gpu")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    paths = _cli_inputs(tmp_path, 10)
    out = str(tmp_path / "out.fits")
    for rows in ("11", "100"):
        r = subprocess.run([cli, *paths, "16", rows, "1e-3", out], capture_output=True, text=True)
        assert r.returncode == 1 and "does not fit" in r.stderr, r.stderr
        assert not os.path.exists(out)


def test_cli_fipipe_failure_writes_no_image(tmp_path):
    """When FIpipe fails (here: no GPU), the CLI exits 1 and writes no output file."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    cli = os.path.join(ROOT, "svdimager_amd", "bin", "sharedlibrary_gpu")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    paths = _cli_inputs(tmp_path, 10)
    out = str(tmp_path / "out.fits")
    r = subprocess.run([cli, *paths, "16", "10", "1e-3", out], capture_output=True, text=True)
    assert r.returncode == 1
    assert not os.path.exists(out)


CXX_CALLER = r"""
#include <cstddef>
#include <cstdio>
// archive/pipe_interface.cpp:8, verbatim in meaning: C++ linkage, no header of this library.
int FIpipe(float* Visreal, float* Visimag, float* Bin, float* Vin, float* dirty_image,
           size_t num_baselines, size_t image_size, float cell_size);
int main() {
    float re[4] = {1, 1, 1, 1}, im[4] = {0, 0, 0, 0};
    float b[12] = {0, 0, 0, 1, 1, 0, -1, 2, 0, 3, -1, 0};
    float v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float dirty[16 * 16];
    const int rc = FIpipe(re, im, b, v, dirty, 4, 16, 1e-3f);
    std::printf("FIpipe returned %d\n", rc);
    return 0;
}
"""


def test_cxx_caller_with_the_reference_prototype_links(tmp_path):
    """A C++ caller that declares FIpipe exactly as archive/pipe_interface.cpp:8 compiles and links
    against libsvdimager and the symbol resolves at run time (the call itself fails without a
    GPU and returns nonzero; with one it returns 0)."""
    src = tmp_path / "caller.cpp"
    src.write_text(CXX_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", str(src), "-o", str(exe), f"-L{libdir}", "-lsvdimager",
                    f"-Wl,-rpath,{libdir}"], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "FIpipe returned" in r.stdout


PROBE = r"""
#include "svdi_math.hpp"
extern "C" {
void p_bessel(double x, int n, double* j) { svdi::bessel_sequence_f64(x, n, j); }
void p_bessel_fixed4(double x, double* j) { svdi::bessel_sequence_fixed<4>(x, j); }
void p_bessel_fixed8(double x, double* j) { svdi::bessel_sequence_fixed<8>(x, j); }
void p_bessel_fixed16(double x, double* j) { svdi::bessel_sequence_fixed<16>(x, j); }
void p_bessel_fixed32(double x, double* j) { svdi::bessel_sequence_fixed<32>(x, j); }
double p_bessel_series(int n, double x) { return svdi::bessel_series(n, x); }
void p_eigen(const double* a, double* ev, double* vec) { svdi::eigen_sym3(a, ev, vec); }
void p_basis(const double* vec, double acx, double* basis) { svdi::oriented_basis(vec, acx, basis); }
float p_es(float x) { return svdi::es_kernel(x); }
void p_reproject(long di, float cell, const float* V, long x, long y, int archive, float* o) {
    if (archive) svdi::sin_reproject_archive(di, cell, V, x, y, o, o + 1);
    else svdi::sin_reproject(di, cell, V, x, y, o, o + 1);
}
}
"""


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    d = tmp_path_factory.mktemp("probe")
    src = d / "probe.cpp"
    src.write_text(PROBE)
    so = d / "probe.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", CSRC, str(src), "-o",
                    str(so)], check=True)
    P = ctypes.CDLL(str(so))
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    P.p_bessel.argtypes = [ctypes.c_double, ctypes.c_int, dp]
    P.p_bessel_fixed4.argtypes = [ctypes.c_double, dp]
    P.p_bessel_fixed8.argtypes = [ctypes.c_double, dp]
    P.p_bessel_fixed16.argtypes = [ctypes.c_double, dp]
    P.p_bessel_fixed32.argtypes = [ctypes.c_double, dp]
    P.p_bessel_series.argtypes = [ctypes.c_int, ctypes.c_double]
    P.p_bessel_series.restype = ctypes.c_double
    P.p_eigen.argtypes = [dp, dp, dp]
    P.p_basis.argtypes = [dp, ctypes.c_double, dp]
    P.p_es.argtypes = [ctypes.c_float]
    P.p_es.restype = ctypes.c_float
    P.p_reproject.argtypes = [ctypes.c_long, ctypes.c_float, fp, ctypes.c_long, ctypes.c_long,
                              ctypes.c_int, fp]
    return P


BESSEL_X = [0.0, 1e-9, 0.005, 0.00999, 0.01, 0.02, 0.3, 0.999, 1.0, 4.2, 10.0, 37.5, -6.5, -0.05,
            999.0, 1000.0, 2.5e4]


@pytest.mark.parametrize("x", BESSEL_X)
def test_device_bessel_matches_scipy(probe, x):
    """J_0..J_31 against scipy: absolute error (the image pass weights J_k against moments of
    comparable size), plus relative error for the orders above 1e-30."""
    ref = special.jv(np.arange(32), x)
    for fn, n in ((lambda j: probe.p_bessel(x, 32, j), 32), (probe.p_bessel_fixed32, 32),
                  (probe.p_bessel_fixed8, 8)):
        j = np.zeros(n)
        fn(x, j) if fn in (probe.p_bessel_fixed32, probe.p_bessel_fixed8) else fn(j)
        r = ref[:n]
        assert np.abs(j - r).max() < 1e-12 + 1e-10 * np.abs(r).max(), (x, n)
        big = np.abs(r) > 1e-30
        if abs(x) < 1000:
            assert np.all(np.abs(j[big] - r[big]) <= 1e-10 * np.abs(r[big])), (x, n)


def test_bessel_fixed_start_range(probe):
    """The unrolled recurrences from fixed start orders (orders up to 16: m = max(COUNT, 8) + 12
    for 1e-2 <= |x| < 2, m = 36 for 2 <= |x| < 8), on a dense sweep through both ranges and their
    ends: J_0..J_{COUNT-1} against scipy to 1e-15 absolute, and 1e-10 relative for the orders above
    1e-30 (near a zero of J_n the relative error of a correct value grows).  A start order too low
    for its range (m = 16 at COUNT = 4, measured 5e-14 absolute near x = 2) fails the first bound."""
    xs = np.concatenate([np.linspace(0.01, 7.999, 401), [0.0100001, 1.999, 2.0, 2.001, 7.9999999, 8.0, 8.5],
                         -np.linspace(0.01, 7.999, 37)])
    for count, fn in ((4, probe.p_bessel_fixed4), (8, probe.p_bessel_fixed8), (16, probe.p_bessel_fixed16)):
        for x in xs:
            ref = special.jv(np.arange(count), x)
            j = np.zeros(count)
            fn(float(x), j)
            assert np.abs(j - ref).max() <= 1e-15, (count, x)
            big = np.abs(ref) > 1e-30
            assert np.all(np.abs(j[big] - ref[big]) <= 1e-10 * np.abs(ref[big])), (count, x)


def test_bessel_power_series_below_two(probe):
    """The per-order power series that the term-error kernel uses for |a| < 2 (one order per
    lane): J_0..J_31 against scipy to 5e-14 relative (orders above 1e-200) and 4e-16 absolute."""
    xs = np.concatenate([[0.0, 1e-12, 1e-3, 0.712, 1.0, 1.999, -0.4, -1.999],
                         np.random.default_rng(9).uniform(-2.0, 2.0, 40)])
    for x in xs:
        ref = special.jv(np.arange(32), x)
        got = np.array([probe.p_bessel_series(n, float(x)) for n in range(32)])
        assert np.abs(got - ref).max() <= 4e-16, x
        big = np.abs(ref) > 1e-200
        assert np.all(np.abs(got[big] - ref[big]) <= 5e-14 * np.abs(ref[big])), x


def test_device_eigen_and_basis_match_oracle(probe):
    rng = np.random.default_rng(4)
    for _ in range(10):
        a = rng.normal(size=(3, 3))
        a = a @ a.T
        ev, vec = np.zeros(3), np.zeros(9)
        probe.p_eigen(np.ascontiguousarray(a.reshape(-1)), ev, vec)
        ev_np = np.linalg.eigh(a)[0]
        assert np.allclose(ev, ev_np, rtol=1e-12)
        basis = np.zeros(9)
        probe.p_basis(vec, -2565018.0, basis)
        R = basis.reshape(3, 3)
        assert np.abs(R @ R.T - np.eye(3)).max() < 1e-12 and np.linalg.det(R) > 0
        assert R[0, 0] > 0 and R[1, 1] * -2565018.0 >= 0


def test_device_es_and_reprojection_match_oracle(probe):
    for x in np.linspace(-1.1, 1.1, 45, dtype=np.float32):
        assert probe.p_es(float(x)) == O.es_f32(float(x))
    V = O.preprocess_f64(__import__("conftest").load_small_snapshot())["vin"].astype(np.float32)
    N, cell = 48, 3e-3
    for archive in (0, 1):
        m = O.reproject(N, cell, V, archive)
        o = np.zeros(2, np.float32)
        for y in range(0, N, 5):
            for x in range(0, N, 3):
                probe.p_reproject(N, cell, np.ascontiguousarray(V.reshape(-1)), x, y, archive, o)
                assert np.array_equal(o, m[y, x]), (archive, x, y)
