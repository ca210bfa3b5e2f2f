"""The integrated CLI's MeasurementSet reader (SVDIntegratedImager.cu:99-134,180-378) on columns
exported as .npy files: the library's reader (svdi_ms_read_snapshot) against the oracle's
restatement, bit for bit, on tables with shuffled TIME values, both row-axis layouts, several
spectral windows, and without WEIGHT_SPECTRUM / ARRAY_CENTER; the oracle itself against the
closed form of the synthetic tables (each DATA / FLAG / WEIGHT value encodes its row, channel and
polarisation); the reference's error cases; and the CLI's exit codes before it reaches a GPU."""
import os
import subprocess

import numpy as np
import pytest

import svdimager_amd as S
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "svdimager_amd", "bin", "svd_integrated_imager_gpu")


def write_ms(d, times, nchan=5, npol=4, layout="chan_pol", nspw=1, weights=True, centre=True,
             seed=0):
    """A main table of len(times) rows.  DATA[row, chan, pol] = (1000 row + 10 chan + pol)
    + i (pol - row / 8), FLAG = (row + chan + pol) % 3 == 0, WEIGHT = 0.5 + row + chan / 100 + pol / 1000.
    layout "chan_pol": row arrays (nchan, npol) as python-casacore returns DATA (IPosition
    (npol, nchan)); "pol_chan": (npol, nchan)."""
    os.makedirs(os.path.join(d, "SPECTRAL_WINDOW"), exist_ok=True)
    os.makedirs(os.path.join(d, "OBSERVATION"), exist_ok=True)
    R = len(times)
    r = np.arange(R)[:, None, None]
    c = np.arange(nchan)[None, :, None]
    p = np.arange(npol)[None, None, :]
    data = ((1000 * r + 10 * c + p) + 1j * (p - r / 8.0)).astype(np.complex64)
    flag = ((r + c + p) % 3 == 0)
    weight = (0.5 + r + c / 100.0 + p / 1000.0).astype(np.float32)
    if layout == "pol_chan":
        data, flag, weight = (np.ascontiguousarray(a.transpose(0, 2, 1)) for a in (data, flag, weight))
    rng = np.random.default_rng(seed)
    np.save(os.path.join(d, "TIME.npy"), np.asarray(times, np.float64))
    np.save(os.path.join(d, "UVW.npy"), rng.normal(0, 300.0, (R, 3)))
    np.save(os.path.join(d, "DATA.npy"), data)
    np.save(os.path.join(d, "FLAG.npy"), flag)
    if weights:
        np.save(os.path.join(d, "WEIGHT_SPECTRUM.npy"), weight)
    freq = 1.0e8 + 1.0e6 * np.arange(nspw * nchan, dtype=np.float64).reshape(nspw, nchan)
    np.save(os.path.join(d, "SPECTRAL_WINDOW", "CHAN_FREQ.npy"), freq)
    if centre:
        np.save(os.path.join(d, "OBSERVATION", "ARRAY_CENTER.npy"), np.array([[-2565018.0, 5085414.0, -2861015.0]]))
    return d


TIMES = [5.0, 2.0, 5.0, 9.5, 2.0, 2.0, 9.5, 5.0, 2.0]        # three snapshots, rows out of order


def assert_same(a, b):
    for k in ("uvw", "freq", "vis0_re", "vis0_im", "vis3_re", "vis3_im", "flag0", "flag3", "weight0", "weight3"):
        x, y = a[k], (b["freq_hz"] if k == "freq" else b[k])
        if x is None or y is None:
            assert x is None and y is None, k
            continue
        assert np.array_equal(np.asarray(x).reshape(-1), np.asarray(y).reshape(-1)), k
    for k in ("num_rows", "num_channels", "array_center_x", "time", "index", "total_snapshots"):
        assert a[k] == b[k], k


@pytest.mark.parametrize("layout", ["chan_pol", "pol_chan"])
@pytest.mark.parametrize("nspw", [1, 2])
def test_reader_matches_oracle(tmp_path, layout, nspw):
    d = write_ms(str(tmp_path), TIMES, layout=layout, nspw=nspw)
    for k in range(3):
        got = S.read_measurement_set_snapshot(d, k)
        ref = O.ms_read_snapshot(d, k)
        assert_same(got, ref)


def test_oracle_closed_form(tmp_path):
    """The oracle's snapshot is what the encoding says: rows of the k-th smallest TIME in table
    order, DATA pols 0 and 3, channel c of the concatenated windows reading channel c mod nchan."""
    nchan, nspw = 5, 2
    d = write_ms(str(tmp_path), TIMES, nchan=nchan, nspw=nspw)
    uvw = np.load(os.path.join(d, "UVW.npy"))
    for k, t in enumerate(sorted(set(TIMES))):
        s = O.ms_read_snapshot(d, k)
        rows = [i for i, x in enumerate(TIMES) if x == t]
        R, C = len(rows), nchan * nspw
        assert (s["num_rows"], s["num_channels"], s["time"], s["total_snapshots"]) == (R, C, t, 3)
        assert np.array_equal(s["uvw"], uvw[rows])
        for lr, row in enumerate(rows):
            for c in range(C):
                i, cr = lr + c * R, c % nchan
                for pol, tag in ((0, "0"), (3, "3")):
                    assert s[f"vis{tag}_re"][i] == np.float32(1000 * row + 10 * cr + pol)
                    assert s[f"vis{tag}_im"][i] == np.float32(pol - row / 8.0)
                    assert s[f"flag{tag}"][i] == int((row + cr + pol) % 3 == 0)
                    assert s[f"weight{tag}"][i] == np.float32(0.5 + row + cr / 100.0 + pol / 1000.0)
        assert np.array_equal(s["freq_hz"], 1.0e8 + 1.0e6 * np.arange(C))
        assert s["array_center_x"] == -2565018.0


def test_reader_without_optional_columns(tmp_path):
    d = write_ms(str(tmp_path), TIMES, weights=False, centre=False)
    got = S.read_measurement_set_snapshot(d, 1)
    assert got["weight0"] is None and got["weight3"] is None and got["array_center_x"] == 0.0
    assert_same(got, O.ms_read_snapshot(d, 1))


def test_reader_square_rows_and_wide_pols(tmp_path):
    """Rows whose two axes are neither 1, 2 nor 4 long: the longer one is the channel axis
    (detect_row_axis_layout), here 6 channels x 5 polarisations in both orders."""
    for layout in ("chan_pol", "pol_chan"):
        d = write_ms(str(tmp_path / layout), TIMES, nchan=6, npol=5, layout=layout)
        got = S.read_measurement_set_snapshot(d, 2)
        assert got["num_channels"] == 6
        assert_same(got, O.ms_read_snapshot(d, 2))


def test_reader_errors(tmp_path):
    d = write_ms(str(tmp_path / "a"), TIMES)
    with pytest.raises(S.SvdiError, match="Snapshot_Index is outside available TIME snapshots"):
        S.read_measurement_set_snapshot(d, 3)
    with pytest.raises(ValueError, match="Snapshot_Index"):
        O.ms_read_snapshot(d, 3)
    d2 = write_ms(str(tmp_path / "b"), TIMES, npol=2)
    with pytest.raises(S.SvdiError, match="fewer than 4 polarisations"):
        S.read_measurement_set_snapshot(d2, 0)
    os.remove(os.path.join(d, "DATA.npy"))
    with pytest.raises(S.SvdiError, match="DATA.npy"):
        S.read_measurement_set_snapshot(d, 0)
    d3 = write_ms(str(tmp_path / "c"), TIMES)
    np.save(os.path.join(d3, "UVW.npy"), np.asfortranarray(np.zeros((len(TIMES), 3))))
    with pytest.raises(S.SvdiError, match="Fortran"):
        S.read_measurement_set_snapshot(d3, 0)
    d4 = write_ms(str(tmp_path / "d"), TIMES)
    np.save(os.path.join(d4, "UVW.npy"), np.zeros((len(TIMES), 2)))
    with pytest.raises(S.SvdiError, match="UVW does not have 3 entries"):
        S.read_measurement_set_snapshot(d4, 0)
    d5 = write_ms(str(tmp_path / "e"), TIMES)
    np.save(os.path.join(d5, "DATA.npy"), np.zeros((len(TIMES), 5, 4), np.int32))
    with pytest.raises(S.SvdiError, match="unsupported element type"):
        S.read_measurement_set_snapshot(d5, 0)
    d6 = write_ms(str(tmp_path / "f"), TIMES)
    np.save(os.path.join(d6, "FLAG.npy"), np.zeros((len(TIMES) - 1, 5, 4), bool))
    with pytest.raises(S.SvdiError, match="fewer rows than TIME"):
        S.read_measurement_set_snapshot(d6, 0)             # TIME 2.0 holds the last row
    with pytest.raises(S.SvdiError, match="FLAG's shape differs"):
        S.read_measurement_set_snapshot(d6, 2)             # TIME 9.5: rows 3 and 6 only
    d7 = write_ms(str(tmp_path / "g"), TIMES)
    path = os.path.join(d7, "DATA.npy")
    os.truncate(path, os.path.getsize(path) - 8)
    with pytest.raises(S.SvdiError, match="fewer elements than its shape"):
        S.read_measurement_set_snapshot(d7, 0)


def test_cli_exit_codes_before_the_gpu(tmp_path):
    """Wrong argument count: usage, exit 2.  A failure reading the table or an invalid argument:
    'Integrated SVD imager failed: ...', exit 1 (SVDIntegratedImager.cu:1773-1853); none of
    these reaches the GPU."""
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 2 and "Usage:" in r.stderr
    d = write_ms(str(tmp_path / "ms"), TIMES)
    out = str(tmp_path / "o.fits")
    cases = [([str(tmp_path / "none"), "64", "1e-3", "0", out], "TIME.npy"),
             ([d, "64", "1e-3", "7", out], "Snapshot_Index is outside available TIME snapshots"),
             ([d, "0", "1e-3", "0", out], "Image_Size must be greater than zero"),
             ([d, "64", "-1", "0", out], "Cell_Size must be greater than zero")]
    for args, msg in cases:
        r = subprocess.run([CLI, *args], capture_output=True, text=True)
        assert r.returncode == 1, (args, r.stderr)
        assert "Integrated SVD imager failed: " in r.stderr and msg in r.stderr, r.stderr
        assert not os.path.exists(out)


def test_saved_columns_read_back(tmp_path, small_snap):
    """prepare.save_measurement_set_columns then the reader: the snapshot comes back as it was
    (weights and flags of polarisations 0 and 3, FP64 uvw and frequencies), at its TIME index."""
    from svdimager_amd import prepare
    half = dict(small_snap)
    for k in ("vis0_re", "vis0_im", "vis3_re", "vis3_im"):
        half[k] = (0.5 * small_snap[k]).astype(np.float32)
    d = prepare.save_measurement_set_columns(str(tmp_path), [small_snap, half], [100.0, 50.0])
    for idx, snap, t in ((1, small_snap, 100.0), (0, half, 50.0)):
        got = S.read_measurement_set_snapshot(d, idx)
        assert got["time"] == t and got["total_snapshots"] == 2
        for k in ("vis0_re", "vis0_im", "vis3_re", "vis3_im", "weight0", "weight3", "flag0", "flag3", "freq"):
            assert np.array_equal(got[k], np.asarray(snap[k]).astype(got[k].dtype)), k
        assert np.array_equal(got["uvw"], snap["uvw"]) and got["array_center_x"] == snap["array_center_x"]
