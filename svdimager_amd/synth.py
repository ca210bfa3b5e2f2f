"""Deterministic synthetic snapshots in the layout of the reference's MeasurementSet reader.

The reference reads one TIME snapshot of a MeasurementSet into ``HostMeasurementSetData``
(SVDIntegratedImager.cu:39-54, filled at :180-331): UVW per row, CHAN_FREQ per channel,
polarisations 0 (XX) and 3 (YY) of DATA / FLAG / WEIGHT_SPECTRUM, flattened with
``idx = row + chan * num_rows`` (SVDIntegratedImager.cu:90-92).  No MeasurementSets ship with
the reference, so this module builds snapshots with the same layout from a seeded model:

* stations in ENU, rotated to local equatorial XYZ at the SKA-Low site latitude (-26.82 deg),
  baselines for every pair i < j in row order, uvw at hour angle H and declination -30 deg
  (Thompson, Moran & Swenson, eq. 4.1);
* XX = YY = sum_k S_k exp(-2 pi i (u l + v m + w (n - 1)) nu / c) plus independent complex
  N(0, 0.01) noise per polarisation, stored as float32 real / imaginary arrays;
* row-level flags applied to every channel and both polarisations, WEIGHT_SPECTRUM ~ U(0.5, 1)
  per (row, channel, polarisation);
* ``array_center_x`` fixed at -2,565,018 m so the second principal axis has a pinned sign
  (SVDIntegratedImager.cu:555-559).

uvw and frequencies are float64.  The visibility sum runs on a CUDA device through torch when
one is present (it is 200 sources x 8.4 M samples at the SKA-Low-like size) and in chunked numpy
otherwise; the two agree to float32 rounding.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, Optional

import numpy as np

C_LIGHT = 299792458.0
SITE_LATITUDE_DEG = -26.82
ARRAY_CENTER_X = -2565018.0
BASE_SEED = 20261202


@dataclasses.dataclass(frozen=True)
class SnapshotConfig:
    name: str
    n_stations: int
    layout: str            # "disk" (uniform in a disk) or "lograd" (log-uniform radius)
    r_min: float           # metres (lograd only)
    r_max: float           # metres
    height_sigma: float    # metres
    n_channels: int
    f_lo: float            # Hz, first channel
    f_hi: float            # Hz, last channel (inclusive)
    n_sources: int
    flux: str              # "uniform" (U(0.1, 1) Jy) or "powerlaw" (N(>S) ~ S^-1.8)
    lm_max: float          # sources at |l|, |m| <= lm_max
    flag_fraction: float   # fraction of rows flagged
    image_size: int
    cell: float            # radians
    num_terms: int         # Chebyshev order quoted for the workload
    ha_hours: float = 0.0
    dec_deg: float = -30.0
    seed: int = BASE_SEED

    @property
    def num_rows(self) -> int:
        return self.n_stations * (self.n_stations - 1) // 2

    @property
    def num_samples(self) -> int:
        return self.num_rows * self.n_channels

    @property
    def grid_size(self) -> int:
        return (3 * self.image_size + 1) // 2


# Small plain-method case: 64 stations in a 2 km disk, 16 channels, 128^2 image.
SMALL = SnapshotConfig(
    name="small-64st-16ch-128px", n_stations=64, layout="disk", r_min=0.0, r_max=1000.0,
    height_sigma=5.0, n_channels=16, f_lo=100e6, f_hi=150e6, n_sources=5, flux="uniform",
    lm_max=0.02, flag_fraction=0.05, image_size=128, cell=1e-4, num_terms=4,
    seed=BASE_SEED + 0)

# SKA-Low-like snapshot: 512 log-radial stations within 35 km, 64 channels, 1024^2 image.
SKALOW = SnapshotConfig(
    name="skalow-512st-64ch-1024px", n_stations=512, layout="lograd", r_min=10.0,
    r_max=35000.0, height_sigma=20.0, n_channels=64, f_lo=50e6, f_hi=350e6, n_sources=200,
    flux="powerlaw", lm_max=0.01, flag_fraction=0.10, image_size=1024, cell=2e-5,
    num_terms=8, seed=BASE_SEED + 1)

# Precision form: the same array with 128 channels at 4096^2.
SKALOW_4K = dataclasses.replace(
    SKALOW, name="skalow-512st-128ch-4096px", n_channels=128, image_size=4096,
    lm_max=0.04, seed=BASE_SEED + 2)

CONFIGS = {0: SMALL, 1: SKALOW, 2: SKALOW_4K}


def station_enu(cfg: SnapshotConfig, rng: np.random.Generator) -> np.ndarray:
    theta = rng.uniform(0.0, 2.0 * math.pi, cfg.n_stations)
    if cfg.layout == "disk":
        radius = cfg.r_max * np.sqrt(rng.uniform(0.0, 1.0, cfg.n_stations))
    elif cfg.layout == "lograd":
        radius = np.exp(rng.uniform(math.log(cfg.r_min), math.log(cfg.r_max), cfg.n_stations))
    else:
        raise ValueError(f"unknown layout {cfg.layout!r}")
    east = radius * np.cos(theta)
    north = radius * np.sin(theta)
    up = rng.normal(0.0, cfg.height_sigma, cfg.n_stations)
    return np.stack([east, north, up], axis=1)


def enu_to_xyz(enu: np.ndarray, latitude_rad: float) -> np.ndarray:
    e, n, u = enu[:, 0], enu[:, 1], enu[:, 2]
    sl, cl = math.sin(latitude_rad), math.cos(latitude_rad)
    x = -sl * n + cl * u
    y = e
    z = cl * n + sl * u
    return np.stack([x, y, z], axis=1)


def baseline_uvw(xyz: np.ndarray, ha_rad: float, dec_rad: float) -> np.ndarray:
    n = xyz.shape[0]
    i, j = np.triu_indices(n, k=1)          # row order: (0,1), (0,2), ..., (1,2), ...
    b = xyz[j] - xyz[i]
    x, y, z = b[:, 0], b[:, 1], b[:, 2]
    sh, ch = math.sin(ha_rad), math.cos(ha_rad)
    sd, cd = math.sin(dec_rad), math.cos(dec_rad)
    u = sh * x + ch * y
    v = -sd * ch * x + sd * sh * y + cd * z
    w = cd * ch * x - cd * sh * y + sd * z
    return np.stack([u, v, w], axis=1)


def sky_model(cfg: SnapshotConfig, rng: np.random.Generator):
    l = rng.uniform(-cfg.lm_max, cfg.lm_max, cfg.n_sources)
    m = rng.uniform(-cfg.lm_max, cfg.lm_max, cfg.n_sources)
    if cfg.flux == "uniform":
        flux = rng.uniform(0.1, 1.0, cfg.n_sources)
    elif cfg.flux == "powerlaw":
        # N(>S) ~ S^-1.8 above 0.05 Jy, capped at 20 Jy.
        flux = np.minimum(0.05 * rng.uniform(0.0, 1.0, cfg.n_sources) ** (-1.0 / 1.8), 20.0)
    else:
        raise ValueError(f"unknown flux model {cfg.flux!r}")
    return l, m, flux


def _model_visibilities(uvw, freq, l, m, flux, device: Optional[str]):
    """sum_k S_k exp(-2 pi i (u l + v m + w (n-1)) nu / c), shape (C, R) complex128 -> c64."""
    n = np.sqrt(1.0 - l * l - m * m) - 1.0
    lmn = np.stack([l, m, n], axis=0)                     # (3, K)
    if device is not None:
        import torch
        dev = torch.device(device)
        uvw_t = torch.as_tensor(uvw, dtype=torch.float64, device=dev)
        lmn_t = torch.as_tensor(lmn, dtype=torch.float64, device=dev)
        flux_t = torch.as_tensor(flux, dtype=torch.float64, device=dev)
        delay = uvw_t @ lmn_t                              # (R, K) metres
        out = torch.empty((len(freq), uvw.shape[0]), dtype=torch.complex64, device=dev)
        for c, f in enumerate(freq):
            ph = delay * (-2.0 * math.pi * float(f) / C_LIGHT)
            out[c] = (torch.polar(flux_t.expand_as(ph), ph)).sum(dim=1).to(torch.complex64)
        return out.cpu().numpy()
    delay = uvw @ lmn                                      # (R, K)
    out = np.empty((len(freq), uvw.shape[0]), dtype=np.complex64)
    for c, f in enumerate(freq):
        ph = delay * (-2.0 * math.pi * f / C_LIGHT)
        out[c] = (flux[None, :] * np.exp(1j * ph)).sum(axis=1)
    return out


def make_snapshot(cfg: SnapshotConfig, ha_hours: Optional[float] = None,
                  device: Optional[str] = None, seed_offset: int = 0) -> Dict[str, np.ndarray]:
    """Build one snapshot.  Returns a dict with the HostMeasurementSetData fields.

    Arrays: ``uvw`` (R, 3) f64, ``freq`` (C,) f64, ``vis0_re`` ... ``vis3_im`` (S,) f32,
    ``flag0`` / ``flag3`` (S,) u8, ``weight0`` / ``weight3`` (S,) f32; scalars ``num_rows``,
    ``num_channels``, ``array_center_x``.  Sample index is ``row + chan * R``.
    """
    ss = np.random.SeedSequence(cfg.seed + seed_offset)
    geo, sky, noise, flags, weights = [np.random.default_rng(s) for s in ss.spawn(5)]
    ha = cfg.ha_hours if ha_hours is None else ha_hours
    enu = station_enu(cfg, geo)
    xyz = enu_to_xyz(enu, math.radians(SITE_LATITUDE_DEG))
    uvw = baseline_uvw(xyz, ha * math.pi / 12.0, math.radians(cfg.dec_deg))
    freq = np.linspace(cfg.f_lo, cfg.f_hi, cfg.n_channels)
    R, C = uvw.shape[0], cfg.n_channels
    S = R * C
    l, m, flux = sky_model(cfg, sky)
    model = _model_visibilities(uvw, freq, l, m, flux, device).reshape(S)

    out: Dict[str, np.ndarray] = {"uvw": uvw, "freq": freq}
    for pol in (0, 3):
        nz = noise.normal(0.0, 0.01, size=(2, S)).astype(np.float32)
        out[f"vis{pol}_re"] = (model.real + nz[0]).astype(np.float32)
        out[f"vis{pol}_im"] = (model.imag + nz[1]).astype(np.float32)
    n_flag = int(round(cfg.flag_fraction * R))
    flagged_rows = np.zeros(R, dtype=np.uint8)
    flagged_rows[flags.choice(R, size=n_flag, replace=False)] = 1
    flag = np.tile(flagged_rows, C)                        # idx = row + chan * R
    out["flag0"] = flag.copy()
    out["flag3"] = flag.copy()
    out["weight0"] = weights.uniform(0.5, 1.0, S).astype(np.float32)
    out["weight3"] = weights.uniform(0.5, 1.0, S).astype(np.float32)
    out["num_rows"] = R
    out["num_channels"] = C
    out["array_center_x"] = ARRAY_CENTER_X
    out["sources"] = np.stack([l, m, flux], axis=1)
    return out


# Batching form: the SKA-Low array at hour angles spread over +-4 h.  Snapshot i of n keeps a
# fraction 1 - f of its rows, f ~ U(0, 0.6) (flagged rows removed, so the row counts are uneven:
# 3.3 M to 8.4 M samples), and is imaged at a size drawn from {512, 1024, 2048}.
BATCHING_SIZES = (512, 1024, 2048)
SKALOW_BATCHING = dataclasses.replace(
    SKALOW, name="skalow-512st-64ch-mixed", flag_fraction=0.0, seed=BASE_SEED + 3)
PER_SAMPLE_FIELDS = ("vis0_re", "vis0_im", "vis3_re", "vis3_im", "flag0", "flag3", "weight0",
                     "weight3")


def batching_draw(index: int, count: int, cfg: SnapshotConfig = SKALOW_BATCHING,
                  sizes=BATCHING_SIZES, max_removed: float = 0.6):
    """(hour angle in hours, fraction of rows removed, image size) of snapshot `index` of
    `count`: hour angles evenly over [-4 h, +4 h], the rest drawn from a stream seeded by
    (cfg.seed, index)."""
    rng = np.random.default_rng([cfg.seed, index])
    ha = -4.0 + 8.0 * index / (count - 1) if count > 1 else 0.0
    removed = float(rng.uniform(0.0, max_removed))
    size = int(sizes[int(rng.integers(0, len(sizes)))])
    return ha, removed, size


def drop_rows(snap: Dict, keep: np.ndarray) -> Dict:
    """The snapshot restricted to rows `keep` (ascending), in the same row + chan * R layout."""
    R, C = int(snap["num_rows"]), int(snap["num_channels"])
    out = dict(snap)
    out["uvw"] = np.ascontiguousarray(snap["uvw"][keep])
    for k in PER_SAMPLE_FIELDS:
        if snap.get(k) is not None:
            out[k] = np.ascontiguousarray(snap[k].reshape(C, R)[:, keep].reshape(-1))
    out["num_rows"] = int(len(keep))
    return out


def make_batching_snapshot(index: int, count: int, cfg: SnapshotConfig = SKALOW_BATCHING,
                           sizes=BATCHING_SIZES, device: Optional[str] = None) -> Dict:
    """Snapshot `index` of `count` of the batching form, with its image size under
    ``"image_size"``."""
    ha, removed, size = batching_draw(index, count, cfg, sizes)
    snap = make_snapshot(cfg, ha_hours=ha, device=device, seed_offset=index)
    R = int(snap["num_rows"])
    rng = np.random.default_rng([cfg.seed, index, 1])
    keep = np.sort(rng.choice(R, size=int(round(R * (1.0 - removed))), replace=False))
    out = drop_rows(snap, keep)
    out["image_size"] = size
    return out
