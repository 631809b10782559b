"""Seeded synthetic workloads for the eikonal forward + adjoint path.

Every random number comes from numpy's PCG64 raw 64-bit stream mapped to
[0, 1) by this module (``(u >> 11) * 2**-53``) and Box-Muller normals, so the
inputs do not depend on numpy's distribution implementations.

Coordinates follow the reference's rotated frame (include/sph2loc.h:140-166):
the study area is centred on (0 deg, 0 deg) and "k km spacing" means
``k / (R_EARTH * pi / 180)`` degrees in both lon and lat.

A batch is a plain dict (host numpy arrays) with the layout the C-ABI takes:
  lon (nx,), lat (ny,)                          degrees, ascending, uniform
  svel, m11, m12, m22 (nperiod, ny, nx)         period-major planes, ix fastest
  ev_iper, ev_lon, ev_lat, ev_rec_off, ev_nrec  one entry per (source, period)
  rec_lon, rec_lat, tt_obs, weight              flattened receivers
which is src/preproc.cpp:71-148's input with the reference's [ix][iy][iper]
planes (include/surf_grid.h:63-65) transposed to period-major.
"""
from __future__ import annotations

import math

import numpy as np

R_EARTH = 6371.0
PI = 3.14159265358979323846264338327950288
DEG2RAD = PI / 180.0


def km_to_deg(km: float) -> float:
    return km / (R_EARTH * DEG2RAD)


class Rng:
    """Portable uniform/normal stream over PCG64's raw 64-bit output."""

    def __init__(self, seed: int):
        self._bg = np.random.PCG64(seed)

    def uniform(self, n: int) -> np.ndarray:
        raw = self._bg.random_raw(n).astype(np.uint64)
        return (raw >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        u1 = np.where(u1 <= 0.0, 2.0 ** -53, u1)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * PI * u2), r * np.sin(2 * PI * u2)])
        return z[:n]


def centred_axis(n: int, spacing_deg: float) -> np.ndarray:
    half = (n - 1) / 2.0
    return (np.arange(n, dtype=np.float64) - half) * spacing_deg


def grf(rng: Rng, nx: int, ny: int, spacing_km: float, corr_km: float, amp: float) -> np.ndarray:
    """Gaussian random field (ny, nx), Gaussian covariance of correlation length
    ``corr_km``, scaled so that max |field| == amp.  Periodic FFT smoothing of
    white noise on a padded grid (padding >= 3 correlation lengths)."""
    pad = int(math.ceil(3.0 * corr_km / spacing_km))
    NX, NY = nx + 2 * pad, ny + 2 * pad
    w = rng.normal(NX * NY).reshape(NY, NX)
    kx = np.fft.fftfreq(NX, d=spacing_km) * 2 * PI
    ky = np.fft.fftfreq(NY, d=spacing_km) * 2 * PI
    # kernel exp(-r^2/corr^2) has spectrum ~ exp(-k^2 corr^2 / 4)
    filt = np.exp(-(ky[:, None] ** 2 + kx[None, :] ** 2) * corr_km ** 2 / 8.0)
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))[pad:pad + ny, pad:pad + nx]
    f = f - f.mean()
    return f * (amp / np.max(np.abs(f)))


def interior_stations(rng: Rng, n: int, lon: np.ndarray, lat: np.ndarray, margin_nodes: int = 3):
    """n stations uniform in the interior, at least ``margin_nodes`` nodes from every
    edge (a source within one cell of the boundary is invalid,
    src/eikonal_solver.cpp:123-125)."""
    dx, dy = lon[1] - lon[0], lat[1] - lat[0]
    x0, x1 = lon[0] + margin_nodes * dx, lon[-1] - margin_nodes * dx
    y0, y1 = lat[0] + margin_nodes * dy, lat[-1] - margin_nodes * dy
    u = rng.uniform(2 * n)
    return x0 + (x1 - x0) * u[:n], y0 + (y1 - y0) * u[n:]


def great_circle_time(lon1, lat1, lon0, lat0, vel):
    """Analytic travel time in a homogeneous medium (the factor T0 of
    src/eikonal_solver.cpp:99-102 with fun0 = 1/vel)."""
    l1, p1, l0, p0 = (np.asarray(v, dtype=np.float64) * DEG2RAD for v in (lon1, lat1, lon0, lat0))
    tmp = np.sin(p1) * np.sin(p0) + np.cos(p1) * np.cos(p0) * np.cos(l1 - l0)
    return np.arccos(np.clip(tmp, -1.0, 1.0)) * R_EARTH / vel


def all_pairs_events(sta_lon, sta_lat, nperiod: int, periods_outer: bool = False):
    """Every station is a virtual source, every other station a receiver, for every
    period.  Event order is (source, period) with period fastest, the order of the
    reference's std::map keys "{evtname}_{iper}" for <=10 periods
    (src/src_rec.cpp:149-171)."""
    ns = len(sta_lon)
    ev_src, ev_iper, rec_idx, off, nrec = [], [], [], [], []
    pos = 0
    order = ([(s, p) for p in range(nperiod) for s in range(ns)] if periods_outer
             else [(s, p) for s in range(ns) for p in range(nperiod)])
    for s, p in order:
        recs = [r for r in range(ns) if r != s]
        ev_src.append(s)
        ev_iper.append(p)
        off.append(pos)
        nrec.append(len(recs))
        rec_idx.extend(recs)
        pos += len(recs)
    return (np.array(ev_src, np.int32), np.array(ev_iper, np.int32), np.array(rec_idx, np.int32),
            np.array(off, np.int32), np.array(nrec, np.int32))


def _assemble(lon, lat, svel, m11, m12, m22, sta_lon, sta_lat, ev_src, ev_iper, rec_idx, off, nrec,
              tt_obs, weight, meta):
    return dict(
        lon=np.ascontiguousarray(lon), lat=np.ascontiguousarray(lat),
        svel=np.ascontiguousarray(svel), m11=np.ascontiguousarray(m11),
        m12=np.ascontiguousarray(m12), m22=np.ascontiguousarray(m22),
        ev_iper=ev_iper, ev_lon=sta_lon[ev_src].copy(), ev_lat=sta_lat[ev_src].copy(),
        ev_src=ev_src, ev_rec_off=off, ev_nrec=nrec,
        rec_idx=rec_idx, rec_lon=sta_lon[rec_idx].copy(), rec_lat=sta_lat[rec_idx].copy(),
        tt_obs=np.ascontiguousarray(tt_obs, dtype=np.float64),
        weight=np.ascontiguousarray(weight, dtype=np.float64),
        sta_lon=sta_lon, sta_lat=sta_lat, meta=meta)


def checkerboard(nx: int, ny: int, v0: float, amp: float, nx_w: int, ny_w: int) -> np.ndarray:
    """src/model_grid.cpp:523-554 perturbation pattern with hmargin = 0:
    v = v0 * (1 + amp * sin(nx_w*pi*i/nx) * sin(ny_w*pi*j/ny)), returned (ny, nx)."""
    xp = np.sin(nx_w * PI * np.arange(nx) / nx)
    yp = np.sin(ny_w * PI * np.arange(ny) / ny)
    return v0 * (1.0 + amp * yp[:, None] * xp[None, :])


def config_checkerboard(seed: int = 1, nx: int = 61, ny: int = 71, spacing_km: float = 2.0,
                        nsta: int = 25, v0: float = 3.0, tt_obs=None):
    """61x71 nodes at 2 km, 25 stations, one period (10 s).  Predicted times come from
    the 3.0 km/s background (the model of this batch); observed times are passed in
    (``tt_obs``, e.g. forward times in :func:`checkerboard`) or, when None, filled
    with the checkerboard's straight-ray estimate so the batch is self-contained."""
    rng = Rng(seed)
    d = km_to_deg(spacing_km)
    lon, lat = centred_axis(nx, d), centred_axis(ny, d)
    sta_lon, sta_lat = interior_stations(rng, nsta, lon, lat)
    ones = np.ones((1, ny, nx))
    svel = v0 * ones
    ev_src, ev_iper, rec_idx, off, nrec = all_pairs_events(sta_lon, sta_lat, 1)
    if tt_obs is None:
        tt_obs = great_circle_time(sta_lon[rec_idx], sta_lat[rec_idx], sta_lon[ev_src].repeat(nrec),
                                   sta_lat[ev_src].repeat(nrec), v0 * 1.03)
    weight = np.ones(rec_idx.size)
    meta = dict(name="checkerboard-61x71", spacing_km=spacing_km, nsta=nsta, periods=[10.0])
    return _assemble(lon, lat, svel, ones, 0 * ones, ones, sta_lon, sta_lat, ev_src, ev_iper, rec_idx, off,
                     nrec, tt_obs, weight, meta)


def config_grf(seed: int = 4, nx: int = 201, ny: int = 201, spacing_km: float = 1.0, nsta: int = 100,
               nperiod: int = 30, v0: float = 3.2, amp: float = 0.10, corr_km: float = 15.0):
    """201x201 nodes at 1 km, 100 stations x 30 periods (3,000 forward+adjoint pairs,
    99 receivers each).  Velocity per period is v0*(1+delta) with delta an independent
    seeded Gaussian random field (max |delta| = amp, correlation corr_km).  Observed
    times are the analytic great-circle times at v0 (the homogeneous background), so
    residuals carry the GRF's imprint."""
    rng = Rng(seed)
    d = km_to_deg(spacing_km)
    lon, lat = centred_axis(nx, d), centred_axis(ny, d)
    sta_lon, sta_lat = interior_stations(rng, nsta, lon, lat)
    svel = np.stack([v0 * (1.0 + grf(rng, nx, ny, spacing_km, corr_km, amp)) for _ in range(nperiod)])
    ones = np.ones((nperiod, ny, nx))
    ev_src, ev_iper, rec_idx, off, nrec = all_pairs_events(sta_lon, sta_lat, nperiod)
    tt_obs = great_circle_time(sta_lon[rec_idx], sta_lat[rec_idx], sta_lon[ev_src].repeat(nrec),
                               sta_lat[ev_src].repeat(nrec), v0)
    weight = np.ones(rec_idx.size)
    meta = dict(name=f"grf-{nx}x{ny}", spacing_km=spacing_km, nsta=nsta, nperiod=nperiod, v0=v0,
                amp=amp, corr_km=corr_km, seed=seed)
    return _assemble(lon, lat, svel, ones, 0 * ones, ones, sta_lon, sta_lat, ev_src, ev_iper, rec_idx, off,
                     nrec, tt_obs, weight, meta)


def subset_events(batch: dict, ev_ids) -> dict:
    """A batch holding only the events ``ev_ids`` (same grid and planes)."""
    ev_ids = np.asarray(ev_ids, dtype=np.int64)
    offs, nrecs = batch["ev_rec_off"][ev_ids], batch["ev_nrec"][ev_ids]
    sel = np.concatenate([np.arange(o, o + n) for o, n in zip(offs, nrecs)]) if ev_ids.size else np.zeros(0, int)
    new_off = np.concatenate([[0], np.cumsum(nrecs)[:-1]]).astype(np.int32) if ev_ids.size else np.zeros(0, np.int32)
    out = dict(batch)
    for k in ("ev_iper", "ev_lon", "ev_lat", "ev_src"):
        out[k] = batch[k][ev_ids].copy()
    out["ev_rec_off"] = new_off
    out["ev_nrec"] = nrecs.astype(np.int32).copy()
    for k in ("rec_idx", "rec_lon", "rec_lat", "tt_obs", "weight"):
        out[k] = batch[k][sel].copy()
    return out
