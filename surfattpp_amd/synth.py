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
        if k in batch:
            out[k] = batch[k][ev_ids].copy()
    out["ev_rec_off"] = new_off
    out["ev_nrec"] = nrecs.astype(np.int32).copy()
    for k in ("rec_idx", "rec_lon", "rec_lat", "tt_obs", "weight"):
        if k in batch:
            out[k] = batch[k][sel].copy()
    return out


# ----------------------------------------------------------------------------
# Topography and its metric (an input producer of the path: the reference
# builds it on the host once per period, src/surf_grid.cpp:80-163)
# ----------------------------------------------------------------------------
RAD2DEG = 180.0 / PI


def gaussian_edifices(rng: Rng, lon: np.ndarray, lat: np.ndarray, n: int = 8, h_km=(0.5, 4.0),
                      w_km=(10.0, 40.0)) -> np.ndarray:
    """Surface relief z (ny, nx) in km: a sum of n Gaussian edifices
    h*exp(-r^2/(2 w^2)) with h in U[h_km], w in U[w_km] and centres uniform over
    the grid; r is the local Cartesian distance in km (the grid is centred on
    (0, 0), so lon and lat map to km by R_EARTH*DEG2RAD)."""
    u = rng.uniform(4 * n).reshape(4, n)
    h = h_km[0] + (h_km[1] - h_km[0]) * u[0]
    w = w_km[0] + (w_km[1] - w_km[0]) * u[1]
    cx = lon[0] + (lon[-1] - lon[0]) * u[2]
    cy = lat[0] + (lat[-1] - lat[0]) * u[3]
    km = R_EARTH * DEG2RAD
    X = lon[None, :] * km
    Y = lat[:, None] * km
    z = np.zeros((lat.size, lon.size))
    for k in range(n):
        r2 = (X - cx[k] * km) ** 2 + (Y - cy[k] * km) ** 2
        z += h[k] * np.exp(-r2 / (2.0 * w[k] ** 2))
    return z


def gaussian_smooth_geo(z: np.ndarray, lon: np.ndarray, lat: np.ndarray, sigma_deg: float) -> np.ndarray:
    """Great-circle Gaussian smoothing of z (ny, nx), the operator of
    include/utils.h:505-573 (gaussian_smooth_geo_2): for node (i, j) the window
    is |m - i| <= hx = int(3 sigma / dlon) and |n - j| <= hy_j =
    int(3 sigma / cos(lat_j) / dlat), clipped to the grid; the weight is
    exp(-delta^2 / (2 sigma^2)) with delta = 2*atan(sqrt(a/(1-a))) in degrees and
    a = sin^2(dlat/2) + cos(lat_j) cos(lat_n) sin^2(dlon/2) (haversine, a <= 1);
    the result is the weighted mean over the window.

    The grid is uniform in longitude, so the weights depend on m - i only and
    each output row j is a sum over window rows n of 1-D convolutions along
    longitude, evaluated with FFTs (agrees with the direct sum to rounding)."""
    ny, nx = z.shape
    s3 = 3.0 * sigma_deg
    dx, dy = lon[1] - lon[0], lat[1] - lat[0]
    hx = int(s3 / dx)
    L = 1 << int(math.ceil(math.log2(nx + 2 * hx + 1)))
    dm = np.arange(-hx, hx + 1)
    s2lon = np.sin(dm * dx * (DEG2RAD * 0.5)) ** 2
    Zf = np.fft.rfft(z, n=L, axis=1)                       # rows n, z at [0, nx)
    Of = np.fft.rfft(np.ones(nx), n=L)                     # the clipped window's support
    lat_r = lat * DEG2RAD
    inv2s2 = 1.0 / (2.0 * sigma_deg * sigma_deg)
    out = np.empty_like(z)
    idx = (-dm) % L                                        # kernel kr[t] = K[-t] at t mod L
    for j in range(ny):
        c0 = math.cos(lat_r[j])
        hy = int(s3 / c0 / dy)
        n1, n2 = max(0, j - hy), min(ny - 1, j + hy)
        s2lat = np.sin((lat[n1:n2 + 1] - lat[j]) * (DEG2RAD * 0.5)) ** 2
        c1 = np.cos(lat_r[n1:n2 + 1])
        a = s2lat[:, None] + (c0 * c1)[:, None] * s2lon[None, :]
        a = np.minimum(a, 1.0)
        w = np.exp(-((2.0 * RAD2DEG) * np.arctan(np.sqrt(a / (1.0 - a)))) ** 2 * inv2s2)
        kr = np.zeros((n2 - n1 + 1, L))
        kr[:, idx] = w
        Kf = np.fft.rfft(kr, axis=1)
        num = np.fft.irfft((Kf * Zf[n1:n2 + 1]).sum(axis=0), n=L)[:nx]
        den = np.fft.irfft(Kf.sum(axis=0) * Of, n=L)[:nx]
        out[j] = num / den
    return out


def gradient_geo(f: np.ndarray, lon: np.ndarray, lat: np.ndarray):
    """Geographic gradient (df/dx, df/dy) in f-units per km of f (ny, nx), the
    differences of include/utils.h:595-649 (gradient_2_geo): central in the
    interior, one-sided at the edges; dx_km = R*dlon*DEG2RAD*cos(lat),
    dy_km = R*dlat*DEG2RAD."""
    ny, nx = f.shape
    dx1 = np.zeros(nx)
    dy1 = np.zeros(ny)
    dx1[1:] = (lon[1:] - lon[:-1]) * (R_EARTH * DEG2RAD)
    dy1[1:] = (lat[1:] - lat[:-1]) * (R_EARTH * DEG2RAD)
    dxm = np.cos(lat * DEG2RAD)[:, None] * dx1[None, :]   # (ny, nx)
    dym = np.broadcast_to(dy1[:, None], (ny, nx))
    tx = np.empty_like(f)
    ty = np.empty_like(f)
    tx[:, 1:-1] = (f[:, 2:] - f[:, :-2]) / (2.0 * dxm[:, 1:-1])
    tx[:, 0] = (f[:, 1] - f[:, 0]) / dxm[:, 1]
    tx[:, -1] = (f[:, -1] - f[:, -2]) / dxm[:, -1]
    ty[1:-1, :] = (f[2:, :] - f[:-2, :]) / (2.0 * dym[1:-1, :])
    ty[0, :] = (f[1, :] - f[0, :]) / dym[1, :]
    ty[-1, :] = (f[-1, :] - f[-2, :]) / dym[-1, :]
    return tx, ty


def topo_metric(z: np.ndarray, lon: np.ndarray, lat: np.ndarray, sigma_deg: float | None):
    """Per-period metric planes (m11, m12, m22) of the surface z (ny, nx, km),
    src/surf_grid.cpp:115-145 and :184-190: z smoothed with sigma_deg (None: not
    smoothed), fx, fy its geographic gradient, D = 1 + fx^2 + fy^2,
    a = (1 + fy^2)/D, b = (1 + fx^2)/D, c = fx*fy/D; m11 = a, m22 = b, m12 = -c."""
    zs = z if sigma_deg is None else gaussian_smooth_geo(z, lon, lat, sigma_deg)
    fx, fy = gradient_geo(zs, lon, lat)
    fx2, fy2 = fx * fx, fy * fy
    den = 1.0 + fx2 + fy2
    return (1.0 + fy2) / den, -(fx * fy / den), (1.0 + fx2) / den


def periods_log(n: int, t0: float = 4.0, t1: float = 40.0) -> np.ndarray:
    """n periods log-spaced from t0 to t1 s."""
    return t0 * (t1 / t0) ** (np.arange(n) / max(n - 1, 1))


def mean_phase_velocity(periods: np.ndarray) -> np.ndarray:
    """A smooth Rayleigh-like dispersion curve (km/s), 2.9 at 4 s rising to 3.8 at
    40 s, linear in log-period; stands in for surfdisp96's average velocity."""
    x = np.log(periods / 4.0) / np.log(10.0)
    return 2.9 + 0.9 * np.clip(x, 0.0, 1.0)


def config_topo_grid(seed: int, n: int, spacing_km: float = 1.5, nsta: int | None = None, nperiod: int = 12,
                     wavelen_factor: float = 2.5, smooth: bool = True):
    """One grid of the topography batch: n x n nodes at spacing_km, 8 seeded Gaussian
    edifices (0.5-4 km high, 10-40 km wide) whose per-period smoothed slopes give the
    anisotropic metric (topo_metric, sigma = v(T)*T*wavelen_factor/(2 pi) km,
    src/surf_grid.cpp:115-118); velocity per period is the mean dispersion velocity
    v(T) times the 2x3-cell +-6 % checkerboard; nsta (default: seeded in 10..60)
    stations, every one a virtual source; observed times are great-circle times at
    v(T)."""
    rng = Rng(seed)
    d = km_to_deg(spacing_km)
    lon, lat = centred_axis(n, d), centred_axis(n, d)
    if nsta is None:
        nsta = 10 + int(rng.uniform(1)[0] * 51)
    sta_lon, sta_lat = interior_stations(rng, nsta, lon, lat)
    z = gaussian_edifices(rng, lon, lat)
    periods = periods_log(nperiod)
    vbar = mean_phase_velocity(periods)
    cb = checkerboard(n, n, 1.0, 0.06, 2, 3)
    svel = np.stack([vbar[p] * cb for p in range(nperiod)])
    m11 = np.empty_like(svel)
    m12 = np.empty_like(svel)
    m22 = np.empty_like(svel)
    for p in range(nperiod):
        sig = km_to_deg(vbar[p] * periods[p] * wavelen_factor / (2.0 * PI)) if smooth else None
        m11[p], m12[p], m22[p] = topo_metric(z, lon, lat, sig)
    ev_src, ev_iper, rec_idx, off, nrec = all_pairs_events(sta_lon, sta_lat, nperiod)
    tt_obs = great_circle_time(sta_lon[rec_idx], sta_lat[rec_idx], sta_lon[ev_src].repeat(nrec),
                               sta_lat[ev_src].repeat(nrec), vbar[ev_iper].repeat(nrec))
    weight = np.ones(rec_idx.size)
    meta = dict(name=f"topo-{n}x{n}", spacing_km=spacing_km, nsta=nsta, nperiod=nperiod, seed=seed,
                periods=periods.tolist(), vbar=vbar.tolist(), wavelen_factor=wavelen_factor)
    b = _assemble(lon, lat, svel, m11, m12, m22, sta_lon, sta_lat, ev_src, ev_iper, rec_idx, off, nrec,
                  tt_obs, weight, meta)
    b["z"] = z
    return b


TOPO_SIZES = (101, 161, 241, 321)


def config_topo(seed: int = 30, sizes=TOPO_SIZES, nperiod: int = 12, **kw):
    """The uneven topography batch: one grid per size (1.5 km spacing), seeds
    seed+k, 10-60 stations each, nperiod periods; a list of batch dicts."""
    return [config_topo_grid(seed + k, n, nperiod=nperiod, **kw) for k, n in enumerate(sizes)]


def config_large(seed: int = 5, nx: int = 401, ny: int = 481, nsta: int = 400, nperiod: int = 30, **kw):
    """401x481 nodes at 1 km, 400 stations x 30 periods: 12,000 forward+adjoint pairs
    with 399 receivers each; velocity as config_grf (3.2 km/s, independent +-10 %
    GRF of 15 km correlation per period)."""
    b = config_grf(seed=seed, nx=nx, ny=ny, nsta=nsta, nperiod=nperiod, **kw)
    b["meta"]["name"] = f"large-{nx}x{ny}"
    return b


def layered_model(nx: int, ny: int, rng: Rng, spacing_km: float = 2.0, amp: float = 0.05, corr_km: float = 30.0,
                  nz: int = 30, z_max: float = 60.0, vs_top: float = 2.8, vs_bottom: float = 4.6):
    """The 121x141 workload's 3-D model: nz depth nodes over 0..z_max km, Vs
    linear vs_top -> vs_bottom km/s times (1 + a seeded +-amp lateral GRF of
    corr_km correlation), Vp = 1.73 Vs, Brocher (2005) density of Vp
    (include/utils.h:223-225).  Returns (zgrids, vs3d, vp3d, rho3d), the 3-D
    arrays (nx, ny, nz) like the reference's RowMajor model tensors."""
    from .surfdisp import vp2rho
    dep = np.linspace(0.0, z_max, nz)
    d = grf(rng, nx, ny, spacing_km, corr_km, amp)          # (ny, nx)
    vs = np.linspace(vs_top, vs_bottom, nz)[None, None, :] * (1.0 + d.T[:, :, None])
    vp = 1.73 * vs
    return dep, np.ascontiguousarray(vs), np.ascontiguousarray(vp), np.ascontiguousarray(vp2rho(vp))


def config_dispersion(seed: int = 2, nx: int = 121, ny: int = 141, spacing_km: float = 2.0, nperiod: int = 20,
                      amp: float = 0.05, corr_km: float = 30.0):
    """121x141 nodes at 2 km, the 25 stations of config_checkerboard, 20 periods
    log-spaced 4-40 s (500 forward+adjoint pairs), with per-period Rayleigh
    phase-velocity maps computed on the GPU by surfdisp96 from the layered
    model of layered_model (spherical earth, fundamental mode, as
    SurfGrid::fwdsurf calls it, src/surf_grid.cpp:194-245).  Observed times:
    great-circle times at each period's mean map velocity."""
    from .surfdisp import fwdsurf
    rng = Rng(seed)
    d = km_to_deg(spacing_km)
    lon, lat = centred_axis(nx, d), centred_axis(ny, d)
    cb = config_checkerboard()
    sta_lon, sta_lat = cb["sta_lon"], cb["sta_lat"]
    periods = periods_log(nperiod)
    dep, vs3d, vp3d, rho3d = layered_model(nx, ny, rng, spacing_km, amp, corr_km)
    svel = fwdsurf(dep, vs3d, periods, vp3d=vp3d, rho3d=rho3d, iflsph=1, iwave=2, mode=1, igr=0)
    vbar = svel.reshape(nperiod, -1).mean(axis=1)
    ones = np.ones((nperiod, ny, nx))
    ev_src, ev_iper, rec_idx, off, nrec = all_pairs_events(sta_lon, sta_lat, nperiod)
    tt_obs = great_circle_time(sta_lon[rec_idx], sta_lat[rec_idx], sta_lon[ev_src].repeat(nrec),
                               sta_lat[ev_src].repeat(nrec), vbar[ev_iper].repeat(nrec))
    weight = np.ones(rec_idx.size)
    meta = dict(name=f"disp-{nx}x{ny}", spacing_km=spacing_km, nsta=len(sta_lon), nperiod=nperiod,
                periods=periods.tolist(), vbar=vbar.tolist(), amp=amp, corr_km=corr_km, seed=seed,
                model="30 nodes 0-60 km, Vs 2.8-4.6 km/s, Vp 1.73 Vs, Brocher density, +-5 % GRF 30 km")
    b = _assemble(lon, lat, svel, ones, 0 * ones, ones, sta_lon, sta_lat, ev_src, ev_iper, rec_idx, off,
                  nrec, tt_obs, weight, meta)
    b["model"] = (dep, vs3d, vp3d, rho3d)
    return b


def config_dispersion_shape(seed: int = 2, nx: int = 121, ny: int = 141, spacing_km: float = 2.0,
                            nperiod: int = 20, amp: float = 0.05, corr_km: float = 30.0):
    """121x141 nodes at 2 km with the 25 stations of config_checkerboard (same seed,
    same degrees) and 20 periods log-spaced 4-40 s: 500 forward+adjoint pairs.
    The per-period phase-velocity maps are the mean dispersion curve
    (mean_phase_velocity) times (1 + a seeded +-5 % GRF of 30 km correlation),
    one realisation shared by all periods: a host-only stand-in for
    config_dispersion's GPU-computed maps, for runs without a GPU.
    Observed times: great-circle times at the period's mean velocity."""
    rng = Rng(seed)
    d = km_to_deg(spacing_km)
    lon, lat = centred_axis(nx, d), centred_axis(ny, d)
    cb = config_checkerboard()
    sta_lon, sta_lat = cb["sta_lon"], cb["sta_lat"]
    periods = periods_log(nperiod)
    vbar = mean_phase_velocity(periods)
    pert = grf(rng, nx, ny, spacing_km, corr_km, amp)
    svel = np.stack([vbar[p] * (1.0 + pert) for p in range(nperiod)])
    ones = np.ones((nperiod, ny, nx))
    ev_src, ev_iper, rec_idx, off, nrec = all_pairs_events(sta_lon, sta_lat, nperiod)
    tt_obs = great_circle_time(sta_lon[rec_idx], sta_lat[rec_idx], sta_lon[ev_src].repeat(nrec),
                               sta_lat[ev_src].repeat(nrec), vbar[ev_iper].repeat(nrec))
    weight = np.ones(rec_idx.size)
    meta = dict(name=f"disp-{nx}x{ny}", spacing_km=spacing_km, nsta=len(sta_lon), nperiod=nperiod,
                periods=periods.tolist(), vbar=vbar.tolist(), amp=amp, corr_km=corr_km, seed=seed)
    return _assemble(lon, lat, svel, ones, 0 * ones, ones, sta_lon, sta_lat, ev_src, ev_iper, rec_idx, off,
                     nrec, tt_obs, weight, meta)
