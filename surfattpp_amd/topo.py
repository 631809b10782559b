"""The topography metric producer on the GPU: the per-period anisotropic
metric planes of a surface (SurfGrid::build_media_matrix_with_topo,
src/surf_grid.cpp:80-145, :184-190) through ``sat_topo_metric``.

``topo_metric(z, lon, lat, sigma_deg)`` returns (m11, m12, m22), each
(nperiod, ny, nx) -- the spha, sphc, sphb planes an EikonalBatch takes.
``sigma_for_periods`` gives the reference's smoothing width per period.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib

R_EARTH = 6371.0
PI = 3.14159265358979323846264338327950288
DEG2RAD = PI / 180.0


def sigma_for_periods(mean_vel, periods, wavelen_factor: float = 2.5) -> np.ndarray:
    """Smoothing width in degrees per period: v(T)*T*wavelen_factor/(2 pi) km
    over R_EARTH*DEG2RAD km per degree (src/surf_grid.cpp:115-118)."""
    v = np.asarray(mean_vel, np.float64)
    t = np.asarray(periods, np.float64)
    return (v * t * wavelen_factor / (2.0 * PI)) / (R_EARTH * DEG2RAD)


def topo_metric(z, lon, lat, sigma_deg, return_smoothed: bool = False, timing: dict | None = None):
    """Metric planes of surface z (ny, nx; km) for each sigma in sigma_deg
    (degrees; <= 0: not smoothed).  Returns (m11, m12, m22) of shape
    (nperiod, ny, nx), plus the smoothed surfaces with return_smoothed.
    ``timing``, if given, receives the device kernels' milliseconds."""
    z = np.ascontiguousarray(z, dtype=np.float64)
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    sig = np.ascontiguousarray(np.atleast_1d(sigma_deg), dtype=np.float64)
    ny, nx = z.shape
    if lon.size != nx or lat.size != ny:
        raise ValueError(f"z is {z.shape}, axes ({lat.size}, {lon.size})")
    if nx < 3 or ny < 3:
        raise ValueError("Grid must be at least 3x3")
    nper = sig.size
    m11, m12, m22 = (np.empty((nper, ny, nx)) for _ in range(3))
    zs = np.empty((nper, ny, nx)) if return_smoothed else None
    ms = C.c_double(0.0)
    rc = _lib.load().sat_topo_metric(nx, ny, nper, lon.ctypes.data, lat.ctypes.data, z.ctypes.data,
                                     sig.ctypes.data, m11.ctypes.data, m12.ctypes.data, m22.ctypes.data,
                                     zs.ctypes.data if zs is not None else None, C.byref(ms))
    _lib.check("sat_topo_metric", rc)
    if timing is not None:
        timing["kernel_ms"] = ms.value
    return (m11, m12, m22, zs) if return_smoothed else (m11, m12, m22)


def window_nodes(nx: int, ny: int, lon, lat, sigma_deg) -> int:
    """Window nodes the smoothing visits for one period (sum over output nodes
    of the clipped window sizes): the work count behind its throughput."""
    dx, dy = lon[1] - lon[0], lat[1] - lat[0]
    if not sigma_deg > 0:
        return nx * ny
    hx = min(int(3.0 * sigma_deg / dx), nx - 1)
    i = np.arange(nx)
    wx = (np.minimum(nx - 1, i + hx) - np.maximum(0, i - hx) + 1).sum()
    total = 0
    for j in range(ny):
        hy = int(3.0 * sigma_deg / math.cos(lat[j] * DEG2RAD) / dy)
        total += (min(ny - 1, j + hy) - max(0, j - hy) + 1) * wx
    return int(total)
