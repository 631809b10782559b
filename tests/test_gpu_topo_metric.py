"""The topography metric producer on the GPU (sat_topo_metric,
surfattpp_amd.topo) against the host restatement of
src/surf_grid.cpp:109-145 (surfattpp_amd.synth.topo_metric, itself pinned by
tests/test_topo_metric.py) and the direct window sum written from the
reference's per-node loop (include/utils.h:505-573).

Tolerances (normalised infinity norm): smoothed surface 1e-12 (summation
order and the rounding of lon(m) - lon(i) differ); metric planes 1e-11 (the
central differences divide rounding of the smoothed surface by neighbour
differences ~100x smaller than the surface).
"""
import time

import numpy as np
import pytest

from surfattpp_amd import _lib
from surfattpp_amd import synth as S
from surfattpp_amd import topo as TG
from tests.test_topo_metric import _grid, _smooth_direct

pytestmark = pytest.mark.gpu


def rel(a, r):
    return float(np.abs(a - r).max() / np.abs(r).max())


def test_smoothing_matches_direct_window_sum():
    lon, lat = _grid(23, 19, lat0=1.7)
    z = S.Rng(7).uniform(19 * 23).reshape(19, 23) * 3.0
    sig = [S.km_to_deg(2.0), S.km_to_deg(4.5), S.km_to_deg(40.0)]  # the last: window beyond the grid
    m11, m12, m22, zs = TG.topo_metric(z, lon, lat, sig, return_smoothed=True)
    for p, s in enumerate(sig):
        ref = _smooth_direct(z, lon, lat, s)
        e = rel(zs[p], ref)
        print(f"sigma {s:.5f} deg: smoothed rel err {e:.2e}", flush=True)
        assert e <= 1e-12


def test_unsmoothed_and_flat_cases():
    lon, lat = _grid(30, 25)
    z = S.Rng(3).uniform(25 * 30).reshape(25, 30)
    m11, m12, m22, zs = TG.topo_metric(z, lon, lat, [0.0], return_smoothed=True)
    assert np.array_equal(zs[0], z)
    h11, h12, h22 = S.topo_metric(z, lon, lat, None)
    assert rel(m11[0], h11) <= 1e-13 and rel(m22[0], h22) <= 1e-13 and rel(m12[0], h12) <= 1e-12
    f11, f12, f22 = TG.topo_metric(np.full((25, 30), 1.2), lon, lat, [S.km_to_deg(5.0)])
    assert np.abs(f11 - 1.0).max() <= 1e-14 and np.abs(f22 - 1.0).max() <= 1e-14
    assert np.abs(f12).max() <= 1e-14


def test_planar_slope_closed_form():
    lon, lat = _grid(17, 13, lat0=2.0)
    al, be = 0.08, -0.05
    km = S.R_EARTH * S.DEG2RAD
    z = al * lon[None, :] * km + be * lat[:, None] * km
    m11, m12, m22 = TG.topo_metric(z, lon, lat, [0.0])
    gx, gy = al / np.cos(lat * S.DEG2RAD)[:, None], be
    D = 1.0 + gx * gx + gy * gy
    assert np.abs(m11[0] - (1.0 + gy * gy) / D).max() <= 1e-12
    assert np.abs(m22[0] - (1.0 + gx * gx) / D).max() <= 1e-12
    assert np.abs(m12[0] + gx * gy / D).max() <= 1e-12


@pytest.mark.parametrize("n", [101, 321])
def test_topography_workload_planes_match_host(n):
    """The uneven topography workload's surface (8 Gaussian edifices at 1.5 km)
    with its 12 per-period widths."""
    b = S.config_topo_grid(31, n, nsta=10, nperiod=12, smooth=False)
    lon, lat, z = b["lon"], b["lat"], b["z"]
    periods = np.array(b["meta"]["periods"])
    sig = TG.sigma_for_periods(b["meta"]["vbar"], periods, b["meta"]["wavelen_factor"])
    tm = {}
    m11, m12, m22, zs = TG.topo_metric(z, lon, lat, sig, return_smoothed=True, timing=tm)
    t0 = time.time()
    worst = [0.0, 0.0, 0.0, 0.0]
    for p in range(periods.size):
        h11, h12, h22 = S.topo_metric(z, lon, lat, sig[p])
        hz = S.gaussian_smooth_geo(z, lon, lat, sig[p])
        for q, (g, h) in enumerate(((zs[p], hz), (m11[p], h11), (m12[p], h12), (m22[p], h22))):
            worst[q] = max(worst[q], rel(g, h))
    host_s = time.time() - t0
    nodes = sum(TG.window_nodes(n, n, lon, lat, s) for s in sig)
    print(f"{n}x{n} x 12 periods: GPU {tm['kernel_ms']:.3f} ms ({nodes / tm['kernel_ms'] / 1e6:.1f} G window "
          f"nodes/s), host FFT restatement {host_s:.2f} s; rel err zs {worst[0]:.2e} m11 {worst[1]:.2e} "
          f"m12 {worst[2]:.2e} m22 {worst[3]:.2e}", flush=True)
    assert worst[0] <= 1e-12
    assert max(worst[1:]) <= 1e-11


def test_invalid_arguments():
    lon, lat = _grid(30, 25)
    z = np.ones((25, 30))
    with pytest.raises(ValueError):
        TG.topo_metric(np.ones((2, 30)), lon, lat[:2], [0.1])
    with pytest.raises(_lib.SatError):
        TG.topo_metric(z, lon[::-1].copy(), lat, [0.1])    # descending axis
    with pytest.raises(_lib.SatError):
        TG.topo_metric(z, lon, lat, [float("nan")])
    wide = np.ones((3, 2049))                              # more columns than the kernel takes
    with pytest.raises(_lib.SatError):
        TG.topo_metric(wide, np.arange(2049) * 0.01, lat[:3], [0.1])
    # a width beyond the grid is clipped to it: every node averages the whole grid
    m11, m12, m22, zs = TG.topo_metric(z * 2.0, lon, lat, [1e3], return_smoothed=True)
    assert np.abs(zs - 2.0).max() <= 1e-14
