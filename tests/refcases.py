"""Inputs and checks of the reference's eikonal unit tests
(tests/eikonal_test/test_eikonal.cpp tests 1-10), restated so the same checks
run against the FP64 oracle (CPU) and against the HIP path (GPU).

``fwd(xx, yy, spha, sphb, sphc, fun, x0, y0) -> T`` and
``adj(xx, yy, spha, sphb, sphc, T, xr, yr, sa) -> Ta`` are the solver under test.
"""
from __future__ import annotations

import numpy as np

PI = 3.14159265358979323846264338327950288
R_EARTH = 6371.0


def linspaced(n, a, b):
    # Eigen::VectorXd::LinSpaced(n, a, b)
    return np.linspace(a, b, n)


def t_exact(lon, lat, lon0, lat0, s):
    # test_eikonal.cpp:48-56
    tmp = np.sin(lat) * np.sin(lat0) + np.cos(lat) * np.cos(lat0) * np.cos(lon - lon0)
    return np.arccos(np.clip(tmp, -1.0, 1.0)) * R_EARTH * s


def iso(nx, ny, fun_val):
    return np.ones((nx, ny)), np.ones((nx, ny)), np.zeros((nx, ny)), np.full((nx, ny), fun_val)


def test1_isotropic_homogeneous(fwd):
    """test_eikonal.cpp:63-128: T ~ great-circle time; max rel < 5 %, none > 2 %."""
    nx, ny = 61, 41
    xx, yy = linspaced(nx, -30, 30), linspaced(ny, -20, 20)
    s = 1.0 / 3.0
    T = fwd(xx, yy, *iso(nx, ny, s), 0.0, 0.0)
    xr, yr = xx * (PI / 180.0), yy * (PI / 180.0)
    dl = np.abs(xr[:, None] / (xr[1] - xr[0]))
    dt = np.abs(yr[None, :] / (yr[1] - yr[0]))
    keep = ~((dl <= 2.0) & (dt <= 2.0))
    Tref = t_exact(xr[:, None], yr[None, :], 0.0, 0.0, s)
    keep &= Tref >= 1e-10
    rel = np.abs(T - Tref)[keep] / Tref[keep]
    assert rel.max() < 0.05
    assert int((rel > 0.02).sum()) == 0
    return T, rel.max()


def test2_near_source_small(fwd):
    """test_eikonal.cpp:133-162."""
    nx = ny = 41
    xx, yy = linspaced(nx, -10, 10), linspaced(ny, -10, 10)
    T = fwd(xx, yy, *iso(nx, ny, 1.0), 0.0, 0.0)
    d = xx[:, None] ** 2 + yy[None, :] ** 2
    ix, iy = np.unravel_index(np.argmin(d), d.shape)
    assert T[ix, iy] < 1.0
    return T


def test3_monotone_along_row(fwd):
    """test_eikonal.cpp:168-200."""
    nx, ny = 81, 51
    xx, yy = linspaced(nx, -40, 40), linspaced(ny, -25, 25)
    T = fwd(xx, yy, *iso(nx, ny, 1.0), 0.0, 0.0)
    iy, isrc = ny // 2, nx // 2
    bad = sum(1 for ix in range(isrc + 1, nx) if T[ix, iy] < T[ix - 1, iy] - 1.0)
    bad += sum(1 for ix in range(isrc - 1, -1, -1) if T[ix, iy] < T[ix + 1, iy] - 1.0)
    assert bad == 0
    return T


def test4_anisotropic_runs(fwd):
    """test_eikonal.cpp:205-231."""
    nx = ny = 31
    xx, yy = linspaced(nx, -15, 15), linspaced(ny, -15, 15)
    T = fwd(xx, yy, np.full((nx, ny), 1.2), np.full((nx, ny), 0.9), np.full((nx, ny), 0.1),
            np.full((nx, ny), 1.0 / 3.0), 0.0, 0.0)
    assert np.isfinite(T).all()
    assert (T >= 0.0).all()
    return T


def test5_grid_refinement(fwd):
    """test_eikonal.cpp:246-321: error(21x21) < error(11x11) against 81x81."""
    lon_min, lon_max, lat_min, lat_max = -20.0, 20.0, -15.0, 15.0

    def solve(nx, ny):
        xx, yy = linspaced(nx, lon_min, lon_max), linspaced(ny, lat_min, lat_max)
        lat = yy * PI / 180.0
        fun = np.tile(1.0 / (3.0 + np.cos(2.0 * lat)), (nx, 1))
        return fwd(xx, yy, np.ones((nx, ny)), np.ones((nx, ny)), np.zeros((nx, ny)), fun, 0.0, 0.0)

    def node(T, lon, lat):
        nx, ny = T.shape
        dx = (lon_max - lon_min) / (nx - 1)
        dy = (lat_max - lat_min) / (ny - 1)
        return T[int(round((lon - lon_min) / dx)), int(round((lat - lat_min) / dy))]

    lons = [-16.0, -8.0, 8.0, 16.0, -12.0, 12.0, 0.0]
    lats = [9.0, -6.0, 6.0, -9.0, 0.0, 0.0, 12.0]
    Tr, Tc, Tf = solve(81, 81), solve(11, 11), solve(21, 21)
    ec = ef = 0.0
    for lo, la in zip(lons, lats):
        r = node(Tr, lo, la)
        if r < 1e-8:
            continue
        ec = max(ec, abs(node(Tc, lo, la) - r) / r)
        ef = max(ef, abs(node(Tf, lo, la) - r) / r)
    assert ef < ec
    return ec, ef


def iso_grid_T(fwd, nx, ny):
    """test_eikonal.cpp:336-345 make_iso_grid."""
    xx, yy = linspaced(nx, -10, 10), linspaced(ny, -10, 10)
    spha, sphb, sphc, fun = iso(nx, ny, 1.0 / 3.0)
    T = fwd(xx, yy, spha, sphb, sphc, fun, 0.0, 0.0)
    return xx, yy, spha, sphb, sphc, T


def test6_adjoint_zero_source(fwd, adj):
    """test_eikonal.cpp:348-366."""
    xx, yy, a, b, c, T = iso_grid_T(fwd, 21, 21)
    Ta = adj(xx, yy, a, b, c, T, np.zeros(0), np.zeros(0), np.zeros(0))
    assert np.abs(Ta).max() < 1e-8
    return Ta


def test7_adjoint_finite_boundary(fwd, adj):
    """test_eikonal.cpp:371-403."""
    xx, yy, a, b, c, T = iso_grid_T(fwd, 31, 31)
    Ta = adj(xx, yy, a, b, c, T, np.array([5.0]), np.array([0.0]), np.array([1.0]))
    assert np.isfinite(Ta).all()
    bd = max(np.abs(Ta[:, 0]).max(), np.abs(Ta[:, -1]).max(), np.abs(Ta[0, :]).max(), np.abs(Ta[-1, :]).max())
    assert bd < 1e-14
    return Ta


def test8_adjoint_linearity(fwd, adj):
    """test_eikonal.cpp:410-440."""
    xx, yy, a, b, c, T = iso_grid_T(fwd, 21, 21)
    Ta_a = adj(xx, yy, a, b, c, T, np.array([5.0]), np.array([0.0]), np.array([2.0]))
    Ta_b = adj(xx, yy, a, b, c, T, np.array([-5.0]), np.array([3.0]), np.array([-1.5]))
    Ta_ab = adj(xx, yy, a, b, c, T, np.array([5.0, -5.0]), np.array([0.0, 3.0]), np.array([2.0, -1.5]))
    err = np.abs(Ta_a + Ta_b - Ta_ab).max()
    assert err < 1e-8
    return err


def case9_inputs():
    """test_eikonal.cpp:445-462: 121x91 smooth anisotropic medium, source (0.3, -0.4)."""
    nx, ny = 121, 91
    xx, yy = linspaced(nx, -30, 30), linspaced(ny, -20, 20)
    lon = (xx * PI / 180.0)[:, None]
    lat = (yy * PI / 180.0)[None, :]
    spha = 1.0 + 0.10 * np.sin(2.0 * lon) * np.cos(lat)
    sphb = 1.0 + 0.08 * np.cos(1.5 * lon) * np.sin(lat)
    sphc = 0.03 * np.sin(lon + lat)
    fun = 1.0 / (3.2 + 0.3 * np.cos(lat)) + 0.0 * lon
    return xx, yy, spha, sphb, sphc, fun, 0.3, -0.4


def case10_inputs():
    """test_eikonal.cpp:538-581: 161x121 complex medium, source (1.2, -2.4)."""
    nx, ny = 161, 121
    xx, yy = linspaced(nx, -35, 35), linspaced(ny, -25, 25)
    lon = xx[:, None] + 0.0 * yy[None, :]
    lat = yy[None, :] + 0.0 * xx[:, None]
    lonr, latr = lon * PI / 180.0, lat * PI / 180.0
    g1 = np.exp(-((lon - 10.0) ** 2 / (2.0 * 8.0 * 8.0) + (lat + 6.0) ** 2 / (2.0 * 5.0 * 5.0)))
    g2 = np.exp(-((lon + 14.0) ** 2 / (2.0 * 10.0 * 10.0) + (lat - 8.0) ** 2 / (2.0 * 7.0 * 7.0)))
    checker = np.sin(3.0 * lonr) * np.cos(2.5 * latr)
    vel = 3.35 + 0.10 * np.cos(1.7 * latr) + 0.08 * np.sin(1.3 * lonr) + 0.14 * g1 - 0.12 * g2 + 0.05 * checker
    vel = np.maximum(2.8, np.minimum(4.2, vel))
    fun = 1.0 / vel
    spha = 1.0 + 0.12 * np.sin(2.0 * lonr) * np.cos(latr) + 0.03 * g1
    sphb = 1.0 + 0.10 * np.cos(1.5 * lonr) * np.sin(latr) - 0.02 * g2
    sphc = 0.04 * np.sin(lonr + latr) + 0.01 * checker
    return xx, yy, spha, sphb, sphc, fun, 1.2, -2.4


def receivers_for(xx, yy, n, seed):
    rng = np.random.default_rng(seed)
    dx, dy = xx[1] - xx[0], yy[1] - yy[0]
    xr = rng.uniform(xx[0] + 2 * dx, xx[-1] - 2 * dx, n)
    yr = rng.uniform(yy[0] + 2 * dy, yy[-1] - 2 * dy, n)
    return xr, yr, rng.normal(size=n)


def rel_inf(a, b):
    """||a - b||_inf / ||b||_inf (the error norm used throughout)."""
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))
