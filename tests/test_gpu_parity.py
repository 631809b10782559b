"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

FP64 bar: ||x - x_oracle||_inf / ||x_oracle||_inf <= 1e-12 for T, Ta, syn and the
per-period kernels, chi within 1e-12 relative, identical iteration counts.
"""
import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import eikonal as E
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth
from tests import refcases as RC

pytestmark = pytest.mark.gpu
TOL64 = 1e-12

gfwd = lambda *a: E.FSM_UW_PS_lonlat_2d(*a)  # noqa: E731
gadj = lambda *a: E.FSM_O1_JSE_lonlat_2d(*a)  # noqa: E731


def test_device_present():
    from surfattpp_amd import _lib
    info = _lib.device_info()
    print(info, flush=True)
    assert "gfx950" in info


@pytest.mark.parametrize("case", ["case9", "case10"])
def test_forward_f64_matches_oracle(case):
    xx, yy, spha, sphb, sphc, fun, x0, y0 = getattr(RC, case + "_inputs")()
    T, it, _ = O.forward(xx, yy, spha, sphb, sphc, fun, x0, y0)
    Tg = gfwd(xx, yy, spha, sphb, sphc, fun, x0, y0)
    assert E.last_info["iters"] == it
    assert RC.rel_inf(Tg, T) <= TOL64


@pytest.mark.parametrize("case", ["case9", "case10"])
def test_adjoint_f64_matches_oracle(case):
    xx, yy, spha, sphb, sphc, fun, x0, y0 = getattr(RC, case + "_inputs")()
    T, _, _ = O.forward(xx, yy, spha, sphb, sphc, fun, x0, y0)
    xr, yr, sa = RC.receivers_for(xx, yy, 30, 5)
    Ta, ita = O.adjoint(xx, yy, spha, sphb, sphc, T, xr, yr, sa)
    Tag = gadj(xx, yy, spha, sphb, sphc, T, xr, yr, sa)
    assert E.last_info["iters"] == ita
    assert RC.rel_inf(Tag, Ta) <= TOL64


def test_golden_twin_vectors_gpu():
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "twin_vectors.npz"))
    for i in range(int(z["ncase"])):
        g = lambda k: z[f"c{i}_{k}"]  # noqa: E731
        Tg = gfwd(g("xx"), g("yy"), g("spha"), g("sphb"), g("sphc"), g("fun"), float(g("x0")), float(g("y0")))
        assert RC.rel_inf(Tg, g("T")) < 1e-12
        Tag = gadj(g("xx"), g("yy"), g("spha"), g("sphb"), g("sphc"), g("T"), g("xr"), g("yr"), g("adj"))
        assert RC.rel_inf(Tag, g("Ta")) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8])
def test_ref1_homogeneous_gpu(n):
    RC.test1_isotropic_homogeneous(gfwd) if n == 1 else None
    if n == 2:
        RC.test2_near_source_small(gfwd)
    if n == 3:
        RC.test3_monotone_along_row(gfwd)
    if n == 4:
        RC.test4_anisotropic_runs(gfwd)
    if n == 5:
        RC.test5_grid_refinement(gfwd)
    if n == 6:
        RC.test6_adjoint_zero_source(gfwd, gadj)
    if n == 7:
        RC.test7_adjoint_finite_boundary(gfwd, gadj)
    if n == 8:
        RC.test8_adjoint_linearity(gfwd, gadj)


@pytest.mark.parametrize("nx,ny", [(3, 3), (5, 4), (64, 9), (65, 7), (129, 11), (200, 13), (512, 5)])
def test_forward_ragged_shapes(nx, ny):
    # every lane layout (columns per lane C = 1..8, partial last lane)
    xx, yy = np.linspace(-1, 1, nx), np.linspace(-0.5, 0.5, ny)
    rng = np.random.default_rng(nx * 1000 + ny)
    fun = 1.0 / (3.0 + 0.3 * rng.uniform(size=(nx, ny)))
    a = 1.0 + 0.05 * rng.uniform(size=(nx, ny))
    b = 1.0 + 0.05 * rng.uniform(size=(nx, ny))
    c = 0.02 * rng.uniform(-1, 1, size=(nx, ny))
    x0, y0 = xx[nx // 2] + 0.1 * (xx[1] - xx[0]), yy[ny // 2] - 0.2 * (yy[1] - yy[0])
    T, it, _ = O.forward(xx, yy, a, b, c, fun, x0, y0)
    Tg = gfwd(xx, yy, a, b, c, fun, x0, y0)
    assert E.last_info["iters"] == it
    assert RC.rel_inf(Tg, T) <= TOL64
    xr, yr, sa = RC.receivers_for(xx, yy, 7, nx) if nx > 5 and ny > 5 else (np.array([xx[1]]), np.array([yy[1]]), np.array([1.0]))
    Ta, ita = O.adjoint(xx, yy, a, b, c, T, xr, yr, sa)
    Tag = gadj(xx, yy, a, b, c, T, xr, yr, sa)
    assert E.last_info["iters"] == ita
    assert RC.rel_inf(Tag, Ta) <= TOL64 or np.abs(Ta).max() == 0.0


def test_errors_follow_reference():
    with pytest.raises(ValueError):
        gfwd(np.array([0.0, 1.0]), np.linspace(0, 1, 5), np.ones((2, 5)), np.ones((2, 5)), np.zeros((2, 5)),
             np.ones((2, 5)), 0.5, 0.5)
    nx = ny = 21
    xx, yy = np.linspace(0, 2, nx), np.linspace(0, 2, ny)
    gfwd(xx, yy, *RC.iso(nx, ny, 0.3), 0.0, 1.0)
    assert E.last_info["flags"] & 2  # source on the boundary: a warning flag, not an error


def _batch_vs_oracle(b, tol=TOL64):
    ro = O.forward_for_event(b)
    with PP.EikonalBatch(b) as bt:
        bt.run()
        rg = bt.results()
        bt.run(mode=1)  # one launch per pair: same kernels, same bits
        rp = bt.results()
    np.testing.assert_array_equal(rg["iters_fwd"], ro["iters_fwd"])
    np.testing.assert_array_equal(rg["iters_adj"], ro["iters_adj"])
    assert RC.rel_inf(rg["syn"], ro["syn"]) <= tol
    assert RC.rel_inf(rg["adj_s"], ro["adj_s"]) <= tol
    assert abs(rg["chi"] - ro["chi"]) <= tol * abs(ro["chi"])
    np.testing.assert_array_equal(rp["adj_s"], rg["adj_s"])
    np.testing.assert_array_equal(rp["syn"], rg["syn"])
    return rg, ro


def test_batch_checkerboard_61x71_f64():
    cb = synth.checkerboard(61, 71, 3.0, 0.06, 2, 3)
    b0 = synth.config_checkerboard()
    # observed = forward times in the checkerboard (computed by the oracle)
    bt = dict(b0, svel=cb[None])
    obs = O.forward_for_event(bt, is_calc_adj=False)["syn"]
    b = synth.config_checkerboard(tt_obs=obs)
    rg, ro = _batch_vs_oracle(b)
    print("chi", rg["chi"], "iters", rg["iters_fwd"].tolist(), rg["iters_adj"].tolist(), flush=True)


def test_batch_grf_201_subset_f64():
    b = synth.config_grf(nperiod=3, nsta=12)
    _batch_vs_oracle(b)


def test_batch_multi_grid_uneven_f64():
    g1 = synth.config_grf(seed=11, nx=41, ny=33, nsta=6, nperiod=2, spacing_km=1.5)
    g2 = synth.config_grf(seed=12, nx=131, ny=57, nsta=5, nperiod=2, spacing_km=1.5)
    with PP.EikonalBatch([g1, g2]) as bt:
        bt.run()
        r = bt.results()
    for gi, g in enumerate([g1, g2]):
        ro = O.forward_for_event(g)
        sl = slice(sum(bt.rec_counts[:gi]), sum(bt.rec_counts[:gi + 1]))
        assert RC.rel_inf(r["syn"][sl], ro["syn"]) <= TOL64
        assert RC.rel_inf(r["adj_s"][gi], ro["adj_s"]) <= TOL64
        np.testing.assert_array_equal(r["iters_fwd"][bt.ev_slices[gi]], ro["iters_fwd"])


def test_batch_zero_receivers_and_forward_only():
    b = synth.config_grf(seed=5, nx=33, ny=29, nsta=4, nperiod=1)
    b["ev_nrec"] = b["ev_nrec"].copy()
    b["ev_nrec"][1] = 0  # an event without receivers: Ta == 0 (test_eikonal.cpp:348)
    ro = O.forward_for_event(b)
    with PP.EikonalBatch(b) as bt:
        bt.run()
        rg = bt.results()
    assert RC.rel_inf(rg["adj_s"], ro["adj_s"]) <= TOL64
    with PP.EikonalBatch(b, calc_adj=False) as bt:
        bt.run()
        rf = bt.results()
    assert RC.rel_inf(rf["syn"], ro["syn"]) <= TOL64
    assert rf["chi"] == 0.0 and not rf["adj_s"].any()


def test_batch_update_media_and_obs():
    b = synth.config_grf(seed=6, nx=45, ny=41, nsta=5, nperiod=2)
    b2 = dict(b, svel=b["svel"] * 1.02, tt_obs=b["tt_obs"] + 0.1)
    ro = O.forward_for_event(b2)
    with PP.EikonalBatch(b) as bt:
        bt.run()
        bt.update_media(0, b2["svel"], b2["m11"], b2["m12"], b2["m22"])
        bt.update_obs(b2["tt_obs"], b2["weight"])
        bt.run()
        rg = bt.results()
    assert RC.rel_inf(rg["adj_s"], ro["adj_s"]) <= TOL64
    assert abs(rg["chi"] - ro["chi"]) <= TOL64 * abs(ro["chi"])


def test_receiver_spread_shared_cells_and_many_receivers_f64():
    """The receiver spread in receiver order (src/eikonal_solver.cpp:356-375):
    receivers that share delta cells (repeated positions, clusters within a
    cell or two) and an event with more receivers (300) than the one-wave LDS
    spread takes, in FP64 against the oracle and in per-pair mode."""
    b = synth.config_checkerboard()
    lon, lat = b["lon"], b["lat"]
    rng = synth.Rng(21)
    dx, dy = lon[1] - lon[0], lat[1] - lat[0]
    # event 0: 300 receivers anywhere inside; event 1: 60 receivers in a 2x2-cell
    # cluster, every fifth one repeated; event 2: 8 copies of one position
    u = rng.uniform(600)
    r0 = (lon[3] + u[:300] * (lon[-4] - lon[3]), lat[3] + u[300:] * (lat[-4] - lat[3]))
    v = rng.uniform(120)
    cx, cy = lon[30] + 0.3 * dx, lat[35] + 0.6 * dy
    c_lon, c_lat = cx + 2 * dx * v[:60], cy + 2 * dy * v[60:]
    c_lon[::5], c_lat[::5] = c_lon[0], c_lat[0]
    r2 = (np.full(8, lon[20] + 0.25 * dx), np.full(8, lat[40] + 0.75 * dy))
    rec_lon = np.concatenate([r0[0], c_lon, r2[0]])
    rec_lat = np.concatenate([r0[1], c_lat, r2[1]])
    nrec = np.array([300, 60, 8], np.int32)
    off = np.array([0, 300, 360], np.int32)
    src = [(lon[10], lat[12]), (lon[45], lat[50]), (lon[33], lat[20])]
    tt = synth.great_circle_time(rec_lon, rec_lat, np.repeat([s[0] for s in src], nrec),
                                 np.repeat([s[1] for s in src], nrec), 3.09)
    bb = dict(b, ev_iper=np.zeros(3, np.int32), ev_lon=np.array([s[0] for s in src]),
              ev_lat=np.array([s[1] for s in src]), ev_rec_off=off, ev_nrec=nrec, rec_lon=rec_lon,
              rec_lat=rec_lat, tt_obs=tt, weight=1.0 + 0.5 * rng.uniform(368))
    rg, ro = _batch_vs_oracle(bb)
    print("chi", rg["chi"], "oracle", ro["chi"], flush=True)


def test_batch_takes_device_inputs_and_outputs():
    """Planes, receivers and observations given in device memory (the caller's
    hipMalloc), results written to device memory: the same bits as host
    pointers (the ownership contract of the replaced interface: host or
    device pointers)."""
    from surfattpp_amd import _lib
    from tests.devmem import DevArray
    import ctypes as C
    b = synth.config_grf(seed=21, nx=47, ny=41, nsta=5, nperiod=2, spacing_km=1.5)
    with PP.EikonalBatch(b) as bt:
        bt.run()
        ref = bt.results()
    nper, ny, nx = b["svel"].shape
    keep = {k: DevArray(np.ascontiguousarray(b[k]).ravel()) for k in ("svel", "m11", "m12", "m22", "rec_lon",
                                                                    "rec_lat", "tt_obs", "weight")}
    try:
        lon, lat = np.ascontiguousarray(b["lon"]), np.ascontiguousarray(b["lat"])
        g = _lib.sat_grid_desc(nx, ny, nper, lon.ctypes.data, lat.ctypes.data, keep["svel"].ptr, keep["m11"].ptr,
                               keep["m12"].ptr, keep["m22"].ptr, 0)
        gd = (_lib.sat_grid_desc * 1)(g)
        ne = len(b["ev_iper"])
        ed = (_lib.sat_event_desc * ne)()
        for e in range(ne):
            ed[e] = _lib.sat_event_desc(0, int(b["ev_iper"][e]), float(b["ev_lon"][e]), float(b["ev_lat"][e]),
                                        int(b["ev_rec_off"][e]), int(b["ev_nrec"][e]))
        d = _lib.sat_batch_desc()
        d.ngrid, d.grids, d.nevent, d.events = 1, gd, ne, ed
        d.nrec_total = len(b["rec_lon"])
        d.rec_lon_deg, d.rec_lat_deg = keep["rec_lon"].ptr, keep["rec_lat"].ptr
        d.tt_obs, d.weight = keep["tt_obs"].ptr, keep["weight"].ptr
        d.precision, d.calc_adj, d.workspace_cap = _lib.PREC_F64, 1, 0
        L = _lib.load()
        h = C.c_void_p()
        assert L.sat_batch_create(C.byref(d), C.byref(h)) == _lib.SAT_OK
        try:
            assert L.sat_batch_run(h, _lib.RUN_BATCHED) == _lib.SAT_OK
            out_syn, out_k = DevArray(np.zeros(d.nrec_total)), DevArray(np.zeros(nper * ny * nx))
            chi = DevArray(np.zeros(1))
            assert L.sat_batch_results(h, out_syn.ptr, out_k.ptr, chi.ptr, None, None, None) == _lib.SAT_OK
            np.testing.assert_array_equal(out_syn.get(d.nrec_total), ref["syn"])
            np.testing.assert_array_equal(out_k.get(nper * ny * nx).reshape(nper, ny, nx), ref["adj_s"])
            assert chi.get(1)[0] == ref["chi"]
            # a new model given in device memory
            sv2 = DevArray((np.ascontiguousarray(b["svel"]) * 1.02).ravel())
            assert L.sat_batch_update_media(h, 0, sv2.ptr, keep["m11"].ptr, keep["m12"].ptr,
                                            keep["m22"].ptr) == _lib.SAT_OK
            assert L.sat_batch_run(h, _lib.RUN_BATCHED) == _lib.SAT_OK
            syn2 = np.zeros(d.nrec_total)
            assert L.sat_batch_results(h, syn2.ctypes.data, None, None, None, None, None) == _lib.SAT_OK
            b2 = dict(b)
            b2["svel"] = np.ascontiguousarray(b["svel"]) * 1.02
            with PP.EikonalBatch(b2) as bt2:
                bt2.run()
                np.testing.assert_array_equal(syn2, bt2.results()["syn"])
            for a in (out_syn, out_k, chi, sv2):
                a.free()
        finally:
            L.sat_batch_destroy(h)
    finally:
        for a in keep.values():
            a.free()
