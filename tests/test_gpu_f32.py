"""GPU parity of the FP32 path (SAT_PREC_F32) against the FP64 oracle.

Bar (the metric's): ||K - K_oracle||_inf / ||K_oracle||_inf <= 1e-5 for the
per-period kernels; travel times and syn within 1e-6 relative; chi within 1e-5.
"""
import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import PREC_F32
from surfattpp_amd import eikonal as E
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth
from tests import refcases as RC

pytestmark = pytest.mark.gpu
TOL_K = 1e-5
TOL_T = 1e-6


@pytest.mark.parametrize("case", ["case9", "case10"])
def test_forward_f32_close_to_oracle(case):
    xx, yy, spha, sphb, sphc, fun, x0, y0 = getattr(RC, case + "_inputs")()
    T, it, _ = O.forward(xx, yy, spha, sphb, sphc, fun, x0, y0)
    Tg = E.FSM_UW_PS_lonlat_2d(xx, yy, spha, sphb, sphc, fun, x0, y0, precision=PREC_F32)
    err = RC.rel_inf(Tg, T)
    print(case, "T rel err", err, "iters", E.last_info["iters"], "oracle", it, flush=True)
    assert err <= TOL_T


@pytest.mark.parametrize("case", ["case9", "case10"])
def test_adjoint_f32_close_to_oracle(case):
    xx, yy, spha, sphb, sphc, fun, x0, y0 = getattr(RC, case + "_inputs")()
    T, _, _ = O.forward(xx, yy, spha, sphb, sphc, fun, x0, y0)
    xr, yr, sa = RC.receivers_for(xx, yy, 30, 5)
    Ta, _ = O.adjoint(xx, yy, spha, sphb, sphc, T, xr, yr, sa)
    Tag = E.FSM_O1_JSE_lonlat_2d(xx, yy, spha, sphb, sphc, T, xr, yr, sa, precision=PREC_F32)
    err = RC.rel_inf(Tag, Ta)
    print(case, "Ta rel err", err, flush=True)
    assert err <= TOL_K


def test_ref_known_answers_f32():
    f = lambda *a: E.FSM_UW_PS_lonlat_2d(*a, precision=PREC_F32)  # noqa: E731
    g = lambda *a: E.FSM_O1_JSE_lonlat_2d(*a, precision=PREC_F32)  # noqa: E731
    RC.test1_isotropic_homogeneous(f)
    RC.test2_near_source_small(f)
    RC.test3_monotone_along_row(f)
    RC.test4_anisotropic_runs(f)
    RC.test5_grid_refinement(f)
    RC.test6_adjoint_zero_source(f, g)
    RC.test7_adjoint_finite_boundary(f, g)
    # linearity of the FP32 adjoint holds to FP32 rounding of the solution
    xx, yy, a, b, c, T = RC.iso_grid_T(f, 21, 21)
    Ta_a = g(xx, yy, a, b, c, T, np.array([5.0]), np.array([0.0]), np.array([2.0]))
    Ta_b = g(xx, yy, a, b, c, T, np.array([-5.0]), np.array([3.0]), np.array([-1.5]))
    Ta_ab = g(xx, yy, a, b, c, T, np.array([5.0, -5.0]), np.array([0.0, 3.0]), np.array([2.0, -1.5]))
    assert RC.rel_inf(Ta_a + Ta_b, Ta_ab) < 1e-6


def test_batch_grf_201_f32_vs_oracle():
    b = synth.config_grf(nperiod=2, nsta=10)
    ro = O.forward_for_event(b)
    with PP.EikonalBatch(b, precision=PREC_F32) as bt:
        bt.run()
        rg = bt.results()
        _, _, nres = bt.conv()
    ek = RC.rel_inf(rg["adj_s"], ro["adj_s"])
    es = RC.rel_inf(rg["syn"], ro["syn"])
    ec = abs(rg["chi"] - ro["chi"]) / abs(ro["chi"])
    print("K err", ek, "syn err", es, "chi err", ec, flush=True)
    print("FP64 re-solves", nres, flush=True)
    # forward rounds equal the reference's (decisions near the tolerance were re-solved in FP64)
    np.testing.assert_array_equal(rg["iters_fwd"], ro["iters_fwd"])
    for p in range(rg["adj_s"].shape[0]):
        assert RC.rel_inf(rg["adj_s"][p], ro["adj_s"][p]) <= TOL_K
    assert es <= TOL_T
    assert ec <= TOL_K


def test_batch_f32_matches_f64_gpu_full_workload():
    # The metric's workload: 201x201, 100 sources x 30 periods.  Every period's
    # kernel within the bar of the FP64 path (which equals the oracle to 1e-12,
    # test_gpu_parity.py), every forward iteration count equal.
    b = synth.config_grf()
    with PP.EikonalBatch(b) as b64, PP.EikonalBatch(b, precision=PREC_F32) as b32:
        b64.run()
        b32.run()
        r64, r32 = b64.results(), b32.results()
        _, _, nres = b32.conv()
    errs = [RC.rel_inf(r32["adj_s"][p], r64["adj_s"][p]) for p in range(r64["adj_s"].shape[0])]
    print("per-period K err max", max(errs), "syn err", RC.rel_inf(r32["syn"], r64["syn"]), "FP64 re-solves", nres,
          flush=True)
    np.testing.assert_array_equal(r32["iters_fwd"], r64["iters_fwd"])
    assert max(errs) <= TOL_K
    assert RC.rel_inf(r32["syn"], r64["syn"]) <= TOL_T


def test_batch_f32_resolve_pool_and_overflow():
    # A band of 0.99 puts most pairs within the band: 64 of them are re-solved
    # by the FP64 batched kernels (the pool), the rest by the FP64 per-pair
    # kernel.  A re-solved pair is an FP64 result: its iteration counts equal
    # the FP64 batch's and its syn agrees to 1e-12 (the pool's update forms
    # reciprocals in FMA instead of IEEE divisions: rounding-level differences);
    # the others within the FP32 bar.
    band = 0.99
    b = synth.subset_events(synth.config_grf(nperiod=3, nsta=40), list(range(0, 120)))
    with PP.EikonalBatch(b) as b64, PP.EikonalBatch(b, precision=PREC_F32) as b32:
        b32.set_f32_band(band)
        b64.run()
        b32.run()
        r64, r32 = b64.results(), b32.results()
        last, prev, nres = b32.conv()
    tol = 1e-4
    amb = ((last >= (1 - band) * tol) & (r32["iters_fwd"] < 2000)) | ((prev >= 0) & (prev < (1 + band) * tol))
    print("re-solves", nres, "ambiguous", int(amb.sum()), "of", len(b["ev_iper"]), flush=True)
    assert nres == int(amb.sum()) and nres > 64
    np.testing.assert_array_equal(r32["iters_fwd"], r64["iters_fwd"])
    np.testing.assert_array_equal(r32["iters_adj"][amb], r64["iters_adj"][amb])
    off = np.concatenate([[0], np.cumsum(b["ev_nrec"])])
    for e in np.nonzero(amb)[0]:
        ref = r64["syn"][off[e]:off[e + 1]]
        assert RC.rel_inf(r32["syn"][off[e]:off[e + 1]], ref) <= 1e-12, e
    assert RC.rel_inf(r32["syn"], r64["syn"]) <= TOL_T
    for p in range(r64["adj_s"].shape[0]):
        assert RC.rel_inf(r32["adj_s"][p], r64["adj_s"][p]) <= TOL_K


def test_batch_f32p_polished_update_full_workload():
    """SAT_PREC_F32P on the 201x201 workload: the FP32 search with the FP64
    polish of the winning candidate.  Travel times (syn) within 1e-9 of the
    FP64 path (FP32 alone: ~1e-8), every period's kernel within 2.5e-6 (FP32
    alone: ~5e-6; what remains is the FP32 adjoint), iteration counts equal."""
    from surfattpp_amd import PREC_F32P
    b = synth.config_grf()
    with PP.EikonalBatch(b) as b64:
        b64.run()
        r64 = b64.results()
    with PP.EikonalBatch(b, precision=PREC_F32P) as bp:
        bp.run()
        rp = bp.results()
        assert bp.waves() == [1]
    errs = [RC.rel_inf(rp["adj_s"][p], r64["adj_s"][p]) for p in range(r64["adj_s"].shape[0])]
    es = RC.rel_inf(rp["syn"], r64["syn"])
    print(f"F32P: per-period K err max {max(errs):.3e}, syn err {es:.3e}", flush=True)
    np.testing.assert_array_equal(rp["iters_fwd"], r64["iters_fwd"])
    assert max(errs) <= 2.5e-6
    assert es <= 1e-9


def test_batch_f32p_small_vs_oracle_and_topography_fallback():
    """F32P against the oracle on a small two-period batch, and on a grid with
    a topography metric (no polished update there: the same results as
    SAT_PREC_F32)."""
    from surfattpp_amd import PREC_F32P
    b = synth.config_grf(seed=11, nx=61, ny=57, nsta=8, nperiod=2, spacing_km=1.5)
    ro = O.forward_for_event(b)
    with PP.EikonalBatch(b, precision=PREC_F32P) as bp:
        bp.run()
        rp = bp.results()
    np.testing.assert_array_equal(rp["iters_fwd"], ro["iters_fwd"])
    assert RC.rel_inf(rp["syn"], ro["syn"]) <= 1e-9
    for p in range(2):
        assert RC.rel_inf(rp["adj_s"][p], ro["adj_s"][p]) <= 2.5e-6
    t = synth.config_topo_grid(33, 61, nsta=6, nperiod=2)
    with PP.EikonalBatch(t, precision=PREC_F32, waves=1) as b32, PP.EikonalBatch(t, precision=PREC_F32P) as bp:
        b32.run()
        bp.run()
        r32, rp = b32.results(), bp.results()
    np.testing.assert_array_equal(rp["syn"], r32["syn"])
    np.testing.assert_array_equal(rp["adj_s"], r32["adj_s"])
