"""The density kernel (is_kden, src/preproc.cpp:127-136) and the anisotropic
terms (is_anisotropy, :37-65) formed inside the batched runs, against the
oracle and against the post-run terms.

Tolerances: FP64 batches equal the oracle to 1e-12 normalised (their adjoint
fields are the oracle's bits, the batch sums them in the reference's event
order); FP32 batches hold the density field in FP32, each period within
1e-5 normalised of the FP64 batch (the kernel bar of FP32 batches; measured
7e-7 for FP32 and 4e-7 for FP32P).  Bitwise equalities: chunked and unchunked
runs, batched / per-solve / graph runs, and the in-run anisotropic terms with
the post-run ones, since each sums the same values in the same order.
"""
import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import _lib
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth as S
from tests.test_gpu_aniso import _topo_batch
from tests.test_oracle_kden import unit_residual_batch

pytestmark = pytest.mark.gpu


def _rel(g, r):
    return float(np.abs(g - r).max() / np.abs(r).max())


def _run_kden(b, prec=_lib.PREC_F64, mode=_lib.RUN_BATCHED, cap=0, runs=1):
    with PP.EikonalBatch(b, precision=prec, workspace_cap=cap) as bt:
        bt.set_kden(True)
        for _ in range(runs):
            bt.run(mode)
        return bt.results(), bt.kden_results()


def test_kden_fp64_matches_oracle_and_leaves_the_kernel_alone():
    b = S.config_grf(seed=12, nx=57, ny=49, nsta=6, nperiod=3, spacing_km=1.5)
    ro = O.forward_for_event(b, kden=True)
    r, kd = _run_kden(b)
    err = _rel(kd, ro["kden_s"])
    print(f"FP64 kden vs oracle: {err:.2e}", flush=True)
    assert err <= 1e-12
    with PP.EikonalBatch(b) as bt:
        bt.run()
        r0 = bt.results()
        with pytest.raises(_lib.SatError):
            bt.kden_results()  # this batch never formed it
    for k in ("syn", "adj_s", "iters_fwd", "iters_adj", "flags"):
        np.testing.assert_array_equal(r[k], r0[k])
    assert r["chi"] == r0["chi"]
    np.testing.assert_array_equal(r["iters_fwd"], ro["iters_fwd"])


def test_kden_unit_residuals_is_the_negated_unmasked_kernel():
    b = unit_residual_batch()
    r, kd = _run_kden(b)
    ro = O.forward_for_event(b, kden=True)
    assert _rel(kd, ro["kden_s"]) <= 1e-12
    lon, lat = b["lon"], b["lat"]
    for p in range(3):
        (e,) = np.nonzero(b["ev_iper"] == p)[0]
        boxed = O.mask(lon, lat, np.ones((lon.size, lat.size)), float(b["ev_lon"][e]), float(b["ev_lat"][e])).T == 0
        np.testing.assert_array_equal(kd[p][~boxed], -r["adj_s"][p][~boxed])
        assert np.abs(kd[p][boxed]).max() > 0.0


def test_kden_run_modes_and_chunks_same_bits():
    g1 = S.config_grf(seed=31, nx=41, ny=33, nsta=6, nperiod=2, spacing_km=1.5)
    g2 = S.config_grf(seed=32, nx=131, ny=57, nsta=5, nperiod=2, spacing_km=1.5)
    _, whole = _run_kden([g1, g2])
    for mode, cap, runs in ((_lib.RUN_PER_SOLVE, 0, 1), (_lib.RUN_GRAPH, 0, 3), (_lib.RUN_BATCHED, 1 << 20, 2),
                            (_lib.RUN_GRAPH, 1 << 20, 2)):
        _, kd = _run_kden([g1, g2], mode=mode, cap=cap, runs=runs)
        for a, c in zip(kd, whole):
            np.testing.assert_array_equal(a, c, err_msg=f"mode {mode} cap {cap}")
    for gi, g in enumerate((g1, g2)):
        assert _rel(whole[gi], O.forward_for_event(g, kden=True)["kden_s"]) <= 1e-12


@pytest.mark.parametrize("prec", [_lib.PREC_F32, _lib.PREC_F32P])
def test_kden_fp32(prec):
    b = S.config_grf(seed=12, nx=57, ny=49, nsta=6, nperiod=3, spacing_km=1.5)
    _, k64 = _run_kden(b)
    _, k32 = _run_kden(b, prec=prec)
    _, k32c = _run_kden(b, prec=prec, cap=1 << 20)
    errs = [_rel(k32[p], k64[p]) for p in range(3)]
    print(f"precision {prec}: kden per-period error vs FP64 {max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-5
    np.testing.assert_array_equal(k32c, k32)


def test_kden_switch_and_errors():
    b = S.config_grf(seed=5, nx=33, ny=29, nsta=4, nperiod=1)
    with PP.EikonalBatch(b) as bt:
        bt.set_kden(True)
        bt.run()
        kd = bt.kden_results()
        bt.set_kden(False)
        bt.run()
        with pytest.raises(_lib.SatError):
            bt.kden_results()
        bt.set_kden(True)
        bt.run(_lib.RUN_GRAPH)
        np.testing.assert_array_equal(bt.kden_results(), kd)
    with PP.EikonalBatch(b, calc_adj=False) as bt:
        with pytest.raises(_lib.SatError):
            bt.set_kden(True)


@pytest.mark.parametrize("prec", [_lib.PREC_F64, _lib.PREC_F32])
def test_aniso_in_run_equals_post_run(prec):
    b = _topo_batch()
    with PP.EikonalBatch(b, precision=prec) as bt:
        bt.set_aniso_metric(b["m11"], b["m12"], b["m22"])
        with pytest.raises(_lib.SatError):
            bt.aniso_kernels()  # no run has formed them yet
        bt.run()
        xi, eta = bt.aniso_kernels()
        xp, ep = bt.aniso_kernels(b["m11"], b["m12"], b["m22"])
        bt.run(_lib.RUN_PER_SOLVE)
        xs, es = bt.aniso_kernels()
        bt.run(_lib.RUN_GRAPH)
        bt.run(_lib.RUN_GRAPH)
        xg, eg = bt.aniso_kernels()
        bt.set_aniso_metric()
        bt.run()
        with pytest.raises(_lib.SatError):
            bt.aniso_kernels()
    np.testing.assert_array_equal(xi, xp)
    np.testing.assert_array_equal(eta, ep)
    np.testing.assert_array_equal(xg, xi)
    np.testing.assert_array_equal(eg, eta)
    if prec == _lib.PREC_F64:  # the per-solve kernel leaves the same FP64 fields
        np.testing.assert_array_equal(xs, xi)
        np.testing.assert_array_equal(es, eta)
    else:
        assert _rel(xs, xi) <= 1e-4 and _rel(es, eta) <= 1e-4


def test_aniso_in_run_chunked():
    """Chunked runs, which the post-run terms refuse, form the terms chunk by
    chunk: the same bits as the unchunked run, for both grids of the batch."""
    g1 = _topo_batch()
    g2 = S.config_grf(seed=32, nx=131, ny=57, nsta=5, nperiod=2, spacing_km=1.5)
    out = {}
    for cap in (0, 1 << 20):
        with PP.EikonalBatch([g1, g2], workspace_cap=cap) as bt:
            for gi, g in enumerate((g1, g2)):
                bt.set_aniso_metric(g["m11"], g["m12"], g["m22"], grid=gi)
            bt.run()
            out[cap] = [bt.aniso_kernels(grid=gi) for gi in range(2)]
            if cap == 0:
                ref = [bt.aniso_kernels(g["m11"], g["m12"], g["m22"], grid=gi) for gi, g in enumerate((g1, g2))]
    for gi in range(2):
        for k in range(2):
            np.testing.assert_array_equal(out[1 << 20][gi][k], out[0][gi][k])
            np.testing.assert_array_equal(out[0][gi][k], ref[gi][k])
