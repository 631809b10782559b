"""The reference's failure behaviours on the GPU, against the oracle
(small grids, both precisions):

- a receiver outside the grid: interp2d returns NaN (include/utils.h:379-383),
  so its residual, the event's adjoint source, chi and (through the adjoint's
  NaN convergence norm, src/eikonal_solver.cpp:495-511) the kernel are NaN and
  the adjoint runs to its 1000-iteration cap; here also SAT_FLAG_REC_OUTSIDE
  and SAT_FLAG_ADJ_MAXITER;
- a NaN slowness in the source's cell: T0 is NaN everywhere, the forward's L1
  norm is NaN and never passes the test, so the forward runs its 2000
  iterations (:292-313) with SAT_FLAG_NOT_CONVERGED; a NaN slowness away from
  every source only leaves that node's tau at its initial value;
- a metric with det = a b - c^2 <= 0 in a patch: the one-point candidates
  there are skipped (:228) and the solve still converges;
- receivers on grid lines, grid nodes and the grid's edges (the cell choice of
  interp2d).
FP64 equals the oracle to 1e-12 with the same NaN pattern, iteration counts
and flags; FP32 has the oracle's NaN pattern and iteration counts.
"""
import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import _lib
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth as S
from tests import refcases as RC

pytestmark = pytest.mark.gpu


def _base():
    return S.config_grf(seed=7, nx=23, ny=21, nsta=4, nperiod=1, spacing_km=2.0)


def _outside_receiver():
    b = _base()
    b["rec_lon"] = b["rec_lon"].copy()
    b["rec_lat"] = b["rec_lat"].copy()
    b["rec_lon"][1] = b["lon"][-1] + 0.01           # beyond the east edge
    b["rec_lon"][2], b["rec_lat"][2] = b["lon"][5], b["lat"][0]   # on the south edge, on a node
    b["rec_lon"][4], b["rec_lat"][4] = b["lon"][-1], b["lat"][9]  # on the east edge, on a node
    b["rec_lon"][5] = b["lon"][11]                  # on an interior grid line
    b["rec_lat"][6] = b["lat"][-1]                  # on the north edge
    return b


def _nan_at_source():
    b = _base()
    sv = b["svel"].copy()
    d = b["lon"][1] - b["lon"][0]
    ix = int(np.floor((b["ev_lon"][0] - b["lon"][0]) / d))
    iy = int(np.floor((b["ev_lat"][0] - b["lat"][0]) / (b["lat"][1] - b["lat"][0])))
    sv[0, iy, ix] = np.nan
    b["svel"] = sv
    return b


def _nan_away():
    b = _base()
    sv = b["svel"].copy()
    sv[0, 2, 20] = np.nan
    b["svel"] = sv
    return b


def _det_not_positive():
    b = _base()
    m12 = b["m12"].copy()
    m12[0, 3:7, 15:20] = -1.3    # a = b = 1: det = 1 - 1.69 < 0
    b["m12"] = m12
    return b


CASES = {"outside_receiver": _outside_receiver, "nan_slowness_at_source": _nan_at_source,
         "nan_slowness_away": _nan_away, "det_not_positive": _det_not_positive}


def _same_nan_rel(g, r):
    """NaN at the same places, and the normalised error of the finite rest."""
    g, r = np.asarray(g), np.asarray(r)
    np.testing.assert_array_equal(np.isnan(g), np.isnan(r))
    f = ~np.isnan(r)
    if not f.any():
        return 0.0
    return RC.rel_inf(g[f], r[f])


def _expected_flags(ro, b):
    fl = ro["flags"].copy()
    for e in range(len(b["ev_iper"])):
        o, n = b["ev_rec_off"][e], b["ev_nrec"][e]
        if np.isnan(ro["syn"][o:o + n]).any():
            fl[e] |= _lib.FLAG_REC_OUTSIDE
        if ro["iters_adj"][e] >= 1000:
            fl[e] |= _lib.FLAG_ADJ_MAXITER
    return fl


@pytest.mark.parametrize("case", sorted(CASES))
def test_failure_case_fp64_matches_oracle(case):
    b = CASES[case]()
    ro = O.forward_for_event(b)
    with PP.EikonalBatch(b) as bt:
        bt.run()
        rg = bt.results()
    print(f"{case}: iters fwd {rg['iters_fwd']} adj {rg['iters_adj']} flags {rg['flags']} chi {rg['chi']}",
          flush=True)
    np.testing.assert_array_equal(rg["iters_fwd"], ro["iters_fwd"])
    np.testing.assert_array_equal(rg["iters_adj"], ro["iters_adj"])
    np.testing.assert_array_equal(rg["flags"], _expected_flags(ro, b))
    assert _same_nan_rel(rg["syn"], ro["syn"]) <= 1e-12
    assert _same_nan_rel(rg["adj_s"], ro["adj_s"]) <= 1e-12
    assert (np.isnan(rg["chi"]) and np.isnan(ro["chi"])) or abs(rg["chi"] - ro["chi"]) <= 1e-12 * abs(ro["chi"])
    if case == "outside_receiver":
        assert np.isnan(rg["syn"][1]) and np.isnan(rg["chi"])
        assert rg["flags"][0] & _lib.FLAG_REC_OUTSIDE and rg["iters_adj"][0] == 1000
    if case == "nan_slowness_at_source":
        assert rg["iters_fwd"][0] == 2000 and rg["flags"][0] & _lib.FLAG_NOT_CONVERGED
    if case == "nan_slowness_away":
        assert np.isfinite(rg["syn"]).all()


@pytest.mark.parametrize("case", sorted(CASES))
def test_failure_case_fp32_iterations_and_nan_pattern(case):
    b = CASES[case]()
    ro = O.forward_for_event(b)
    with PP.EikonalBatch(b, precision=_lib.PREC_F32) as bt:
        bt.run()
        rg = bt.results()
    np.testing.assert_array_equal(rg["iters_fwd"], ro["iters_fwd"])
    np.testing.assert_array_equal(rg["flags"] & ~_lib.FLAG_ADJ_MAXITER,
                                  _expected_flags(ro, b) & ~_lib.FLAG_ADJ_MAXITER)
    es = _same_nan_rel(rg["syn"], ro["syn"])
    ek = _same_nan_rel(rg["adj_s"], ro["adj_s"])
    print(f"{case} FP32: syn err {es:.2e} kernel err {ek:.2e}", flush=True)
    assert es <= 1e-6 and ek <= 1e-4
