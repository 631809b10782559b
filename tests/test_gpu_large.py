"""The large batched mixed-precision workload: 401x481 nodes at 1 km, 400
stations x 30 periods = 12,000 forward+adjoint pairs (399 receivers each) in
one batch, FP64, FP32 and FP32 with the FP64-polished update (SAT_PREC_F32P).

- FP64: a seeded subset of pairs spread over the periods equals the oracle
  (the reference's arithmetic, src/eikonal_solver.cpp, src/preproc.cpp:71-148)
  to 1e-12 with equal iteration counts;
- FP32 and FP32P: every period's kernel within 1e-5 (FP32) and 5e-6 (FP32P)
  of the FP64 batch, normalised infinity norm (the FP64 batch equals the
  oracle on the subset), and every pair's forward iteration count equal to
  the FP64 one (the FP64 re-solve of pairs whose stop decision is near the
  tolerance).
"""
import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import _lib
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth
from tests import refcases as RC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def large():
    b = synth.config_large()
    with PP.EikonalBatch(b, precision=_lib.PREC_F64) as bt:
        bt.run()
        r64 = bt.results()
    return b, r64


def _period_errors(r, r64):
    return np.array([RC.rel_inf(r["adj_s"][p], r64["adj_s"][p]) for p in range(r64["adj_s"].shape[0])])


def test_large_fp64_subset_matches_oracle(large):
    b, r64 = large
    ne = len(b["ev_iper"])
    assert ne == 12000 and b["lon"].size == 401 and b["lat"].size == 481
    rng = np.random.default_rng(2024)
    # one pair from each of 8 periods spread over the 30
    periods = np.linspace(0, 29, 8).astype(int)
    ids = [int(rng.choice(np.nonzero(b["ev_iper"] == p)[0])) for p in periods]
    sub = synth.subset_events(b, ids)
    ro = O.forward_for_event(sub)
    with PP.EikonalBatch(sub, precision=_lib.PREC_F64) as bt:
        bt.run()
        rs = bt.results()
    # the batch of 12,000 and the 8-pair batch give the same pairs the same values
    for k, e in enumerate(ids):
        o, n = int(b["ev_rec_off"][e]), int(b["ev_nrec"][e])
        so, sn = int(sub["ev_rec_off"][k]), int(sub["ev_nrec"][k])
        np.testing.assert_array_equal(r64["syn"][o:o + n], rs["syn"][so:so + sn])
        assert r64["iters_fwd"][e] == ro["iters_fwd"][k] and r64["iters_adj"][e] == ro["iters_adj"][k]
    assert RC.rel_inf(rs["syn"], ro["syn"]) <= 1e-12
    for p in np.unique(sub["ev_iper"]):
        assert RC.rel_inf(rs["adj_s"][p], ro["adj_s"][p]) <= 1e-12
    assert abs(rs["chi"] - ro["chi"]) <= 1e-12 * abs(ro["chi"])


@pytest.mark.parametrize("prec,bar", [(_lib.PREC_F32, 1e-5), (_lib.PREC_F32P, 5e-6)])
def test_large_fp32_per_period_error(large, prec, bar):
    b, r64 = large
    with PP.EikonalBatch(b, precision=prec) as bt:
        bt.run()
        r = bt.results()
        _, _, nres = bt.conv()
    errs = _period_errors(r, r64)
    es = RC.rel_inf(r["syn"], r64["syn"])
    print(f"precision {prec}: per-period kernel error max {errs.max():.3e} median {np.median(errs):.3e}, "
          f"syn {es:.3e}, {nres} FP64 re-solves", flush=True)
    np.testing.assert_array_equal(r["iters_fwd"], r64["iters_fwd"])
    assert errs.max() <= bar
    assert es <= 1e-7
