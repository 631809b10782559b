"""GPU parity of the uneven topography batch: grids of 101^2, 161^2, 241^2 and
321^2 nodes at 1.5 km in one batch, anisotropic metric from the smoothed
relief (synth.config_topo), 12 periods each.

FP64: a subset of events of every grid in one heterogeneous batch against the
FP64 oracle (1e-12, identical iteration counts).  FP32: the whole batch against
the FP64 GPU path (per-period kernel 1e-5, syn 1e-6, forward iterations equal).
"""
import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import PREC_F32
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth
from tests import refcases as RC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def topo():
    return synth.config_topo()


def test_topo_uneven_subset_f64_vs_oracle(topo):
    subs = []
    for g in topo:
        ne = len(g["ev_iper"])
        subs.append(synth.subset_events(g, [0, ne // 3, (2 * ne) // 3 + 1, ne - 1]))
    with PP.EikonalBatch(subs) as bt:
        bt.run()
        r = bt.results()
    off = 0
    for gi, g in enumerate(subs):
        ro = O.forward_for_event(g)
        nr = len(g["rec_lon"])
        print(g["meta"]["name"], "iters", ro["iters_fwd"].tolist(), ro["iters_adj"].tolist(), flush=True)
        assert RC.rel_inf(r["syn"][off:off + nr], ro["syn"]) <= 1e-12
        np.testing.assert_array_equal(r["iters_fwd"][bt.ev_slices[gi]], ro["iters_fwd"])
        np.testing.assert_array_equal(r["iters_adj"][bt.ev_slices[gi]], ro["iters_adj"])
        used = np.unique(g["ev_iper"])
        for p in used:
            assert RC.rel_inf(r["adj_s"][gi][p], ro["adj_s"][p]) <= 1e-12
        off += nr


def test_topo_uneven_full_f32_vs_f64(topo):
    with PP.EikonalBatch(topo) as b64, PP.EikonalBatch(topo, precision=PREC_F32) as b32:
        b64.run()
        b32.run()
        r64, r32 = b64.results(), b32.results()
        _, _, nres = b32.conv()
        ms64, ms32 = b64.last_ms()[0], b32.last_ms()[0]
    ne = sum(len(g["ev_iper"]) for g in topo)
    errs = []
    for gi in range(len(topo)):
        for p in range(r64["adj_s"][gi].shape[0]):
            errs.append(RC.rel_inf(r32["adj_s"][gi][p], r64["adj_s"][gi][p]))
    print(f"{ne} pairs: FP64 {ms64:.1f} ms, FP32 {ms32:.1f} ms, K err max {max(errs):.3e}, "
          f"syn err {RC.rel_inf(r32['syn'], r64['syn']):.3e}, FP64 re-solves {nres}", flush=True)
    np.testing.assert_array_equal(r32["iters_fwd"], r64["iters_fwd"])
    assert max(errs) <= 1e-5
    assert RC.rel_inf(r32["syn"], r64["syn"]) <= 1e-6


def test_dispersion_shape_500_pairs():
    # 121x141 at 2 km, 25 stations x 20 periods: FP64 = oracle on a subset, FP32 within the bar
    b = synth.config_dispersion_shape()
    sub = synth.subset_events(b, [0, 137, 311, 499])
    with PP.EikonalBatch(sub) as bt:
        bt.run()
        r = bt.results()
    ro = O.forward_for_event(sub)
    assert RC.rel_inf(r["syn"], ro["syn"]) <= 1e-12
    np.testing.assert_array_equal(r["iters_fwd"], ro["iters_fwd"])
    for p in np.unique(sub["ev_iper"]):
        assert RC.rel_inf(r["adj_s"][p], ro["adj_s"][p]) <= 1e-12
    with PP.EikonalBatch(b) as b64, PP.EikonalBatch(b, precision=PREC_F32) as b32:
        b64.run()
        b32.run()
        r64, r32 = b64.results(), b32.results()
    errs = [RC.rel_inf(r32["adj_s"][p], r64["adj_s"][p]) for p in range(20)]
    print("500 pairs: K err max", max(errs), flush=True)
    np.testing.assert_array_equal(r32["iters_fwd"], r64["iters_fwd"])
    assert max(errs) <= 1e-5
