"""The batched run captured as one hipGraph (SAT_RUN_GRAPH): every replay gives
the same bits as the eager run, with eager runs, synchronisations and input
updates between the replays; a chunked batch and the FP32 re-solve streams are
inside the graph too."""
import numpy as np
import pytest

from surfattpp_amd import PREC_F32, PREC_F64, RUN_BATCHED, RUN_GRAPH
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth

pytestmark = pytest.mark.gpu


def _same(r, q):
    for k in ("syn", "iters_fwd", "iters_adj", "flags"):
        np.testing.assert_array_equal(r[k], q[k], err_msg=k)
    np.testing.assert_array_equal(r["adj_s"], q["adj_s"])
    assert r["chi"] == q["chi"]


@pytest.mark.parametrize("precision,cap", [(PREC_F64, 0), (PREC_F32, 0), (PREC_F32, 3 << 20)])
def test_graph_replays_equal_eager(precision, cap):
    b = synth.config_grf(seed=61, nx=121, ny=101, nsta=12, nperiod=3)
    b2 = dict(b, svel=b["svel"] * 1.01)
    with PP.EikonalBatch(b, precision=precision, workspace_cap=cap) as bt:
        bt.run(mode=RUN_BATCHED)
        eager = bt.results()
        for k in range(3):
            bt.run(mode=RUN_GRAPH)
            _same(bt.results(), eager)
            bt.run(mode=RUN_BATCHED)  # eager work between the replays
            _same(bt.results(), eager)
        ms = bt.last_ms()[0]
        bt.run(mode=RUN_GRAPH)
        assert np.isnan(bt.phase_ms()["fwd_sweeps"]) and bt.last_ms()[0] > 0
        # a new model in place: the graph reads it without a new capture
        bt.update_media(0, b2["svel"], b2["m11"], b2["m12"], b2["m22"])
        bt.run(mode=RUN_BATCHED)
        eager2 = bt.results()
        bt.run(mode=RUN_GRAPH)
        _same(bt.results(), eager2)
        assert not np.array_equal(eager2["syn"], eager["syn"])
    print("eager run ms", ms, flush=True)
