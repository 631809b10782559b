"""Speculative FP64 solves (SAT_OPT_SPECULATE): the pairs the band rule
re-solved in a recent run are solved in FP64 from the start of the next run,
beside the FP32 forward sweeps.  A speculated pair's result is its FP64
solution, so every forward iteration count still equals the FP64 batch's (which
equals the oracle) and the results stay within the FP32 bar; when the inputs
repeat, the speculated pairs are exactly the band rule's and the run gives the
same bits as a run without speculation."""
import numpy as np
import pytest

from surfattpp_amd import PREC_F32, PREC_F64, RUN_BATCHED, RUN_GRAPH
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth

pytestmark = pytest.mark.gpu

TOL_K, TOL_T = 1e-5, 1e-6


def _rel(a, r):
    return float(np.abs(a - r).max() / np.abs(r).max())


def _same(r, q):
    for k in ("syn", "iters_fwd", "iters_adj", "flags"):
        np.testing.assert_array_equal(r[k], q[k], err_msg=k)
    np.testing.assert_array_equal(r["adj_s"], q["adj_s"])
    assert r["chi"] == q["chi"]


def _within(r, ref):
    np.testing.assert_array_equal(r["iters_fwd"], ref["iters_fwd"])
    for p in range(ref["adj_s"].shape[0]):
        assert _rel(r["adj_s"][p], ref["adj_s"][p]) <= TOL_K
    assert _rel(r["syn"], ref["syn"]) <= TOL_T


@pytest.fixture(scope="module")
def batch():
    # 201x201, 30 stations (sources) x 10 periods: a few pairs fall within the band
    return synth.config_grf(seed=4, nx=201, ny=201, nsta=30, nperiod=10)


@pytest.fixture(scope="module")
def ref64(batch):
    with PP.EikonalBatch(batch, precision=PREC_F64) as b:
        b.run()
        return b.results()


def test_repeated_inputs_same_bits(batch, ref64):
    with PP.EikonalBatch(batch, precision=PREC_F32) as b:
        assert b.get_option("speculate") == 1
        b.run()
        r1 = b.results()
        n1 = b.conv()[2]
        outs = []
        for _ in range(3):
            b.run()
            outs.append((b.results(), b.conv()[2], b.phase_ms()))
    print("band re-solves", n1, "then", [o[1] for o in outs], "fp64 phase ms",
          [round(o[2]["fp64_resolves"], 3) for o in outs], flush=True)
    assert n1 > 0, "the workload has no pair within the band"
    _within(r1, ref64)
    for r, n, _ in outs:
        assert n == n1  # the speculated pairs are the band rule's
        _same(r, r1)
    with PP.EikonalBatch(batch, precision=PREC_F32) as b:
        b.set_option("speculate", 0)
        for _ in range(2):
            b.run()
            _same(b.results(), r1)
            assert b.conv()[2] == n1


def test_after_model_update(batch):
    """The band pairs of a perturbed model are not those of the first: the
    speculated pairs that no longer need FP64 keep their FP64 result, the band
    rule re-solves the new ones, and everything stays within the FP32 bar of an
    FP64 batch on the new model."""
    rng = np.random.default_rng(5)
    sv2 = batch["svel"] * (1.0 + 0.01 * rng.standard_normal(batch["svel"].shape))
    b2 = dict(batch, svel=sv2)
    with PP.EikonalBatch(b2, precision=PREC_F64) as b:
        b.run()
        ref2 = b.results()
    with PP.EikonalBatch(b2, precision=PREC_F32) as b:
        b.run()
        fresh = b.results()
        n_fresh = b.conv()[2]
    with PP.EikonalBatch(batch, precision=PREC_F32) as b:
        b.run()
        b.run()
        n_old = b.conv()[2]
        b.update_media(0, b2["svel"], b2["m11"], b2["m12"], b2["m22"])
        b.run()
        r = b.results()
        n_new = b.conv()[2]
    print("band pairs: old model", n_old, "new model", n_fresh, "speculated + band after the update", n_new,
          flush=True)
    _within(r, ref2)
    _within(fresh, ref2)
    assert n_new >= n_fresh


def test_pool_overflow(batch, ref64):
    """A band of 0.99 marks most pairs: the speculated ones fill the FP64 pool,
    the band rule sends the rest to the pool's remainder and the per-pair
    kernel."""
    sub = synth.subset_events(batch, list(range(0, 120)))
    with PP.EikonalBatch(sub, precision=PREC_F64) as b:
        b.run()
        ref = b.results()
    with PP.EikonalBatch(sub, precision=PREC_F32) as b:
        b.set_f32_band(0.99)
        outs = []
        for _ in range(3):
            b.run()
            outs.append((b.results(), b.conv()[2]))
    print("re-solved or speculated per run", [n for _, n in outs], flush=True)
    for r, n in outs:
        assert n > 64
        _within(r, ref)


def test_graph_replays_equal_eager(batch):
    """The speculated solves are inside the captured graph: the first replay
    speculates the eager run's band pairs, and every replay gives its bits."""
    with PP.EikonalBatch(batch, precision=PREC_F32) as b:
        b.run(mode=RUN_BATCHED)
        first = b.results()
        for _ in range(3):
            b.run(mode=RUN_GRAPH)
            _same(b.results(), first)
            b.run(mode=RUN_BATCHED)
            _same(b.results(), first)
