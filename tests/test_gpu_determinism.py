"""Batched runs are deterministic: the same batch run twice (and in a second
batch object) gives the same bits for syn, the per-period kernels and chi, in
FP64 and in FP32 (where the FP64 re-solves take pool slots in whatever order
the pairs finish).  A batch's second run pops its queues in a different order
(longest pairs of the first run first, order_kernel) than its first run
(batch order): the results must not depend on it."""
import numpy as np
import pytest

from surfattpp_amd import PREC_F32, PREC_F64
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [PREC_F64, PREC_F32])
def test_repeat_runs_identical(prec):
    b = synth.config_grf(seed=4, nx=201, ny=201, nsta=30, nperiod=10)
    outs = []
    for _ in range(2):
        with PP.EikonalBatch(b, precision=prec) as bt:
            for _ in range(2):
                bt.run()
                outs.append(bt.results())
    print(f"prec {prec}: chi {outs[0]['chi']!r}", flush=True)
    for r in outs[1:]:
        assert np.array_equal(r["syn"], outs[0]["syn"])
        assert np.array_equal(r["adj_s"], outs[0]["adj_s"])
        assert r["chi"] == outs[0]["chi"]
        assert np.array_equal(r["iters_fwd"], outs[0]["iters_fwd"])
