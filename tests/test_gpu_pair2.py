"""The two-pairs-per-wave FP32 forward sweeps (isotropic grids of at most 32
lanes x 7 columns, one wave per pair otherwise): each half of a wave sweeps
its own pair in the reference's Gauss-Seidel order, so the results are held to
the FP32 path's bar against the FP64 path (which equals the oracle,
test_gpu_parity.py): per-period kernels 1e-5, syn 1e-6, chi 1e-5 relative,
every forward iteration count equal -- and to the same bar against the
one-pair-per-wave sweeps (SURFATT_PAIR2=0).

Cases: odd pair counts (a half idles at the end), chunked runs, graph replays,
and few sweep slots (SURFATT_SWEEP_WAVES_PER_CU=1), where every half takes
several pairs from the work queue in turn while the other half is mid-pair.
"""
import os

import numpy as np
import pytest

from surfattpp_amd import PREC_F32, PREC_F64, RUN_GRAPH
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth
from tests.test_gpu_waves import _close

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _pair2_on(monkeypatch):
    # the library takes two pairs per wave where its model expects them to be
    # faster; these cases take them wherever the grid allows
    monkeypatch.setenv("SURFATT_PAIR2", "1")


def _run(b, precision=PREC_F32, waves=None, cap=0, mode=0, runs=1):
    with PP.EikonalBatch(b, precision=precision, waves=waves, workspace_cap=cap) as bt:
        for _ in range(runs):
            bt.run(mode)
        r = bt.results()
        _, _, nres = bt.conv()
    return r, nres


@pytest.mark.parametrize("nx,ny,nsta", [(201, 57, 5), (61, 71, 4), (224, 31, 3), (33, 29, 3)])
def test_pair2_matches_fp64_and_one_pair_layout(nx, ny, nsta):
    b = synth.config_grf(seed=nx + 3 * ny, nx=nx, ny=ny, nsta=nsta, nperiod=3, spacing_km=1.0)
    n = len(b["ev_iper"])
    sub = synth.subset_events(b, list(range(n - 1 if n % 2 == 0 else n - 2)))  # an odd number of pairs
    assert len(sub["ev_iper"]) % 2 == 1
    for batch in (b, sub):
        r64, _ = _run(batch, PREC_F64)
        r32, n = _run(batch, waves=1)
        _close(r32, r64, f"{nx}x{ny} pair2 vs FP64")
        os.environ["SURFATT_PAIR2"] = "0"
        r1, _ = _run(batch, waves=1)
        os.environ["SURFATT_PAIR2"] = "1"
        _close(r32, r1, f"{nx}x{ny} pair2 vs one pair per wave")
        print(f"{nx}x{ny}, {len(batch['ev_iper'])} pairs: {n} FP64 re-solves", flush=True)


def test_pair2_few_slots_chunks_and_graph(monkeypatch):
    b = synth.config_grf(seed=77, nx=121, ny=101, nsta=24, nperiod=25, spacing_km=1.0)
    assert len(b["ev_iper"]) == 600
    r64, _ = _run(b, PREC_F64)
    whole, _ = _run(b)
    _close(whole, r64, "pair2")
    monkeypatch.setenv("SURFATT_SWEEP_WAVES_PER_CU", "1")  # 256 slots, 128 waves of two halves, 600 pairs
    few, _ = _run(b)
    for k in ("syn", "iters_fwd", "iters_adj", "flags"):
        np.testing.assert_array_equal(few[k], whole[k], err_msg=k)
    np.testing.assert_array_equal(few["adj_s"], whole["adj_s"])
    assert few["chi"] == whole["chi"]
    monkeypatch.delenv("SURFATT_SWEEP_WAVES_PER_CU")
    chunked, _ = _run(b, cap=3 << 20)
    np.testing.assert_array_equal(chunked["adj_s"], whole["adj_s"])
    np.testing.assert_array_equal(chunked["iters_fwd"], whole["iters_fwd"])
    graph, _ = _run(b, mode=RUN_GRAPH, runs=3)
    np.testing.assert_array_equal(graph["adj_s"], whole["adj_s"])
    assert graph["chi"] == whole["chi"]
