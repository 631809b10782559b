"""The FP32 adjoint's delta row masks (delta_kernel marks and zeroes the lane
segments that hold a receiver share; the sweeps read the delta planes only
there): receivers on the grid's edges and corners, whose shares fall in the
last, partly padded lane segment, and a batch whose masks would not fit the
LDS, which reads every delta instead.  Either way the FP32 batch keeps the FP64
batch's (= the oracle's) iteration counts and stays within the FP32 bar."""
import numpy as np
import pytest

from surfattpp_amd import PREC_F32, PREC_F64
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth

pytestmark = pytest.mark.gpu


def _rel(a, r):
    return float(np.abs(a - r).max() / np.abs(r).max())


def _compare(b):
    with PP.EikonalBatch(b, precision=PREC_F64) as t:
        t.run()
        r64 = t.results()
    with PP.EikonalBatch(b, precision=PREC_F32) as t:
        outs = []
        for _ in range(2):  # the second run also speculates the band rule's pairs
            t.run()
            outs.append(t.results())
    for r in outs:
        np.testing.assert_array_equal(r["iters_fwd"], r64["iters_fwd"])
        k = r["adj_s"] if r["adj_s"].ndim == 3 else r["adj_s"][None]
        k64 = r64["adj_s"] if r64["adj_s"].ndim == 3 else r64["adj_s"][None]
        for p in range(k64.shape[0]):
            assert _rel(k[p], k64[p]) <= 1e-5
        assert _rel(r["syn"], r64["syn"]) <= 1e-6
    return r64, outs


def test_receivers_on_edges_and_corners():
    # 201 columns: eight per lane, the last lane's segment half padding
    b = synth.config_grf(seed=7, nx=201, ny=151, nsta=12, nperiod=2)
    b["rec_lon"] = b["rec_lon"].copy()
    b["rec_lat"] = b["rec_lat"].copy()
    lon, lat = b["lon"], b["lat"]
    b["rec_lon"][0], b["rec_lat"][0] = lon[-1], lat[-1]   # north-east corner
    b["rec_lon"][1], b["rec_lat"][1] = lon[0], lat[0]     # south-west corner
    b["rec_lon"][2], b["rec_lat"][2] = lon[-1], lat[40]   # east edge, on a node
    b["rec_lon"][3], b["rec_lat"][3] = lon[-2], lat[-2]   # the last interior cell
    b["rec_lon"][4], b["rec_lat"][4] = lon[100], lat[0]   # south edge
    r64, outs = _compare(b)
    assert np.all(np.isfinite(outs[-1]["syn"]))


def test_planes_read_whole_same_bits(monkeypatch):
    """A batch whose masks would not fit the LDS limit (SURFATT_MASK_LDS=0 here)
    reads every delta, and the set-up zeroes the planes: the same values reach
    the sweeps, so the results are the same bits as with the masks."""
    b = synth.config_grf(seed=8, nx=201, ny=121, nsta=10, nperiod=2)
    with PP.EikonalBatch(b, precision=PREC_F32) as t:
        t.run()
        t.run()
        masked = t.results()
    monkeypatch.setenv("SURFATT_MASK_LDS", "0")
    with PP.EikonalBatch(b, precision=PREC_F32) as t:
        t.run()
        t.run()
        whole = t.results()
    for key in ("syn", "iters_fwd", "iters_adj"):
        np.testing.assert_array_equal(masked[key], whole[key], err_msg=key)
    np.testing.assert_array_equal(masked["adj_s"], whole["adj_s"])
    _compare(b)
