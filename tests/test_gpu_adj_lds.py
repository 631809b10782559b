"""The LDS tier of the one-wave per-pair adjoint: grids whose swept planes
(and, in FP32, coefficient planes) fit in a work-group's LDS sweep them there
instead of in the slot in global memory -- the per-solve backward
(FSM_O1_JSE_lonlat_2d, src/eikonal_solver.cpp:323-511), the per-pair batch
kernel (RUN_PER_SOLVE) and the FP64 re-solves of FP32 batches.

The tier performs the same updates in the same order, so its results are held
bit for bit to those of the global-memory sweeps (SURFATT_ADJ_LDS=0), and the
FP64 ones to the oracle at 1e-12 with equal iteration counts.  One grid per
case is too large for the tier and takes the global sweeps in the same batch.
"""
import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import PREC_F32, PREC_F64, RUN_PER_SOLVE
from surfattpp_amd import eikonal as E
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth
from tests import refcases as RC

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def _medium(nx, ny, seed):
    xx, yy = np.linspace(-4, 4, nx), np.linspace(-3, 3, ny)
    rng = np.random.default_rng(seed)
    fun = 1.0 / (3.0 + 0.3 * rng.uniform(size=(nx, ny)))
    a = 1.0 + 0.05 * rng.uniform(size=(nx, ny))
    b = 1.0 + 0.05 * rng.uniform(size=(nx, ny))
    c = 0.02 * rng.uniform(-1, 1, size=(nx, ny))
    return xx, yy, a, b, c, fun


@pytest.mark.parametrize("nx,ny", [(61, 71), (33, 29), (130, 20), (90, 140)])
def test_per_solve_backward_lds_matches_global_and_oracle(nx, ny, monkeypatch):
    xx, yy, a, b, c, fun = _medium(nx, ny, nx * ny)
    x0, y0 = xx[nx // 3] + 0.3 * (xx[1] - xx[0]), yy[ny // 2]
    T, _, _ = O.forward(xx, yy, a, b, c, fun, x0, y0)
    xr, yr, sa = RC.receivers_for(xx, yy, 7, nx)
    Ta, ita = O.adjoint(xx, yy, a, b, c, T, xr, yr, sa)
    out = {}
    for lds in ("1", "0"):
        monkeypatch.setenv("SURFATT_ADJ_LDS", lds)
        for prec in (PREC_F64, PREC_F32):
            out[lds, prec] = E.FSM_O1_JSE_lonlat_2d(xx, yy, a, b, c, T, xr, yr, sa, precision=prec)
            out[lds, prec, "it"] = E.last_info["iters"]
    assert out["1", PREC_F64, "it"] == ita
    assert RC.rel_inf(out["1", PREC_F64], Ta) <= TOL64
    for prec in (PREC_F64, PREC_F32):
        np.testing.assert_array_equal(out["1", prec], out["0", prec])
        assert out["1", prec, "it"] == out["0", prec, "it"]
    print(f"{nx}x{ny}: {ita} adjoint rounds", flush=True)


def _grids():
    small = synth.config_grf(seed=21, nx=61, ny=71, nsta=5, nperiod=2, spacing_km=1.0)
    tiny = synth.config_grf(seed=22, nx=33, ny=29, nsta=4, nperiod=2, spacing_km=1.0)
    big = synth.config_grf(seed=23, nx=121, ny=101, nsta=4, nperiod=1, spacing_km=1.0)  # global sweeps
    return [small, tiny, big]


def _run(grids, precision, mode, lds, monkeypatch, band=None):
    monkeypatch.setenv("SURFATT_ADJ_LDS", lds)
    with PP.EikonalBatch(grids, precision=precision) as bt:
        if band is not None:
            bt.set_f32_band(band)
        bt.run(mode)
        r = bt.results()
        _, _, nres = bt.conv()
    return r, nres


def _same(r, q):
    for k in ("syn", "iters_fwd", "iters_adj", "flags"):
        np.testing.assert_array_equal(r[k], q[k], err_msg=k)
    for a, c in zip(r["adj_s"], q["adj_s"]):
        np.testing.assert_array_equal(a, c)
    assert r["chi"] == q["chi"]


def test_per_pair_batch_lds_matches_global_and_oracle(monkeypatch):
    grids = _grids()
    lds, _ = _run(grids, PREC_F64, RUN_PER_SOLVE, "1", monkeypatch)
    glob, _ = _run(grids, PREC_F64, RUN_PER_SOLVE, "0", monkeypatch)
    _same(lds, glob)
    for gi, b in enumerate(grids[:2]):
        ro = O.forward_for_event(b)
        r, _ = _run([b], PREC_F64, RUN_PER_SOLVE, "1", monkeypatch)
        np.testing.assert_array_equal(r["iters_adj"], ro["iters_adj"])
        assert RC.rel_inf(r["adj_s"], ro["adj_s"]) <= TOL64  # a one-grid batch: its kernels
        np.testing.assert_array_equal(r["adj_s"], lds["adj_s"][gi])


def test_fp32_resolves_lds_match_global(monkeypatch):
    grids = _grids()
    # a wide band: many pairs re-solved in FP64 by the per-pair kernel
    lds, nres = _run(grids, PREC_F32, 0, "1", monkeypatch, band=0.9)
    glob, nres0 = _run(grids, PREC_F32, 0, "0", monkeypatch, band=0.9)
    assert nres > 0 and nres == nres0
    _same(lds, glob)
    ps, _ = _run(grids, PREC_F32, RUN_PER_SOLVE, "1", monkeypatch)
    ps0, _ = _run(grids, PREC_F32, RUN_PER_SOLVE, "0", monkeypatch)
    _same(ps, ps0)
    print(f"{nres} of {len(lds['iters_fwd'])} pairs re-solved", flush=True)
