"""Grids wider than 512 columns (the reference accepts any nx >= 3,
src/eikonal_solver.cpp:35-36): every path sweeps them with 8 columns per lane
on a work-group of 2, 4 or 8 waves per pair (wide_waves), the lanes' edge
values crossing through LDS.

FP64 bar (per-solve calls and FP64 batches, the reference's operation order):
1e-12 normalised against the oracle with equal iteration counts, as for the
one-wave grids (test_gpu_parity.py).  FP32 batches: the FP32 bar against the
FP64 batch (syn 1e-6, equal forward iteration counts; kernels 1e-5 where the
source is within ~500 cells of every node, 5e-5 on the 700- and 1030-column
grids, whose FP32 error grows with the cell count, ~1.5e-5 at 1,000 cells).
Above 4096 columns the calls are refused (SAT_EINVAL / ValueError).
"""
import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import PREC_F32, PREC_F32P, PREC_F64, RUN_GRAPH, RUN_PER_SOLVE
from surfattpp_amd import _lib
from surfattpp_amd import eikonal as E
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth
from tests import refcases as RC

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def _medium(nx, ny, seed):
    xx, yy = np.linspace(-8, 8, nx), np.linspace(-1.5, 1.5, ny)
    rng = np.random.default_rng(seed)
    fun = 1.0 / (3.0 + 0.3 * rng.uniform(size=(nx, ny)))
    a = 1.0 + 0.05 * rng.uniform(size=(nx, ny))
    b = 1.0 + 0.05 * rng.uniform(size=(nx, ny))
    c = 0.02 * rng.uniform(-1, 1, size=(nx, ny))
    return xx, yy, a, b, c, fun


@pytest.mark.parametrize("nx,ny", [(600, 200), (1024, 64), (2100, 9)])
def test_per_solve_wide_matches_oracle(nx, ny):
    xx, yy, a, b, c, fun = _medium(nx, ny, nx + ny)
    x0, y0 = xx[nx // 3] + 0.3 * (xx[1] - xx[0]), yy[ny // 2] - 0.2 * (yy[1] - yy[0])
    T, it, _ = O.forward(xx, yy, a, b, c, fun, x0, y0)
    Tg = E.FSM_UW_PS_lonlat_2d(xx, yy, a, b, c, fun, x0, y0)
    assert E.last_info["iters"] == it
    assert RC.rel_inf(Tg, T) <= TOL64
    xr, yr, sa = RC.receivers_for(xx, yy, 9, nx)
    Ta, ita = O.adjoint(xx, yy, a, b, c, T, xr, yr, sa)
    Tag = E.FSM_O1_JSE_lonlat_2d(xx, yy, a, b, c, T, xr, yr, sa)
    assert E.last_info["iters"] == ita
    assert RC.rel_inf(Tag, Ta) <= TOL64
    Tf = E.FSM_UW_PS_lonlat_2d(xx, yy, a, b, c, fun, x0, y0, precision=PREC_F32)
    assert RC.rel_inf(Tf, T) <= 1e-6
    print(f"{nx}x{ny}: {it} forward, {ita} adjoint rounds", flush=True)


def _wide_batch(nx, ny, nsta, nper, seed):
    b = synth.config_grf(seed=seed, nx=nx, ny=ny, nsta=nsta, nperiod=nper, spacing_km=2.0)
    return dict(b, m12=b["m12"] + 0.004, m11=b["m11"] * 0.99)  # a general metric


def test_fp64_batch_wide_matches_oracle_every_mode():
    b = _wide_batch(600, 31, 5, 2, 11)
    ro = O.forward_for_event(b)
    with PP.EikonalBatch(b) as bt:
        assert bt.waves() == [2]
        out = {}
        for mode in (0, RUN_PER_SOLVE, RUN_GRAPH):
            bt.run(mode)
            out[mode] = bt.results()
    r = out[0]
    np.testing.assert_array_equal(r["iters_fwd"], ro["iters_fwd"])
    np.testing.assert_array_equal(r["iters_adj"], ro["iters_adj"])
    assert RC.rel_inf(r["syn"], ro["syn"]) <= TOL64
    assert RC.rel_inf(r["adj_s"], ro["adj_s"]) <= TOL64
    assert abs(r["chi"] - ro["chi"]) <= TOL64 * abs(ro["chi"])
    for mode in (RUN_PER_SOLVE, RUN_GRAPH):  # the same kernels' arithmetic: the same bits
        np.testing.assert_array_equal(out[mode]["adj_s"], r["adj_s"])
        np.testing.assert_array_equal(out[mode]["syn"], r["syn"])


@pytest.mark.parametrize("prec", [PREC_F32, PREC_F32P])
def test_fp32_batch_wide_and_narrow_grids(prec):
    wide = _wide_batch(1030, 23, 5, 2, 12)
    iso = synth.config_grf(seed=13, nx=700, ny=19, nsta=4, nperiod=2, spacing_km=2.0)
    narrow = synth.config_grf(seed=14, nx=61, ny=41, nsta=4, nperiod=2, spacing_km=2.0)
    grids = [wide, iso, narrow]
    with PP.EikonalBatch(grids, precision=PREC_F64) as bt:
        bt.run()
        r64 = bt.results()
    with PP.EikonalBatch(grids, precision=prec) as bt:
        assert bt.waves()[:2] == [4, 2]
        bt.run()
        r32 = bt.results()
        # a wide band: many pairs re-solved in FP64 (on the wide grids' own geometry)
        bt.set_f32_band(0.9)
        bt.run()
        rres = bt.results()
        _, _, nres = bt.conv()
    # iteration counts equal, syn 1e-6; per-period kernels 1e-5 on the grid of
    # 61 columns, 5e-5 on the 700- and 1030-column grids: the FP32 sweeps'
    # error grows with the number of cells between source and node (about
    # 1.5e-5 at 1,000 cells)
    np.testing.assert_array_equal(r32["iters_fwd"], r64["iters_fwd"])
    assert RC.rel_inf(r32["syn"], r64["syn"]) <= 1e-6
    for gi, bar in enumerate((5e-5, 5e-5, 1e-5)):
        errs = [RC.rel_inf(r32["adj_s"][gi][p], r64["adj_s"][gi][p]) for p in range(r64["adj_s"][gi].shape[0])]
        print(f"precision {prec} grid {gi}: kernel error {max(errs):.2e}", flush=True)
        assert max(errs) <= bar
    print(f"precision {prec}: {nres} of {len(r64['iters_fwd'])} pairs re-solved", flush=True)
    assert nres > 0
    np.testing.assert_array_equal(rres["iters_fwd"], r64["iters_fwd"])
    for a, c in zip(rres["adj_s"], r64["adj_s"]):
        assert RC.rel_inf(a, c) <= 1e-6  # the FP64 re-solves' Ta is stored in FP32


def test_wider_than_4096_columns_is_refused():
    nx, ny = 4097, 9
    xx, yy, a, b, c, fun = _medium(nx, ny, 1)
    with pytest.raises(_lib.SatError):
        E.FSM_UW_PS_lonlat_2d(xx, yy, a, b, c, fun, 0.0, 0.0)
    g = synth.config_grf(seed=3, nx=nx, ny=ny, nsta=2, nperiod=1, spacing_km=1.0)
    with pytest.raises(RuntimeError):
        PP.EikonalBatch(g)
