"""The anisotropic kernel terms of accumulate_kernels (src/preproc.cpp:37-65)
on the GPU (sat_batch_aniso_kernels) against the oracle composed per event
the way preproc::forward_for_event runs it (src/preproc.cpp:71-142): the FP64
forward solve, receiver interpolation, adjoint solve and mask of the oracle,
then gradient_2_geo (include/utils.h:595-649) and the reference's xi/eta
terms on the topography metric of an uneven surface.

Tolerance: FP64 batches the same bits as the oracle (the GPU's T and Ta are
the oracle's bits and the terms follow its operations in its order); FP32
batches 1e-4 normalised (their masked Ta planes are held in FP32, each
per-period kernel within 1e-5 of FP64).
"""
import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import _lib
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth as S

pytestmark = pytest.mark.gpu


def _topo_batch():
    b = S.config_grf(seed=9, nx=45, ny=39, nsta=5, nperiod=2, spacing_km=1.5)
    lon, lat = b["lon"], b["lat"]
    km = S.R_EARTH * S.DEG2RAD
    x = (lon[None, :] - lon[0]) * km
    y = (lat[:, None] - lat[0]) * km
    z = 4.0 * np.sin(x / 9.0) * np.cos(y / 13.0) + 0.4 * np.exp(-((x - 30.0) ** 2 + (y - 25.0) ** 2) / 60.0)
    m11, m12, m22 = S.topo_metric(z, lon, lat, None)
    nper = b["svel"].shape[0]
    for k, m in (("m11", m11), ("m12", m12), ("m22", m22)):
        b[k] = np.repeat(m[None], nper, axis=0)
    return b


def _oracle_aniso(b, rot=None):
    """rot (optional list): receives per period the sums of Ta |grad T|^2 cos(2 phi)
    and Ta |grad T|^2 sin(2 phi), phi the direction of grad T."""
    lon, lat = b["lon"], b["lat"]
    nper, ny, nx = b["svel"].shape
    xi, eta = np.zeros((nper, ny, nx)), np.zeros((nper, ny, nx))
    if rot is not None:
        rot[:] = [np.zeros((nper, ny, nx)), np.zeros((nper, ny, nx))]
    for e in range(len(b["ev_iper"])):
        p = int(b["ev_iper"][e])
        m11, m12, m22, sv = (b[k][p] for k in ("m11", "m12", "m22", "svel"))
        x0, y0 = float(b["ev_lon"][e]), float(b["ev_lat"][e])
        T, _, _ = O.forward(lon, lat, m11.T, m22.T, (-m12).T, (1.0 / sv).T, x0, y0)
        o, n = int(b["ev_rec_off"][e]), int(b["ev_nrec"][e])
        rl, rb = b["rec_lon"][o:o + n], b["rec_lat"][o:o + n]
        syn = O.interp2d(lon, lat, T, rl, rb)
        adj = b["weight"][o:o + n] * (syn - b["tt_obs"][o:o + n])
        Ta, _ = O.adjoint(lon, lat, m11.T, m22.T, (-m12).T, T, rl, rb, adj)
        Ta = O.mask(lon, lat, Ta, x0, y0).T
        tx, ty = S.gradient_geo(np.ascontiguousarray(T.T), lon, lat)
        if rot is not None:
            phi, g2 = np.arctan2(ty, tx), np.hypot(tx, ty) ** 2
            rot[0][p] += Ta * g2 * np.cos(2.0 * phi)
            rot[1][p] += Ta * g2 * np.sin(2.0 * phi)
        a, bb, c = m11, m22, -m12
        tx2, ty2, txty = tx * tx, ty * ty, tx * ty
        xi[p] += Ta * (tx2 * (-a * a + c * c) + 2.0 * txty * (a * c - bb * c) + ty2 * (-c * c + bb * bb))
        eta[p] += Ta * (2.0 * (a * c) * tx2 - 2.0 * (a * bb + c * c) * txty + 2.0 * (bb * c) * ty2)
    return xi, eta


def _rel(g, r):
    return float(np.abs(g - r).max() / np.abs(r).max())


@pytest.mark.parametrize("prec,tol", [(_lib.PREC_F64, 1e-9), (_lib.PREC_F32, 1e-4)])
def test_aniso_terms_match_oracle(prec, tol):
    b = _topo_batch()
    rx, re = _oracle_aniso(b)
    with PP.EikonalBatch(b, precision=prec) as bt:
        bt.run()
        xi, eta = bt.aniso_kernels(b["m11"], b["m12"], b["m22"])
        xi2, eta2 = bt.aniso_kernels(b["m11"], b["m12"], b["m22"])
    ex, ee = _rel(xi, rx), _rel(eta, re)
    print(f"precision {prec}: adj_xi rel err {ex:.2e}, adj_eta rel err {ee:.2e}", flush=True)
    assert ex <= tol and ee <= tol
    if prec == _lib.PREC_F64:  # the oracle's operations in its order: the same bits
        np.testing.assert_array_equal(xi, rx)
        np.testing.assert_array_equal(eta, re)
    np.testing.assert_array_equal(xi2, xi)
    np.testing.assert_array_equal(eta2, eta)


def test_aniso_isotropic_metric_and_errors():
    """a = b = 1, c = 0 (no topography): the terms reduce to the 2-phi rotation
    of the gradient, adj_xi = -Ta |grad T|^2 cos(2 phi), adj_eta =
    -Ta |grad T|^2 sin(2 phi), which pins the composed oracle's signs and the
    roles of Tx and Ty."""
    b = S.config_grf(seed=3, nx=33, ny=29, nsta=4, nperiod=1, spacing_km=1.5)
    assert np.all(b["m11"] == 1.0) and np.all(b["m22"] == 1.0) and np.all(b["m12"] == 0.0)
    rot = []
    rx, re = _oracle_aniso(b, rot)
    assert _rel(rx, -rot[0]) <= 1e-12 and _rel(re, -rot[1]) <= 1e-12
    with PP.EikonalBatch(b) as bt:
        bt.run()
        xi, eta = bt.aniso_kernels(b["m11"], b["m12"], b["m22"])
        with pytest.raises(ValueError):
            bt.aniso_kernels(b["m11"][:, :-1], b["m12"], b["m22"])
        with pytest.raises(ValueError):
            bt.aniso_kernels(b["m11"], b["m12"], b["m22"], grid=1)
        out = np.empty_like(b["m11"])
        ptrs = [np.ascontiguousarray(b[k]).ctypes.data for k in ("m11", "m22", "m12")]
        assert _lib.load().sat_batch_aniso_kernels(bt._h, 1, *ptrs, out.ctypes.data, out.ctypes.data) == _lib.SAT_EINVAL
    assert _rel(xi, rx) <= 1e-9 and _rel(eta, re) <= 1e-9
    with PP.EikonalBatch(b, calc_adj=False) as bt:
        bt.run()
        with pytest.raises(_lib.SatError):
            bt.aniso_kernels(b["m11"], b["m12"], b["m22"])


def test_aniso_refuses_chunked_batches():
    """A run split into chunks reuses the workspace: the earlier chunks' fields
    are gone at its end, so the terms are refused rather than built from them."""
    g1 = S.config_grf(seed=31, nx=41, ny=33, nsta=6, nperiod=2, spacing_km=1.5)
    g2 = S.config_grf(seed=32, nx=131, ny=57, nsta=5, nperiod=2, spacing_km=1.5)
    with PP.EikonalBatch([g1, g2], workspace_cap=1 << 20) as bt:
        bt.run()
        with pytest.raises(_lib.SatError):
            bt.aniso_kernels(g1["m11"], g1["m12"], g1["m22"])


def test_aniso_full_workload_timed():
    """The 201x201 workload (100 stations x 30 periods, 3,000 pairs), FP32
    batch: the terms over all 3,000 resident field pairs, twice (same bits)."""
    import time
    b = S.config_grf()
    with PP.EikonalBatch(b, precision=_lib.PREC_F32) as bt:
        bt.run()
        xi, eta = bt.aniso_kernels(b["m11"], b["m12"], b["m22"])
        t0 = time.perf_counter()
        xi2, eta2 = bt.aniso_kernels(b["m11"], b["m12"], b["m22"])
        dt = time.perf_counter() - t0
    print(f"201x201 x 30 periods, 3000 events: {dt * 1e3:.2f} ms per call (host copies included)", flush=True)
    assert np.isfinite(xi).all() and np.isfinite(eta).all() and np.abs(xi).max() > 0.0
    np.testing.assert_array_equal(xi2, xi)
    np.testing.assert_array_equal(eta2, eta)


def test_aniso_after_per_solve_run_and_before_any_run():
    """The per-pair kernel (SAT_RUN_PER_SOLVE) keeps T in its own scratch; it
    must also leave each event's FP64 T where the terms read it.  Refused before
    the first run (no field is on the device yet)."""
    b = _topo_batch()
    rx, re = _oracle_aniso(b)
    with PP.EikonalBatch(b, precision=_lib.PREC_F64) as bt:
        with pytest.raises(_lib.SatError):
            bt.aniso_kernels(b["m11"], b["m12"], b["m22"])
        bt.run()
        xb, eb = bt.aniso_kernels(b["m11"], b["m12"], b["m22"])
        bt.run(_lib.RUN_PER_SOLVE)
        xs, es = bt.aniso_kernels(b["m11"], b["m12"], b["m22"])
    print(f"per-solve vs batched: xi equal {np.array_equal(xs, xb)}, eta equal {np.array_equal(es, eb)}; "
          f"vs oracle {_rel(xs, rx):.2e} {_rel(es, re):.2e}", flush=True)
    assert _rel(xs, rx) <= 1e-9 and _rel(es, re) <= 1e-9
    assert _rel(xs, xb) <= 1e-9 and _rel(es, eb) <= 1e-9


def test_aniso_after_resolve_pool_overflow():
    """FP32 batch whose band sends every pair to an FP64 re-solve: beyond the
    re-solve pool (64 entries) the per-pair kernel re-solves them, and its T
    must reach the terms as the pool's does."""
    b = S.config_grf(seed=5, nx=31, ny=29, nsta=36, nperiod=2, spacing_km=2.0)
    lon, lat = b["lon"], b["lat"]
    km = S.R_EARTH * S.DEG2RAD
    x = (lon[None, :] - lon[0]) * km
    y = (lat[:, None] - lat[0]) * km
    z = 2.0 * np.sin(x / 11.0) * np.cos(y / 7.0)
    m11, m12, m22 = S.topo_metric(z, lon, lat, None)
    for k, m in (("m11", m11), ("m12", m12), ("m22", m22)):
        b[k] = np.repeat(m[None], 2, axis=0)
    ne = len(b["ev_iper"])
    assert ne > 64
    rx, re = _oracle_aniso(b)
    with PP.EikonalBatch(b, precision=_lib.PREC_F32) as bt:
        bt.set_f32_band(0.999)
        for mode in (_lib.RUN_BATCHED, _lib.RUN_PER_SOLVE):
            bt.run(mode)
            _, _, nres = bt.conv()
            xi, eta = bt.aniso_kernels(b["m11"], b["m12"], b["m22"])
            print(f"mode {mode}: {nres} of {ne} pairs re-solved in FP64; "
                  f"xi {_rel(xi, rx):.2e} eta {_rel(eta, re):.2e}", flush=True)
            assert nres > 64
            assert _rel(xi, rx) <= 1e-5 and _rel(eta, re) <= 1e-5
