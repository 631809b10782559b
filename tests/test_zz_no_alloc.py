"""sat_eik_forward_noalloc, the allocation-free variant of sat_eik_forward
(device inputs and outputs, planes in a static device array), called through
the C ABI against sat_eik_forward on the same inputs: the same kernels, so the
same bits, on two calls in a row (the second on a smaller grid, which must
not see the first call's planes), in both precisions; a grid beyond the
array's capacity and a non-uniform axis are refused."""
import numpy as np
import pytest

from surfattpp_amd import _lib
from tests import refcases as RC
from tests.devmem import DevArray

pytestmark = pytest.mark.gpu


def _host_forward(prec, xx, yy, spha, sphb, sphc, fun, x0, y0):
    nx, ny = xx.size, yy.size
    args = [np.ascontiguousarray(xx, dtype=np.float64), np.ascontiguousarray(yy, dtype=np.float64)]
    args += [np.asfortranarray(m, dtype=np.float64) for m in (spha, sphb, sphc, fun)]
    T = np.empty((nx, ny), order="F")
    info = _lib.sat_info()
    rc = _lib.load().sat_eik_forward(prec, nx, ny, *(a.ctypes.data for a in args), x0, y0, T.ctypes.data, info)
    _lib.check("sat_eik_forward", rc)
    return T, (info.iters, info.flags, info.l1, info.linf)


def _noalloc_forward(prec, xx, yy, spha, sphb, sphc, fun, x0, y0, expect_rc=_lib.SAT_OK):
    nx, ny = xx.size, yy.size
    dev = [DevArray(a) for a in (xx, yy, spha, sphb, sphc, fun)]
    dT = DevArray(np.full((nx, ny), -1.0))
    try:
        info = _lib.sat_info()
        rc = _lib.load().sat_eik_forward_noalloc(prec, nx, ny, *(d.ptr for d in dev), x0, y0, dT.ptr, info)
        assert rc == expect_rc
        return (dT.get((nx, ny)), (info.iters, info.flags, info.l1, info.linf)) if rc == _lib.SAT_OK else None
    finally:
        for d in dev + [dT]:
            d.free()


def _grf_case(nx, ny, seed):
    rng = np.random.default_rng(seed)
    xx, yy = RC.linspaced(nx, -0.6, 0.6), RC.linspaced(ny, -0.5, 0.5)
    lon, lat = np.meshgrid(xx, yy, indexing="ij")
    fun = 1.0 / (3.2 + 0.2 * np.sin(7.0 * lon + rng.uniform()) * np.cos(5.0 * lat))
    spha = 1.0 + 0.1 * np.sin(3.0 * lon) * np.cos(4.0 * lat)
    sphb = 1.0 + 0.05 * np.cos(2.0 * lon)
    sphc = 0.02 * np.sin(lon + lat)
    return xx, yy, spha, sphb, sphc, fun, float(xx[nx // 3] + 0.3 * (xx[1] - xx[0])), float(yy[ny // 2])


@pytest.mark.parametrize("prec", [_lib.PREC_F64, _lib.PREC_F32])
def test_noalloc_forward_equals_forward_on_two_calls(prec):
    first = RC.case9_inputs()               # 121 x 91
    second = _grf_case(61, 47, 3)            # smaller: must not read the first call's planes
    for case in (first, second):
        T_ref, info_ref = _host_forward(prec, *case)
        T, info = _noalloc_forward(prec, *case)
        print(f"prec {prec}: {case[0].size}x{case[1].size} iters {info[0]} (ref {info_ref[0]}), "
              f"max |dT| {np.abs(T - T_ref).max():.3e}", flush=True)
        np.testing.assert_array_equal(T, T_ref)
        assert info == info_ref


def test_noalloc_refuses_oversized_and_nonuniform_grids():
    big = _grf_case(512, 400, 5)             # 204,800 nodes: beyond the static array
    _noalloc_forward(_lib.PREC_F64, *big, expect_rc=_lib.SAT_EINVAL)
    xx, yy, spha, sphb, sphc, fun, x0, y0 = _grf_case(31, 29, 6)
    bent = xx.copy()
    bent[7] += 0.3 * (xx[1] - xx[0])
    _noalloc_forward(_lib.PREC_F64, bent, yy, spha, sphb, sphc, fun, x0, y0, expect_rc=_lib.SAT_EINVAL)
    # still usable after the refusals
    T, _ = _noalloc_forward(_lib.PREC_F64, xx, yy, spha, sphb, sphc, fun, x0, y0)
    T_ref, _ = _host_forward(_lib.PREC_F64, xx, yy, spha, sphb, sphc, fun, x0, y0)
    np.testing.assert_array_equal(T, T_ref)
