"""GPU combine_kernels (sat_combine_kernels) against the oracle: the same FP64
operations in the same order, so the same bits, in all three modes; and the
201x201 x 30 depths x 30 periods size, timed (device kernel)."""
import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import preproc as PP
from tests.test_combine_oracle import _inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_combine_equals_oracle(mode):
    adj, s, vs, vp = _inputs(23, 17, 9, 6, 11 + mode)
    got = PP.combine_kernels(adj, *s, vs, vp, mode=mode)
    want = O.combine_kernels(adj, *s, vs, vp, mode=mode)
    np.testing.assert_array_equal(got[0], want[0])
    if mode:
        np.testing.assert_array_equal(got[1], want[1])
        np.testing.assert_array_equal(got[2], want[2])
    else:
        assert got[1] is None and got[2] is None


def test_combine_full_size_timed():
    nx = ny = 201
    nz, nper = 30, 30
    adj, s, vs, vp = _inputs(nx, ny, nz, nper, 5)
    t = {}
    k0, _, _ = PP.combine_kernels(adj, *s, vs, vp, mode=0, timing=t)
    want, _, _ = O.combine_kernels(adj, *s, vs, vp, mode=0)
    np.testing.assert_array_equal(k0, want)
    nbytes = 8 * (3 * s[0].size + adj.size + 2 * vs.size + k0.size)
    print(f"combine 201x201x30x30: {t['kernel_ms']:.3f} ms, {nbytes / t['kernel_ms'] / 1e6:.0f} GB/s", flush=True)


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_combine_reads_device_kernels(mode):
    """sat_batch_combine_kernels on a batch's device kernels equals
    sat_combine_kernels on the same kernels copied to the host, bit for bit,
    for each grid of an uneven batch."""
    from surfattpp_amd import synth
    grids = [synth.config_grf(seed=7, nx=41, ny=37, nsta=5, nperiod=3),
             synth.config_grf(seed=8, nx=33, ny=45, nsta=4, nperiod=2)]
    with PP.EikonalBatch(grids) as bt:
        bt.run()
        res = bt.results()
        for gi, g in enumerate(grids):
            K = res["adj_s"][gi]
            nper, ny, nx = K.shape
            _, s, vs, vp = _inputs(nx, ny, 7, nper, 30 + gi + mode)
            want = PP.combine_kernels(K, *s, vs, vp, mode=mode)
            got = bt.combine_kernels(*s, vs, vp, mode=mode, grid=gi)
            for a, b in zip(got, want):
                if b is None:
                    assert a is None
                else:
                    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("aniso,kden", [(True, False), (False, True), (True, True)])
def test_combine_ex_equals_oracle(mode, aniso, kden):
    """sat_combine_kernels_ex (anisotropic Gc/Gs and density branches of
    combine_kernels, src/preproc.cpp:246-338) equals the oracle's restatement
    bit for bit in every mode."""
    from tests.test_combine_oracle import _aniso_inputs
    nx, ny, nz, nper = 9, 7, 6, 5
    adj, s, vs, vp = _inputs(nx, ny, nz, nper, 30 + mode)
    kw = _aniso_inputs(nx, ny, nz, nper, 40 + mode) if aniso else {}
    if kden:
        kw["adj_den"] = np.random.default_rng(50 + mode).normal(size=(nper, ny, nx))
    ref = O.combine_kernels_ex(adj, *s, vs, vp, mode=mode, **kw)
    got = PP.combine_kernels_ex(adj, *s, vs, vp, mode=mode, **kw)
    assert set(got) == set(ref)
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
