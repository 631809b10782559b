"""The isotropic combine_kernels restatement (oracle/eik_oracle.c,
src/preproc.cpp:196-342), pinned without the oracle's own formula:

* layout: one nonzero sensitivity at (ix, iy, k, iper) of the reference's
  RowMajor [nx][ny][nz][nper] tensor lands at (ix, iy, k) of the RowMajor
  [nx][ny][nz] output, scaled by -adj_s[iper][iy][ix];
* the chain-rule factors of the vs-only mode equal central differences of the
  Brocher relations vs -> vp and vp -> rho (include/utils.h, vs2vp / vp2rho);
* the sums over periods equal a vectorised numpy evaluation (different order:
  to rounding);
* rho_scaling: rho = 0.33 * vs.
"""
import numpy as np

from oracle import oracle as O


def vs2vp(vs):  # include/utils.h (Brocher)
    return 0.9409 + 2.0947 * vs - 0.8206 * vs**2 + 0.2683 * vs**3 - 0.0251 * vs**4


def vp2rho(vp):
    return 1.6612 * vp - 0.4721 * vp**2 + 0.0671 * vp**3 - 0.0043 * vp**4 + 0.000106 * vp**5


def _inputs(nx, ny, nz, nper, seed):
    r = np.random.default_rng(seed)
    adj = r.normal(size=(nper, ny, nx))
    s = [r.normal(size=(nx, ny, nz, nper)) for _ in range(3)]
    vs = 2.5 + 2.0 * r.uniform(size=(nx, ny, nz))
    return adj, s, vs, vs2vp(vs)


def test_layout_single_spike():
    nx, ny, nz, nper = 5, 4, 3, 2
    adj = np.arange(1, nper * ny * nx + 1, dtype=float).reshape(nper, ny, nx)
    s = [np.zeros((nx, ny, nz, nper)) for _ in range(3)]
    s[0][3, 1, 2, 1] = 2.0
    vs = np.full((nx, ny, nz), 3.0)
    kvs, kvp, krho = O.combine_kernels(adj, *s, vs, vs2vp(vs), mode=1)
    want = np.zeros((nx, ny, nz))
    want[3, 1, 2] = -adj[1, 1, 3] * 2.0
    np.testing.assert_array_equal(kvs, want)
    assert not kvp.any() and not krho.any()


def test_chain_rule_factors_match_brocher_derivatives():
    nx, ny, nz = 3, 2, 4
    vs = np.linspace(2.0, 4.6, nx * ny * nz).reshape(nx, ny, nz)
    vp = vs2vp(vs)
    adj = -np.ones((1, ny, nx))
    one, zero = np.ones((nx, ny, nz, 1)), np.zeros((nx, ny, nz, 1))
    h = 1e-6
    dab_fd = (vs2vp(vs + h) - vs2vp(vs - h)) / (2 * h)
    dra_fd = (vp2rho(vp + h) - vp2rho(vp - h)) / (2 * h)
    kvp, _, _ = O.combine_kernels(adj, zero, one, zero, vs, vp, mode=0)   # K = s_vp * dab
    krho, _, _ = O.combine_kernels(adj, zero, zero, one, vs, vp, mode=0)  # K = s_rho * dra * dab
    assert np.abs(kvp - dab_fd).max() <= 1e-7 * np.abs(dab_fd).max()  # central differences: ~1e-9
    assert np.abs(krho - dra_fd * dab_fd).max() <= 1e-7 * np.abs(dra_fd * dab_fd).max()


def test_sums_match_vectorised_formula():
    nx, ny, nz, nper = 7, 6, 5, 4
    adj, s, vs, vp = _inputs(nx, ny, nz, nper, 3)
    att = np.transpose(adj, (2, 1, 0))[:, :, None, :]  # (nx, ny, 1, nper)
    dab = (vs2vp(vs + 1e-6) - vs2vp(vs - 1e-6)) / 2e-6
    dra = (vp2rho(vp + 1e-6) - vp2rho(vp - 1e-6)) / 2e-6
    S = s[0] + s[1] * dab[..., None] + s[2] * (dra * dab)[..., None]
    want0 = -(att * S).sum(axis=3)
    k0, _, _ = O.combine_kernels(adj, *s, vs, vp, mode=0)
    assert np.abs(k0 - want0).max() <= 1e-7 * np.abs(want0).max()
    k1 = O.combine_kernels(adj, *s, vs, vp, mode=1)
    for kk, sk in zip(k1, s):
        w = -(att * sk).sum(axis=3)
        assert np.abs(kk - w).max() <= 1e-13 * np.abs(w).max()
    kv, kp, kr = O.combine_kernels(adj, *s, vs, vp, mode=2)
    np.testing.assert_array_equal(kv, k1[0])
    np.testing.assert_array_equal(kp, k1[1])
    np.testing.assert_array_equal(kr, k1[0] * 0.33)
