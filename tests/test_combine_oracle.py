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


def _aniso_inputs(nx, ny, nz, nper, seed):
    r = np.random.default_rng(seed)
    return dict(r1=0.3 * r.normal(size=(nx, ny, nper)), r2=0.3 * r.normal(size=(nx, ny, nper)),
                adj_xi=r.normal(size=(nper, ny, nx)), adj_eta=r.normal(size=(nper, ny, nx)),
                svel=3.0 + r.uniform(size=(nper, ny, nx)), sen_gc=r.normal(size=(nx, ny, nz, nper)),
                sen_gs=r.normal(size=(nx, ny, nz, nper)))


def test_ex_without_options_equals_isotropic():
    nx, ny, nz, nper = 6, 5, 4, 3
    adj, s, vs, vp = _inputs(nx, ny, nz, nper, 5)
    for mode in (0, 1, 2):
        ref = O.combine_kernels(adj, *s, vs, vp, mode=mode)
        ex = O.combine_kernels_ex(adj, *s, vs, vp, mode=mode)
        for k, v in zip(("vs", "vp", "rho"), ref):
            if k in ex:
                np.testing.assert_array_equal(ex[k], v)


def test_gc_gs_are_derivatives_of_the_scaled_phase_terms():
    """Per node, the Gc and Gs terms (src/preproc.cpp:319-336) are d/dr1 and
    d/dr2 of F = att/sqrt(D) + (axi r1 + aeta r2)/(sv D), D = 1 + r1^2 + r2^2:
    with unit sensitivities, central differences of F pin every sign and
    power of D and the roles of r1 / r2 and xi / eta."""
    nx, ny, nz, nper = 4, 3, 1, 1
    adj, s, vs, vp = _inputs(nx, ny, nz, nper, 6)
    a = _aniso_inputs(nx, ny, nz, nper, 7)
    a["sen_gc"] = np.ones((nx, ny, nz, nper))
    a["sen_gs"] = np.ones((nx, ny, nz, nper))
    out = O.combine_kernels_ex(adj, *s, vs, vp, mode=1, **a)
    att = adj[0].T[:, :, None]
    axi, aeta, sv = a["adj_xi"][0].T[:, :, None], a["adj_eta"][0].T[:, :, None], a["svel"][0].T[:, :, None]

    def F(r1, r2):
        D = 1.0 + r1 * r1 + r2 * r2
        return att / np.sqrt(D) + (axi * r1 + aeta * r2) / (sv * D)
    r1, r2, h = a["r1"], a["r2"], 1e-6
    dF1 = (F(r1 + h, r2) - F(r1 - h, r2)) / (2 * h)
    dF2 = (F(r1, r2 + h) - F(r1, r2 - h)) / (2 * h)
    assert np.abs(out["gc"] - dF1).max() <= 1e-8 * np.abs(dF1).max()
    assert np.abs(out["gs"] - dF2).max() <= 1e-8 * np.abs(dF2).max()


def test_aniso_isotropic_limit_and_scale():
    """r1 = r2 = 0: Gc = sum_p s_gc axi / sv, Gs = sum_p s_gs aeta / sv and the
    vs/vp/rho kernels are the isotropic ones; nonzero r1, r2 scale the vs, vp
    and rho terms by 1/sqrt(1 + r1^2 + r2^2) (src/preproc.cpp:253-256)."""
    nx, ny, nz, nper = 5, 4, 3, 2
    adj, s, vs, vp = _inputs(nx, ny, nz, nper, 8)
    a = _aniso_inputs(nx, ny, nz, nper, 9)
    z = dict(a, r1=np.zeros_like(a["r1"]), r2=np.zeros_like(a["r2"]))
    out = O.combine_kernels_ex(adj, *s, vs, vp, mode=1, **z)
    iso = O.combine_kernels(adj, *s, vs, vp, mode=1)
    np.testing.assert_array_equal(out["vs"], iso[0])
    T = lambda p: np.transpose(p, (2, 1, 0))[:, :, None, :]  # (nper, ny, nx) -> (nx, ny, 1, nper)  # noqa: E731
    want_gc = (a["sen_gc"] * T(a["adj_xi"]) / T(a["svel"])).sum(axis=3)
    want_gs = (a["sen_gs"] * T(a["adj_eta"]) / T(a["svel"])).sum(axis=3)
    assert np.abs(out["gc"] - want_gc).max() <= 1e-13 * np.abs(want_gc).max()
    assert np.abs(out["gs"] - want_gs).max() <= 1e-13 * np.abs(want_gs).max()
    sc = 1.0 / np.sqrt(1.0 + a["r1"] ** 2 + a["r2"] ** 2)[:, :, None, :]
    out = O.combine_kernels_ex(adj, *s, vs, vp, mode=1, **a)
    for k, sk in (("vs", s[0]), ("vp", s[1]), ("rho", s[2])):
        w = -(T(adj) * sc * sk).sum(axis=3)
        assert np.abs(out[k] - w).max() <= 1e-13 * np.abs(w).max()


def test_aniso_layout_single_spike():
    nx, ny, nz, nper = 5, 4, 3, 2
    adj = np.zeros((nper, ny, nx))
    s = [np.zeros((nx, ny, nz, nper)) for _ in range(3)]
    vs = np.full((nx, ny, nz), 3.0)
    a = dict(r1=np.zeros((nx, ny, nper)), r2=np.zeros((nx, ny, nper)), adj_xi=np.zeros((nper, ny, nx)),
             adj_eta=np.zeros((nper, ny, nx)), svel=np.full((nper, ny, nx), 2.0),
             sen_gc=np.zeros((nx, ny, nz, nper)), sen_gs=np.zeros((nx, ny, nz, nper)))
    a["adj_xi"][1, 2, 3] = 4.0       # period 1, iy 2, ix 3
    a["sen_gc"][3, 2, 1, 1] = 0.5    # ix 3, iy 2, k 1, period 1
    out = O.combine_kernels_ex(adj, *s, vs, vs2vp(vs), mode=1, **a)
    want = np.zeros((nx, ny, nz))
    want[3, 2, 1] = 4.0 / 2.0 * 0.5
    np.testing.assert_array_equal(out["gc"], want)
    assert not out["gs"].any()


def test_density_kernel_is_the_vs_expression_with_adj_den():
    nx, ny, nz, nper = 6, 5, 4, 3
    adj, s, vs, vp = _inputs(nx, ny, nz, nper, 10)
    den = np.random.default_rng(11).normal(size=(nper, ny, nx))
    a = _aniso_inputs(nx, ny, nz, nper, 12)
    for mode, kw in ((0, {}), (1, {}), (1, a)):
        out = O.combine_kernels_ex(adj, *s, vs, vp, mode=mode, adj_den=den, **kw)
        ref = O.combine_kernels_ex(den, *s, vs, vp, mode=mode, **kw)
        np.testing.assert_array_equal(out["den"], ref["vs"])
