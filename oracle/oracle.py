"""ctypes bindings for the checker libraries in oracle/.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package surfattpp_amd.

* ``liboracle.so``  -- FP64 C restatement of src/eikonal_solver.cpp and the
  isotropic part of src/preproc.cpp (eik_oracle.c).
* ``_ref/libtwin_r8.so`` / ``_ref/libtwin_r4.so`` -- the reference's Fortran
  twin tests/eikonal_test/eikon_solver_2d.f90 compiled from its own source
  (promoted to 8-byte default REAL, and as shipped).
* ``_ref/libsurfdisp96.so`` -- the reference's surfdisp96 (src/surfker/surfdisp96.f
  with its C binding src/surfker/surfdisp96_c_api.f90) compiled from its own
  source: the checker of the GPU dispersion solver.

All 2-D fields are Fortran-ordered numpy arrays of shape (nx, ny), i.e. the
reference's column-major (ix, iy) storage.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_BENCH = None
_TWIN = {}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

FLAG_NOT_CONVERGED = 1
FLAG_SRC_ON_BOUNDARY = 2


def build() -> None:
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def _p(a):
    return a.ctypes.data_as(_dp)


def _ip_(a):
    return a.ctypes.data_as(_ip)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = _bind(C.CDLL(path))
    return _LIB


def _bind(L):
    L.sat_oracle_forward.restype = C.c_int
    L.sat_oracle_forward.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                     C.c_double, C.c_double, C.c_int, _dp, _ip, _ip]
    L.sat_oracle_adjoint.restype = C.c_int
    L.sat_oracle_adjoint.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                     C.c_int, _dp, _dp, _dp, C.c_int, _dp, _ip]
    L.sat_oracle_mask.restype = None
    L.sat_oracle_mask.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_double]
    L.sat_oracle_interp2d.restype = None
    L.sat_oracle_interp2d.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp]
    L.sat_oracle_combine_kernels.restype = C.c_int
    L.sat_oracle_combine_kernels.argtypes = [C.c_int] * 5 + [_dp] * 9
    L.sat_oracle_combine_kernels_ex.restype = C.c_int
    L.sat_oracle_combine_kernels_ex.argtypes = [C.c_int] * 5 + [_dp] * 20
    L.sat_oracle_forward_for_event.restype = C.c_int
    L.sat_oracle_forward_for_event.argtypes = [
        C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp,
        C.c_int, _ip, _dp, _dp, _ip, _ip, _dp, _dp, _dp, _dp,
        C.c_int, _dp, _dp, _dp, _ip, _ip, _ip]
    L.sat_oracle_forward_for_event_kden.restype = C.c_int
    L.sat_oracle_forward_for_event_kden.argtypes = L.sat_oracle_forward_for_event.argtypes + [_dp]
    return L


def bench_lib():
    """The same restatement built with -O3 -mavx2 -mfma (the reference's Release
    flags, CMakeLists.txt:41, with a portable ISA level): timed as the CPU
    baseline, never used as the parity checker (it forms FMAs)."""
    global _BENCH
    if _BENCH is None:
        path = os.path.join(_HERE, "liboracle_bench.so")
        if not os.path.exists(path):
            build()
        _BENCH = _bind(C.CDLL(path))
    return _BENCH


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _v64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def mutant_lib(n: int):
    """The restatement with planted mistake n (liboracle_mutN.so, eik_oracle.c)."""
    path = os.path.join(_HERE, f"liboracle_mut{n}.so")
    if not os.path.exists(path):
        build()
    return _bind(C.CDLL(path))


def forward(xx_deg, yy_deg, spha, sphb, sphc, fun, x0_deg, y0_deg, max_iter=0, lib_=None):
    """FSM_UW_PS_lonlat_2d (src/eikonal_solver.cpp:22).  Returns (T, iters, flags)."""
    xx, yy = _v64(xx_deg), _v64(yy_deg)
    nx, ny = xx.size, yy.size
    A, B, Cc, F = _f64(spha), _f64(sphb), _f64(sphc), _f64(fun)
    for m in (A, B, Cc, F):
        assert m.shape == (nx, ny)
    T = np.zeros((nx, ny), dtype=np.float64, order="F")
    it = np.zeros(1, dtype=np.int32)
    fl = np.zeros(1, dtype=np.int32)
    rc = (lib_ or lib()).sat_oracle_forward(nx, ny, _p(xx), _p(yy), _p(A), _p(B), _p(Cc), _p(F),
                                  float(x0_deg), float(y0_deg), int(max_iter), _p(T), _ip_(it), _ip_(fl))
    if rc == 22:
        raise ValueError("Grid must be at least 3x3")
    if rc:
        raise MemoryError("oracle forward failed")
    return T, int(it[0]), int(fl[0])


def adjoint(xx_deg, yy_deg, spha, sphb, sphc, T, xrec_deg, yrec_deg, source_adj, max_iter=0, lib_=None):
    """FSM_O1_JSE_lonlat_2d (src/eikonal_solver.cpp:323).  Returns (Ta, iters)."""
    xx, yy = _v64(xx_deg), _v64(yy_deg)
    nx, ny = xx.size, yy.size
    A, B, Cc, TT = _f64(spha), _f64(sphb), _f64(sphc), _f64(T)
    xr, yr, sa = _v64(xrec_deg), _v64(yrec_deg), _v64(source_adj)
    nr = xr.size
    if nr == 0:
        xr = yr = sa = np.zeros(1)
    Ta = np.zeros((nx, ny), dtype=np.float64, order="F")
    it = np.zeros(1, dtype=np.int32)
    rc = (lib_ or lib()).sat_oracle_adjoint(nx, ny, _p(xx), _p(yy), _p(A), _p(B), _p(Cc), _p(TT), nr,
                                  _p(xr), _p(yr), _p(sa), int(max_iter), _p(Ta), _ip_(it))
    if rc == 22:
        raise ValueError("Grid must be at least 3x3")
    if rc:
        raise MemoryError("oracle adjoint failed")
    return Ta, int(it[0])


def mask(xx, yy, Ta, x0, y0):
    """mask_uniform_grid (src/eikonal_solver.cpp:514), in place on a copy."""
    xv, yv = _v64(xx), _v64(yy)
    out = _f64(Ta).copy(order="F")
    lib().sat_oracle_mask(xv.size, yv.size, _p(xv), _p(yv), _p(out), float(x0), float(y0))
    return out


def interp2d(xgrid, ygrid, z, xq, yq):
    """interp2d (include/utils.h:358)."""
    xg, yg = _v64(xgrid), _v64(ygrid)
    Z = _f64(z)
    xqv, yqv = _v64(xq), _v64(yq)
    out = np.zeros(xqv.size)
    if xqv.size:
        lib().sat_oracle_interp2d(xg.size, yg.size, _p(xg), _p(yg), _p(Z), xqv.size, _p(xqv), _p(yqv), _p(out))
    return out


def forward_for_event(batch, is_calc_adj=True, lib=None, kden=False):
    """preproc::forward_for_event (src/preproc.cpp:71) over one grid's events.

    ``batch`` is a surfattpp_amd-style dict (see tests/inputs) with keys
    lon, lat, svel/m11/m12/m22 (nperiod, ny, nx), ev_iper, ev_lon, ev_lat,
    ev_rec_off, ev_nrec, rec_lon, rec_lat, tt_obs, weight.
    Returns dict(syn, adj_s (nperiod, ny, nx), chi, iters_fwd, iters_adj, flags),
    with kden also kden_s (nperiod, ny, nx): is_kden's density kernel
    (src/preproc.cpp:127-136).
    """
    lon, lat = _v64(batch["lon"]), _v64(batch["lat"])
    nx, ny = lon.size, lat.size
    planes = {k: np.ascontiguousarray(batch[k], dtype=np.float64) for k in ("svel", "m11", "m12", "m22")}
    nper = planes["svel"].shape[0]
    for v in planes.values():
        assert v.shape == (nper, ny, nx)
    ev_iper = np.ascontiguousarray(batch["ev_iper"], dtype=np.int32)
    ne = ev_iper.size
    ev_lon, ev_lat = _v64(batch["ev_lon"]), _v64(batch["ev_lat"])
    off = np.ascontiguousarray(batch["ev_rec_off"], dtype=np.int32)
    nrec = np.ascontiguousarray(batch["ev_nrec"], dtype=np.int32)
    rl, rb = _v64(batch["rec_lon"]), _v64(batch["rec_lat"])
    tt, w = _v64(batch["tt_obs"]), _v64(batch["weight"])
    nrt = max(rl.size, 1)
    syn = np.zeros(nrt)
    adj_s = np.zeros((nper, ny, nx))
    chi = np.zeros(1)
    itf = np.zeros(ne, dtype=np.int32)
    ita = np.zeros(ne, dtype=np.int32)
    fl = np.zeros(ne, dtype=np.int32)
    L = lib if lib is not None else globals()["lib"]()
    kden_s = np.zeros((nper, ny, nx)) if kden else None
    rc = L.sat_oracle_forward_for_event_kden(
        nx, ny, _p(lon), _p(lat), nper, _p(planes["svel"]), _p(planes["m11"]), _p(planes["m12"]),
        _p(planes["m22"]), ne, _ip_(ev_iper), _p(ev_lon), _p(ev_lat), _ip_(off), _ip_(nrec),
        _p(rl), _p(rb), _p(tt), _p(w), int(bool(is_calc_adj)), _p(syn), _p(adj_s), _p(chi),
        _ip_(itf), _ip_(ita), _ip_(fl), _p(kden_s) if kden else None)
    if rc:
        raise RuntimeError(f"oracle forward_for_event failed rc={rc}")
    out = dict(syn=syn[: rl.size], adj_s=adj_s, chi=float(chi[0]), iters_fwd=itf, iters_adj=ita, flags=fl)
    if kden:
        out["kden_s"] = kden_s
    return out


# ---------------------------------------------------------------------------
# The reference's Fortran twin (tests/eikonal_test/eikon_solver_2d.f90),
# called through its external-subroutine symbols (all arguments by reference).
# ---------------------------------------------------------------------------

def twin_available(kind="r8"):
    return os.path.exists(os.path.join(_HERE, "_ref", f"libtwin_{kind}.so"))


def _twin(kind):
    if kind not in _TWIN:
        L = C.CDLL(os.path.join(_HERE, "_ref", f"libtwin_{kind}.so"))
        L.fsm_uw_ps_lonlat_2d_.restype = None
        L.fsm_o1_jse_lonlat_2d_.restype = None
        L.mask_.restype = None
        _TWIN[kind] = L
    return _TWIN[kind]


def twin_forward(xx_deg, yy_deg, spha, sphb, sphc, fun, x0_deg, y0_deg, kind="r8"):
    """eikon_solver_2d.f90:764 FSM_UW_PS_lonlat_2d(xx,yy,nx,ny,spha,sphb,sphc,T,fun,x0,y0,u)."""
    xx, yy = _v64(xx_deg), _v64(yy_deg)
    nx, ny = C.c_int(xx.size), C.c_int(yy.size)
    A, B, Cc, F = _f64(spha), _f64(sphb), _f64(sphc), _f64(fun)
    T = np.zeros((xx.size, yy.size), order="F")
    u = np.zeros((xx.size, yy.size), order="F")
    x0, y0 = C.c_double(x0_deg), C.c_double(y0_deg)
    _twin(kind).fsm_uw_ps_lonlat_2d_(_p(xx), _p(yy), C.byref(nx), C.byref(ny), _p(A), _p(B), _p(Cc),
                                     _p(T), _p(F), C.byref(x0), C.byref(y0), _p(u))
    return T


def twin_adjoint(xx_deg, yy_deg, spha, sphb, sphc, T, xrec_deg, yrec_deg, source_adj, kind="r8"):
    """eikon_solver_2d.f90:1276 FSM_O1_JSE_lonlat_2d(xx,yy,nx,ny,spha,sphb,sphc,T,Ta,xrec,yrec,adj,nr)."""
    xx, yy = _v64(xx_deg), _v64(yy_deg)
    nx, ny = C.c_int(xx.size), C.c_int(yy.size)
    A, B, Cc, TT = _f64(spha), _f64(sphb), _f64(sphc), _f64(T)
    xr, yr, sa = _v64(xrec_deg), _v64(yrec_deg), _v64(source_adj)
    nr = C.c_int(xr.size)
    if xr.size == 0:
        xr = yr = sa = np.zeros(1)
    Ta = np.zeros((xx.size, yy.size), order="F")
    _twin(kind).fsm_o1_jse_lonlat_2d_(_p(xx), _p(yy), C.byref(nx), C.byref(ny), _p(A), _p(B), _p(Cc),
                                      _p(TT), _p(Ta), _p(xr), _p(yr), _p(sa), C.byref(nr))
    return Ta


_SURFDISP = None


def surfdisp_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libsurfdisp96.so"))


def surfdisp96(thk, vp, vs, rho, periods, iflsph=1, iwave=2, mode=1, igr=0):
    """The reference's surfdisp96_c (src/surfker/surfdisp96_c_api.f90:5) on one
    layered column: REAL*4 thickness/vp/vs/rho (the last layer the half-space),
    FP64 periods; returns the phase (igr=0) or group velocities per period."""
    global _SURFDISP
    if _SURFDISP is None:
        L = C.CDLL(os.path.join(_HERE, "_ref", "libsurfdisp96.so"))
        L.surfdisp96_c.restype = None
        L.surfdisp96_c.argtypes = [C.c_void_p] * 4 + [C.c_int] * 6 + [C.c_void_p] * 2
        _SURFDISP = L
    a = [np.ascontiguousarray(x, dtype=np.float32) for x in (thk, vp, vs, rho)]
    t = np.ascontiguousarray(periods, dtype=np.float64)
    cg = np.zeros(t.size)
    _SURFDISP.surfdisp96_c(*(x.ctypes.data for x in a), a[0].size, iflsph, iwave, mode, igr, t.size,
                           t.ctypes.data, cg.ctypes.data)
    return cg


def combine_kernels(adj_s, sen_vs, sen_vp, sen_rho, vs3d, vp3d, mode=0):
    """preproc::combine_kernels, isotropic part (src/preproc.cpp:196-342): the
    oracle's loop, periods outermost.  Shapes as surfattpp_amd.preproc.combine_kernels."""
    adj = np.ascontiguousarray(adj_s, dtype=np.float64)
    nper, ny, nx = adj.shape
    s = [np.ascontiguousarray(a, dtype=np.float64) for a in (sen_vs, sen_vp, sen_rho)]
    nz = s[0].shape[2]
    v = [np.ascontiguousarray(a, dtype=np.float64) for a in (vs3d, vp3d)]
    out = [np.zeros((nx, ny, nz)) for _ in range(3)]
    rc = lib().sat_oracle_combine_kernels(nx, ny, nz, nper, int(mode), _p(adj), _p(s[0]), _p(s[1]), _p(s[2]),
                                          _p(v[0]), _p(v[1]), _p(out[0]), _p(out[1]), _p(out[2]))
    if rc:
        raise RuntimeError(f"oracle combine_kernels failed rc={rc}")
    return tuple(out)


def combine_kernels_ex(adj_s, sen_vs, sen_vp, sen_rho=None, vs3d=None, vp3d=None, mode=0, r1=None, r2=None,
                       adj_xi=None, adj_eta=None, svel=None, sen_gc=None, sen_gs=None, adj_den=None):
    """preproc::combine_kernels with the anisotropic (Gc, Gs) and density
    branches (src/preproc.cpp:196-342).  r1, r2: (nx, ny, nper) or None
    (isotropic); adj_xi, adj_eta, svel, adj_den: (nper, ny, nx); sen_*:
    (nx, ny, nz, nper).  Returns dict of the kernels (nx, ny, nz) that the
    options form: vs, vp, rho, gc, gs, den."""
    c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    adj = c(adj_s)
    nper, ny, nx = adj.shape
    s_vs, s_vp, s_rho = c(sen_vs), c(sen_vp), c(sen_rho)
    nz = s_vs.shape[2]
    aniso, kden = r1 is not None, adj_den is not None
    out = {k: np.zeros((nx, ny, nz)) for k in ("vs", "vp", "rho", "gc", "gs", "den")}
    args = [adj, s_vs, s_vp, s_rho, c(vs3d), c(vp3d), c(r1), c(r2), c(adj_xi), c(adj_eta), c(svel), c(sen_gc),
            c(sen_gs), c(adj_den)]
    keep = [a for a in args if a is not None]
    ptr = lambda a: None if a is None else _p(a)  # noqa: E731
    rc = lib().sat_oracle_combine_kernels_ex(nx, ny, nz, nper, int(mode), *(ptr(a) for a in args),
                                             *(_p(out[k]) for k in ("vs", "vp", "rho", "gc", "gs", "den")))
    del keep
    if rc:
        raise RuntimeError(f"oracle combine_kernels_ex failed rc={rc}")
    res = {"vs": out["vs"]}
    if mode:
        res.update(vp=out["vp"], rho=out["rho"])
    if aniso:
        res.update(gc=out["gc"], gs=out["gs"])
    if kden:
        res["den"] = out["den"]
    return res
