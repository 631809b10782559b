"""Surface-wave dispersion on the GPU: the reference's surfdisp96
(src/surfker/surfdisp96.f) for many layered columns at once, and the
phase/group-velocity maps of a 3-D model (SurfGrid::fwdsurf,
src/surf_grid.cpp:194-245), through ``sat_surfdisp96`` and ``sat_fwdsurf``.

``build_disp_req`` restates surfker::build_disp_req (src/surfdisp.cpp:60-139):
nz depth nodes become nz + 1 REAL*4 layers (thickness to the next node, the
last node's layer repeated, then a zero-thickness half-space).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

IFLSPH = 1  # include/config.h:79: spherical earth
IMODE = 1   # include/config.h:80: fundamental mode


class DispersionError(RuntimeError):
    """No root of the fundamental mode: the reference prints a message and stops
    (src/surfker/surfdisp96.f:343-372)."""


def vs2vp(vs):
    """Brocher (2005) vp(vs), include/utils.h:185-187 (scalar form, element-wise)."""
    vs = np.asarray(vs, np.float64)
    return 0.9409 + 2.0947 * vs - 0.8206 * vs * vs + 0.2683 * vs * vs * vs - 0.0251 * vs * vs * vs * vs


def vp2rho(vp):
    """Brocher (2005) density, include/utils.h:223-225 (scalar form, element-wise)."""
    vp = np.asarray(vp, np.float64)
    return (1.6612 * vp - 0.4721 * vp * vp + 0.0671 * vp * vp * vp - 0.0043 * vp * vp * vp * vp
            + 0.000106 * vp * vp * vp * vp * vp)


def build_disp_req(dep, vs, vp=None, rho=None):
    """(thk, vp, vs, rho) REAL*4 layers of one column (or of columns along the
    last axis) from nz depth nodes; Brocher vp and density when not given."""
    dep = np.asarray(dep, np.float64)
    vs = np.asarray(vs, np.float64)
    if vp is None:
        vp = vs2vp(vs)
        rho = vp2rho(vp)
    vp, rho = np.asarray(vp, np.float64), np.asarray(rho, np.float64)
    nz = dep.size
    if nz < 2 or vs.shape[-1] != nz or vp.shape != vs.shape or rho.shape != vs.shape:
        raise ValueError("build_disp_req: dep, vs, vp, and rho must have the same length (>= 2)")
    thk = np.zeros(nz + 1)
    thk[: nz - 1] = dep[1:] - dep[:-1]
    thk[nz - 1] = thk[nz - 2]
    ext = lambda a: np.concatenate([a, a[..., -1:]], axis=-1)  # noqa: E731
    th = np.broadcast_to(thk, vs.shape[:-1] + (nz + 1,))
    return tuple(np.ascontiguousarray(x, dtype=np.float32) for x in (th, ext(vp), ext(vs), ext(rho)))


def surfdisp96(thk, vp, vs, rho, periods, iflsph: int = IFLSPH, iwave: int = 2, mode: int = IMODE, igr: int = 0,
               return_flags: bool = False, timing: dict | None = None):
    """Velocities (ncol, kmax) of layered columns (ncol, nlayer) -- or (kmax,)
    for one column -- at the periods.  Raises DispersionError where a column has
    no fundamental-mode root, unless return_flags (then (cg, flags))."""
    arrs = [np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32) for x in (thk, vp, vs, rho)]
    one = np.ndim(thk) == 1
    ncol, nlayer = arrs[0].shape
    if any(a.shape != (ncol, nlayer) for a in arrs):
        raise ValueError("thk, vp, vs, rho must have the same shape")
    t = np.ascontiguousarray(periods, dtype=np.float64)
    cg = np.zeros((ncol, t.size))
    flags = np.zeros(ncol, dtype=np.int32)
    ms = C.c_double(0.0)
    rc = _lib.load().sat_surfdisp96(ncol, nlayer, *(a.ctypes.data for a in arrs), int(iflsph), int(iwave),
                                    int(mode), int(igr), t.size, t.ctypes.data, cg.ctypes.data, flags.ctypes.data,
                                    C.byref(ms))
    _lib.check("sat_surfdisp96", rc)
    if timing is not None:
        timing["kernel_ms"] = ms.value
    if not return_flags and flags.any():
        raise DispersionError(f"Dispersion calculation failed for {int(flags.sum())} column(s)")
    out = cg[0] if one else cg
    return (out, flags) if return_flags else out


def fwdsurf(zgrids, vs3d, periods, vp3d=None, rho3d=None, iflsph: int = IFLSPH, iwave: int = 2,
            mode: int = IMODE, igr: int = 0, return_flags: bool = False, timing: dict | None = None):
    """Velocity maps (nper, ny, nx) of the 3-D model vs3d (nx, ny, nz) at the
    depth nodes zgrids; vp3d and rho3d (nx, ny, nz) or Brocher's relations."""
    z = np.ascontiguousarray(zgrids, dtype=np.float64)
    vs = np.ascontiguousarray(vs3d, dtype=np.float64)
    if vs.ndim != 3 or vs.shape[2] != z.size:
        raise ValueError("vs3d must be (nx, ny, nz) with nz = len(zgrids)")
    if (vp3d is None) != (rho3d is None):
        raise ValueError("give both vp3d and rho3d or neither")
    vp = None if vp3d is None else np.ascontiguousarray(vp3d, dtype=np.float64)
    rho = None if rho3d is None else np.ascontiguousarray(rho3d, dtype=np.float64)
    for a in (vp, rho):
        if a is not None and a.shape != vs.shape:
            raise ValueError("vp3d and rho3d must match vs3d")
    nx, ny, nz = vs.shape
    t = np.ascontiguousarray(periods, dtype=np.float64)
    svel = np.zeros((t.size, ny, nx))
    flags = np.zeros((ny, nx), dtype=np.int32)
    ms = C.c_double(0.0)
    rc = _lib.load().sat_fwdsurf(nx, ny, nz, z.ctypes.data, vs.ctypes.data, vp.ctypes.data if vp is not None else None,
                                 rho.ctypes.data if rho is not None else None, t.size, t.ctypes.data, int(iflsph),
                                 int(iwave), int(mode), int(igr), svel.ctypes.data, flags.ctypes.data, C.byref(ms))
    _lib.check("sat_fwdsurf", rc)
    if timing is not None:
        timing["kernel_ms"] = ms.value
    if not return_flags and flags.any():
        raise DispersionError(f"Dispersion calculation failed for {int(flags.sum())} column(s)")
    return (svel, flags) if return_flags else svel
