"""Python mirror of the reference's ``namespace eikonal`` (include/eikonal_solver.h).

Same names, argument order and meaning as the C++ free functions; fields are
numpy arrays of shape (nx, ny) indexed [ix, iy] (the reference's Eigen
column-major matrices).  Each call runs on the GPU through the C ABI.

  FSM_UW_PS_lonlat_2d   include/eikonal_solver.h:32-40  (src/eikonal_solver.cpp:22)
  FSM_O1_JSE_lonlat_2d  include/eikonal_solver.h:74-83  (src/eikonal_solver.cpp:323)
  mask_uniform_grid     include/eikonal_solver.h:85-89  (src/eikonal_solver.cpp:514)

Errors follow the reference: a grid smaller than 3x3 raises ValueError
(std::invalid_argument at src/eikonal_solver.cpp:35-36); non-convergence and a
source on the boundary are warnings (returned in ``last_info``).
"""
from __future__ import annotations

import numpy as np

from . import _lib

last_info = {"iters": 0, "flags": 0, "l1": 0.0, "linf": 0.0}


def _f(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _v(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def _ptr(a):
    return a.ctypes.data


def _store_info(info):
    last_info.update(iters=info.iters, flags=info.flags, l1=info.l1, linf=info.linf)


def FSM_UW_PS_lonlat_2d(xx_deg, yy_deg, spha, sphb, sphc, fun, x0_deg, y0_deg,
                        precision: int = _lib.PREC_F64):
    """Travel time T (nx, ny) from a point source at (x0_deg, y0_deg)."""
    xx, yy = _v(xx_deg), _v(yy_deg)
    nx, ny = xx.size, yy.size
    if nx < 3 or ny < 3:
        raise ValueError("Grid must be at least 3x3")
    A, B, Cc, F = _f(spha), _f(sphb), _f(sphc), _f(fun)
    for m in (A, B, Cc, F):
        if m.shape != (nx, ny):
            raise ValueError(f"field shape {m.shape} != ({nx}, {ny})")
    T = np.empty((nx, ny), dtype=np.float64, order="F")
    info = _lib.sat_info()
    rc = _lib.load().sat_eik_forward(precision, nx, ny, _ptr(xx), _ptr(yy), _ptr(A), _ptr(B), _ptr(Cc),
                                     _ptr(F), float(x0_deg), float(y0_deg), _ptr(T), info)
    _lib.check("sat_eik_forward", rc)
    _store_info(info)
    return T


def FSM_O1_JSE_lonlat_2d(xx_deg, yy_deg, spha, sphb, sphc, T, xrec_deg, yrec_deg, sourceAdj,
                         precision: int = _lib.PREC_F64):
    """Adjoint-state field Ta (nx, ny) for receiver adjoint sources sourceAdj."""
    xx, yy = _v(xx_deg), _v(yy_deg)
    nx, ny = xx.size, yy.size
    if nx < 3 or ny < 3:
        raise ValueError("Grid must be at least 3x3")
    A, B, Cc, TT = _f(spha), _f(sphb), _f(sphc), _f(T)
    for m in (A, B, Cc, TT):
        if m.shape != (nx, ny):
            raise ValueError(f"field shape {m.shape} != ({nx}, {ny})")
    xr, yr, sa = _v(xrec_deg), _v(yrec_deg), _v(sourceAdj)
    if not (xr.size == yr.size == sa.size):
        raise ValueError("receiver arrays differ in length")
    Ta = np.empty((nx, ny), dtype=np.float64, order="F")
    info = _lib.sat_info()
    rc = _lib.load().sat_eik_backward(precision, nx, ny, _ptr(xx), _ptr(yy), _ptr(A), _ptr(B), _ptr(Cc),
                                      _ptr(TT), xr.size, _ptr(xr) if xr.size else None,
                                      _ptr(yr) if yr.size else None, _ptr(sa) if sa.size else None,
                                      _ptr(Ta), info)
    _lib.check("sat_eik_backward", rc)
    _store_info(info)
    return Ta


def mask_uniform_grid(xx, yy, tableAdj, x0, y0):
    """Zero tableAdj within 1.5 cells of (x0, y0), in place (Fortran-ordered float64)."""
    xv, yv = _v(xx), _v(yy)
    if not (isinstance(tableAdj, np.ndarray) and tableAdj.dtype == np.float64 and tableAdj.flags.f_contiguous):
        raise TypeError("tableAdj must be a Fortran-ordered float64 array (modified in place)")
    rc = _lib.load().sat_eik_mask(xv.size, yv.size, _ptr(xv), _ptr(yv), _ptr(tableAdj), float(x0), float(y0))
    _lib.check("sat_eik_mask", rc)
    return tableAdj
