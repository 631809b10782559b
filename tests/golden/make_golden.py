"""Writes tests/golden/twin_vectors.npz from the reference's Fortran twin.

The twin (tests/eikonal_test/eikon_solver_2d.f90 of the reference, compiled
from its own source into oracle/_ref/ by oracle/Makefile, default REAL
promoted to 8 bytes) is run on small seeded cases; inputs and its outputs are
stored so the oracle stays pinned where the twin cannot be built.

  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "twin_vectors.npz")


def cases():
    rng = np.random.default_rng(20240611)
    out = []
    # (nx, ny, lon range, lat range, anisotropy amplitude, velocity contrast, source, receivers)
    specs = [
        (41, 31, (-10.0, 10.0), (-7.5, 7.5), 0.0, 0.0, (0.37, -0.81), 6),
        (37, 45, (-3.0, 3.0), (-4.0, 4.0), 0.08, 0.10, (-0.55, 0.92), 9),
        (33, 33, (-1.0, 1.0), (-1.0, 1.0), 0.05, 0.06, (0.11, 0.07), 12),
    ]
    for nx, ny, xr, yr, an, dv, (x0, y0), nrec in specs:
        xx = np.linspace(*xr, nx)
        yy = np.linspace(*yr, ny)
        lon = np.deg2rad(xx)[:, None]
        lat = np.deg2rad(yy)[None, :]
        spha = 1.0 + an * np.sin(2.0 * lon / np.deg2rad(xr[1] - xr[0]) * 6) * np.cos(lat)
        sphb = 1.0 + an * np.cos(1.5 * lon / np.deg2rad(xr[1] - xr[0]) * 6) * np.sin(lat * 3)
        sphc = 0.3 * an * np.sin(lon + lat) + 0.0 * lon
        vel = 3.2 * (1.0 + dv * np.sin(5.0 * lon / np.deg2rad(xr[1] - xr[0])) * np.cos(4.0 * lat / np.deg2rad(yr[1] - yr[0])))
        fun = 1.0 / vel
        dx, dy = xx[1] - xx[0], yy[1] - yy[0]
        rx = rng.uniform(xx[0] + 2 * dx, xx[-1] - 2 * dx, nrec)
        ry = rng.uniform(yy[0] + 2 * dy, yy[-1] - 2 * dy, nrec)
        ra = rng.normal(size=nrec)
        out.append(dict(xx=xx, yy=yy, spha=spha, sphb=sphb, sphc=sphc, fun=fun, x0=x0, y0=y0,
                        xr=rx, yr=ry, adj=ra))
    return out


def main():
    if not O.twin_available("r8"):
        raise SystemExit("oracle/_ref/libtwin_r8.so missing: run `make -C oracle ref`")
    arrays = {}
    for i, c in enumerate(cases()):
        T = O.twin_forward(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"], "r8")
        Ta = O.twin_adjoint(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], T, c["xr"], c["yr"], c["adj"], "r8")
        for k, v in c.items():
            arrays[f"c{i}_{k}"] = np.asarray(v)
        arrays[f"c{i}_T"] = T
        arrays[f"c{i}_Ta"] = Ta
    arrays["ncase"] = np.array(len(cases()))
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
