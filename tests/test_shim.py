"""The C++20 shim (include/surfatt_eik.hpp): compiles and links against the
library on the CPU; on the GPU its results equal the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests import refcases as RC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "surfattpp_amd")


def _build(tmp_path):
    exe = str(tmp_path / "shim_driver")
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "shim_driver.cpp"), "-o", exe,
           "-L", LIBDIR, "-lsurfatt_eik", f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_shim_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libsurfatt_eik.so")):
        pytest.skip("library not built")
    _build(tmp_path)


@pytest.mark.gpu
def test_shim_matches_oracle(tmp_path):
    xx, yy, spha, sphb, sphc, fun, x0, y0 = RC.case9_inputs()
    xr, yr, sa = RC.receivers_for(xx, yy, 20, 3)
    nx, ny = xx.size, yy.size
    inp = tmp_path / "case.bin"
    with open(inp, "wb") as f:
        f.write(np.array([nx, ny], dtype=np.int32).tobytes())
        f.write(np.array([x0, y0], dtype=np.float64).tobytes())
        f.write(np.array([xr.size], dtype=np.int32).tobytes())
        for a in (xx, yy):
            f.write(np.asarray(a, dtype=np.float64).tobytes())
        for m in (spha, sphb, sphc, fun):
            f.write(np.asfortranarray(m, dtype=np.float64).tobytes(order="F"))
        for a in (xr, yr, sa):
            f.write(np.asarray(a, dtype=np.float64).tobytes())
    out = tmp_path / "out.bin"
    exe = _build(tmp_path)
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = np.fromfile(out, dtype=np.float64)
    T = res[: nx * ny].reshape((nx, ny), order="F")
    Ta = res[nx * ny:].reshape((nx, ny), order="F")
    To, _, _ = O.forward(xx, yy, spha, sphb, sphc, fun, x0, y0)
    Tao, _ = O.adjoint(xx, yy, spha, sphb, sphc, To, xr, yr, sa)
    Tao = O.mask(xx, yy, Tao, x0, y0)
    assert RC.rel_inf(T, To) <= 1e-12
    assert RC.rel_inf(Ta, Tao) <= 1e-12
