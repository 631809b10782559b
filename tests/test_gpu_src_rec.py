"""The reference's example observation tables through the GPU path: a table
loaded by surfattpp_amd.src_rec (Python) or include/surfatt_srcrec.hpp (C++),
its events run in one batch, against the FP64 oracle.

Grid: the station box widened by 5 cells (examples/*/input_params.yml,
grid_method 0); velocity: each period's mean observed velocity from the table
(the reference's periods_info.meanvel), identity metric.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import PREC_F32, PREC_F64
from surfattpp_amd import preproc as PP
from surfattpp_amd import src_rec as S
from tests import refcases as RC

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
TABLES = [(os.path.join(GOLD, "src_rec_checkerboard.csv"), 0.02), (os.path.join(GOLD, "src_rec_hawaii.csv"), 0.01)]


def _batch(path, interval):
    sr = S.SrcRec().load(path)
    lon, lat = S.build_model_axes(S.build_stas(sr, None), [interval, interval], 5)
    nper = sr.periods_info.nperiod
    svel = np.repeat(sr.periods_info.meanvel, lat.size * lon.size).reshape(nper, lat.size, lon.size)
    return sr, sr.to_batch(lon, lat, svel)


@pytest.mark.parametrize("path,interval", TABLES)
def test_table_batch_matches_oracle(path, interval):
    sr, b = _batch(path, interval)
    print("table", os.path.basename(path), "events", len(b["ev_iper"]), "grid", b["lon"].size, b["lat"].size,
          flush=True)
    ro = O.forward_for_event(b)
    with PP.EikonalBatch(b, precision=PREC_F64) as bt:
        bt.run()
        r64 = bt.results()
    assert RC.rel_inf(r64["syn"], ro["syn"]) <= 1e-12
    assert RC.rel_inf(r64["adj_s"], ro["adj_s"]) <= 1e-12
    assert abs(r64["chi"] - ro["chi"]) <= 1e-12 * abs(ro["chi"])
    with PP.EikonalBatch(b, precision=PREC_F32) as bt:
        bt.run()
        r32 = bt.results()
    assert RC.rel_inf(r32["syn"], r64["syn"]) <= 1e-5
    assert RC.rel_inf(r32["adj_s"], r64["adj_s"]) <= 1e-5
    # syn back into the table: every observation row gets its event's time
    sr.scatter_syn(r64["syn"])
    sr.gather_syn_tt()
    assert np.all(sr.tt_fwd > 0)


@pytest.mark.parametrize("path,interval", TABLES)
def test_cpp_srcrec_forward_for_event(tmp_path, path, interval):
    libdir = os.path.join(ROOT, "surfattpp_amd")
    exe = str(tmp_path / "srcrec_driver")
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "srcrec_driver.cpp"), "-o", exe,
           "-L", libdir, "-lsurfatt_eik", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    pre = str(tmp_path / "o")
    r = subprocess.run([exe, path, pre, str(interval), str(interval), "5", "run"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    vals = {}
    with open(pre + ".txt") as f:
        for line in f:
            k, *v = line.split()
            vals.setdefault(k, []).append(v)
    chi = float(vals["chi"][0][0])
    tt_fwd = np.array([float(v[0]) for v in vals["tt_fwd"]])
    adj = np.array([float(v[0]) for v in vals["adj_s"]])
    sr, b = _batch(path, interval)
    ro = O.forward_for_event(b)
    assert abs(chi - ro["chi"]) <= 1e-12 * abs(ro["chi"])
    assert RC.rel_inf(adj, np.asarray(ro["adj_s"]).ravel()) <= 1e-12
    sr.scatter_syn(ro["syn"])
    sr.gather_syn_tt()
    assert RC.rel_inf(tt_fwd, sr.tt_fwd) <= 1e-12
    mean, sd = (float(x) for x in vals["residual"][0])
    m_o, s_o = sr.compute_residual_stats()
    assert abs(mean - m_o) <= 1e-9 * abs(m_o) and abs(sd - s_o) <= 1e-9 * abs(s_o)
    assert os.path.getsize(pre + "_fwd.csv") > 0
