"""The src_rec wire format (surfattpp_amd/src_rec.py, include/surfatt_srcrec.hpp)
against src/src_rec.cpp:20-474 on the reference's own example tables
(tests/golden/src_rec_*.csv, rows copied from examples/00_checkerboard_iso and
examples/04_hawaii_topo).

Grouping and statistics are checked against an independent evaluation
(pandas); the parser against Python's csv module and hand-made documents with
rapidcsv's rules; the C++ header against the Python mirror, value for value and
byte for byte.
"""
import csv
import io
import math
import os
import subprocess

import numpy as np
import pandas as pd
import pytest

from surfattpp_amd import src_rec as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CHECKER = os.path.join(GOLD, "src_rec_checkerboard.csv")
HAWAII = os.path.join(GOLD, "src_rec_hawaii.csv")
LIBDIR = os.path.join(ROOT, "surfattpp_amd")


# ----------------------------------------------------------------------------
# parser and number conversion
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("path", [CHECKER, HAWAII])
def test_parser_matches_csv_module_on_example_tables(path):
    with open(path, newline="") as f:
        want = list(csv.reader(f))
    with open(path, "rb") as f:
        got = S.parse_csv(f.read())
    assert got == want


def test_parser_follows_rapidcsv_rules():
    doc = b'\xef\xbb\xbfa,b,c\r\n"x, y","he said ""hi""",3\n\n1,2'
    rows = S.parse_csv(doc)
    assert rows[0] == ["a", "b", "c"]                  # BOM skipped, CR dropped
    assert rows[1] == ["x, y", 'he said "hi"', "3"]   # quoted separator, doubled quotes
    assert rows[2] == [""]                             # an empty line is a row
    assert rows[3] == ["1", "2"]                       # last row without LF
    assert S.parse_csv(b"a,b\n") == [["a", "b"]]       # no trailing empty row
    assert S.parse_csv(b' a ,b"c"\n') == [[" a ", 'b"c"']]  # no trimming; quote inside a cell is text


@pytest.mark.parametrize("s,v", [("1.5", 1.5), ("  2", 2.0), ("1.5abc", 1.5), ("-3e2x", -300.0),
                                 ("0x1p-2", 0.25), ("inf", math.inf), ("-Infinity", -math.inf),
                                 (".5", 0.5), ("7.", 7.0), ("1e-400", 0.0)])
def test_stod_prefix_semantics(s, v):
    assert S.stod(s) == v


@pytest.mark.parametrize("s", ["", "abc", " ", "e5", "1e999"])
def test_stod_rejects(s):
    with pytest.raises(ValueError):
        S.stod(s)


def test_stod_nan():
    assert math.isnan(S.stod("nan")) and math.isnan(S.stod("NaN(1)"))


# ----------------------------------------------------------------------------
# load, periods, events
# ----------------------------------------------------------------------------
def _expected_events(path):
    """Events from pandas: one per (evtname, distinct period) with rows in file
    order, keyed "{evt}_{index of the period}", in byte order of the keys."""
    df = pd.read_csv(path, dtype={"evtname": str, "staname": str}, keep_default_na=False)
    periods = np.unique(df["period"].to_numpy())
    pidx = {p: i for i, p in enumerate(periods)}
    out = {}
    for (evt, per), g in df.groupby(["evtname", "period"], sort=False):
        rows = g.index.to_numpy()
        out[f"{evt}_{pidx[per]}"] = (pidx[per], rows, g["evla"].iloc[-1], g["evlo"].iloc[-1], per)
    keys = sorted(out, key=lambda k: k.encode())
    return df, periods, [(k, *out[k]) for k in keys]


@pytest.mark.parametrize("path", [CHECKER, HAWAII])
def test_events_match_independent_grouping(path):
    sr = S.SrcRec().load(path)
    df, periods, want = _expected_events(path)
    assert sr.n_obs() == len(df)
    np.testing.assert_array_equal(sr.periods_info.periods, periods)
    assert list(sr.events_local) == [w[0] for w in want]
    assert sr.nsrc_total() == len(want)
    for key, iper, rows, evla, evlo, per in want:
        e = sr.events_local[key]
        assert e.iper == iper and e.evla == evla and e.evlo == evlo and e.period == per
        np.testing.assert_array_equal(e.rec_indices, rows)
        assert e.syn_data.shape == rows.shape
    # src_name_list: event names in byte order, periods ascending within each
    names = [k.rsplit("_", 1) for k in sr.src_name_list]
    assert names == sorted(names, key=lambda t: (t[0].encode(), int(t[1])))


def test_map_order_puts_period_10_before_2():
    sr = S.SrcRec().load(CHECKER)
    keys = list(sr.events_local)
    assert sr.periods_info.nperiod == 16
    assert keys[:4] == ["MX00_0", "MX00_1", "MX00_10", "MX00_11"]
    assert keys.index("MX00_2") == 8


@pytest.mark.parametrize("path", [CHECKER, HAWAII])
def test_period_statistics(path):
    sr = S.SrcRec().load(path)
    df = pd.read_csv(path)
    g = df.groupby("period")["vel"]
    np.testing.assert_array_equal(sr.periods_info.n_obs, g.count().to_numpy())
    np.testing.assert_allclose(sr.periods_info.meanvel, g.mean().to_numpy(), rtol=1e-14)


def _table(cols, rows):
    return ",".join(cols) + "\n" + "".join(",".join(map(str, r)) + "\n" for r in rows)


def test_optional_columns_fall_back():
    cols = ["tt", "staname", "stla", "stlo", "evtname", "evla", "evlo", "period"]
    rows = [[10.0, "A", 1.0, 0.0, "E", 0.0, 0.0, 5.0], [0.0, "B", 2.0, 0.0, "E", 0.0, 0.0, 5.0]]
    sr = S.SrcRec().loads(_table(cols, rows))
    np.testing.assert_array_equal(sr.weight, [1.0, 1.0])
    # same meridian: the haversine distance is R * dlat exactly up to rounding
    np.testing.assert_allclose(sr.dist, [S.R_EARTH * S.DEG2RAD * 1.0, S.R_EARTH * S.DEG2RAD * 2.0], rtol=1e-14)
    assert sr.vel[0] == sr.dist[0] / 10.0 and sr.vel[1] == 0.0  # tt <= 0 -> 0


def test_a_column_that_does_not_convert_counts_as_absent():
    cols = ["tt", "staname", "stla", "stlo", "evtname", "evla", "evlo", "period", "weight"]
    rows = [[10.0, "A", 1.0, 0.0, "E", 0.0, 0.0, 5.0, 0.5], [9.0, "B", 2.0, 0.0, "E", 0.0, 0.0, 5.0, "x"]]
    sr = S.SrcRec().loads(_table(cols, rows))
    np.testing.assert_array_equal(sr.weight, [1.0, 1.0])


def test_required_column_errors():
    cols = ["tt", "staname", "stla", "stlo", "evtname", "evla", "evlo"]
    with pytest.raises(KeyError):
        S.SrcRec().loads(_table(cols, [[1, "A", 0, 0, "E", 0, 0]]))
    cols.append("period")
    with pytest.raises(IndexError):
        S.SrcRec().loads(_table(cols, [[1, "A", 0, 0, "E", 0, 0, 5], [1, "A"]]))
    with pytest.raises(ValueError):
        S.SrcRec().loads(_table(cols, [[1, "A", 0, 0, "E", 0, 0, "x"]]))


def test_periods_within_1e6_group_together():
    cols = ["tt", "staname", "stla", "stlo", "evtname", "evla", "evlo", "period", "weight", "dist", "vel"]
    rows = [[1, "A", 0, 0, "E", 0, 0, 5.0, 1, 3, 3.0], [1, "B", 0, 0, "E", 0, 0, 5.0 + 5e-7, 1, 3, 4.0]]
    sr = S.SrcRec().loads(_table(cols, rows))
    # two distinct periods (exact), each within 1e-6 of both rows
    assert sr.periods_info.nperiod == 2 and sr.periods_info.n_obs == [2, 2]
    assert list(sr.events_local) == ["E_0", "E_1"]
    np.testing.assert_array_equal(sr.events_local["E_0"].rec_indices, [0, 1])


@pytest.mark.parametrize("size", [1, 2, 3, 7])
def test_rank_deal(size):
    full = S.SrcRec().load(CHECKER)
    seen = []
    for rank in range(size):
        sr = S.SrcRec(rank, size).load(CHECKER)
        assert sr.src_name_list == full.src_name_list
        assert sr.src_name_list_local == [k for i, k in enumerate(full.src_name_list) if i % size == rank]
        assert list(sr.events_local) == sorted(sr.src_name_list_local, key=lambda k: k.encode())
        seen += list(sr.events_local)
    assert sorted(seen) == sorted(full.src_name_list)


def test_build_stas_intersection_and_first_row():
    a = S.SrcRec().loads(_table(["tt", "staname", "stla", "stlo", "evtname", "evla", "evlo", "period"],
                                [[1, "S2", 2, 20, "E", 0, 0, 5], [1, "S1", 1, 10, "E", 0, 0, 5],
                                 [1, "S2", 9, 90, "F", 0, 0, 5]]))
    b = S.SrcRec().loads(_table(["tt", "staname", "stla", "stlo", "evtname", "evla", "evlo", "period"],
                                [[1, "S2", 3, 30, "E", 0, 0, 5]]))
    st = S.build_stas(a, None)
    assert st.stnm == ["S1", "S2"] and st.stla == [1.0, 2.0] and st.stlo == [10.0, 20.0]
    st = S.build_stas(a, b)
    assert st.stnm == ["S2"] and st.stla == [2.0]
    assert S.build_stas(None, None).stnm == []


def test_linspaced_and_model_axes():
    v = S.linspaced(11, 102.05, 102.25)
    assert v[0] == 102.05 and v[-1] == 102.25
    np.testing.assert_allclose(np.diff(v), 0.02, rtol=1e-11)
    w = S.linspaced(5, -156.0, -155.0)   # |high| < |low|: counted back from high
    assert w[0] == -156.0 and w[-1] == -155.0
    np.testing.assert_allclose(np.diff(w), 0.25, rtol=1e-13)
    sr = S.SrcRec().load(CHECKER)
    st = S.build_stas(sr, None)
    lon, lat = S.build_model_axes(st, [0.02, 0.02], 5)
    assert lon[0] == min(st.stlo) - 5 * 0.02 and lon[-1] == max(st.stlo) + 5 * 0.02
    assert lon.size == round((max(st.stlo) - min(st.stlo)) / 0.02) + 11
    assert lat.size == round((max(st.stla) - min(st.stla)) / 0.02) + 11


def test_write_round_trip(tmp_path):
    sr = S.SrcRec().load(CHECKER)
    p = tmp_path / "w.csv"
    sr.write(str(p))
    back = S.SrcRec().load(str(p))
    for k, prec in (("tt", 6), ("stla", 4), ("stlo", 4), ("evla", 4), ("evlo", 4), ("period_all", 2),
                    ("weight", 4), ("dist", 3), ("vel", 6)):
        np.testing.assert_allclose(getattr(back, k), getattr(sr, k), rtol=0, atol=0.51 * 10.0 ** -prec)
    assert back.staname == sr.staname and back.evtname == sr.evtname
    assert list(back.events_local) == list(sr.events_local)
    with pytest.raises(RuntimeError):
        sr.write(str(tmp_path / "f.csv"), is_fwd=True)  # no synthetic times yet


def test_to_batch_and_scatter():
    sr = S.SrcRec().load(CHECKER)
    lon, lat = np.linspace(101.9, 102.5, 31), np.linspace(24.0, 24.6, 31)
    svel = np.ones((sr.periods_info.nperiod, lat.size, lon.size)) * 3.0
    b = sr.to_batch(lon, lat, svel)
    evs = list(sr.events_local.values())
    assert b["ev_nrec"].sum() == b["rec_lon"].size == sum(e.rec_indices.size for e in evs)
    e3 = evs[3]
    sl = slice(b["ev_rec_off"][3], b["ev_rec_off"][3] + b["ev_nrec"][3])
    np.testing.assert_array_equal(b["rec_lon"][sl], sr.stlo[e3.rec_indices])
    np.testing.assert_array_equal(b["tt_obs"][sl], sr.tt[e3.rec_indices])
    assert b["ev_iper"][3] == e3.iper and b["ev_lon"][3] == e3.evlo
    syn = np.arange(b["rec_lon"].size, dtype=np.float64)
    sr.scatter_syn(syn)
    sr.gather_syn_tt()
    np.testing.assert_array_equal(sr.tt_fwd[e3.rec_indices], syn[sl])
    assert sorted(sr.tt_fwd.tolist()) == syn.tolist()  # every row written once
    mean, sd = sr.compute_residual_stats()
    r = np.concatenate([e.syn_data - sr.tt[e.rec_indices] for e in evs])
    assert mean == pytest.approx(r.mean(), rel=1e-12) and sd == pytest.approx(r.std(), rel=1e-9)
    with pytest.raises(ValueError):
        sr.to_batch(lon, lat, svel[:2])


# ----------------------------------------------------------------------------
# the C++ header equals the Python mirror
# ----------------------------------------------------------------------------
def _driver(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libsurfatt_eik.so")):
        pytest.skip("library not built")
    exe = str(tmp_path / "srcrec_driver")
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "srcrec_driver.cpp"), "-o", exe,
           "-L", LIBDIR, "-lsurfatt_eik", f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def python_report(sr, interval, margin):
    """The driver's report, written from the Python mirror."""
    out = io.StringIO()
    g = lambda x: "%.17g" % x  # noqa: E731
    out.write(f"n_obs {sr.n_obs()}\nnsrc_total {sr.nsrc_total()}\nnperiod {sr.periods_info.nperiod}\n")
    for p, m, n in zip(sr.periods_info.periods, sr.periods_info.meanvel, sr.periods_info.n_obs):
        out.write(f"period {g(p)} {g(m)} {n}\n")
    for k, e in sr.events_local.items():
        out.write(f"event {k} {e.iper} {g(e.evla)} {g(e.evlo)} {g(e.period)}"
                  + "".join(f" {r}" for r in e.rec_indices) + "\n")
    for i in range(sr.n_obs()):
        out.write(f"row {g(sr.weight[i])} {g(sr.dist[i])} {g(sr.vel[i])}\n")
    st = S.build_stas(sr, None)
    for n, la, lo in zip(st.stnm, st.stla, st.stlo):
        out.write(f"station {n} {g(la)} {g(lo)}\n")
    lon, lat = S.build_model_axes(st, interval, margin)
    out.write("".join(f"x {g(x)}\n" for x in lon))
    out.write("".join(f"y {g(y)}\n" for y in lat))
    return out.getvalue()


@pytest.mark.parametrize("path,interval", [(CHECKER, 0.02), (HAWAII, 0.01)])
def test_cpp_header_equals_python_mirror(tmp_path, path, interval):
    exe = _driver(tmp_path)
    pre = str(tmp_path / "o")
    r = subprocess.run([exe, path, pre, str(interval), str(interval), "5"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    sr = S.SrcRec().load(path)
    with open(pre + ".txt") as f:
        assert f.read() == python_report(sr, [interval, interval], 5)
    sr.write(str(tmp_path / "py.csv"))
    with open(pre + ".csv", "rb") as a, open(tmp_path / "py.csv", "rb") as b:
        assert a.read() == b.read()
