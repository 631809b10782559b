"""Source-receiver observation table: the reference's ``SrcRec``
(include/src_rec.h:44-136, src/src_rec.cpp:20-474) on one process.

The table is the input wire format of the eikonal path: one CSV row per
(event, station, period) observation.  ``load`` reads it the way the
reference does (rapidcsv with its default parameters, include/rapidcsv.h;
numbers through ``std::stod``), derives the per-period statistics
(``get_periods``, src/src_rec.cpp:235-266) and groups the rows into events
keyed ``"{evtname}_{iper}"`` (``get_events``, :141-231).  The events a rank
holds are kept in the order of the reference's ``std::map`` keys, which is
the order ``preproc::forward_for_event`` visits them (src/preproc.cpp:79) and
therefore the order of the chi sum and of the per-period kernel sums.

``to_batch`` turns the events into the batch dict that
``surfattpp_amd.preproc.EikonalBatch`` takes (one grid), ``scatter_syn``
writes a run's synthetic times back into ``events_local[key].syn_data``.
``build_model_axes`` restates the station-box grid of ``ModelGrid``
(src/model_grid.cpp:127-165, grid_method 0) so a table can be run without
further input.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

import numpy as np

R_EARTH = 6371.0
DEG2RAD = 3.14159265358979323846264338327950288 / 180.0

# Columns written by SrcRec::write (src/src_rec.cpp:420-447): name, decimals
# (None for text columns).
WRITE_COLUMNS = (("tt", 6), ("staname", None), ("stla", 4), ("stlo", 4), ("evtname", None), ("evla", 4),
                 ("evlo", 4), ("period", 2), ("weight", 4), ("dist", 3), ("vel", 6))

# std::stod: leading white space, then the longest prefix strtod accepts
# (decimal or hexadecimal floating point, inf/infinity, nan[(chars)]).
_STOD = re.compile(
    r"[ \t\n\v\f\r]*([+-]?(?:0[xX](?:[0-9a-fA-F]+\.?[0-9a-fA-F]*|\.[0-9a-fA-F]+)(?:[pP][+-]?\d+)?"
    r"|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?"
    r"|[iI][nN][fF](?:[iI][nN][iI][tT][yY])?"
    r"|[nN][aA][nN](?:\([0-9A-Za-z_]*\))?))")


def stod(s: str) -> float:
    """std::stod (strtod of the longest valid prefix; an error when nothing
    converts or the value overflows)."""
    m = _STOD.match(s)
    if not m:
        raise ValueError(f"stod: no conversion of {s!r}")
    t = m.group(1)
    low = t.lower().lstrip("+-")
    if low.startswith("0x"):
        sign = -1.0 if t.startswith("-") else 1.0
        mant = low[2:]
        if "p" not in mant:
            mant += "p0"
        v = sign * float.fromhex("0x" + mant)
    elif low.startswith("nan"):
        v = math.copysign(math.nan, -1.0 if t.startswith("-") else 1.0)
    else:
        v = float(t)
    if math.isinf(v) and not low.startswith("inf"):
        raise ValueError(f"stod: {s!r} out of range")
    return v


def parse_csv(data: bytes) -> list[list[str]]:
    """Rows of a CSV file as rapidcsv's parser splits them with its default
    parameters (include/rapidcsv.h:1557-1760): separator ',', quote '"'
    toggling only at the start of a cell, CR dropped, every LF ends a row
    (empty lines included), cells not trimmed, a cell wrapped in quotes
    unquoted with doubled quotes collapsed, a UTF-8 byte-order mark skipped."""
    if data[:3] == b"\xef\xbb\xbf":
        data = data[3:]
    text = data.decode("utf-8", errors="surrogateescape")
    rows: list[list[str]] = []
    row: list[str] = []
    cell: list[str] = []
    quoted = False

    def unquote(c: str) -> str:
        if len(c) >= 2 and c[0] == '"' and c[-1] == '"':
            return c[1:-1].replace('""', '"')
        return c

    for ch in text:
        if ch == '"':
            if not cell or cell[0] == '"':
                quoted = not quoted
            cell.append(ch)
        elif ch == ",":
            if not quoted:
                row.append(unquote("".join(cell)))
                cell = []
            else:
                cell.append(ch)
        elif ch == "\r":
            pass
        elif ch == "\n":
            row.append(unquote("".join(cell)))
            rows.append(row)
            row, cell, quoted = [], [], False
        else:
            cell.append(ch)
    if row or cell:
        row.append(unquote("".join(cell)))
        rows.append(row)
    return rows


def gps2dist(lat0, lon0, lat1, lon1):
    """Haversine distance in km (include/utils.h:286-299), element-wise, with
    the C library's sin/cos/atan2 in the reference's operation order."""
    out = []
    for la0, lo0, la1, lo1 in zip(*(np.atleast_1d(np.asarray(v, np.float64)).tolist()
                                     for v in (lat0, lon0, lat1, lon1))):
        dlat = (la0 - la1) * DEG2RAD
        dlon = (lo1 - lo0) * DEG2RAD
        a = (math.sin(dlat * 0.5) * math.sin(dlat * 0.5)
             + math.sin(dlon * 0.5) * math.sin(dlon * 0.5) * math.cos(la0 * DEG2RAD) * math.cos(la1 * DEG2RAD))
        out.append(R_EARTH * (2.0 * math.atan2(math.sqrt(a), math.sqrt(1.0 - a))))
    return np.array(out, np.float64)


@dataclass
class PeriodInfo:
    """include/src_rec.h:30-35."""
    periods: np.ndarray = field(default_factory=lambda: np.zeros(0))
    meanvel: np.ndarray = field(default_factory=lambda: np.zeros(0))
    n_obs: list = field(default_factory=list)
    nperiod: int = 0


@dataclass
class EventInfo:
    """include/src_rec.h:15-22."""
    evla: float = 0.0
    evlo: float = 0.0
    period: float = 0.0
    iper: int = 0
    rec_indices: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    syn_data: np.ndarray = field(default_factory=lambda: np.zeros(0))


def real_t_equal(a, b, tol: float = 1e-6):
    """include/utils.h:23-25."""
    return np.abs(np.asarray(a) - b) < tol


class SrcRec:
    """One observation table (phase or group velocity) on one process.

    ``rank`` / ``size`` select the events this process holds the way the
    reference deals them to MPI ranks (``select_rank_for_src``,
    include/parallel.h:39-41): event i of ``src_name_list`` goes to rank
    i % size."""

    def __init__(self, rank: int = 0, size: int = 1):
        if size < 1 or not 0 <= rank < size:
            raise ValueError("rank must lie in [0, size)")
        self.rank, self.size = rank, size
        self._n_obs = 0
        self._nsrc_total = 0
        empty = np.zeros(0)
        self.stla = self.stlo = self.evla = self.evlo = self.dist = empty
        self.period_all = self.tt = self.vel = self.weight = empty
        self.tt_fwd = np.zeros(0)
        self.vel_fwd = np.zeros(0)
        self.evtname: list[str] = []
        self.staname: list[str] = []
        self.events_local: dict[str, EventInfo] = {}
        self.src_name_list_local: list[str] = []
        self.src_name_list: list[str] = []
        self.periods_info = PeriodInfo()

    def n_obs(self) -> int:
        return self._n_obs

    def nsrc_total(self) -> int:
        return self._nsrc_total

    # -- src/src_rec.cpp:20-131 ------------------------------------------------
    def load(self, filepath: str) -> "SrcRec":
        """Read the table (columns by header name; ``weight``, ``dist`` and
        ``vel`` optional: weight 1, haversine distance, dist/tt or 0 when
        tt <= 0), then the period statistics and the events."""
        try:
            with open(filepath, "rb") as f:
                rows = parse_csv(f.read())
        except OSError as e:
            raise OSError(f"SrcRec::load: failed to open file {filepath}: {e}") from e
        self._from_rows(rows)
        return self

    def loads(self, text: str) -> "SrcRec":
        """``load`` from the text of a table."""
        self._from_rows(parse_csv(text.encode()))
        return self

    def _from_rows(self, rows):
        header = rows[0] if rows else []
        colidx = {}
        for i, name in enumerate(header):
            colidx[name] = i  # a repeated name maps to its last column, as in rapidcsv
        body = rows[1:]
        n = len(body)

        def column(name: str, numeric: bool):
            if name not in colidx:
                raise KeyError(f"column not found: {name}")
            j = colidx[name]
            out = []
            for r, row in enumerate(body):
                if j >= len(row):
                    raise IndexError(f"requested column index {j} >= {len(row)} (number of columns on row "
                                     f"index {r})")
                out.append(stod(row[j]) if numeric else row[j])
            return np.array(out, dtype=np.float64) if numeric else out

        def optional(name: str):
            try:
                return column(name, True)
            except (KeyError, IndexError, ValueError):
                return None

        self._n_obs = n
        self.stla = column("stla", True)
        self.stlo = column("stlo", True)
        self.evla = column("evla", True)
        self.evlo = column("evlo", True)
        self.period_all = column("period", True)
        self.tt = column("tt", True)
        self.evtname = column("evtname", False)
        self.staname = column("staname", False)
        w = optional("weight")
        self.weight = w if w is not None else np.ones(n)
        d = optional("dist")
        self.dist = d if d is not None else gps2dist(self.stla, self.stlo, self.evla, self.evlo)
        v = optional("vel")
        if v is None:
            with np.errstate(divide="ignore", invalid="ignore"):
                v = np.where(self.tt > 0, self.dist / np.where(self.tt > 0, self.tt, 1.0), 0.0)
        self.vel = v
        self.get_periods()
        self.get_events()

    # -- src/src_rec.cpp:235-266 -----------------------------------------------
    def get_periods(self):
        """Sorted distinct periods; per period the count and mean velocity of
        the rows within 1e-6 of it (summed in row order)."""
        per = np.unique(self.period_all)
        info = PeriodInfo(periods=per, meanvel=np.zeros(per.size), n_obs=[], nperiod=int(per.size))
        for i, p in enumerate(per):
            sel = np.nonzero(np.abs(self.period_all - p) < 1e-6)[0]
            s = 0.0
            for j in sel:
                s += float(self.vel[j])
            info.n_obs.append(int(sel.size))
            info.meanvel[i] = s / sel.size
        self.periods_info = info

    # -- src/src_rec.cpp:141-231 -----------------------------------------------
    def get_events(self):
        """Events keyed "{evtname}_{iper}" for every distinct event name (in
        byte order) and period index with at least one row; a record's rows in
        file order; evla/evlo/period from the event's last row.  The rank keeps
        events i % size == rank of that list, visited in key order."""
        names = sorted(set(self.evtname), key=lambda s: s.encode("utf-8", "surrogateescape"))
        ev_arr = np.array(self.evtname, dtype=object)
        events: dict[str, EventInfo] = {}
        self.src_name_list = []
        for src in names:
            is_src = ev_arr == src
            for iper in range(self.periods_info.nperiod):
                sel = np.nonzero(is_src & real_t_equal(self.periods_info.periods[iper], self.period_all))[0]
                if sel.size == 0:
                    continue
                last = sel[-1]
                key = f"{src}_{iper:d}"
                events[key] = EventInfo(evla=float(self.evla[last]), evlo=float(self.evlo[last]),
                                        period=float(self.period_all[last]), iper=iper,
                                        rec_indices=sel.astype(np.int64), syn_data=np.zeros(sel.size))
                self.src_name_list.append(key)
        self._nsrc_total = len(self.src_name_list)
        self.src_name_list_local = [k for i, k in enumerate(self.src_name_list) if i % self.size == self.rank]
        local = {k: events[k] for k in self.src_name_list_local}
        self.events_local = {k: local[k] for k in sorted(local, key=lambda s: s.encode("utf-8", "surrogateescape"))}

    # -- src/src_rec.cpp:325-368 -----------------------------------------------
    def gather_syn_tt(self):
        """tt_fwd[row] = the event's synthetic time at that row; vel_fwd =
        dist / tt_fwd (this process's events)."""
        self.tt_fwd = np.zeros(self._n_obs)
        for key in self.src_name_list:
            e = self.events_local.get(key)
            if e is not None:
                self.tt_fwd[e.rec_indices] = e.syn_data
        with np.errstate(divide="ignore", invalid="ignore"):
            self.vel_fwd = self.dist / self.tt_fwd

    # -- src/src_rec.cpp:371-398 -----------------------------------------------
    def compute_residual_stats(self):
        """(mean, stddev) of syn - tt over this process's events, summed in
        event-key then receiver order."""
        n = s = sq = 0.0
        for e in self.events_local.values():
            for i, r in enumerate(e.rec_indices):
                res = float(e.syn_data[i]) - float(self.tt[r])
                n += 1.0
                s += res
                sq += res * res
        if n == 0.0:
            return 0.0, 0.0
        mean = s / n
        var = sq / n - mean * mean
        return mean, math.sqrt(max(var, 0.0))

    # -- src/src_rec.cpp:403-459 -----------------------------------------------
    def write(self, filepath: str, is_fwd: bool = False):
        """The table as CSV in the reference's column order and decimals
        (fixed notation); tt and vel are the synthetic ones with is_fwd."""
        if self._n_obs == 0:
            raise RuntimeError("SrcRec::write: no observations to write")
        if is_fwd and (self.tt_fwd.size == 0 or self.vel_fwd.size == 0):
            raise RuntimeError("Forward-modeled travel times (tt_fwd) are not assigned")
        src = dict(tt=self.tt_fwd if is_fwd else self.tt, staname=self.staname, stla=self.stla, stlo=self.stlo,
                   evtname=self.evtname, evla=self.evla, evlo=self.evlo, period=self.period_all,
                   weight=self.weight, dist=self.dist, vel=self.vel_fwd if is_fwd else self.vel)
        cols = []
        for name, prec in WRITE_COLUMNS:
            vals = src[name]
            cols.append(list(vals) if prec is None else [_fixed(float(x), prec) for x in vals])
        lines = [",".join(_csv_cell(n) for n, _ in WRITE_COLUMNS)]
        for r in range(self._n_obs):
            lines.append(",".join(_csv_cell(c[r]) for c in cols))
        with open(filepath, "wb") as f:
            f.write(("\n".join(lines) + "\n").encode("utf-8", "surrogateescape"))

    # -- the path's input -----------------------------------------------------
    def to_batch(self, lon, lat, svel, m11=None, m12=None, m22=None) -> dict:
        """The batch dict of this process's events over one grid, in the order
        forward_for_event visits them.  lon (nx,), lat (ny,) in degrees;
        svel and the metric planes (nperiod, ny, nx) period-major, the period
        index being the events' iper (m11 = m22 = 1, m12 = 0 when omitted)."""
        lon = np.ascontiguousarray(lon, dtype=np.float64)
        lat = np.ascontiguousarray(lat, dtype=np.float64)
        svel = np.ascontiguousarray(svel, dtype=np.float64)
        if svel.ndim != 3 or svel.shape[1:] != (lat.size, lon.size):
            raise ValueError(f"svel must be (nperiod, {lat.size}, {lon.size})")
        nper = svel.shape[0]
        if nper < self.periods_info.nperiod:
            raise ValueError(f"svel holds {nper} periods, the table {self.periods_info.nperiod}")
        ones = np.ones_like(svel)
        m11 = ones if m11 is None else np.ascontiguousarray(m11, dtype=np.float64)
        m22 = ones if m22 is None else np.ascontiguousarray(m22, dtype=np.float64)
        m12 = np.zeros_like(svel) if m12 is None else np.ascontiguousarray(m12, dtype=np.float64)
        evs = list(self.events_local.values())
        nrec = np.array([e.rec_indices.size for e in evs], np.int32)
        off = np.concatenate([[0], np.cumsum(nrec)[:-1]]).astype(np.int32) if evs else np.zeros(0, np.int32)
        rec = np.concatenate([e.rec_indices for e in evs]) if evs else np.zeros(0, np.int64)
        return dict(lon=lon, lat=lat, svel=svel, m11=m11, m12=m12, m22=m22,
                    ev_iper=np.array([e.iper for e in evs], np.int32),
                    ev_lon=np.array([e.evlo for e in evs], np.float64),
                    ev_lat=np.array([e.evla for e in evs], np.float64),
                    ev_rec_off=off, ev_nrec=nrec, rec_idx=rec,
                    rec_lon=self.stlo[rec].copy(), rec_lat=self.stla[rec].copy(),
                    tt_obs=self.tt[rec].copy(), weight=self.weight[rec].copy(),
                    meta=dict(name="src_rec", nevent=len(evs), nperiod=self.periods_info.nperiod))

    def scatter_syn(self, syn) -> None:
        """Write a run's syn (receivers in to_batch order) into syn_data."""
        syn = np.asarray(syn, dtype=np.float64)
        total = sum(e.rec_indices.size for e in self.events_local.values())
        if syn.size != total:
            raise ValueError(f"syn holds {syn.size} values, the events {total}")
        off = 0
        for e in self.events_local.values():
            e.syn_data = syn[off:off + e.rec_indices.size].copy()
            off += e.rec_indices.size


def _fixed(x: float, prec: int) -> str:
    """std::fixed << std::setprecision(prec) of a double (glibc's %.*f)."""
    if math.isnan(x):
        return "-nan" if math.copysign(1.0, x) < 0 else "nan"
    if math.isinf(x):
        return "-inf" if x < 0 else "inf"
    return f"{x:.{prec}f}"


def _csv_cell(s: str) -> str:
    """rapidcsv's writer quotes a cell holding the separator, a space or a
    line feed, doubling its quotes (include/rapidcsv.h:1815-1844)."""
    if "," in s or " " in s or "\n" in s:
        return '"' + s.replace('"', '""') + '"'
    return s


class Stations:
    """include/src_rec.h:24-28."""

    def __init__(self):
        self.stnm: list[str] = []
        self.stla: list[float] = []
        self.stlo: list[float] = []


def build_stas(sr_ph: SrcRec | None, sr_gr: SrcRec | None) -> Stations:
    """SrcRec::build_stas (src/src_rec.cpp:274-323): station name -> the
    coordinates of its first row; the names of both tables when both hold
    rows, else those of the one that does; in name order."""

    def make_map(sr: SrcRec):
        m: dict[str, tuple[float, float]] = {}
        for i in range(sr.n_obs()):
            m.setdefault(sr.staname[i], (float(sr.stla[i]), float(sr.stlo[i])))
        return m

    have_ph = sr_ph is not None and sr_ph.n_obs() > 0
    have_gr = sr_gr is not None and sr_gr.n_obs() > 0
    if have_ph and have_gr:
        mp, mg = make_map(sr_ph), make_map(sr_gr)
        merged = {k: v for k, v in mp.items() if k in mg}
    elif have_ph:
        merged = make_map(sr_ph)
    elif have_gr:
        merged = make_map(sr_gr)
    else:
        merged = {}
    st = Stations()
    for name in sorted(merged, key=lambda s: s.encode("utf-8", "surrogateescape")):
        st.stnm.append(name)
        st.stla.append(merged[name][0])
        st.stlo.append(merged[name][1])
    return st


def linspaced(n: int, low: float, high: float) -> np.ndarray:
    """Eigen 3.4 ``VectorX<double>::LinSpaced(n, low, high)`` for n > 1:
    low + i*step, the last node exactly high, step = (high-low)/(n-1); when
    |high| < |low| it counts back from high (high - (n-1-i)*step, the first
    node exactly low)."""
    if n == 1:
        return np.array([high], np.float64)
    step = (high - low) / float(n - 1)
    i = np.arange(n, dtype=np.float64)
    if abs(high) < abs(low):
        out = high - (float(n - 1) - i) * step
        out[0] = low
    else:
        out = low + i * step
        out[-1] = high
    return out


def build_model_axes(st: Stations, interval, num_grid_margin: int):
    """ModelGrid's lon/lat axes for grid_method 0 (src/model_grid.cpp:127-165):
    the station box widened by num_grid_margin cells, llround((end-beg)/d)+1
    nodes per axis, Eigen LinSpaced."""
    if not st.stlo:
        raise ValueError("no stations")
    di, dj = float(interval[0]), float(interval[1])
    xbeg = min(st.stlo) - num_grid_margin * di
    xend = max(st.stlo) + num_grid_margin * di
    ybeg = min(st.stla) - num_grid_margin * dj
    yend = max(st.stla) + num_grid_margin * dj
    nx = int(_llround((xend - xbeg) / di) + 1)
    ny = int(_llround((yend - ybeg) / dj) + 1)
    return linspaced(nx, xbeg, xend), linspaced(ny, ybeg, yend)


def _llround(x: float) -> int:
    """C llround: halves away from zero."""
    a = abs(x)
    r = math.floor(a)
    if a - r >= 0.5:  # exact: a and r share the binade or r = 0
        r += 1
    return int(r) * (1 if x >= 0 else -1)
