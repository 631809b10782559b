"""Batched mirror of ``preproc::forward_for_event`` (include/preproc.h:14,
src/preproc.cpp:71-148).

The reference loops serially over a rank's events (src/preproc.cpp:79) after
an MPI round-robin deal (include/parallel.h:39-41) and reduces chi and the
per-period kernels with MPI_Allreduce (:146, :238).  Here one ``EikonalBatch``
holds every (source, period) pair of an inversion iteration on the GPU and one
``run()`` executes forward -> receiver interpolation -> residual -> adjoint ->
mask -> kernel accumulation -> chi for all of them.

A batch dict (see surfattpp_amd.synth) has: lon (nx,), lat (ny,), svel/m11/
m12/m22 (nperiod, ny, nx), ev_iper, ev_lon, ev_lat, ev_rec_off, ev_nrec,
rec_lon, rec_lat, tt_obs, weight.  A list of such dicts (one per grid) is an
uneven multi-grid batch.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib


def _c64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class EikonalBatch:
    """Device-resident batch of forward+adjoint pairs over one or more grids."""

    def __init__(self, grids: dict | Sequence[dict], precision: int = _lib.PREC_F64, calc_adj: bool = True,
                 workspace_cap: int = 0, waves=None):
        """workspace_cap: bytes of per-event device workspace held at once (0: the
        library's default); above it each run processes the events in consecutive
        chunks of the batch order, with the same results.
        waves: waves per pair of each grid's batched FP32 sweeps (an int for every
        grid or a sequence, each 1, 2 or 4; None or 0: the library's choice; see
        surfattpp_amd.tune for measured choices).  Results do not depend on it."""
        if isinstance(grids, dict):
            grids = [grids]
        if waves is None or isinstance(waves, int):
            waves = [int(waves or 0)] * len(grids)
        if len(waves) != len(grids):
            raise ValueError("one waves value per grid")
        self._keep = []
        gdesc = (_lib.sat_grid_desc * len(grids))()
        events, rl, rb, tt, w = [], [], [], [], []
        rec_base = 0
        self.shapes = []
        self.ev_slices = []
        for gi, b in enumerate(grids):
            lon, lat = _c64(b["lon"]), _c64(b["lat"])
            planes = [_c64(b[k]) for k in ("svel", "m11", "m12", "m22")]
            nper = planes[0].shape[0]
            for p in planes:
                if p.shape != (nper, lat.size, lon.size):
                    raise ValueError(f"plane shape {p.shape} != ({nper}, {lat.size}, {lon.size})")
            self._keep += [lon, lat, *planes]
            gdesc[gi] = _lib.sat_grid_desc(lon.size, lat.size, nper, lon.ctypes.data, lat.ctypes.data,
                                           planes[0].ctypes.data, planes[1].ctypes.data,
                                           planes[2].ctypes.data, planes[3].ctypes.data, int(waves[gi]))
            self.shapes.append((nper, lat.size, lon.size))
            n0 = len(events)
            for e in range(len(b["ev_iper"])):
                events.append((gi, int(b["ev_iper"][e]), float(b["ev_lon"][e]), float(b["ev_lat"][e]),
                               int(b["ev_rec_off"][e]) + rec_base, int(b["ev_nrec"][e])))
            self.ev_slices.append(slice(n0, len(events)))
            nrec_g = len(b["rec_lon"])
            rl.append(_c64(b["rec_lon"]))
            rb.append(_c64(b["rec_lat"]))
            tt.append(_c64(b["tt_obs"]))
            w.append(_c64(b["weight"]))
            rec_base += nrec_g
        self.nevent = len(events)
        self.nrec = rec_base
        edesc = (_lib.sat_event_desc * max(self.nevent, 1))()
        for i, ev in enumerate(events):
            edesc[i] = _lib.sat_event_desc(*ev)
        self._rec = [np.concatenate(x) if x else np.zeros(0) for x in (rl, rb, tt, w)]
        self.rec_counts = [len(x) for x in rl]
        d = _lib.sat_batch_desc()
        d.ngrid = len(grids)
        d.grids = gdesc
        d.nevent = self.nevent
        d.events = edesc
        d.nrec_total = self.nrec
        d.rec_lon_deg, d.rec_lat_deg, d.tt_obs, d.weight = (a.ctypes.data if a.size else None for a in self._rec)
        d.precision = int(precision)
        d.calc_adj = 1 if calc_adj else 0
        d.workspace_cap = int(workspace_cap)
        self.precision = precision
        self.calc_adj = calc_adj
        h = C.c_void_p()
        rc = _lib.load().sat_batch_create(C.byref(d), C.byref(h))
        _lib.check("sat_batch_create", rc)
        self._h = h
        self._keep = None  # inputs are on the device now

    def run(self, mode: int = _lib.RUN_BATCHED, sync: bool = True):
        _lib.check("sat_batch_run", _lib.load().sat_batch_run(self._h, int(mode)))
        if sync:
            self.sync()

    def sync(self):
        _lib.check("sat_batch_sync", _lib.load().sat_batch_sync(self._h))

    def last_ms(self):
        t, s = C.c_double(), C.c_double()
        _lib.check("sat_batch_last_ms", _lib.load().sat_batch_last_ms(self._h, C.byref(t), C.byref(s)))
        return t.value, s.value

    PHASES = ("init", "fwd_sweeps", "adj_setup", "adj_sweeps", "fp64_resolves", "reduce", "total")

    def phase_ms(self) -> dict:
        """Per-phase ms of the last batched run (HIP events on the kernels' streams)."""
        buf = (C.c_double * len(self.PHASES))()
        _lib.check("sat_batch_phase_ms", _lib.load().sat_batch_phase_ms(self._h, buf, len(self.PHASES)))
        return dict(zip(self.PHASES, list(buf)))

    def update_media(self, grid: int, svel, m11, m12, m22):
        planes = [_c64(x) for x in (svel, m11, m12, m22)]
        for p in planes:
            if p.shape != self.shapes[grid]:
                raise ValueError("plane shape mismatch")
        rc = _lib.load().sat_batch_update_media(self._h, grid, *(p.ctypes.data for p in planes))
        _lib.check("sat_batch_update_media", rc)

    def update_obs(self, tt_obs, weight):
        tt, w = _c64(tt_obs), _c64(weight)
        if tt.size != self.nrec or w.size != self.nrec:
            raise ValueError("receiver count mismatch")
        rc = _lib.load().sat_batch_update_obs(self._h, tt.ctypes.data, w.ctypes.data)
        _lib.check("sat_batch_update_obs", rc)

    def results(self):
        syn = np.zeros(max(self.nrec, 1))
        K = [np.zeros(s) for s in self.shapes]
        Kflat = np.zeros(sum(int(np.prod(s)) for s in self.shapes))
        chi = np.zeros(1)
        ne = max(self.nevent, 1)
        itf, ita, fl = (np.zeros(ne, dtype=np.int32) for _ in range(3))
        rc = _lib.load().sat_batch_results(self._h, syn.ctypes.data, Kflat.ctypes.data, chi.ctypes.data,
                                           itf.ctypes.data, ita.ctypes.data, fl.ctypes.data)
        _lib.check("sat_batch_results", rc)
        off = 0
        for i, s in enumerate(self.shapes):
            n = int(np.prod(s))
            K[i] = Kflat[off:off + n].reshape(s)
            off += n
        return dict(syn=syn[: self.nrec], adj_s=K if len(K) > 1 else K[0], chi=float(chi[0]),
                    iters_fwd=itf[: self.nevent], iters_adj=ita[: self.nevent], flags=fl[: self.nevent])

    def waves(self) -> list:
        """Waves per pair of each grid's batched sweeps, as chosen at creation."""
        buf = (C.c_int * len(self.shapes))()
        _lib.check("sat_batch_waves", _lib.load().sat_batch_waves(self._h, buf))
        return list(buf)

    OPTIONS = {"pair2": _lib.OPT_PAIR2, "adj_cols": _lib.OPT_ADJ_COLS, "order": _lib.OPT_ORDER,
               "speculate": _lib.OPT_SPECULATE}

    def set_option(self, name: str, value: int):
        """A scheduling option (sat_batch_set_option): pair2 (-1 the library's
        rule, 0, 1), adj_cols (0, 1), order (0, 1), speculate (0, 1).  Results
        do not depend on them beyond FP32 rounding (speculate: pairs re-solved
        in FP64 in recent runs are solved in FP64 from the start of the run)."""
        _lib.check("sat_batch_set_option", _lib.load().sat_batch_set_option(self._h, self.OPTIONS[name], int(value)))

    def get_option(self, name: str) -> int:
        v = C.c_int()
        _lib.check("sat_batch_get_option", _lib.load().sat_batch_get_option(self._h, self.OPTIONS[name], C.byref(v)))
        return v.value

    def set_f32_band(self, band: float):
        _lib.check("sat_batch_set_f32_band", _lib.load().sat_batch_set_f32_band(self._h, float(band)))

    def conv(self):
        """Per event Linf of the last two forward rounds, and the FP64 re-solve count."""
        ne = max(self.nevent, 1)
        last, prev = np.zeros(ne), np.zeros(ne)
        n = C.c_int()
        rc = _lib.load().sat_batch_conv(self._h, last.ctypes.data, prev.ctypes.data, C.byref(n))
        _lib.check("sat_batch_conv", rc)
        return last[: self.nevent], prev[: self.nevent], n.value

    def combine_kernels(self, sen_vs, sen_vp, sen_rho=None, vs3d=None, vp3d=None, mode: int = 0, grid: int = 0,
                        timing: dict | None = None):
        """combine_kernels (below) on this batch's per-period kernels of its last
        run, read on the device in place (grid `grid`)."""
        nper, ny, nx = self.shapes[grid]
        s_vs = _c64(sen_vs)
        nz = s_vs.shape[2]
        shape4 = (nx, ny, nz, nper)
        for a in (s_vs, sen_vp) + ((sen_rho,) if mode != COMBINE_VS_VP_RHO_SCALED else ()):
            if a is None or np.shape(a) != shape4:
                raise ValueError(f"sensitivities must be {shape4}")
        s_vp = _c64(sen_vp)
        s_rho = _c64(sen_rho) if sen_rho is not None else None
        v_s = _c64(vs3d) if vs3d is not None else None
        v_p = _c64(vp3d) if vp3d is not None else None
        if mode == COMBINE_VS and (s_rho is None or v_s is None or v_p is None
                                   or v_s.shape != (nx, ny, nz) or v_p.shape != (nx, ny, nz)):
            raise ValueError("mode 0 needs sen_rho and vs3d, vp3d of shape (nx, ny, nz)")
        k_vs = np.empty((nx, ny, nz))
        k_vp = np.empty((nx, ny, nz)) if mode else None
        k_rho = np.empty((nx, ny, nz)) if mode else None
        ms = C.c_double(0.0)
        ptr = lambda a: a.ctypes.data if a is not None else None  # noqa: E731
        rc = _lib.load().sat_batch_combine_kernels(self._h, int(grid), nz, int(mode), ptr(s_vs), ptr(s_vp),
                                                   ptr(s_rho), ptr(v_s), ptr(v_p), ptr(k_vs), ptr(k_vp),
                                                   ptr(k_rho), C.byref(ms))
        _lib.check("sat_batch_combine_kernels", rc)
        if timing is not None:
            timing["kernel_ms"] = ms.value
        return k_vs, k_vp, k_rho

    def _metric_planes(self, m11, m12, m22, grid):
        if not 0 <= grid < len(self.shapes):
            raise ValueError(f"grid {grid} is not in this batch")
        shape = tuple(self.shapes[grid])
        planes = [_c64(m) for m in (m11, m22, -np.asarray(m12, dtype=np.float64))]
        for m in planes:
            if m.shape != shape:
                raise ValueError(f"metric planes must be {shape}")
        return planes

    def aniso_kernels(self, m11=None, m12=None, m22=None, grid: int = 0):
        """The anisotropic terms of accumulate_kernels (src/preproc.cpp:37-65)
        for this batch's last run on grid `grid`: (adj_xi, adj_eta), each
        (nperiod, ny, nx).  m11, m12, m22 (nperiod, ny, nx): the grid's
        topography metric (SurfGrid a = m11, b = m22, c = -m12,
        src/surf_grid.cpp:186-189); all None: the terms the last run formed
        for the metric set_aniso_metric registered (chunked runs included)."""
        if not 0 <= grid < len(self.shapes):
            raise ValueError(f"grid {grid} is not in this batch")
        shape = tuple(self.shapes[grid])
        xi, eta = np.empty(shape), np.empty(shape)
        if m11 is None and m12 is None and m22 is None:
            rc = _lib.load().sat_batch_aniso_kernels(self._h, int(grid), None, None, None,
                                                     xi.ctypes.data, eta.ctypes.data)
        else:
            planes = self._metric_planes(m11, m12, m22, grid)
            rc = _lib.load().sat_batch_aniso_kernels(self._h, int(grid), *(m.ctypes.data for m in planes),
                                                     xi.ctypes.data, eta.ctypes.data)
        _lib.check("sat_batch_aniso_kernels", rc)
        return xi, eta

    def set_aniso_metric(self, m11=None, m12=None, m22=None, grid: int = 0):
        """Register grid `grid`'s topography metric: every later run forms the
        anisotropic terms itself (aniso_kernels() with no metric returns them).
        All None unregisters."""
        if m11 is None and m12 is None and m22 is None:
            rc = _lib.load().sat_batch_set_aniso_metric(self._h, int(grid), None, None, None)
        else:
            planes = self._metric_planes(m11, m12, m22, grid)
            rc = _lib.load().sat_batch_set_aniso_metric(self._h, int(grid), *(m.ctypes.data for m in planes))
        _lib.check("sat_batch_set_aniso_metric", rc)

    def set_kden(self, on: bool = True):
        """Form the density kernel (is_kden, src/preproc.cpp:127-136) in every
        later run: a second adjoint per pair from the receiver sources -weight."""
        _lib.check("sat_batch_set_kden", _lib.load().sat_batch_set_kden(self._h, int(bool(on))))

    def kden_results(self):
        """The last run's density kernel kden_s, per grid (nperiod, ny, nx), as
        adj_s in results()."""
        flat = np.zeros(max(sum(int(np.prod(s)) for s in self.shapes), 1))
        _lib.check("sat_batch_kden_results", _lib.load().sat_batch_kden_results(self._h, flat.ctypes.data))
        out, off = [], 0
        for s in self.shapes:
            n = int(np.prod(s))
            out.append(flat[off:off + n].reshape(s))
            off += n
        return out if len(out) > 1 else out[0]

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().sat_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def forward_for_event(batch: dict | Sequence[dict], is_calc_adj: bool = True,
                      precision: int = _lib.PREC_F64) -> dict:
    """One-shot: create a batch, run it, return results (syn, adj_s, chi, iters, flags)."""
    with EikonalBatch(batch, precision=precision, calc_adj=is_calc_adj) as b:
        b.run()
        return b.results()


COMBINE_VS, COMBINE_VS_VP_RHO, COMBINE_VS_VP_RHO_SCALED = 0, 1, 2


def combine_kernels(adj_s, sen_vs, sen_vp, sen_rho=None, vs3d=None, vp3d=None, mode: int = COMBINE_VS,
                    timing: dict | None = None):
    """Isotropic preproc::combine_kernels (src/preproc.cpp:196-342) on the GPU.

    adj_s (nper, ny, nx) per-period kernels (EikonalBatch.results()["adj_s"]);
    sen_vs, sen_vp, sen_rho (nx, ny, nz, nper) depth sensitivities in the
    reference's RowMajor layout; vs3d, vp3d (nx, ny, nz) for mode 0 (vs only:
    the vp and rho sensitivities enter through dvp/dvs and drho/dvp); mode 1:
    vs, vp, rho kernels; mode 2: rho = 0.33 * vs.  Returns (ker_vs, ker_vp,
    ker_rho), each (nx, ny, nz) (None where the mode has no such kernel).
    `timing`, if given, receives the device kernel's milliseconds."""
    adj = _c64(adj_s)
    nper, ny, nx = adj.shape
    s_vs = _c64(sen_vs)
    nz = s_vs.shape[2]
    shape4 = (nx, ny, nz, nper)
    for a in (s_vs, sen_vp) + ((sen_rho,) if mode != COMBINE_VS_VP_RHO_SCALED else ()):
        if a is None or np.shape(a) != shape4:
            raise ValueError(f"sensitivities must be {shape4}")
    s_vp = _c64(sen_vp)
    s_rho = _c64(sen_rho) if sen_rho is not None else None
    v_s = _c64(vs3d) if vs3d is not None else None
    v_p = _c64(vp3d) if vp3d is not None else None
    if mode == COMBINE_VS and (s_rho is None or v_s is None or v_p is None
                               or v_s.shape != (nx, ny, nz) or v_p.shape != (nx, ny, nz)):
        raise ValueError("mode 0 needs sen_rho and vs3d, vp3d of shape (nx, ny, nz)")
    k_vs = np.empty((nx, ny, nz))
    k_vp = np.empty((nx, ny, nz)) if mode else None
    k_rho = np.empty((nx, ny, nz)) if mode else None
    ms = C.c_double(0.0)
    ptr = lambda a: a.ctypes.data if a is not None else None  # noqa: E731
    rc = _lib.load().sat_combine_kernels(nx, ny, nz, nper, int(mode), ptr(adj), ptr(s_vs), ptr(s_vp), ptr(s_rho),
                                         ptr(v_s), ptr(v_p), ptr(k_vs), ptr(k_vp), ptr(k_rho), C.byref(ms))
    _lib.check("sat_combine_kernels", rc)
    if timing is not None:
        timing["kernel_ms"] = ms.value
    return k_vs, k_vp, k_rho


def combine_kernels_ex(adj_s, sen_vs, sen_vp, sen_rho=None, vs3d=None, vp3d=None, mode: int = COMBINE_VS,
                       r1=None, r2=None, adj_xi=None, adj_eta=None, svel=None, sen_gc=None, sen_gs=None,
                       adj_den=None, timing: dict | None = None) -> dict:
    """preproc::combine_kernels with its anisotropic and density branches
    (src/preproc.cpp:196-342) on the GPU (sat_combine_kernels_ex).

    As combine_kernels, plus: r1, r2 (nx, ny, nper) -- SurfGrid r1_loc, r2_loc
    (is_anisotropy; None: isotropic) with adj_xi, adj_eta (EikonalBatch.
    aniso_kernels), svel (nper, ny, nx) and sen_gc, sen_gs (nx, ny, nz, nper);
    adj_den (nper, ny, nx): the density accumulators (is_kden).  Returns a dict
    of (nx, ny, nz) kernels: vs, and vp, rho (modes 1, 2), gc, gs, den as the
    options form them."""
    c = lambda a: None if a is None else _c64(a)  # noqa: E731
    adj = _c64(adj_s)
    nper, ny, nx = adj.shape
    s_vs = _c64(sen_vs)
    nz = s_vs.shape[2]
    aniso, kden = r1 is not None, adj_den is not None
    ins = dict(adj_s=adj, sen_vs=s_vs, sen_vp=c(sen_vp), sen_rho=c(sen_rho), vs3d=c(vs3d), vp3d=c(vp3d),
               r1=c(r1), r2=c(r2), adj_xi=c(adj_xi), adj_eta=c(adj_eta), svel=c(svel), sen_gc=c(sen_gc),
               sen_gs=c(sen_gs), adj_den=c(adj_den))
    shapes = {"sen_vp": (nx, ny, nz, nper), "sen_rho": (nx, ny, nz, nper), "sen_gc": (nx, ny, nz, nper),
              "sen_gs": (nx, ny, nz, nper), "vs3d": (nx, ny, nz), "vp3d": (nx, ny, nz), "r1": (nx, ny, nper),
              "r2": (nx, ny, nper), "adj_xi": (nper, ny, nx), "adj_eta": (nper, ny, nx), "svel": (nper, ny, nx),
              "adj_den": (nper, ny, nx)}
    for k, shp in shapes.items():
        if ins[k] is not None and ins[k].shape != shp:
            raise ValueError(f"{k} must be {shp}")
    out = {"vs": np.empty((nx, ny, nz))}
    if mode:
        out.update(vp=np.empty((nx, ny, nz)), rho=np.empty((nx, ny, nz)))
    if aniso:
        out.update(gc=np.empty((nx, ny, nz)), gs=np.empty((nx, ny, nz)))
    if kden:
        out["den"] = np.empty((nx, ny, nz))
    d = _lib.sat_combine_desc()
    d.nx, d.ny, d.nz, d.nper, d.mode = nx, ny, nz, nper, int(mode)
    for k, a in ins.items():
        setattr(d, k, a.ctypes.data if a is not None else None)
    for k in ("vs", "vp", "rho", "gc", "gs", "den"):
        setattr(d, "ker_" + k, out[k].ctypes.data if k in out else None)
    ms = C.c_double(0.0)
    _lib.check("sat_combine_kernels_ex", _lib.load().sat_combine_kernels_ex(C.byref(d), C.byref(ms)))
    if timing is not None:
        timing["kernel_ms"] = ms.value
    return out
