"""Autotuner of the batched kernels' schedule.

A grid's batched FP32 sweeps run one pair per work-group of W waves (W = 1, 2
or 4; 64*W lanes, ceil(ceil(nx/64)/W) columns per lane).  More waves shorten a
pair's sweeps but take more of the device: the best W per grid depends on the
batch (how many pairs of which grids share the GPU).  The library picks W by a
rule (more waves for the largest grids while the batch stays within three
waves per SIMD); `autotune` measures instead.  It times the whole batched run
for the library's choice, then, grid by grid from the largest, for every
other W with the remaining grids held, and keeps any change that shortens the
run by more than 2 %.  Then, with the waves fixed, it tries the batch's
scheduling options (sat_batch_set_option) the same way, one at a time:
  pair2     two pairs per wave in the forward sweeps of grids of at most 64
            columns (on / off against the library's rule; tried only when the
            batch has such a grid),
  adj_cols  the FP32 adjoint of one-wave grids on 8 columns per lane or on
            the forward's columns per lane,
  order     the queues popped longest pairs (of the previous run) first, or
            in batch order,
  speculate the pairs re-solved in FP64 in recent runs solved in FP64 from
            the start of the run, beside the FP32 forward sweeps, or only
            after them (the band rule alone).
Every choice performs the same Gauss-Seidel order (results agree to FP32
rounding; a speculated pair has its FP64 result, which the band rule gives it
when the inputs repeat, as they do here), so only time is compared.  The FP32 re-solve band is not tuned: it
is the margin of the iteration-count parity, not a speed knob; neither is the
number of resident waves, which sizes the workspace at creation.  The choice
is stored in a JSON cache keyed by the device, the precision and the batch's
shape (grid sizes, pairs per grid, isotropic or not), so a batch of the same
shape -- every iteration of an inversion -- uses it without measuring again.

The cache file is $SURFATT_TUNE_CACHE, or ~/.cache/surfattpp_amd/tune.json.
"""
from __future__ import annotations

import json
import os
import numpy as np

from . import _lib
from . import preproc as PP

CACHE_ENV = "SURFATT_TUNE_CACHE"
WAVES = (1, 2, 4)


def cache_path() -> str:
    return os.environ.get(CACHE_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "surfattpp_amd",
                                                     "tune.json")


def _grids(grids) -> list:
    return [grids] if isinstance(grids, dict) else list(grids)


def _is_iso(g) -> bool:
    return bool(np.all(g["m11"] == 1.0) and np.all(g["m22"] == 1.0) and np.all(g["m12"] == 0.0))


def signature(grids, precision: int, calc_adj: bool = True) -> str:
    """The cache key: device, precision, adjoint or not, and per grid (nx, ny,
    pairs, isotropic)."""
    shapes = [[int(g["lon"].size), int(g["lat"].size), int(len(g["ev_iper"])), int(_is_iso(g))]
              for g in _grids(grids)]
    return json.dumps([_lib.device_info(), int(precision), bool(calc_adj), shapes])


def _load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path: str, key: str, entry: dict) -> None:
    cache = _load_cache(path)
    cache[key] = entry
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1)
    os.replace(tmp, path)


def time_batch(grids, waves, precision: int = _lib.PREC_F32, calc_adj: bool = True, runs: int = 3,
               options: dict | None = None):
    """(best ms of `runs` timed runs after a warm-up, waves as created), with
    the scheduling `options` set on the batch."""
    with PP.EikonalBatch(_grids(grids), precision=precision, calc_adj=calc_adj, waves=waves) as b:
        for k, v in (options or {}).items():
            b.set_option(k, v)
        b.run()  # warm-up; also the rounds the next runs order their queues by
        ms = []
        for _ in range(runs):
            b.run()
            ms.append(b.last_ms()[0])
        return min(ms), b.waves()


def _option_trials(grids) -> list:
    """(option, value) pairs worth timing for this batch, against the defaults."""
    trials = []
    if any(g["lon"].size <= 64 for g in grids):
        trials += [("pair2", 0), ("pair2", 1)]
    trials += [("adj_cols", 0), ("order", 0), ("speculate", 0)]
    return trials


def autotune(grids, precision: int = _lib.PREC_F32, calc_adj: bool = True, runs: int = 3,
             gain: float = 0.02, path: str | None = None, retune: bool = False) -> dict:
    """Waves per grid and scheduling options for this batch: from the cache, or
    measured and cached.  Returns dict(waves=[...], options={...}, ms=...,
    default_waves=[...], default_ms=..., trials=[[waves, options, ms], ...])."""
    grids = _grids(grids)
    path = path or cache_path()
    key = signature(grids, precision, calc_adj)
    if not retune:
        hit = _load_cache(path).get(key)
        if hit is not None:
            hit.setdefault("options", {})
            return hit
    if precision != _lib.PREC_F32:  # FP64 batches run one wave per pair; the options are FP32 ones
        entry = dict(waves=[1] * len(grids), options={}, ms=None, default_waves=[1] * len(grids),
                     default_ms=None, trials=[])
        _store(path, key, entry)
        return entry
    ms0, cur = time_batch(grids, 0, precision, calc_adj, runs)
    best = ms0
    opts: dict = {}
    trials = [[list(cur), {}, ms0]]
    order = sorted(range(len(grids)), key=lambda i: -grids[i]["lon"].size * grids[i]["lat"].size)
    for gi in order:
        for w in WAVES:
            if w == cur[gi] or (w > 1 and grids[gi]["lon"].size <= 64 * (w // 2)):
                continue
            trial = list(cur)
            trial[gi] = w
            ms, _ = time_batch(grids, trial, precision, calc_adj, runs)
            trials.append([trial, {}, ms])
            if ms < best * (1.0 - gain):
                best, cur = ms, trial
    for name, value in _option_trials(grids):
        if opts.get(name) == value:
            continue
        trial = dict(opts)
        trial[name] = value
        ms, _ = time_batch(grids, cur, precision, calc_adj, runs, trial)
        trials.append([list(cur), dict(trial), ms])
        if ms < best * (1.0 - gain):
            best, opts = ms, trial
    entry = dict(waves=list(cur), options=opts, ms=best, default_waves=list(trials[0][0]), default_ms=ms0,
                 trials=trials)
    _store(path, key, entry)
    return entry


def tuned_batch(grids, precision: int = _lib.PREC_F32, calc_adj: bool = True, **kw) -> PP.EikonalBatch:
    """An EikonalBatch with the autotuned waves per grid and scheduling options
    (measured on first use)."""
    entry = autotune(grids, precision, calc_adj)
    b = PP.EikonalBatch(_grids(grids), precision=precision, calc_adj=calc_adj, waves=entry["waves"], **kw)
    for k, v in entry.get("options", {}).items():
        b.set_option(k, v)
    return b
