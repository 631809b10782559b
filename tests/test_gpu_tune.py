"""The autotuner (surfattpp_amd.tune): measures waves per grid and the
scheduling options for a batch, caches the choice by batch shape, and reuses it
without measuring."""
import json

import numpy as np
import pytest

from surfattpp_amd import PREC_F32
from surfattpp_amd import synth, tune

pytestmark = pytest.mark.gpu


def test_autotune_measures_caches_and_reuses(tmp_path, monkeypatch):
    path = str(tmp_path / "tune.json")
    monkeypatch.setenv(tune.CACHE_ENV, path)
    g1 = synth.config_grf(seed=51, nx=257, ny=97, nsta=6, nperiod=2, spacing_km=1.5)
    g2 = synth.config_grf(seed=52, nx=97, ny=81, nsta=8, nperiod=2, spacing_km=1.5)
    e = tune.autotune([g1, g2], runs=2)
    print("tuned", e["waves"], e["ms"], "library", e["default_waves"], e["default_ms"], flush=True)
    assert all(w in (1, 2, 4) for w in e["waves"])
    assert e["ms"] <= e["default_ms"]
    # waves of each grid, then the options (no grid of <= 64 columns: adj_cols, order, speculate)
    assert len(e["trials"]) >= 6
    tried = [t[1] for t in e["trials"] if t[1]]
    assert {"adj_cols": 0} in tried and {"order": 0} in tried and {"speculate": 0} in tried
    assert set(e["options"]) <= {"pair2", "adj_cols", "order", "speculate"}
    cache = json.load(open(path))
    assert list(cache.values())[0]["waves"] == e["waves"]
    assert list(cache.values())[0]["options"] == e["options"]
    # the same shape again: served from the cache (no new trials are measured)
    monkeypatch.setattr(tune, "time_batch", lambda *a, **k: pytest.fail("measured again"))
    assert tune.autotune([g1, g2])["waves"] == e["waves"]
    with tune.tuned_batch([g1, g2]) as b:
        assert b.waves() == e["waves"]
        for k, v in e["options"].items():
            assert b.get_option(k) == v
        b.run()
        r = b.results()
    assert np.isfinite(r["chi"]) and r["chi"] > 0
