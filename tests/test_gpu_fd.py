"""Differentiation check on the GPU at the 201x201 workload (1 km, 100 sources x
30 periods, 3.2 km/s with a +-10 % GRF): central finite differences of the FP64
misfit along 16 seeded smooth directions (GRFs of 40 km correlation, unit
inf-norm, h = 1e-4 relative) against the adjoint kernels of the FP64 and FP32
batched paths.  The relation and the bar are those of tests/test_fd_oracle.py:
dchi/dh = -sum K * dA * v * delta to discretisation level (ratio in [0.85, 1.10]);
the FP32 kernel's prediction agrees with the FP64 one to 1e-4.
"""
import numpy as np
import pytest

from surfattpp_amd import PREC_F32
from surfattpp_amd import preproc as PP
from surfattpp_amd import synth as S
from tests.test_fd_oracle import cell_area_km2

pytestmark = pytest.mark.gpu


def test_fd_16_directions_201():
    b = S.config_grf()
    lon, lat = b["lon"], b["lat"]
    dA = cell_area_km2(lon, lat)
    h = 1e-4
    with PP.EikonalBatch(b, precision=PREC_F32) as b32:
        b32.run()
        K32 = b32.results()["adj_s"]
    rng = S.Rng(2024)
    ratios, agree = [], []
    with PP.EikonalBatch(b) as bt:
        bt.run()
        K = bt.results()["adj_s"]
        for k in range(16):
            d = np.stack([S.grf(rng, lon.size, lat.size, 1.0, 40.0, 1.0) for _ in range(30)])
            pred = -float(np.sum(K * dA[None] * b["svel"] * d))
            pred32 = -float(np.sum(K32 * dA[None] * b["svel"] * d))
            chi = []
            for sgn in (1.0, -1.0):
                bt.update_media(0, b["svel"] * (1.0 + sgn * h * d), b["m11"], b["m12"], b["m22"])
                bt.run()
                chi.append(bt.results()["chi"])
            fd = (chi[0] - chi[1]) / (2.0 * h)
            ratios.append(fd / pred)
            agree.append(abs(pred32 - pred) / abs(pred))
            print(f"dir {k:2d}: fd {fd:.6g} adjoint FP64 {pred:.6g} FP32 {pred32:.6g} ratio {fd / pred:.4f}",
                  flush=True)
    print("ratio min/mean/max", min(ratios), float(np.mean(ratios)), max(ratios), flush=True)
    assert all(0.85 <= r <= 1.10 for r in ratios)
    assert max(agree) <= 1e-4
