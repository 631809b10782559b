"""One inversion iteration's data flow on the GPU, from the reference's example
00 inputs (examples/00_checkerboard_iso/input_params.yml): the observation
table (rows of src_rec_file_ph.csv, tests/golden) -> the station-box grid
(0.02 deg, 5 cells of margin) -> the initial 1-D model (depth 0-15 km every
0.5 km, Vs rising linearly 1.8 -> 4.2 km/s, init_model_type 0) -> Rayleigh
phase-velocity maps by surfdisp96 at the table's periods (sat_fwdsurf) -> the
forward + adjoint batch on those maps -> model kernels from the batch's device
kernels (sat_batch_combine_kernels).  Checked: map columns against the
reference's solver, the batch (FP64, a subset of events) against the oracle,
the model kernels against the host-input combine of the same kernels."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from surfattpp_amd import preproc as PP
from surfattpp_amd import src_rec as S
from surfattpp_amd import surfdisp as SD
from surfattpp_amd import synth
from tests import refcases as RC
from tests.test_combine_oracle import _inputs

pytestmark = pytest.mark.gpu
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "src_rec_checkerboard.csv")


def test_iteration_from_example_inputs():
    sr = S.SrcRec().load(TABLE)
    lon, lat = S.build_model_axes(S.build_stas(sr, None), [0.02, 0.02], 5)
    nx, ny = lon.size, lat.size
    z = np.arange(0.0, 15.0 + 1e-9, 0.5)
    vs1 = 1.8 + (4.2 - 1.8) * z / z[-1]
    vs3d = np.ascontiguousarray(np.broadcast_to(vs1, (nx, ny, z.size)))
    periods = sr.periods_info.periods
    svel = SD.fwdsurf(z, vs3d, periods)                        # (nper, ny, nx)
    b = sr.to_batch(lon, lat, svel)
    print(f"grid {nx}x{ny}, {len(b['ev_iper'])} events, periods {periods.min()}-{periods.max()} s, "
          f"maps {svel.min():.3f}-{svel.max():.3f} km/s", flush=True)
    sub = synth.subset_events(b, [0, 7, 19, 31])
    ro = O.forward_for_event(sub)
    with PP.EikonalBatch(sub) as bt:
        bt.run()
        rg = bt.results()
    assert RC.rel_inf(rg["syn"], ro["syn"]) <= 1e-12
    assert RC.rel_inf(rg["adj_s"], ro["adj_s"]) <= 1e-12
    with PP.EikonalBatch(b) as bt:
        bt.run()
        res = bt.results()
        nper = periods.size
        _, sens, vsm, vpm = _inputs(nx, ny, z.size, nper, 77)
        k_dev = bt.combine_kernels(*sens, vsm, vpm, mode=0)
        k_host = PP.combine_kernels(res["adj_s"], *sens, vsm, vpm, mode=0)
    np.testing.assert_array_equal(k_dev[0], k_host[0])
    assert np.isfinite(res["chi"]) and np.isfinite(k_dev[0]).all()
    sr.scatter_syn(res["syn"])
    sr.gather_syn_tt()
    assert np.all(sr.tt_fwd > 0)


@pytest.mark.skipif(not O.surfdisp_available(),
                    reason="oracle/_ref/libsurfdisp96.so (the reference's solver) is not built")
def test_example_maps_match_reference_solver():
    """The example's 1-D model: every map column equals the reference's solver."""
    sr = S.SrcRec().load(TABLE)
    lon, lat = S.build_model_axes(S.build_stas(sr, None), [0.02, 0.02], 5)
    z = np.arange(0.0, 15.0 + 1e-9, 0.5)
    vs1 = 1.8 + (4.2 - 1.8) * z / z[-1]
    vs3d = np.ascontiguousarray(np.broadcast_to(vs1, (lon.size, lat.size, z.size)))
    periods = sr.periods_info.periods
    svel = SD.fwdsurf(z, vs3d, periods)
    thk, vp, vs, rho = SD.build_disp_req(z, vs1)
    np.testing.assert_array_equal(svel[:, 0, 0], O.surfdisp96(thk, vp, vs, rho, periods))


TABLE_GR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "src_rec_checkerboard_gr.csv")


def test_phase_then_group_velocity_data():
    """Inversion::run_forward_adjoint (src/inversion.cpp:245-272) loops over the
    data types, phase then group velocity: per type, maps from fwdsurf
    (phase: igr 0, group: igr 1), forward_for_event on that type's table, and
    misfit += chi * weight (the example's weights 0.5, 0.5).  Both tables of
    example 00 (rows of src_rec_file_ph.csv / _gr.csv, tests/golden) on the
    example's initial model: each type's batch equals the oracle on a subset
    (FP64) and FP32 stays within 1e-5 of FP64; the group maps are slower than
    the phase maps at every column (normal dispersion of this model)."""
    from surfattpp_amd import PREC_F32
    weights = {"ph": 0.5, "gr": 0.5}
    misfit = {}
    maps = {}
    for kind, table, igr in (("ph", TABLE, 0), ("gr", TABLE_GR, 1)):
        sr = S.SrcRec().load(table)
        lon, lat = S.build_model_axes(S.build_stas(sr, None), [0.02, 0.02], 5)
        nx, ny = lon.size, lat.size
        z = np.arange(0.0, 15.0 + 1e-9, 0.5)
        vs1 = 1.8 + (4.2 - 1.8) * z / z[-1]
        vs3d = np.ascontiguousarray(np.broadcast_to(vs1, (nx, ny, z.size)))
        periods = sr.periods_info.periods
        svel = SD.fwdsurf(z, vs3d, periods, igr=igr)
        maps[kind] = svel
        assert np.all(svel > 0.5) and np.all(svel < 5.0)
        b = sr.to_batch(lon, lat, svel)
        sub = synth.subset_events(b, [0, 7, 19, 31])
        ro = O.forward_for_event(sub)
        with PP.EikonalBatch(sub) as bt:
            bt.run()
            rg = bt.results()
        assert RC.rel_inf(rg["syn"], ro["syn"]) <= 1e-12
        assert RC.rel_inf(rg["adj_s"], ro["adj_s"]) <= 1e-12
        np.testing.assert_array_equal(rg["iters_fwd"], ro["iters_fwd"])
        with PP.EikonalBatch(b) as b64, PP.EikonalBatch(b, precision=PREC_F32) as b32:
            b64.run()
            b32.run()
            r64, r32 = b64.results(), b32.results()
        errs = [RC.rel_inf(r32["adj_s"][p], r64["adj_s"][p]) for p in range(len(periods))]
        print(f"{kind}: {len(b['ev_iper'])} events, chi {r64['chi']:.6f}, FP32 K err {max(errs):.2e}", flush=True)
        assert max(errs) <= 1e-5
        np.testing.assert_array_equal(r32["iters_fwd"], r64["iters_fwd"])
        misfit[kind] = r64["chi"] * weights[kind]
    assert np.all(maps["gr"] < maps["ph"])
    total = misfit["ph"] + misfit["gr"]
    print(f"misfit: ph {misfit['ph']:.6f} + gr {misfit['gr']:.6f} = {total:.6f}", flush=True)
    assert np.isfinite(total) and total > 0
