#!/usr/bin/env python3
"""Benchmark: forward + adjoint solve pairs per second on one MI355X.

Workload: 201 x 201 lon/lat grid at 1 km, 100 virtual sources x 30 periods =
3,000 (source, period) pairs with 99 receivers each; velocity 3.2 km/s with an
independent seeded +-10 % Gaussian random field (15 km correlation) per
period; observed times from the 3.2 km/s background (synthetic data).

A step is one pass of the whole hot path over the batch -- forward solve,
receiver interpolation, residual, adjoint solve, mask, per-period kernel
accumulation and chi for all 3,000 pairs -- with inputs resident in HBM
(src/preproc.cpp:71-148 for every event of an inversion iteration).

`value` is the batched FP32 path.  The line also carries the FP64 batched and
the plain per-solve FP64 GPU numbers, the FP32 kernel error against the FP64
path (which equals the FP64 oracle to 1e-12), the roofline of the dominant
kernel (the FP32 forward sweeps), per-phase times and the CPU baseline (the
FP64 oracle on the host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+adjoint solves/s @201² grid: batched FP32 vs plain FP64, ≤1e-5 rel err"
HBM_PEAK_GBS = 8000.0


# Algorithmic bytes of the forward sweeps per node and sweep: 9 words (slowness,
# a, b, c, T0, T0x, T0y and tau read; tau written) of w = 4 bytes (FP32 sweeps).
# This kernel's own streams per node and sweep -- sigma read and written (8 + 8),
# T0, T0x, T0y, slowness (4 each), the round-start sigma once per round (8) --
# come to 34 bytes, within 6 % of the same figure.
FWD_WORDS_PER_NODE_SWEEP = 9
FWD_WORD_BYTES = 4


def fwd_sweep_bytes(nx, ny, iters_f):
    """Algorithmic HBM bytes of one launch of the FP32 forward sweep kernel over
    the batch: w*N*(4 sweeps * 9 words * I_f) summed over pairs, I_f the pair's
    forward rounds (the FP64 path's; the FP32 path's are equal)."""
    N = nx * ny
    return float(FWD_WORD_BYTES * N * 4 * FWD_WORDS_PER_NODE_SWEEP * iters_f.astype(np.float64).sum())


def host_cpu_info() -> dict:
    """What the host offers: the CPUs this process may run on, the online CPUs,
    the share a job is given (OMP_NUM_THREADS, set per GPU on a shared box) and
    the CPU model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or affinity
    return {"affinity_cpus": affinity, "nproc": os.cpu_count() or 1, "job_cpu_share": share, "model": model}


def cpu_baseline(batch, seconds: float, threads: int, info: dict):
    """The FP64 oracle (-O3 -mavx2 -mfma) on `threads` host threads of this
    process: pairs dealt round-robin i % threads (include/parallel.h:39-41), one
    pair at a time per thread, each thread stopping at the deadline.  ctypes
    releases the GIL for the duration of each solve, so the threads run on as
    many cores."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O
    from surfattpp_amd import synth
    lib = O.bench_lib()
    ne = len(batch["ev_iper"])
    subsets = [synth.subset_events(batch, [e]) for e in range(ne)]
    deadline = time.time() + seconds

    def work(rank):
        done, busy = 0, 0.0
        for e in range(rank, ne, threads):
            if time.time() > deadline:
                break
            t = time.perf_counter()
            O.forward_for_event(subsets[e], lib=lib)
            busy += time.perf_counter() - t
            done += 1
        return done, busy

    t0 = time.time()
    with ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(work, range(threads)))
    dt = time.time() - t0
    n = int(sum(d for d, _ in res))
    per_pair = sum(b for _, b in res) / max(n, 1)
    return dict(value=n / dt, unit="fwd+adjoint solves/s", cores=threads, kind="port",
                s_per_pair_one_core=round(per_pair, 4),
                all_cpus_linear=round(info["affinity_cpus"] / per_pair, 1),
                host=info,
                sample=f"{n} of {ne} pairs of the same workload (FP64 oracle, -O3 -mavx2 -mfma), "
                       f"round-robin over {threads} threads of one process ({info['job_cpu_share']} CPUs "
                       f"is this job's share of the {info['affinity_cpus']}-CPU host), {dt:.1f} s wall; "
                       f"all_cpus_linear: the measured per-pair time scaled to every CPU of the host")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-pairs", type=int, default=30, help="pairs timed in per-solve FP64 mode")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the uneven-topography and 401x481 workloads (reported under 'workloads')")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device twice (the plain copy "
                         "kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    from surfattpp_amd import PREC_F32, PREC_F32P, PREC_F64, RUN_PER_SOLVE, _lib, synth
    from surfattpp_amd import preproc as PP

    dev = _lib.device_info()
    b = synth.config_grf()
    nx, ny = b["lon"].size, b["lat"].size
    ne = len(b["ev_iper"])

    # FP64 batched (reference arithmetic): iteration counts and the FP64 kernels.
    with PP.EikonalBatch(b, precision=PREC_F64) as b64:
        b64.run()
        b64.run()
        f64_ms = b64.last_ms()
        r64 = b64.results()
    iters_f, iters_a = r64["iters_fwd"], r64["iters_adj"]

    # plain per-solve FP64: one launch per pair, on a subset
    sub = synth.subset_events(b, np.linspace(0, ne - 1, args.plain_pairs).astype(int))
    with PP.EikonalBatch(sub, precision=PREC_F64) as bp:
        bp.run(mode=RUN_PER_SOLVE)
        bp.run(mode=RUN_PER_SOLVE)
        plain_ms = bp.last_ms()[0]
    plain_rate = args.plain_pairs / (plain_ms * 1e-3)

    # batched FP32: the measured path.  The first run of a batch has no history:
    # its queues are in batch order and it speculates no FP64 solves (both
    # follow the previous run's rounds and re-solves); reported beside `value`.
    bt = PP.EikonalBatch(b, precision=PREC_F32)
    first_run_ms = None
    for k in range(args.warmup):
        bt.run()
        if k == 0:
            first_run_ms = round(float(bt.last_ms()[0]), 3)
    bt.sync()
    phases = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bt.run(sync=False)
        bt.sync()  # per step, so the per-launch event timing is read for every step
        phases.append(bt.phase_ms())
    bt.sync()
    wall = time.perf_counter() - t0
    r32 = bt.results()
    bt.close()
    ms_per_step = wall / args.steps * 1e3
    value = ne * args.steps / wall

    # the same workload with the FP64-polished update (SAT_PREC_F32P)
    bp = PP.EikonalBatch(b, precision=PREC_F32P)
    for _ in range(args.warmup):
        bp.run()
    bp.sync()
    t0 = time.perf_counter()
    php = []
    for _ in range(args.steps):
        bp.run(sync=False)
        bp.sync()
        php.append(bp.phase_ms())
    wall_p = time.perf_counter() - t0
    rp = bp.results()
    bp.close()
    f32p = {"value": round(ne * args.steps / wall_p, 2), "ms_per_step": round(wall_p / args.steps * 1e3, 3),
            "fwd_sweeps_ms": round(float(np.mean([p["fwd_sweeps"] for p in php])), 3),
            "err_vs_fp64": {"kernel_max_per_period": _kernel_err(rp, r64, 1),
                            "syn": float(np.abs(rp["syn"] - r64["syn"]).max() / np.abs(r64["syn"]).max())},
            "iters_fwd_equal": bool(np.array_equal(rp["iters_fwd"], r64["iters_fwd"]))}

    err_k = max(float(np.abs(r32["adj_s"][p] - r64["adj_s"][p]).max() / np.abs(r64["adj_s"][p]).max())
                for p in range(r64["adj_s"].shape[0]))
    err_syn = float(np.abs(r32["syn"] - r64["syn"]).max() / np.abs(r64["syn"]).max())
    err_chi = abs(r32["chi"] - r64["chi"]) / abs(r64["chi"])

    phase_mean = {k: round(float(np.mean([p[k] for p in phases])), 3) for k in phases[0]}
    kern_ms = phase_mean["fwd_sweeps"]  # HIP events around the kernel on its stream
    bytes_launch = fwd_sweep_bytes(nx, ny, iters_f)
    achieved = bytes_launch / (kern_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "fwd_sweep_f32_grf201_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    # The kernel is bound by VALU issue, not HBM: its wave64 VALU instruction
    # count per launch (rocprofv3 SQ_INSTS_VALU, committed under profiles/) at 4
    # cycles each against the device's SIMD-cycles over this run's kernel time
    # at the peak engine clock; and the share of the lane-steps that update a
    # node (the lane-skewed sweep's ramps and idle lanes issue too).
    valu = None
    vf = os.path.join(ROOT, "profiles", "fwd_sweep_f32_grf201_valu.json")
    if os.path.exists(vf):
        with open(vf) as f:
            valu = json.load(f)
    simds, clock_hz = 4 * cu_count(dev), 2.4e9
    issue_frac = (round(valu["valu_insts_per_launch"] * 4 / (simds * clock_hz * kern_ms * 1e-3), 4)
                  if valu else None)
    busy = busy_lane_fraction(nx, ny, waves=1)

    extra = None if args.no_extra else extra_workloads()

    cpu = None
    if not args.no_cpu:
        info = host_cpu_info()
        cpu = cpu_baseline(b, args.cpu_seconds, max(1, min(info["affinity_cpus"], info["job_cpu_share"])), info)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "fwd+adjoint solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "grf-201x201-1km-100src-30per", "grid": [nx, ny], "pairs": ne,
                   "receivers_per_pair": int(b["ev_nrec"][0]), "sweep_precision": "fp32 (sigma = tau - 1)",
                   "fp64_parts": "T0 init, T, adjoint coefficients, convergence norm, misfit, kernel accumulation",
                   "device": dev},
        "roofline": {"bound": "valu", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "fwd_sweep_kernel<SAT_PREC_F32,4,iso>", "kernel_ms": round(kern_ms, 3),
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "hbm_frac": round(achieved / HBM_PEAK_GBS, 4),
                     "valu_issue_frac": issue_frac,
                     "valu_insts_per_launch": valu["valu_insts_per_launch"] if valu else None,
                     "busy_lane_frac": round(busy, 4),
                     "note": "achieved/frac: algorithmic HBM bytes over the kernel time against 8 TB/s; the "
                             "kernel is bound by VALU issue: valu_issue_frac = SQ_INSTS_VALU x 4 cycles / "
                             f"({simds} SIMDs x 2.4 GHz x kernel time)"},
        "phases_ms": phase_mean,
        "ms_first_run_no_history": first_run_ms,
        "cpu_baseline": cpu,
        "err_vs_fp64": {"kernel_max_per_period": err_k, "syn": err_syn, "chi": err_chi},
        "batched_f64": {"value": round(ne / (f64_ms[0] * 1e-3), 2), "ms_per_step": round(f64_ms[0], 3)},
        "batched_f32p": f32p,
        "plain_f64_per_solve": {"value": round(plain_rate, 3), "pairs_timed": args.plain_pairs},
        "speedup": {"vs_batched_f64": round(value / (ne / (f64_ms[0] * 1e-3)), 3),
                    "vs_plain_f64_per_solve": round(value / plain_rate, 1),
                    "vs_cpu": (round(value / cpu["value"], 1) if cpu else None),
                    "note": "vs_batched_f64 is the precision speed-up (the same batched kernels in FP64); "
                            "the per-solve path is one pair per launch, one wave on the device"},
        "iters": {"fwd_mean": float(iters_f.mean()), "adj_mean": float(iters_a.mean()),
                  "fwd_hist": np.bincount(iters_f).tolist(), "adj_hist": np.bincount(iters_a).tolist()},
        "workloads": extra,
    }
    print(json.dumps(line), flush=True)
    if args.calib_copy_mib > 0:
        calib_copy(args.calib_copy_mib << 20)


def cu_count(dev: str) -> int:
    """Compute units from the library's device string ("... | CUs=256 | ...")."""
    for part in dev.split("|"):
        part = part.strip()
        if part.startswith("CUs="):
            return int(part[4:])
    return 256


def busy_lane_fraction(nx: int, ny: int, waves: int = 1) -> float:
    """Share of the lane-steps of one forward sweep that update a node: the
    sweep runs ny + La - 1 steps on 64*waves lanes of C columns each (C =
    ceil(ceil(nx/64)/waves), La = ceil(nx/C))."""
    C = -(-(-(-nx // 64)) // waves)
    La = -(-nx // C)
    return nx * ny / (64 * waves * C * (ny + La - 1))


def _kernel_err(r32, r64, ngrid):
    k32 = r32["adj_s"] if ngrid > 1 else [r32["adj_s"]]
    k64 = r64["adj_s"] if ngrid > 1 else [r64["adj_s"]]
    return max(float(np.abs(a[p] - c[p]).max() / np.abs(c[p]).max())
               for a, c in zip(k32, k64) for p in range(c.shape[0]))


def _time_pair(grids, runs=2):
    """FP64 (one run after a warm-up), FP32 and FP32 with the polished update
    (`runs` runs after a warm-up each) of one batch: pairs/s of each, kernel and
    syn error against FP64, FP32 phases."""
    from surfattpp_amd import PREC_F32, PREC_F32P, PREC_F64
    from surfattpp_amd import preproc as PP
    ngrid = len(grids) if isinstance(grids, list) else 1
    with PP.EikonalBatch(grids, precision=PREC_F64) as b64:
        b64.run()
        b64.run()
        ms64 = b64.last_ms()[0]
        r64 = b64.results()
    def timed(prec, graph=False):
        with PP.EikonalBatch(grids, precision=prec) as bt:
            bt.run()
            t = []
            for _ in range(runs):
                bt.run()
                t.append(bt.last_ms()[0])
            out = [float(np.mean(t)), bt.phase_ms(), bt.results(), bt.conv()[2]]
            if graph:
                # the whole run (every launch group on its streams, the re-solves,
                # the accumulation) as one captured graph, one launch per run
                from surfattpp_amd import _lib
                bt.run(mode=_lib.RUN_GRAPH)
                tg = []
                for _ in range(runs):
                    bt.run(mode=_lib.RUN_GRAPH)
                    tg.append(bt.last_ms()[0])
                out.append(float(np.mean(tg)))
            return out

    ms32, ph, r32, nres, ms32g = timed(PREC_F32, graph=True)
    msp, php, rp, _ = timed(PREC_F32P)
    ne = len(r64["iters_fwd"])

    def err(r):
        return {"kernel_max_per_period": _kernel_err(r, r64, ngrid),
                "syn": float(np.abs(r["syn"] - r64["syn"]).max() / np.abs(r64["syn"]).max())}
    return {"pairs": ne, "value_f32": round(ne / (ms32 * 1e-3), 2), "ms_f32": round(ms32, 3),
            "value_f64": round(ne / (ms64 * 1e-3), 2), "ms_f64": round(ms64, 3),
            "ms_f32_one_graph_launch": round(ms32g, 3),
            "err_vs_fp64": err(r32),
            "iters_fwd_equal": bool(np.array_equal(r32["iters_fwd"], r64["iters_fwd"])),
            "fp64_resolves": int(nres), "phases_ms_f32": {k: round(v, 3) for k, v in ph.items()},
            "f32p": {"value": round(ne / (msp * 1e-3), 2), "ms": round(msp, 3), "err_vs_fp64": err(rp),
                     "iters_fwd_equal": bool(np.array_equal(rp["iters_fwd"], r64["iters_fwd"])),
                     "fwd_sweeps_ms": round(php["fwd_sweeps"], 3)}}


def extra_workloads():
    """The uneven topography batch (101^2, 161^2, 241^2, 321^2 at 1.5 km, 12 periods,
    anisotropic metric), the 121x141 batch (25 sources x 20 periods) and the 401x481
    batch (400 sources x 30 periods = 12,000 pairs)."""
    from surfattpp_amd import synth
    topo = synth.config_topo()
    out = {"topo-uneven-101..321-1.5km-12per": dict(
        _time_pair(topo), grids=[[g["lon"].size, g["lat"].size, len(g["ev_iper"])] for g in topo])}
    del topo
    disp = synth.config_dispersion()
    out["disp-121x141-2km-25src-20per"] = dict(_time_pair(disp), maps="surfdisp96 on the GPU (sat_fwdsurf)")
    del disp
    out["large-401x481-1km-400src-30per"] = _time_pair(synth.config_large())
    out["combine-201x201-30z-30per"] = combine_workload()
    out["topo-metric-321x321-1.5km-12per"] = topo_metric_workload()
    out["fwdsurf-201x201-30z-30per"] = fwdsurf_workload()
    return out


def fwdsurf_workload(nx=201, ny=201, nper=30):
    """Rayleigh phase-velocity maps of a 201x201-column layered model (30 depth
    nodes, 0-60 km) at 30 periods 4-40 s by surfdisp96 on the GPU
    (SurfGrid::fwdsurf, src/surf_grid.cpp:194-245), beside the reference's
    surfdisp96 (oracle/_ref, built from its Fortran source) timed on a sample of
    columns on one host core and scaled to all columns over the host's cores."""
    from surfattpp_amd import surfdisp as SD
    from surfattpp_amd import synth
    dep, vs3d, vp3d, rho3d = synth.layered_model(nx, ny, synth.Rng(2))
    periods = synth.periods_log(nper)
    ms = []
    for _ in range(3):
        t = {}
        svel = SD.fwdsurf(dep, vs3d, periods, vp3d=vp3d, rho3d=rho3d, timing=t)
        ms.append(t["kernel_ms"])
    best = min(ms)
    out = {"kernel_ms": round(best, 2), "columns": nx * ny, "periods": nper,
           "columns_per_s": round(nx * ny / best * 1e3, 1)}
    from oracle import oracle as O
    if O.surfdisp_available():
        thk, vp, vs, rho = SD.build_disp_req(dep, vs3d.reshape(nx * ny, -1), vp3d.reshape(nx * ny, -1),
                                             rho3d.reshape(nx * ny, -1))
        sample = np.arange(0, nx * ny, (nx * ny) // 200)
        t0 = time.time()
        worst = 0.0
        for i in sample:
            ref = O.surfdisp96(thk[i], vp[i], vs[i], rho[i], periods)
            ix, iy = divmod(int(i), ny)
            worst = max(worst, float(np.abs(svel[:, iy, ix] / ref - 1.0).max()))
        per_col = (time.time() - t0) / sample.size
        info = host_cpu_info()
        cores = max(1, min(info["affinity_cpus"], info["job_cpu_share"]))
        out["cpu_reference"] = {"s_per_column_one_core": round(per_col, 6), "cores": cores,
                                "columns_per_s": round(cores / per_col, 1),
                                "all_cpus_linear": round(info["affinity_cpus"] / per_col, 1),
                                "sample": f"{sample.size} columns, reference surfdisp96 (amdflang -O2)"}
        out["max_rel_err_vs_reference"] = worst
    return out


def topo_metric_workload(n=321, nperiod=12):
    """The anisotropic metric planes of the topography workload's largest grid
    (321x321 at 1.5 km, 8 Gaussian edifices, 12 per-period smoothing widths,
    src/surf_grid.cpp:109-145) on the GPU: device time of the smoothing and
    metric kernels, window nodes visited per second and the FP64 rate of the
    window sums (3 flops per window node: w*z + num, den + w)."""
    from surfattpp_amd import synth
    from surfattpp_amd import topo as TG
    b = synth.config_topo_grid(33, n, nsta=10, nperiod=nperiod, smooth=False)
    lon, lat, z = b["lon"], b["lat"], b["z"]
    sig = TG.sigma_for_periods(b["meta"]["vbar"], b["meta"]["periods"], b["meta"]["wavelen_factor"])
    ms = []
    for _ in range(3):
        t = {}
        TG.topo_metric(z, lon, lat, sig, timing=t)
        ms.append(t["kernel_ms"])
    best = min(ms)
    nodes = sum(TG.window_nodes(n, n, lon, lat, s) for s in sig)
    return {"kernel_ms": round(best, 3), "window_nodes": nodes, "Gnodes_per_s": round(nodes / best / 1e6, 1),
            "fp64_TFLOPs": round(3 * nodes / best / 1e9, 2)}


def combine_workload(nx=201, ny=201, nz=30, nper=30):
    """The model kernels from 30 per-period kernels (vs-only, src/preproc.cpp:280-301)
    at 201x201 nodes x 30 depths: device kernel time and its HBM rate (each
    sensitivity read once, FP64)."""
    from surfattpp_amd import preproc as PP
    r = np.random.default_rng(7)
    adj = r.normal(size=(nper, ny, nx))
    sens = [r.normal(size=(nx, ny, nz, nper)) for _ in range(3)]
    vs = 2.5 + 2.0 * r.uniform(size=(nx, ny, nz))
    vp = 0.9409 + 2.0947 * vs - 0.8206 * vs**2 + 0.2683 * vs**3 - 0.0251 * vs**4
    ms = []
    for _ in range(3):
        t = {}
        PP.combine_kernels(adj, *sens, vs, vp, mode=0, timing=t)
        ms.append(t["kernel_ms"])
    nbytes = 8 * (3 * sens[0].size + adj.size + 2 * vs.size + vs.size)
    best = min(ms)
    return {"kernel_ms": round(best, 4), "bytes": nbytes, "GBps": round(nbytes / best / 1e6, 1),
            "frac_hbm": round(nbytes / best / 1e6 / HBM_PEAK_GBS, 4)}


def calib_copy(nbytes: int) -> None:
    """Two device-to-device copies of nbytes (HIP's blit kernel)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMemset failed")
        for _ in range(2):
            if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


if __name__ == "__main__":
    main()
