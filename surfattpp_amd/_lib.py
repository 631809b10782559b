"""ctypes binding of libsurfatt_eik.so (include/surfatt_eik.h).

The library is the only compute path: if it is missing or cannot be loaded,
every entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsurfatt_eik.so")

SAT_OK = 0
SAT_EINVAL = 22
SAT_ENOMEM = 12
SAT_EDEVICE = 5
PREC_F64 = 0
PREC_F32 = 1
PREC_F32P = 2  # FP32 batched sweeps with the FP64-polished isotropic update
RUN_BATCHED = 0
RUN_PER_SOLVE = 1
RUN_GRAPH = 2
FLAG_NOT_CONVERGED = 1
FLAG_SRC_ON_BOUNDARY = 2
FLAG_ADJ_MAXITER = 4
FLAG_REC_OUTSIDE = 8
OPT_PAIR2 = 1      # scheduling options of a batch (sat_batch_set_option)
OPT_ADJ_COLS = 2
OPT_ORDER = 3
OPT_SPECULATE = 4

# Every symbol declared in include/surfatt_eik.h.
EXPORTS = (
    "sat_eik_forward", "sat_eik_forward_noalloc", "sat_eik_backward", "sat_eik_mask",
    "sat_batch_create", "sat_batch_update_media", "sat_batch_update_obs", "sat_batch_run",
    "sat_batch_sync", "sat_batch_results", "sat_batch_last_ms", "sat_batch_destroy",
    "sat_batch_set_f32_band", "sat_batch_conv", "sat_batch_phase_ms", "sat_batch_waves",
    "sat_batch_set_option", "sat_batch_get_option",
    "sat_combine_kernels", "sat_combine_kernels_ex", "sat_batch_combine_kernels", "sat_batch_aniso_kernels", "sat_batch_set_aniso_metric",
    "sat_batch_set_kden", "sat_batch_kden_results",
    "sat_topo_metric", "sat_surfdisp96", "sat_fwdsurf", "sat_device_info",
)


class SatError(RuntimeError):
    def __init__(self, fn: str, code: int):
        names = {SAT_EINVAL: "SAT_EINVAL", SAT_ENOMEM: "SAT_ENOMEM", SAT_EDEVICE: "SAT_EDEVICE"}
        super().__init__(f"{fn} failed: {names.get(code, code)}")
        self.code = code


class sat_info(C.Structure):
    _fields_ = [("iters", C.c_int), ("flags", C.c_int), ("l1", C.c_double), ("linf", C.c_double)]


class sat_grid_desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nperiod", C.c_int),
                ("lon_deg", C.c_void_p), ("lat_deg", C.c_void_p),
                ("svel", C.c_void_p), ("m11", C.c_void_p), ("m12", C.c_void_p), ("m22", C.c_void_p),
                ("waves", C.c_int)]


class sat_event_desc(C.Structure):
    _fields_ = [("grid", C.c_int), ("iper", C.c_int), ("src_lon_deg", C.c_double),
                ("src_lat_deg", C.c_double), ("rec_off", C.c_int), ("nrec", C.c_int)]


class sat_combine_desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nper", C.c_int), ("mode", C.c_int)] + \
        [(n, C.c_void_p) for n in ("adj_s", "sen_vs", "sen_vp", "sen_rho", "vs3d", "vp3d", "r1", "r2", "adj_xi",
                                   "adj_eta", "svel", "sen_gc", "sen_gs", "adj_den", "ker_vs", "ker_vp", "ker_rho",
                                   "ker_gc", "ker_gs", "ker_den")]


class sat_batch_desc(C.Structure):
    _fields_ = [("ngrid", C.c_int), ("grids", C.POINTER(sat_grid_desc)),
                ("nevent", C.c_int), ("events", C.POINTER(sat_event_desc)),
                ("nrec_total", C.c_int), ("rec_lon_deg", C.c_void_p), ("rec_lat_deg", C.c_void_p),
                ("tt_obs", C.c_void_p), ("weight", C.c_void_p),
                ("precision", C.c_int), ("calc_adj", C.c_int), ("workspace_cap", C.c_size_t)]


_LIB = None


def load(path: str = LIB_PATH):
    """Load the library (once).  Raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m surfattpp_amd.build` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.sat_eik_forward.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp,
                                  C.c_double, C.c_double, vp, C.POINTER(sat_info)]
    L.sat_eik_forward_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp,
                                          C.c_double, C.c_double, vp, C.POINTER(sat_info)]
    L.sat_eik_backward.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp,
                                   C.c_int, vp, vp, vp, vp, C.POINTER(sat_info)]
    L.sat_eik_mask.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.c_double, C.c_double]
    L.sat_batch_create.argtypes = [C.POINTER(sat_batch_desc), C.POINTER(C.c_void_p)]
    L.sat_batch_update_media.argtypes = [vp, C.c_int, vp, vp, vp, vp]
    L.sat_batch_update_obs.argtypes = [vp, vp, vp]
    L.sat_batch_run.argtypes = [vp, C.c_int]
    L.sat_batch_sync.argtypes = [vp]
    L.sat_batch_results.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.sat_batch_last_ms.argtypes = [vp, dp, dp]
    L.sat_batch_destroy.argtypes = [vp]
    L.sat_batch_set_f32_band.argtypes = [vp, C.c_double]
    L.sat_batch_conv.argtypes = [vp, vp, vp, vp]
    L.sat_batch_phase_ms.argtypes = [vp, dp, C.c_int]
    L.sat_batch_waves.argtypes = [vp, vp]
    L.sat_batch_set_option.argtypes = [vp, C.c_int, C.c_int]
    L.sat_batch_get_option.argtypes = [vp, C.c_int, vp]
    L.sat_combine_kernels.argtypes = [C.c_int] * 5 + [vp] * 9 + [dp]
    L.sat_combine_kernels_ex.argtypes = [C.POINTER(sat_combine_desc), dp]
    L.sat_batch_combine_kernels.argtypes = [vp, C.c_int, C.c_int, C.c_int] + [vp] * 8 + [dp]
    L.sat_batch_aniso_kernels.argtypes = [vp, C.c_int] + [vp] * 5
    L.sat_batch_set_aniso_metric.argtypes = [vp, C.c_int, vp, vp, vp]
    L.sat_batch_set_kden.argtypes = [vp, C.c_int]
    L.sat_batch_kden_results.argtypes = [vp, vp]
    L.sat_topo_metric.argtypes = [C.c_int] * 3 + [vp] * 8 + [dp]
    L.sat_surfdisp96.argtypes = [C.c_int, C.c_int] + [vp] * 4 + [C.c_int] * 5 + [vp] * 3 + [dp]
    L.sat_fwdsurf.argtypes = [C.c_int] * 3 + [vp] * 4 + [C.c_int, vp] + [C.c_int] * 4 + [vp, vp, dp]
    L.sat_device_info.argtypes = [C.c_char_p, C.c_int]
    for name in EXPORTS:
        getattr(L, name).restype = C.c_int
    _LIB = L
    return L


def check(fn: str, rc: int) -> None:
    if rc != SAT_OK:
        raise SatError(fn, rc)


def device_info() -> str:
    buf = C.create_string_buffer(256)
    rc = load().sat_device_info(buf, 256)
    check("sat_device_info", rc)
    return buf.value.decode()
