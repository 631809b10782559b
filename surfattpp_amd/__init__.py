"""surfattpp_amd -- MI355X-native (gfx950) eikonal forward + adjoint hot path of SurfATT.

Python mirror of the reference's call surface over the C ABI in
include/surfatt_eik.h (library: libsurfatt_eik.so, built in-tree):

  surfattpp_amd.eikonal.FSM_UW_PS_lonlat_2d / FSM_O1_JSE_lonlat_2d / mask_uniform_grid
  surfattpp_amd.preproc.EikonalBatch / forward_for_event
  surfattpp_amd.tune.autotune / tuned_batch   (work-group shape per grid)
  surfattpp_amd.synth                          (seeded workloads, topography metric)
"""
from . import _lib  # noqa: F401
from ._lib import PREC_F32, PREC_F32P, PREC_F64, RUN_BATCHED, RUN_GRAPH, RUN_PER_SOLVE  # noqa: F401

__all__ = ["eikonal", "preproc", "synth", "tune", "PREC_F32", "PREC_F32P", "PREC_F64", "RUN_BATCHED",
           "RUN_GRAPH", "RUN_PER_SOLVE"]
