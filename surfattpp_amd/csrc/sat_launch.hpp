// sat_launch.hpp -- host-side launchers of the kernel templates.  Each is
// defined, with its instantiations over the columns-per-lane C, in one of the
// k_*.hip translation units, so the kernels build in parallel.
#pragma once
#include "sat_fused.hpp"
#include "sat_phases.hpp"

namespace sat {

// pair_kernel<PREC, C, TaT, W>; ta_f32: masked Ta stored as float (FP32 batches).
// W = 1 (C in [1, kMaxC]) or, for grids wider than 512 columns, W = wide_waves(nx)
// with C = kMaxC (the same for launch_solo).
template <int PREC>
hipError_t launch_pair(int ta_f32, int C, int W, dim3 grid, hipStream_t s, const BatchDev& bd, int first,
                       int count, int group, const int* list, const int* dcount);
template <int PREC>
hipError_t launch_solo(int forward, int C, int W, hipStream_t s, const SoloDev& sd);
// Their wide-grid instantiations (k_wide64.hip, k_wide32.hip): W in {2, 4, 8}, C = kMaxC.
template <int PREC>
hipError_t launch_pair_wide(int ta_f32, int W, dim3 grid, hipStream_t s, const BatchDev& bd, int first, int count,
                            int group, const int* list, const int* dcount);
template <int PREC>
hipError_t launch_solo_wide(int forward, int W, hipStream_t s, const SoloDev& sd);
// The batched sweeps of wide grids: W in {2, 4, 8}, C = kMaxC; forward mode
// kMetricExact (FP64 batches), 0 (the FP64 re-solves), kMetricGeneral or
// kMetricIso (FP32 batches).
template <int PREC>
hipError_t launch_fwd_sweep_wide(int mode, int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first,
                                 int count, int group);
template <int PREC>
hipError_t launch_adj_sweep_wide(int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first, int count);

template <int PREC>
hipError_t launch_trig(dim3 grid, hipStream_t s, const PhaseDev& pd);
template <int PREC>
hipError_t launch_init(dim3 grid, hipStream_t s, const PhaseDev& pd);
// Sweep kernels: W = 1 (one wave per event, C in [1, kMaxC]) or a work-group
// of W waves per event: FP64 W = kMultiWave or kMultiWaveWide, C in [1, 2]
// (the re-solves of FP32 batches); FP32 W = 2 (C in [1, 4]) or 4 (C in [1, 2]).
constexpr int kMultiWave = 4;
constexpr int kMultiWaveWide = 8;
// Waves per pair of the FP64 re-solves of a grid whose one-wave sweeps take C1
// columns per lane (C1 = ceil(nx / 64)): four up to 256 columns; eight for
// 257..384 columns, which then sweep one column per lane instead of two.  The
// re-solves are latency-bound; one column per lane halves a step but makes a
// sweep longer by a third and its barrier span eight waves: measured, the
// 321-column grid's re-solves gain 10 %, the 401-column grid's lose 11 %.
inline int resolve_waves(int C1) { return (C1 > kMultiWave && C1 <= 6) ? kMultiWaveWide : kMultiWave; }
template <int PREC>
hipError_t launch_fwd_sweep(int iso, int C, int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first,
                            int count, int group);
template <int PREC>
hipError_t launch_fwd_sweep_mw(int iso, int C, int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first,
                               int count, int group);
// FP32 isotropic forward sweeps, two pairs per wave (C columns per lane, C <= kPair2MaxC)
hipError_t launch_fwd_sweep_pair2(int C, dim3 grid, hipStream_t s, const PhaseDev& pd, int first, int count,
                                  int group, int hstride);
hipError_t launch_recv(dim3 grid, hipStream_t s, const PhaseDev& pd, int nrec);
template <int PREC>
hipError_t launch_adj_setup(dim3 grid, hipStream_t s, const PhaseDev& pd);
template <int PREC>
hipError_t launch_delta(dim3 grid, hipStream_t s, const PhaseDev& pd);
template <int PREC>
hipError_t launch_adj_sweep(int C, int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first, int count);
template <int PREC>
hipError_t launch_adj_sweep_mw(int C, int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first, int count);

// Run f.template operator()<C>() for C in [1, kMaxC].
template <class F>
hipError_t with_C(int C, F&& f);

// Lets a kernel take `bytes` of dynamic LDS beyond the 64 KiB a launch gets
// without asking (the LDS tier of the per-pair adjoint).
template <class K>
hipError_t allow_dyn_lds(K* kernel, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}

template <class F>
hipError_t with_C(int C, F&& f) {
  switch (C) {
    case 1: f.template operator()<1>(); break;
    case 2: f.template operator()<2>(); break;
    case 3: f.template operator()<3>(); break;
    case 4: f.template operator()<4>(); break;
    case 5: f.template operator()<5>(); break;
    case 6: f.template operator()<6>(); break;
    case 7: f.template operator()<7>(); break;
    case 8: f.template operator()<8>(); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace sat
