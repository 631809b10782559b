// k_wide_sweeps.hpp -- the batched sweep kernels of grids wider than 512
// columns (kMaxC columns per lane on W = 2, 4 or 8 waves per pair), included by
// k_wide64.hip and k_wide32.hip.
#pragma once
#include "sat_launch_impl.hpp"

namespace sat {

template <int PREC>
hipError_t launch_fwd_sweep_wide(int mode, int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first,
                                 int count, int group) {
  auto go = [&]<int WW, int MODE>() {
    hipLaunchKernelGGL((fwd_sweep_kernel<PREC, kMaxC, MODE, WW>), grid, dim3(kWave * WW), 0, s, pd, first, count,
                       group);
  };
  auto by_w = [&]<int MODE>() -> hipError_t {
    if (W == 2) go.template operator()<2, MODE>();
    else if (W == 4) go.template operator()<4, MODE>();
    else if (W == 8) go.template operator()<8, MODE>();
    else return hipErrorInvalidValue;
    return hipGetLastError();
  };
  if constexpr (PREC == SAT_PREC_F64) {
    // FP64 batches: the reference-order update; the FP64 re-solves: the reciprocal-based
    // one (general metric) or its isotropic form in scaled coordinates
    if (mode == kMetricExact) return by_w.template operator()<kMetricExact>();
    if (mode == kMetricGeneral) return by_w.template operator()<kMetricGeneral>();
    if (mode == kMetricIso) return by_w.template operator()<kMetricIso>();
    return hipErrorInvalidValue;
  } else {
    if (mode == kMetricIso) return by_w.template operator()<kMetricIso>();
    if (mode == kMetricGeneral) return by_w.template operator()<kMetricGeneral>();
    return hipErrorInvalidValue;  // polished sweeps: one wave per pair
  }
}

template <int PREC>
hipError_t launch_adj_sweep_wide(int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first, int count) {
  auto go = [&]<int WW>() {
    hipLaunchKernelGGL((adj_sweep_kernel<PREC, kMaxC, WW>), grid, dim3(kWave * WW), adj_dyn_lds<PREC>(pd), s, pd,
                       first, count);
  };
  if (W == 2) go.template operator()<2>();
  else if (W == 4) go.template operator()<4>();
  else if (W == 8) go.template operator()<8>();
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int PREC>
hipError_t launch_pair_wide(int ta_f32, int W, dim3 grid, hipStream_t s, const BatchDev& bd, int first, int count,
                            int group, const int* list, const int* dcount) {
  auto go = [&]<int WW>() {
    if (ta_f32)
      hipLaunchKernelGGL((pair_kernel<PREC, kMaxC, float, WW>), grid, dim3(kWave * WW), 0, s, bd, first, count,
                         group, list, dcount);
    else if constexpr (PREC == SAT_PREC_F64)
      hipLaunchKernelGGL((pair_kernel<PREC, kMaxC, double, WW>), grid, dim3(kWave * WW), 0, s, bd, first, count,
                         group, list, dcount);
  };
  if (W == 2) go.template operator()<2>();
  else if (W == 4) go.template operator()<4>();
  else if (W == 8) go.template operator()<8>();
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int PREC>
hipError_t launch_solo_wide(int forward, int W, hipStream_t s, const SoloDev& sd) {
  auto go = [&]<int WW>() {
    if (forward) hipLaunchKernelGGL((solo_forward<PREC, kMaxC, WW>), dim3(1), dim3(kWave * WW), 0, s, sd);
    else hipLaunchKernelGGL((solo_backward<PREC, kMaxC, WW>), dim3(1), dim3(kWave * WW), 0, s, sd);
  };
  if (W == 2) go.template operator()<2>();
  else if (W == 4) go.template operator()<4>();
  else if (W == 8) go.template operator()<8>();
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace sat
