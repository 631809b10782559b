// sat_launch_impl.hpp -- definitions of the launchers declared in sat_launch.hpp;
// each k_*.hip includes this and instantiates its share explicitly.
#pragma once
#include "sat_launch.hpp"

namespace sat {

template <int PREC>
hipError_t launch_pair(int ta_f32, int C, dim3 grid, hipStream_t s, const BatchDev& bd, int first, int count,
                       int group, const int* list, const int* dcount) {
  return with_C(C, [&]<int CC>() {
    if (ta_f32)
      hipLaunchKernelGGL((pair_kernel<PREC, CC, float>), grid, dim3(kWave), 0, s, bd, first, count, group, list,
                         dcount);
    else if constexpr (PREC == SAT_PREC_F64)
      hipLaunchKernelGGL((pair_kernel<PREC, CC, double>), grid, dim3(kWave), 0, s, bd, first, count, group, list,
                         dcount);
  });
}

template <int PREC>
hipError_t launch_solo(int forward, int C, hipStream_t s, const SoloDev& sd) {
  return with_C(C, [&]<int CC>() {
    if (forward) hipLaunchKernelGGL((solo_forward<PREC, CC>), dim3(1), dim3(kWave), 0, s, sd);
    else hipLaunchKernelGGL((solo_backward<PREC, CC>), dim3(1), dim3(kWave), 0, s, sd);
  });
}

template <int PREC>
hipError_t launch_trig(dim3 grid, hipStream_t s, const PhaseDev& pd) {
  hipLaunchKernelGGL((trig_kernel<PREC>), grid, dim3(kWave), 0, s, pd);
  return hipGetLastError();
}

template <int PREC>
hipError_t launch_init(dim3 grid, hipStream_t s, const PhaseDev& pd) {
  hipLaunchKernelGGL((init_kernel<PREC>), grid, dim3(256), 0, s, pd);
  return hipGetLastError();
}

template <int PREC>
hipError_t launch_fwd_sweep(int iso, int C, int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first,
                            int count, int group) {
  if (W == kMultiWave) {
    if constexpr (PREC == SAT_PREC_F64) {
      if (C == 1)
        hipLaunchKernelGGL((fwd_sweep_kernel<PREC, 1, false, kMultiWave>), grid, dim3(kWave * kMultiWave), 0, s,
                           pd, first, count, group);
      else if (C == 2)
        hipLaunchKernelGGL((fwd_sweep_kernel<PREC, 2, false, kMultiWave>), grid, dim3(kWave * kMultiWave), 0, s,
                           pd, first, count, group);
      else
        return hipErrorInvalidValue;
      return hipGetLastError();
    }
    return hipErrorInvalidValue;
  }
  if (W != 1) return hipErrorInvalidValue;
  return with_C(C, [&]<int CC>() {
    if constexpr (PREC == SAT_PREC_F32) {
      if (iso) {
        hipLaunchKernelGGL((fwd_sweep_kernel<PREC, CC, true, 1>), grid, dim3(kWave), 0, s, pd, first, count, group);
        return;
      }
    }
    hipLaunchKernelGGL((fwd_sweep_kernel<PREC, CC, false, 1>), grid, dim3(kWave), 0, s, pd, first, count, group);
  });
}

template <int PREC>
hipError_t launch_adj_setup(dim3 grid, hipStream_t s, const PhaseDev& pd) {
  hipLaunchKernelGGL((adj_setup_kernel<PREC>), grid, dim3(256), 0, s, pd);
  return hipGetLastError();
}

template <int PREC>
hipError_t launch_delta(dim3 grid, hipStream_t s, const PhaseDev& pd) {
  hipLaunchKernelGGL((delta_kernel<PREC>), grid, dim3(kWave), 0, s, pd);
  return hipGetLastError();
}

template <int PREC>
hipError_t launch_adj_sweep(int C, int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first, int count) {
  if (W == kMultiWave) {
    if constexpr (PREC == SAT_PREC_F64) {
      if (C == 1)
        hipLaunchKernelGGL((adj_sweep_kernel<PREC, 1, kMultiWave>), grid, dim3(kWave * kMultiWave), 0, s, pd, first,
                           count);
      else if (C == 2)
        hipLaunchKernelGGL((adj_sweep_kernel<PREC, 2, kMultiWave>), grid, dim3(kWave * kMultiWave), 0, s, pd, first,
                           count);
      else
        return hipErrorInvalidValue;
      return hipGetLastError();
    }
    return hipErrorInvalidValue;
  }
  if (W != 1) return hipErrorInvalidValue;
  return with_C(C, [&]<int CC>() {
    hipLaunchKernelGGL((adj_sweep_kernel<PREC, CC, 1>), grid, dim3(kWave), 0, s, pd, first, count);
  });
}

}  // namespace sat

#define SAT_INSTANTIATE_PAIR(P)                                                                            \
  template hipError_t sat::launch_pair<P>(int, int, dim3, hipStream_t, const BatchDev&, int, int, int,     \
                                          const int*, const int*);                                         \
  template hipError_t sat::launch_solo<P>(int, int, hipStream_t, const SoloDev&);

#define SAT_INSTANTIATE_PHASE(P)                                                                           \
  template hipError_t sat::launch_trig<P>(dim3, hipStream_t, const PhaseDev&);                             \
  template hipError_t sat::launch_init<P>(dim3, hipStream_t, const PhaseDev&);                             \
  template hipError_t sat::launch_fwd_sweep<P>(int, int, int, dim3, hipStream_t, const PhaseDev&, int, int, int); \
  template hipError_t sat::launch_adj_setup<P>(dim3, hipStream_t, const PhaseDev&);                        \
  template hipError_t sat::launch_delta<P>(dim3, hipStream_t, const PhaseDev&);                            \
  template hipError_t sat::launch_adj_sweep<P>(int, int, dim3, hipStream_t, const PhaseDev&, int, int);
