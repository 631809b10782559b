// sat_launch_impl.hpp -- definitions of the launchers declared in sat_launch.hpp;
// each k_*.hip includes this and instantiates its share explicitly.
#pragma once
#include "sat_launch.hpp"

namespace sat {

template <int PREC>
hipError_t launch_pair(int ta_f32, int C, int W, dim3 grid, hipStream_t s, const BatchDev& bd, int first,
                       int count, int group, const int* list, const int* dcount) {
  if (W != 1) {
    if (C != kMaxC) return hipErrorInvalidValue;
    return launch_pair_wide<PREC>(ta_f32, W, grid, s, bd, first, count, group, list, dcount);
  }
  const size_t lds = bd.adj_lds_bytes;  // the LDS tier of the adjoint (sat_fused.hpp)
  hipError_t ea = hipSuccess;
  const hipError_t e = with_C(C, [&]<int CC>() {
    if (ta_f32) {
      ea = allow_dyn_lds(pair_kernel<PREC, CC, float>, lds);
      if (ea == hipSuccess)
        hipLaunchKernelGGL((pair_kernel<PREC, CC, float>), grid, dim3(kWave), lds, s, bd, first, count, group, list,
                           dcount);
    } else if constexpr (PREC == SAT_PREC_F64) {
      ea = allow_dyn_lds(pair_kernel<PREC, CC, double>, lds);
      if (ea == hipSuccess)
        hipLaunchKernelGGL((pair_kernel<PREC, CC, double>), grid, dim3(kWave), lds, s, bd, first, count, group,
                           list, dcount);
    }
  });
  return ea != hipSuccess ? ea : e;
}

template <int PREC>
hipError_t launch_solo(int forward, int C, int W, hipStream_t s, const SoloDev& sd) {
  if (W != 1) {
    if (C != kMaxC) return hipErrorInvalidValue;
    return launch_solo_wide<PREC>(forward, W, s, sd);
  }
  hipError_t ea = hipSuccess;
  const hipError_t e = with_C(C, [&]<int CC>() {
    if (forward) {
      hipLaunchKernelGGL((solo_forward<PREC, CC>), dim3(1), dim3(kWave), 0, s, sd);
    } else {
      ea = allow_dyn_lds(solo_backward<PREC, CC>, sd.adj_lds_bytes);
      if (ea == hipSuccess)
        hipLaunchKernelGGL((solo_backward<PREC, CC>), dim3(1), dim3(kWave), sd.adj_lds_bytes, s, sd);
    }
  });
  return ea != hipSuccess ? ea : e;
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
  if (W != 1) return launch_fwd_sweep_mw<PREC>(iso, C, W, grid, s, pd, first, count, group);
  // iso: kMetricGeneral, kMetricIso or kMetricIsoPolish (FP32 only)
  if (iso == kMetricIsoPolish && PREC != SAT_PREC_F32) return hipErrorInvalidValue;
  return with_C(C, [&]<int CC>() {
    if constexpr (PREC == SAT_PREC_F32) {
      if (iso == kMetricIso) {
        hipLaunchKernelGGL((fwd_sweep_kernel<PREC, CC, kMetricIso, 1>), grid, dim3(kWave), 0, s, pd, first, count,
                           group);
        return;
      }
      if (iso == kMetricIsoPolish) {
        hipLaunchKernelGGL((fwd_sweep_kernel<PREC, CC, kMetricIsoPolish, 1>), grid, dim3(kWave), 0, s, pd, first,
                           count, group);
        return;
      }
    }
    hipLaunchKernelGGL((fwd_sweep_kernel<PREC, CC, kMetricGeneral, 1>), grid, dim3(kWave), 0, s, pd, first, count,
                       group);
  });
}

// Multi-wave sweeps (a work-group of W waves per event): FP64 W = kMultiWave
// or kMultiWaveWide, C <= 2, general metric (the re-solves of FP32 batches); FP32 W = 2 (C <= 4)
// and W = 4 (C <= 2), both metrics.
template <int PREC, int W, int C, int MODE>
static void launch_fwd_mw1(dim3 grid, hipStream_t s, const PhaseDev& pd, int first, int count, int group) {
  hipLaunchKernelGGL((fwd_sweep_kernel<PREC, C, MODE, W>), grid, dim3(kWave * W), 0, s, pd, first, count, group);
}

template <int PREC>
hipError_t launch_fwd_sweep_mw(int iso, int C, int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first,
                               int count, int group) {
  if (C == kMaxC) return launch_fwd_sweep_wide<PREC>(iso, W, grid, s, pd, first, count, group);
  if constexpr (PREC == SAT_PREC_F64) {
    // the re-solves of FP32 batches: the general metric or the isotropic one in
    // scaled coordinates (FwdF64RIso)
    if (iso != kMetricGeneral && iso != kMetricIso) return hipErrorInvalidValue;
    auto go = [&]<int WW, int CC>() {
      if (iso == kMetricIso) launch_fwd_mw1<PREC, WW, CC, kMetricIso>(grid, s, pd, first, count, group);
      else launch_fwd_mw1<PREC, WW, CC, kMetricGeneral>(grid, s, pd, first, count, group);
    };
    if (W == kMultiWave && C == 1) go.template operator()<kMultiWave, 1>();
    else if (W == kMultiWave && C == 2) go.template operator()<kMultiWave, 2>();
    else if (W == kMultiWaveWide && C == 1)  // grids of 257..384 columns (resolve_waves)
      go.template operator()<kMultiWaveWide, 1>();
    else return hipErrorInvalidValue;
  } else {
    if (iso == kMetricIsoPolish) return hipErrorInvalidValue;  // polished sweeps: one wave per pair
    auto go = [&]<int WW, int CC>() {
      if (iso) launch_fwd_mw1<PREC, WW, CC, kMetricIso>(grid, s, pd, first, count, group);
      else launch_fwd_mw1<PREC, WW, CC, kMetricGeneral>(grid, s, pd, first, count, group);
    };
    if (W == 2 && C == 1) go.template operator()<2, 1>();
    else if (W == 2 && C == 2) go.template operator()<2, 2>();
    else if (W == 2 && C == 3) go.template operator()<2, 3>();
    else if (W == 2 && C == 4) go.template operator()<2, 4>();
    else if (W == 4 && C == 1) go.template operator()<4, 1>();
    else if (W == 4 && C == 2) go.template operator()<4, 2>();
    else return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int PREC>
hipError_t launch_adj_setup(dim3 grid, hipStream_t s, const PhaseDev& pd) {
  hipLaunchKernelGGL((adj_setup_kernel<PREC>), grid, dim3(256), 0, s, pd);
  return hipGetLastError();
}

template <int PREC>
hipError_t launch_delta(dim3 grid, hipStream_t s, const PhaseDev& pd) {
  hipLaunchKernelGGL((delta_kernel<PREC>), grid, dim3(kWave), delta_lds_bytes(pd.delta_cap), s, pd);
  return hipGetLastError();
}

template <int PREC>
hipError_t launch_adj_sweep(int C, int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first, int count) {
  if (W != 1) return launch_adj_sweep_mw<PREC>(C, W, grid, s, pd, first, count);
  return with_C(C, [&]<int CC>() {
    hipLaunchKernelGGL((adj_sweep_kernel<PREC, CC, 1>), grid, dim3(kWave), adj_dyn_lds<PREC>(pd), s, pd, first,
                       count);
  });
}

template <int PREC>
hipError_t launch_adj_sweep_mw(int C, int W, dim3 grid, hipStream_t s, const PhaseDev& pd, int first, int count) {
  if (C == kMaxC) return launch_adj_sweep_wide<PREC>(W, grid, s, pd, first, count);
  auto go = [&]<int WW, int CC>() {
    hipLaunchKernelGGL((adj_sweep_kernel<PREC, CC, WW>), grid, dim3(kWave * WW), adj_dyn_lds<PREC>(pd), s, pd,
                       first, count);
  };
  if constexpr (PREC == SAT_PREC_F64) {
    if (W == kMultiWave && C == 1) go.template operator()<kMultiWave, 1>();
    else if (W == kMultiWave && C == 2) go.template operator()<kMultiWave, 2>();
    else if (W == kMultiWaveWide && C == 1) go.template operator()<kMultiWaveWide, 1>();
    else return hipErrorInvalidValue;
  } else {
    if (W == 2 && C == 1) go.template operator()<2, 1>();
    else if (W == 2 && C == 2) go.template operator()<2, 2>();
    else if (W == 2 && C == 3) go.template operator()<2, 3>();
    else if (W == 2 && C == 4) go.template operator()<2, 4>();
    else if (W == 4 && C == 1) go.template operator()<4, 1>();
    else if (W == 4 && C == 2) go.template operator()<4, 2>();
    else return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace sat

#define SAT_INSTANTIATE_PAIR(P)                                                                            \
  template hipError_t sat::launch_pair<P>(int, int, int, dim3, hipStream_t, const BatchDev&, int, int, int, \
                                          const int*, const int*);                                         \
  template hipError_t sat::launch_solo<P>(int, int, int, hipStream_t, const SoloDev&);

// the multi-wave launchers are instantiated in their own translation units
#define SAT_INSTANTIATE_PHASE(P)                                                                           \
  extern template hipError_t sat::launch_fwd_sweep_mw<P>(int, int, int, dim3, hipStream_t, const PhaseDev&, \
                                                         int, int, int);                                   \
  extern template hipError_t sat::launch_adj_sweep_mw<P>(int, int, dim3, hipStream_t, const PhaseDev&, int, \
                                                         int);                                             \
  template hipError_t sat::launch_trig<P>(dim3, hipStream_t, const PhaseDev&);                             \
  template hipError_t sat::launch_init<P>(dim3, hipStream_t, const PhaseDev&);                             \
  template hipError_t sat::launch_fwd_sweep<P>(int, int, int, dim3, hipStream_t, const PhaseDev&, int, int, int); \
  template hipError_t sat::launch_adj_setup<P>(dim3, hipStream_t, const PhaseDev&);                        \
  template hipError_t sat::launch_delta<P>(dim3, hipStream_t, const PhaseDev&);                            \
  template hipError_t sat::launch_adj_sweep<P>(int, int, dim3, hipStream_t, const PhaseDev&, int, int);

#define SAT_INSTANTIATE_MW(P)                                                                              \
  template hipError_t sat::launch_fwd_sweep_mw<P>(int, int, int, dim3, hipStream_t, const PhaseDev&, int, int, \
                                                  int);                                                    \
  template hipError_t sat::launch_adj_sweep_mw<P>(int, int, dim3, hipStream_t, const PhaseDev&, int, int);
