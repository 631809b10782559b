// k_pair2.hip -- the two-pairs-per-wave FP32 forward sweeps (sat_pair2.hpp).
#include "sat_launch.hpp"
#include "sat_pair2.hpp"

namespace sat {

hipError_t launch_fwd_sweep_pair2(int C, dim3 grid, hipStream_t s, const PhaseDev& pd, int first, int count,
                                  int group, int hstride) {
  switch (C) {
#define SAT_P2(CC)                                                                                             \
  case CC:                                                                                                     \
    hipLaunchKernelGGL((fwd_sweep_kernel_pair2<CC>), grid, dim3(kWave), 0, s, pd, first, count, group, hstride); \
    break;
    SAT_P2(1) SAT_P2(2) SAT_P2(3) SAT_P2(4) SAT_P2(5) SAT_P2(6) SAT_P2(7)
#undef SAT_P2
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace sat
