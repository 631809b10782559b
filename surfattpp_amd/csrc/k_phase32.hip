// k_phase32.hip -- FP32 instantiations of the batched phase kernels (sat_phases.hpp).
#include "sat_launch_impl.hpp"

using namespace sat;
SAT_INSTANTIATE_PHASE(SAT_PREC_F32)
