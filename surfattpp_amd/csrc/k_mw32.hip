// k_mw32.hip -- FP32 multi-wave sweep kernels (several waves per pair).
#include "sat_launch_impl.hpp"

using namespace sat;
SAT_INSTANTIATE_MW(SAT_PREC_F32)
