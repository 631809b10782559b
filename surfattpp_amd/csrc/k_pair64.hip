// k_pair64.hip -- FP64 instantiations of the per-pair and per-solve kernels (sat_fused.hpp).
#include "sat_launch_impl.hpp"

using namespace sat;
SAT_INSTANTIATE_PAIR(SAT_PREC_F64)
