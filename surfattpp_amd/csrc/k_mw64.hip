// k_mw64.hip -- FP64 multi-wave sweep kernels (the re-solves of FP32 batches).
#include "sat_launch_impl.hpp"

using namespace sat;
SAT_INSTANTIATE_MW(SAT_PREC_F64)
