// k_wide64.hip -- FP64 batched sweeps of grids wider than 512 columns.
#include "k_wide_sweeps.hpp"

template hipError_t sat::launch_fwd_sweep_wide<SAT_PREC_F64>(int, int, dim3, hipStream_t, const sat::PhaseDev&, int, int,
                                                         int);
template hipError_t sat::launch_adj_sweep_wide<SAT_PREC_F64>(int, dim3, hipStream_t, const sat::PhaseDev&, int, int);
