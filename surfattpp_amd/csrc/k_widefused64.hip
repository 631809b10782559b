// k_widefused64.hip -- FP64 per-pair and per-solve kernels of grids wider than 512 columns.
#include "k_wide_sweeps.hpp"

template hipError_t sat::launch_pair_wide<SAT_PREC_F64>(int, int, dim3, hipStream_t, const sat::BatchDev&, int, int, int,
                                                    const int*, const int*);
template hipError_t sat::launch_solo_wide<SAT_PREC_F64>(int, int, hipStream_t, const sat::SoloDev&);
