// k_phase64.hip -- FP64 instantiations of the batched phase kernels (sat_phases.hpp).
#include "sat_launch_impl.hpp"

using namespace sat;
SAT_INSTANTIATE_PHASE(SAT_PREC_F64)

namespace sat {

// src/preproc.cpp:105-115 per receiver, and the receiver's delta cell and
// weight (src/eikonal_solver.cpp:357-369); in the density-kernel pass only the
// weight of the source -w.
__global__ void recv_kernel(PhaseDev pd, int nrec) {
  for (int k = pd.rec_lo + (int)(blockIdx.x * blockDim.x + threadIdx.x); k < pd.rec_lo + nrec;
       k += gridDim.x * blockDim.x) {
    const int e = pd.rec_ev[k];
    if (e < pd.ev_base || e >= pd.ev_base + pd.nevent || !pass_event(pd, e)) continue;
    const EventX& ev = pd.ev[e];
    const GridDev& g = pd.grids[ev.grid];
    double adj;
    if (pd.kden_pass) {  // the density kernel's source: -w (src/preproc.cpp:129-130)
      adj = -pd.weight[k];
    } else {
      const double s = interp2d_point(g.lon, g.nx, g.lat, g.ny, pd.T + ev.pl_off, pd.rec_lon[k], pd.rec_lat[k]);
      pd.syn[k] = s;
      adj = pd.weight[k] * (s - pd.tt_obs[k]);
      pd.adjsrc[k] = adj;
      if (s != s) atomicOr(pd.flags + e, FLAG_REC_OUTSIDE);
    }
    const double xrec = pd.rec_lon[k] * kDEG2RAD;
    const double yrec = pd.rec_lat[k] * kDEG2RAD;
    int idx0 = (int)floor((xrec - g.xx[0]) / g.dx);
    int idy0 = (int)floor((yrec - g.yy[0]) / g.dy);
    idx0 = imax(0, imin(idx0, g.nx - 2));
    idy0 = imax(0, imin(idy0, g.ny - 2));
    const double r1 = dmin(1.0, (xrec - g.xx[idx0]) / g.dx);
    const double r2 = dmin(1.0, (yrec - g.yy[idy0]) / g.dy);
    const double area = g.dx * kR * cr::cos_cr(yrec) * g.dy * kR;
    pd.rcell[k] = make_int2(idx0, idy0);
    pd.rwt[3 * (size_t)k] = adj / area;
    pd.rwt[3 * (size_t)k + 1] = r1;
    pd.rwt[3 * (size_t)k + 2] = r2;
  }
}

}  // namespace sat

hipError_t sat::launch_recv(dim3 grid, hipStream_t s, const PhaseDev& pd, int nrec) {
  hipLaunchKernelGGL(recv_kernel, grid, dim3(256), 0, s, pd, nrec);
  return hipGetLastError();
}
