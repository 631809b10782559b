// sat_phases.hpp -- the batched pipeline as a sequence of kernels.
//
//   trig_kernel        per (event, column): sin/cos of the longitude offset; per
//                      event: slowness at the source, sin/cos of its latitude
//   init_kernel        per (event, node): analytic factor T0, T0x, T0y
//                      (src/eikonal_solver.cpp:96-117)
//   fwd_sweep_kernel   persistent, one wave per event: forward iterations, T
//                      (:146-317); FP32: flags pairs for the FP64 re-solve
//   recv_kernel        per receiver: interp2d, residual, adjoint source
//                      (src/preproc.cpp:105-115)
//   adj_setup_kernel   per (event, node): face fluxes and the sweep coefficients
//                      (src/eikonal_solver.cpp:377-441)
//   delta_kernel       per event: receivers spread in receiver order (:356-375),
//                      chi term (src/preproc.cpp:142)
//   adj_sweep_kernel   persistent, one wave per event: adjoint iterations, mask,
//                      masked Ta (src/eikonal_solver.cpp:471-536)
// Node-parallel work runs at full occupancy; the wave-per-event kernels hold only
// the sweep pipeline in registers.
#pragma once
#include <type_traits>

#include "sat_crmath.hpp"
#include "sat_engine.hpp"
#include "sat_policies.hpp"

namespace sat {

struct EventX {
  int grid, iper, rec_off, nrec;
  int fb[4];      // fixed near-source box x0, x1, y0, y1 (src/eikonal_solver.cpp:119-129)
  int mb[4];      // mask box x0, x1, y0, y1 (:528-531)
  double lon, lat;
  size_t seg_off; // element offset of this event's segmented planes
  size_t pl_off;  // element offset of this event's column-major planes
  size_t col_off; // offset into the column trig table
};

struct PhaseDev {
  const GridDev* grids;
  const EventX* ev;
  int nevent;
  double* evs;          // per event: fun0, sin(y0), cos(y0)
  double2* coltab;      // per (event, column): sin, cos of xx - x0
  void* t0;             // per event segmented planes, sweep precision
  void* t0x;
  void* t0y;
  double* T0d;          // FP32 path: FP64 T0, column-major, per event
  double* T;            // travel time, column-major, per event
  void* co[6];          // adjoint coefficients, segmented, per event
  void* ta_out;         // masked Ta, column-major, per event
  const double* rec_lon;
  const double* rec_lat;
  const double* tt_obs;
  const double* weight;
  const int* rec_ev;    // receiver -> event
  double* syn;
  double* adjsrc;
  int2* rcell;          // per receiver: delta cell (idx0, idy0)
  double* rwt;          // per receiver: w, r1, r2 (src/eikonal_solver.cpp:360-369)
  double* chi_ev;
  int* iters_f;
  int* iters_a;
  int* flags;
  double* conv;
  unsigned char* skip;  // FP32: events re-solved in FP64
  int* fb_list;
  int* fb_count;
  double f32_band;
  char* slots;          // per resident wave: 3 swept planes
  size_t slot_bytes, slot_plane;
  const int* order;
  int* counter;
  int fwd_maxiter, adj_maxiter;
};

template <int PREC> struct PhaseT;
template <> struct PhaseT<SAT_PREC_F64> {
  using V = double;     // T0 planes and adjoint
  using U = double;     // forward swept value (tau)
  using Fwd = FwdF64;
  using FwdIso = FwdF64;
  using Adj = AdjF64;
};
template <> struct PhaseT<SAT_PREC_F32> {
  using V = float;
  using U = double;     // sigma = tau - 1 in FP64
  using Fwd = FwdF32;
  using FwdIso = FwdF32Iso;
  using Adj = AdjF32;
};

// ---------------------------------------------------------------------------
template <int PREC>
__global__ void trig_kernel(PhaseDev pd) {
  const int e = blockIdx.y;
  const EventX ev = pd.ev[e];
  const GridDev& g = pd.grids[ev.grid];
  const double x0 = ev.lon * kDEG2RAD;
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  if (ix < g.nx) {
    double s, c;
    cr::sincos_cr(g.xx[ix] - x0, s, c);
    pd.coltab[ev.col_off + ix] = make_double2(s, c);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double y0 = ev.lat * kDEG2RAD;
    // src/eikonal_solver.cpp:71-82 source cell and bilinear slowness
    int idx0 = (int)floor((x0 - g.xx[0]) / g.dx);
    int idy0 = (int)floor((y0 - g.yy[0]) / g.dy);
    idx0 = imax(0, imin(idx0, g.nx - 2));
    idy0 = imax(0, imin(idy0, g.ny - 2));
    const double r1 = dmin(1.0, (x0 - g.xx[idx0]) / g.dx);
    const double r2 = dmin(1.0, (y0 - g.yy[idy0]) / g.dy);
    pd.evs[3 * (size_t)e] = bilinear_cm(g.f_plain + (size_t)ev.iper * g.plane, g.nx, idx0, idy0, r1, r2);
    double s, c;
    cr::sincos_cr(y0, s, c);
    pd.evs[3 * (size_t)e + 1] = s;
    pd.evs[3 * (size_t)e + 2] = c;
  }
}

// src/eikonal_solver.cpp:96-117; acos correctly rounded in the FP64 path.
template <int PREC>
__global__ void init_kernel(PhaseDev pd) {
  using V = typename PhaseT<PREC>::V;
  const int e = blockIdx.y;
  const EventX ev = pd.ev[e];
  const GridDev& g = pd.grids[ev.grid];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.P * g.ny) return;
  const int iy = q / g.P, ix = q - iy * g.P;
  V* t0 = static_cast<V*>(pd.t0) + ev.seg_off;
  V* t0x = static_cast<V*>(pd.t0x) + ev.seg_off;
  V* t0y = static_cast<V*>(pd.t0y) + ev.seg_off;
  if (ix >= g.nx) {
    t0[q] = V(0); t0x[q] = V(0); t0y[q] = V(0);
    return;
  }
  const double fun0 = pd.evs[3 * (size_t)e], sy0 = pd.evs[3 * (size_t)e + 1], cy0 = pd.evs[3 * (size_t)e + 2];
  const double2 rw = g.trig_yy[iy];
  const double2 cl = pd.coltab[ev.col_off + ix];
  const double syy = rw.x, cyy = rw.y, sdl = cl.x, cdl = cl.y;
  const double tmp = syy * sy0 + cyy * cy0 * cdl;
  const double tmp_c = dmax(-1.0, dmin(1.0, tmp));
  const double ac = (PREC == SAT_PREC_F64) ? cr::acos_cr(tmp_c) : acos(tmp_c);
  const double T0 = ac * kR * fun0;
  const double sin_d = sqrt(dmax(0.0, 1.0 - tmp_c * tmp_c));
  double T0x = 0.0, T0y = 0.0;
  if (!(sin_d < 1e-14)) {
    T0x = cyy * cy0 * sdl / sin_d * kR * fun0;
    T0y = -(cyy * sy0 - syy * cy0 * cdl) / sin_d * kR * fun0;
  }
  t0[q] = (V)T0;
  t0x[q] = (V)T0x;
  t0y[q] = (V)T0y;
  if (PREC == SAT_PREC_F32) pd.T0d[ev.pl_off + ix + (size_t)iy * g.nx] = T0;
}

template <int PREC, bool ISO>
__device__ inline auto phase_fwd_ctx(const PhaseDev& pd, const GridDev& g, const EventX& ev) {
  using PT = PhaseT<PREC>;
  using V = typename PT::V;
  using Pol = std::conditional_t<ISO, typename PT::FwdIso, typename PT::Fwd>;
  typename Pol::Ctx cx;
  const size_t po = (size_t)ev.iper * g.seg_plane;
  cx.g = &g;
  cx.t0 = static_cast<const V*>(pd.t0) + ev.seg_off;
  cx.t0x = static_cast<const V*>(pd.t0x) + ev.seg_off;
  cx.t0y = static_cast<const V*>(pd.t0y) + ev.seg_off;
  cx.fb = FixedBox{ev.fb[0], ev.fb[1], ev.fb[2], ev.fb[3]};
  if constexpr (PREC == SAT_PREC_F64) {
    cx.a = static_cast<const V*>(g.seg[0]) + po;
    cx.b = static_cast<const V*>(g.seg[1]) + po;
    cx.c = static_cast<const V*>(g.seg[2]) + po;
    cx.f = static_cast<const V*>(g.seg[3]) + po;
    cx.dx = g.dx;
    cx.dy = g.dy;
  } else if constexpr (ISO) {
    cx.f = static_cast<const V*>(g.seg[3]) + po;
    cx.rowtab = g.rowtab;
    cx.idx = (float)(1.0 / g.dx);
    cx.idy = (float)(1.0 / g.dy);
  } else {
    cx.a = static_cast<const V*>(g.seg[0]) + po;
    cx.b = static_cast<const V*>(g.seg[1]) + po;
    cx.c = static_cast<const V*>(g.seg[2]) + po;
    cx.f = static_cast<const V*>(g.seg[3]) + po;
    cx.dsx = static_cast<const V*>(g.seg[4]) + po;
    cx.dsy = static_cast<const V*>(g.seg[5]) + po;
    cx.idx = (float)(1.0 / g.dx);
    cx.idy = (float)(1.0 / g.dy);
  }
  return cx;
}

// Persistent forward sweeps, one wave per event.
template <int PREC, int C, bool ISO>
__global__ __launch_bounds__(64) void fwd_sweep_kernel(PhaseDev pd, int first, int count, int group) {
  using PT = PhaseT<PREC>;
  using U = typename PT::U;
  using Pol = std::conditional_t<ISO, typename PT::FwdIso, typename PT::Fwd>;
  const int lane = threadIdx.x;
  char* sb = pd.slots + (size_t)blockIdx.x * pd.slot_bytes;
  U* Q0 = reinterpret_cast<U*>(sb);
  U* Q1 = reinterpret_cast<U*>(sb + pd.slot_plane);
  U* Q2 = reinterpret_cast<U*>(sb + 2 * pd.slot_plane);
  constexpr bool kSigma = PREC == SAT_PREC_F32;
  for (;;) {
    int idx = 0;
    if (lane == 0) idx = atomicAdd(pd.counter, 1);
    idx = __shfl(idx, 0);
    if (idx >= count) break;
    const int e = pd.order[first + idx];
    const EventX ev = pd.ev[e];
    const GridDev& g = pd.grids[ev.grid];
    const int P = g.P, nx = g.nx, n = P * g.ny;
    for (int q = lane; q < n; q += kWave) {  // :119-129 initial tau (sigma)
      const int iy = q / P, ix = q - iy * P;
      const bool fixed = ix < nx && ix >= ev.fb[0] && ix <= ev.fb[1] && iy >= ev.fb[2] && iy <= ev.fb[3];
      Q0[q] = kSigma ? (fixed ? U(0) : U(9)) : (fixed ? U(1) : U(10));
    }
    __syncthreads();
    const auto cx = phase_fwd_ctx<PREC, ISO>(pd, g, ev);
    IterOut io;
    const U* u = iterate<Pol, C>(cx, Q0, Q1, Q2, pd.fwd_maxiter, kFwdTol, io);
    if (lane == 0) {
      pd.conv[2 * (size_t)e] = io.linf;
      pd.conv[2 * (size_t)e + 1] = io.linf_prev;
      pd.iters_f[e] = io.iters;
      if (!io.converged) atomicOr(pd.flags + e, FLAG_NOT_CONVERGED);
    }
    if constexpr (kSigma) {
      // A stop/continue decision within f32_band of the tolerance: re-solved in FP64.
      const double b = pd.f32_band;
      const bool near_last = io.converged && io.linf >= (1.0 - b) * kFwdTol;
      const bool near_prev = io.linf_prev >= 0.0 && io.linf_prev < (1.0 + b) * kFwdTol;
      const bool amb = b > 0.0 && (near_last || near_prev);
      if (lane == 0) {
        pd.skip[e] = amb ? 1 : 0;
        if (amb) pd.fb_list[atomicAdd(pd.fb_count + group, 1)] = e;
      }
      if (amb) {
        __syncthreads();
        continue;
      }
    }
    // :317 T = tau .* T0 (FP64) or T0 + sigma*T0 (FP32 sweeps), column-major
    double* T = pd.T + ev.pl_off;
    const typename PT::V* t0 = static_cast<const typename PT::V*>(pd.t0) + ev.seg_off;
    for (int iy = 0; iy < g.ny; iy++)
      for (int ix = lane; ix < nx; ix += kWave) {
        const int o = iy * P + ix;
        const size_t p = ix + (size_t)iy * nx;
        if (kSigma) T[p] = pd.T0d[ev.pl_off + p] + (double)u[o] * pd.T0d[ev.pl_off + p];
        else T[p] = (double)u[o] * (double)t0[o];
      }
    __syncthreads();
  }
}

// src/preproc.cpp:105-115 per receiver, and the receiver's delta cell and
// weight (src/eikonal_solver.cpp:357-369).
__global__ void recv_kernel(PhaseDev pd, int nrec) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nrec; k += gridDim.x * blockDim.x) {
    const int e = pd.rec_ev[k];
    if (e < 0 || pd.skip[e]) continue;
    const EventX& ev = pd.ev[e];
    const GridDev& g = pd.grids[ev.grid];
    const double s = interp2d_point(g.lon, g.nx, g.lat, g.ny, pd.T + ev.pl_off, pd.rec_lon[k], pd.rec_lat[k]);
    pd.syn[k] = s;
    const double adj = pd.weight[k] * (s - pd.tt_obs[k]);
    pd.adjsrc[k] = adj;
    if (s != s) atomicOr(pd.flags + e, FLAG_REC_OUTSIDE);
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

// src/eikonal_solver.cpp:407-441, :446-449 per node.  FP64: a1p, a2m, b1p, b2m,
// dl = 0, d.  FP32: cW, cE, cS, cN (the update divided through by d), cD = 0.
template <int PREC>
__global__ void adj_setup_kernel(PhaseDev pd) {
  const int e = blockIdx.y;
  if (pd.skip[e]) return;
  const EventX ev = pd.ev[e];
  const GridDev& g = pd.grids[ev.grid];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.P * g.ny) return;
  const int iy = q / g.P, ix = q - iy * g.P;
  const bool interior = ix >= 1 && ix < g.nx - 1 && iy >= 1 && iy < g.ny - 1;
  const double* a = g.aA_plain + (size_t)ev.iper * g.plane;
  const double* b = g.b_plain + (size_t)ev.iper * g.plane;
  const double* c = g.c_plain + (size_t)ev.iper * g.plane;
  const size_t o = ev.seg_off + q;
  if constexpr (PREC == SAT_PREC_F64) {
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (interior) {
      const Faces F = faces_at(g, pd.T + ev.pl_off, a, b, c, ix, iy);
      v[0] = F.a1p; v[1] = F.a2m; v[2] = F.b1p; v[3] = F.b2m;
      v[5] = (F.a2p - F.a1m) / g.dx + (F.b2p - F.b1m) / g.dy;  // :479-481
    }
#pragma unroll
    for (int i = 0; i < 6; i++) static_cast<double*>(pd.co[i])[o] = v[i];
  } else {
    float v[5] = {0, 0, 0, 0, 0};
    if (interior) {
      const Faces F = faces_at(g, pd.T + ev.pl_off, a, b, c, ix, iy);
      const double d = (F.a2p - F.a1m) / g.dx + (F.b2p - F.b1m) / g.dy;
      if (!(fabs(d) < kAdjEps)) {
        const double rdx = 1.0 / (g.dx * d), rdy = 1.0 / (g.dy * d);
        v[0] = (float)(F.a1p * rdx);
        v[1] = (float)(-F.a2m * rdx);
        v[2] = (float)(F.b1p * rdy);
        v[3] = (float)(-F.b2m * rdy);
      }
    }
#pragma unroll
    for (int i = 0; i < 5; i++) static_cast<float*>(pd.co[i])[o] = v[i];
  }
}

// src/eikonal_solver.cpp:356-375 and src/preproc.cpp:142, one thread per event.
// delta(node) is the receiver-order sum of the contributions to that node.
template <int PREC>
__global__ void delta_kernel(PhaseDev pd) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= pd.nevent || pd.skip[e]) return;
  const EventX ev = pd.ev[e];
  const GridDev& g = pd.grids[ev.grid];
  const int nx = g.nx, ny = g.ny;
  double s = 0.0;
  for (int r = 0; r < ev.nrec; r++) {
    const int k = ev.rec_off + r;
    const double d = pd.syn[k] - pd.tt_obs[k];
    s += pd.weight[k] * (d * d);
  }
  pd.chi_ev[e] = 0.5 * s;
  // contribution m of receiver k to node (idx0 + (m>>1), idy0 + (m&1)), order of :371-374
  auto contrib = [&](int k, int m) {
    const double w = pd.rwt[3 * (size_t)k], r1 = pd.rwt[3 * (size_t)k + 1], r2 = pd.rwt[3 * (size_t)k + 2];
    return (m == 0) ? w * (1 - r1) * (1 - r2) : (m == 1) ? w * (1 - r1) * r2 : (m == 2) ? w * r1 * (1 - r2)
                                                                                       : w * r1 * r2;
  };
  const int k0 = ev.rec_off, k1 = ev.rec_off + ev.nrec;
  if constexpr (PREC == SAT_PREC_F64) {
    double* dl = static_cast<double*>(pd.co[4]) + ev.seg_off;
    for (int k = k0; k < k1; k++) {
      const int2 c = pd.rcell[k];
      for (int m = 0; m < 4; m++) dl[(c.y + (m & 1)) * g.P + c.x + (m >> 1)] += contrib(k, m);
    }
  } else {
    float* cD = static_cast<float*>(pd.co[4]) + ev.seg_off;
    const double* a = g.aA_plain + (size_t)ev.iper * g.plane;
    const double* b = g.b_plain + (size_t)ev.iper * g.plane;
    const double* cc = g.c_plain + (size_t)ev.iper * g.plane;
    for (int k = k0; k < k1; k++) {
      const int2 c = pd.rcell[k];
      for (int m = 0; m < 4; m++) {
        const int ni = c.x + (m >> 1), nj = c.y + (m & 1);
        bool seen = false;  // an earlier receiver touched this node: already summed
        for (int kk = k0; kk < k && !seen; kk++) {
          const int2 c2 = pd.rcell[kk];
          seen = (ni - c2.x == 0 || ni - c2.x == 1) && (nj - c2.y == 0 || nj - c2.y == 1);
        }
        if (seen) continue;
        double acc = 0.0;
        for (int kk = k; kk < k1; kk++) {
          const int2 c2 = pd.rcell[kk];
          for (int mm = 0; mm < 4; mm++)
            if (c2.x + (mm >> 1) == ni && c2.y + (mm & 1) == nj) acc += contrib(kk, mm);
        }
        const bool interior = ni >= 1 && ni < nx - 1 && nj >= 1 && nj < ny - 1;
        float v = 0.0f;
        if (interior) {
          const Faces F = faces_at(g, pd.T + ev.pl_off, a, b, cc, ni, nj);
          const double d = (F.a2p - F.a1m) / g.dx + (F.b2p - F.b1m) / g.dy;
          if (!(fabs(d) < kAdjEps)) v = (float)(acc / d);
        }
        cD[nj * g.P + ni] = v;
      }
    }
  }
}

// Persistent adjoint sweeps, one wave per event; mask and store.
template <int PREC, int C>
__global__ __launch_bounds__(64) void adj_sweep_kernel(PhaseDev pd, int first, int count) {
  using PT = PhaseT<PREC>;
  using V = typename PT::V;
  using Pol = typename PT::Adj;
  const int lane = threadIdx.x;
  char* sb = pd.slots + (size_t)blockIdx.x * pd.slot_bytes;
  V* Q0 = reinterpret_cast<V*>(sb);
  V* Q1 = reinterpret_cast<V*>(sb + pd.slot_plane);
  V* Q2 = reinterpret_cast<V*>(sb + 2 * pd.slot_plane);
  for (;;) {
    int idx = 0;
    if (lane == 0) idx = atomicAdd(pd.counter, 1);
    idx = __shfl(idx, 0);
    if (idx >= count) break;
    const int e = pd.order[first + idx];
    if (pd.skip[e]) continue;
    const EventX ev = pd.ev[e];
    const GridDev& g = pd.grids[ev.grid];
    const int P = g.P, nx = g.nx, ny = g.ny, n = P * ny;
    for (int q = lane; q < n; q += kWave) {  // :446-449 boundary 0, interior 100
      const int iy = q / P, ix = q - iy * P;
      Q0[q] = (ix >= 1 && ix < nx - 1 && iy >= 1 && iy < ny - 1) ? V(100) : V(0);
    }
    __syncthreads();
    typename Pol::Ctx cx;
    cx.g = &g;
    if constexpr (PREC == SAT_PREC_F64) {
      cx.a1p = static_cast<const double*>(pd.co[0]) + ev.seg_off;
      cx.a2m = static_cast<const double*>(pd.co[1]) + ev.seg_off;
      cx.b1p = static_cast<const double*>(pd.co[2]) + ev.seg_off;
      cx.b2m = static_cast<const double*>(pd.co[3]) + ev.seg_off;
      cx.dl = static_cast<const double*>(pd.co[4]) + ev.seg_off;
      cx.d = static_cast<const double*>(pd.co[5]) + ev.seg_off;
      cx.dx = g.dx;
      cx.dy = g.dy;
    } else {
      cx.cW = static_cast<const float*>(pd.co[0]) + ev.seg_off;
      cx.cE = static_cast<const float*>(pd.co[1]) + ev.seg_off;
      cx.cS = static_cast<const float*>(pd.co[2]) + ev.seg_off;
      cx.cN = static_cast<const float*>(pd.co[3]) + ev.seg_off;
      cx.cD = static_cast<const float*>(pd.co[4]) + ev.seg_off;
    }
    IterOut io;
    const V* ta = iterate<Pol, C>(cx, Q0, Q1, Q2, pd.adj_maxiter, kAdjTol, io);
    if (lane == 0) {
      pd.iters_a[e] = io.iters;
      if (!io.converged) atomicOr(pd.flags + e, FLAG_ADJ_MAXITER);
    }
    V* out = static_cast<V*>(pd.ta_out) + ev.pl_off;  // :514-536 mask, column-major
    for (int iy = 0; iy < ny; iy++)
      for (int ix = lane; ix < nx; ix += kWave) {
        const bool m = ix >= ev.mb[0] && ix <= ev.mb[1] && iy >= ev.mb[2] && iy <= ev.mb[3];
        out[ix + (size_t)iy * nx] = m ? V(0) : ta[iy * P + ix];
      }
    __syncthreads();
  }
}

}  // namespace sat
