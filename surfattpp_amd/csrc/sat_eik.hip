// sat_eik.hip -- kernels and C ABI of the gfx950 eikonal forward + adjoint path.
//
// Kernels
//   media_kernel      per-period metric/slowness planes (src/eikonal_solver.cpp:55-65,
//                     :380-389; src/preproc.cpp:97-103 and :29)
//   pair_kernel<C>    persistent: one 64-lane wave drains a queue of (source, period)
//                     pairs; per pair forward solve -> receiver interp2d -> residual ->
//                     adjoint solve -> mask -> masked Ta out (src/preproc.cpp:79-143)
//   reduce_kernel     per-period kernel planes, adj_s[iper] = sum_e Ta_e * svel^-3 in
//                     event order (src/preproc.cpp:28-31)
//   chi_kernel        chi = sum_e 0.5*sum_r w r^2 in event order (src/preproc.cpp:142)
//   solo_*_kernel<C>  the per-solve entry points (FSM_UW_PS_lonlat_2d, FSM_O1_JSE_lonlat_2d)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <type_traits>
#include <vector>

#include "../../include/surfatt_eik.h"
#include "sat_device.hpp"
#include "sat_engine.hpp"
#include "sat_phases.hpp"
#include "sat_policies.hpp"

namespace sat {

constexpr int kMaxC = 8;  // nx <= 512

template <int PREC> struct Prec;
template <> struct Prec<SAT_PREC_F64> {
  using Fwd = FwdF64;
  using Adj = AdjF64;
  using V = double;
  static constexpr int kSeg = 9;    // T0 T0x T0y u0 u1 u2 | a1p a2m b1p b2m dl d ta0 ta1 ta2
  static constexpr int kPlain = 1;  // T
};
template <> struct Prec<SAT_PREC_F32> {
  using Fwd = FwdF32;
  using Adj = AdjF32;
  using V = float;
  static constexpr int kSeg = 8;    // t0 t0x t0y u0 u1 u2 | cW cE cS cN cD ta0 ta1 ta2
  static constexpr int kPlain = 4;  // T, T0 (FP64), d, dl (kept all-zero between solves)
};

// ----------------------------------------------------------------------------
// Device descriptors
// ----------------------------------------------------------------------------
struct EventDev {
  int grid, iper;
  double lon, lat;
  int rec_off, nrec;
  size_t ta_off;  // element offset of this event's masked Ta plane in the Ta output
};

struct BatchDev {
  const GridDev* grids;
  const EventDev* events;
  const int* order;      // event ids, processed in this order
  int* counter;          // work-queue head (one per launch group)
  const double* rec_lon;
  const double* rec_lat;
  const double* tt_obs;
  const double* weight;
  double* syn;
  double* adjsrc;
  double* chi_ev;
  int* iters_f;
  int* iters_a;
  int* flags;
  void* ta_out;          // masked Ta planes (the batch's sweep precision)
  double* conv;          // per event: Linf of the last two forward rounds
  int* fb_list;          // FP32 batches: events re-solved in FP64 (decision near the tolerance)
  int* fb_count;         // one counter per launch group
  double f32_band;       // relative band around the forward tolerance that triggers the re-solve
  char* ws;              // slot workspaces
  size_t ws_slot;        // bytes per slot
  size_t ws_seg;         // bytes per segmented plane
  size_t ws_plain;       // bytes per plain FP64 plane
  int calc_adj;
  int fwd_maxiter, adj_maxiter;
};

__device__ inline Slot make_slot(char* base, int nseg, int nplain, size_t seg_bytes, size_t plain_bytes) {
  Slot s;
#pragma unroll
  for (int i = 0; i < 9; i++) s.seg[i] = (i < nseg) ? (void*)(base + (size_t)i * seg_bytes) : nullptr;
#pragma unroll
  for (int i = 0; i < 4; i++)
    s.pl[i] = (i < nplain) ? reinterpret_cast<double*>(base + (size_t)nseg * seg_bytes + (size_t)i * plain_bytes)
                           : nullptr;
  return s;
}

// ----------------------------------------------------------------------------
// Media planes.  mode 0: from (svel, m11, m12, m22) as forward_for_event passes
// them (spha = m11, sphb = m22, sphc = -m12, fun = (1/svel)*1,
// src/preproc.cpp:97-103).  mode 1: from (spha, sphb, sphc, fun) directly.
// ----------------------------------------------------------------------------
__global__ void media_kernel(GridDev g, int mode, int prec, const double* __restrict__ in0,
                             const double* __restrict__ in1, const double* __restrict__ in2,
                             const double* __restrict__ in3, double* f_plain, double* aA_plain, double* b_plain,
                             double* c_plain, double* vinv3, void* s0, void* s1, void* s2, void* s3, void* s4,
                             void* s5) {
  const int nx = g.nx, ny = g.ny, P = g.P;
  const size_t total = (size_t)g.nperiod * ny * P;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total;
       id += (size_t)gridDim.x * blockDim.x) {
    const int q = (int)(id % P);
    const size_t rest = id / P;
    const int iy = (int)(rest % ny);
    const int ip = (int)(rest / ny);
    const size_t so = (size_t)ip * g.seg_plane + (size_t)iy * P + q;
    double va = 1.0, vb = 1.0, vc = 0.0, vf = 1.0;
    if (q < nx) {
      const size_t po = (size_t)ip * g.plane + (size_t)iy * nx + q;
      double spha, sphb, sphc, sv = 1.0;
      if (mode == 0) {
        sv = in0[po];
        spha = in1[po];
        sphc = -in2[po];
        sphb = in3[po];
        vf = (1.0 / sv) * 1.0;
      } else {
        spha = in0[po];
        sphb = in1[po];
        sphc = in2[po];
        vf = in3[po];
      }
      const double cos_y = cr::cos_cr(g.yy[iy]);
      const double cos2_y = cos_y * cos_y;
      const double R2 = kR * kR;
      va = spha / (R2 * cos2_y);            // src/eikonal_solver.cpp:61
      vb = sphb / R2;                       // :62
      vc = sphc / (kR * kR * cos_y);        // :63
      f_plain[po] = vf;
      aA_plain[po] = spha / (R2 * cos_y * cos_y);  // :385
      b_plain[po] = sphb / R2;                     // :386
      c_plain[po] = sphc / (R2 * cos_y);           // :387
      if (vinv3) vinv3[po] = (mode == 0) ? cr::pow_m3_cr(sv) : 0.0;  // src/preproc.cpp:29
    }
    if (prec == SAT_PREC_F64) {
      static_cast<double*>(s0)[so] = va;
      static_cast<double*>(s1)[so] = vb;
      static_cast<double*>(s2)[so] = vc;
      static_cast<double*>(s3)[so] = vf;
    } else {
      const double det = va * vb - vc * vc;
      static_cast<float*>(s0)[so] = (float)va;
      static_cast<float*>(s1)[so] = (float)vb;
      static_cast<float*>(s2)[so] = (float)vc;
      static_cast<float*>(s3)[so] = (float)vf;
      static_cast<float*>(s4)[so] = (det > 0.0) ? (float)sqrt(vf * vf * vb / det) : -1.0f;
      static_cast<float*>(s5)[so] = (det > 0.0) ? (float)sqrt(vf * vf * va / det) : -1.0f;
    }
  }
}

// Per row: correctly rounded sin/cos of the latitude; isotropic grids also the
// FP32 row metric {a, b, 1/sqrt(a), 1/sqrt(b)} (src/eikonal_solver.cpp:55-65 with
// spha = sphb = 1, sphc = 0).
__global__ void row_kernel(GridDev g, double2* trig_yy, float4* rowtab) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  if (iy >= g.ny) return;
  double s, c;
  cr::sincos_cr(g.yy[iy], s, c);
  trig_yy[iy] = make_double2(s, c);
  if (rowtab) {
    const double R2 = kR * kR;
    const double a = 1.0 / (R2 * (c * c)), b = 1.0 / R2;
    rowtab[iy] = make_float4((float)a, (float)b, (float)(1.0 / sqrt(a)), (float)(1.0 / sqrt(b)));
  }
}

template <int PREC>
__device__ inline typename Prec<PREC>::Fwd::Ctx fwd_ctx(const GridDev& g, int iper, const Slot& ws,
                                                         const FixedBox& fb) {
  using V = typename Prec<PREC>::V;
  typename Prec<PREC>::Fwd::Ctx cx;
  const size_t po = (size_t)iper * g.seg_plane;
  cx.g = &g;
  cx.t0 = static_cast<const V*>(ws.seg[0]);
  cx.t0x = static_cast<const V*>(ws.seg[1]);
  cx.t0y = static_cast<const V*>(ws.seg[2]);
  cx.a = static_cast<const V*>(g.seg[0]) + po;
  cx.b = static_cast<const V*>(g.seg[1]) + po;
  cx.c = static_cast<const V*>(g.seg[2]) + po;
  if constexpr (PREC == SAT_PREC_F64) {
    cx.f = static_cast<const V*>(g.seg[3]) + po;
    cx.dx = g.dx;
    cx.dy = g.dy;
  } else {
    cx.f = static_cast<const V*>(g.seg[3]) + po;
    cx.dsx = static_cast<const V*>(g.seg[4]) + po;
    cx.dsy = static_cast<const V*>(g.seg[5]) + po;
    cx.idx = (float)(1.0 / g.dx);
    cx.idy = (float)(1.0 / g.dy);
  }
  cx.fb = fb;
  return cx;
}

template <int PREC>
__device__ inline typename Prec<PREC>::Adj::Ctx adj_ctx(const GridDev& g, const Slot& ws) {
  using V = typename Prec<PREC>::V;
  typename Prec<PREC>::Adj::Ctx cx;
  cx.g = &g;
  if constexpr (PREC == SAT_PREC_F64) {
    cx.a1p = static_cast<const V*>(ws.seg[0]);
    cx.a2m = static_cast<const V*>(ws.seg[1]);
    cx.b1p = static_cast<const V*>(ws.seg[2]);
    cx.b2m = static_cast<const V*>(ws.seg[3]);
    cx.dl = static_cast<const V*>(ws.seg[4]);
    cx.d = static_cast<const V*>(ws.seg[5]);
    cx.dx = g.dx;
    cx.dy = g.dy;
  } else {
    cx.cW = static_cast<const V*>(ws.seg[0]);
    cx.cE = static_cast<const V*>(ws.seg[1]);
    cx.cS = static_cast<const V*>(ws.seg[2]);
    cx.cN = static_cast<const V*>(ws.seg[3]);
    cx.cD = static_cast<const V*>(ws.seg[4]);
  }
  return cx;
}

// Forward solve (src/eikonal_solver.cpp:22-317) into the plain FP64 plane T = pl[0].
template <int PREC, int C>
__device__ inline IterOut solve_forward(const GridDev& g, int iper, double lon, double lat, const Slot& ws,
                                        int maxiter, int& flags) {
  using P = Prec<PREC>;
  using V = typename P::V;            // T0 planes
  using U = typename P::Fwd::V;       // swept value (tau, or sigma in FP64 for the FP32 path)
  V* S[3];
  U* Q[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    S[i] = static_cast<V*>(ws.seg[i]);
    Q[i] = static_cast<U*>(ws.seg[3 + i]);
  }
  const FixedBox fb = fwd_init<V, U>(g, iper, lon, lat, S[0], S[1], S[2], Q[0], ws.pl[1], flags);
  __syncthreads();
  const auto cx = fwd_ctx<PREC>(g, iper, ws, fb);
  IterOut io;
  const U* u = iterate<typename P::Fwd, C>(cx, Q[0], Q[1], Q[2], maxiter, kFwdTol, io);
  if (!io.converged) flags |= FLAG_NOT_CONVERGED;
  fwd_finish<V, U>(g, u, S[0], ws.pl[1], ws.pl[0]);
  __syncthreads();
  return io;
}

// Adjoint solve (src/eikonal_solver.cpp:323-511) from T (plain FP64); returns Ta plane.
template <int PREC, int C>
__device__ inline const typename Prec<PREC>::V* solve_adjoint(const GridDev& g, int iper, const double* T, int nr,
                                                               const double* rlon, const double* rlat,
                                                               const double* adj, const Slot& ws, int maxiter,
                                                               IterOut& io, int& flags) {
  using P = Prec<PREC>;
  using V = typename P::V;
  if constexpr (PREC == SAT_PREC_F64) adj_setup_f64(g, iper, T, ws);
  else adj_setup_f32(g, iper, T, ws);
  __syncthreads();
  adj_delta<V>(g, nr, rlon, rlat, adj, ws);
  __syncthreads();
  const auto cx = adj_ctx<PREC>(g, ws);
  const int t0 = (PREC == SAT_PREC_F64) ? 6 : 5;
  const V* ta = iterate<typename P::Adj, C>(cx, static_cast<V*>(ws.seg[t0]), static_cast<V*>(ws.seg[t0 + 1]),
                                            static_cast<V*>(ws.seg[t0 + 2]), maxiter, kAdjTol, io);
  if (!io.converged) flags |= FLAG_ADJ_MAXITER;
  return ta;
}

// ----------------------------------------------------------------------------
// One (source, period) pair: src/preproc.cpp:88-143.
// ----------------------------------------------------------------------------
template <int PREC, int C, typename TaT>
__device__ void run_pair(const BatchDev& bd, int e, const Slot& ws, int group) {
  using V = typename Prec<PREC>::V;
  const int lane = threadIdx.x;
  const EventDev ev = bd.events[e];
  const GridDev& g = bd.grids[ev.grid];
  int flags = 0;
  const IterOut fo = solve_forward<PREC, C>(g, ev.iper, ev.lon, ev.lat, ws, bd.fwd_maxiter, flags);
  if (lane == 0) {
    bd.conv[2 * (size_t)e] = fo.linf;
    bd.conv[2 * (size_t)e + 1] = fo.linf_prev;
  }
  if constexpr (PREC == SAT_PREC_F32) {
    // The FP32 iterate's convergence norm differs from the FP64 one by its
    // rounding; when a stop/continue decision lies within f32_band of the
    // tolerance the FP32 run may stop one round apart from the reference.
    // Such pairs are re-solved by the FP64 path after this launch.
    const double b = bd.f32_band;
    const bool near_last = fo.converged && fo.linf >= (1.0 - b) * kFwdTol;
    const bool near_prev = fo.linf_prev >= 0.0 && fo.linf_prev < (1.0 + b) * kFwdTol;
    if (b > 0.0 && (near_last || near_prev)) {
      if (lane == 0) {
        const int slot = atomicAdd(bd.fb_count + group, 1);
        bd.fb_list[slot] = e;
      }
      __syncthreads();
      return;
    }
  }
  const double* T = ws.pl[0];
  // :105-115 receivers, residuals, adjoint source
  int outside = 0;
  for (int r = lane; r < ev.nrec; r += kWave) {
    const int k = ev.rec_off + r;
    const double s = interp2d_point(g.lon, g.nx, g.lat, g.ny, T, bd.rec_lon[k], bd.rec_lat[k]);
    bd.syn[k] = s;
    bd.adjsrc[k] = bd.weight[k] * (s - bd.tt_obs[k]);
    outside |= (s != s) ? 1 : 0;
  }
  if (wave_or_i(outside)) flags |= FLAG_REC_OUTSIDE;
  __syncthreads();
  int aiters = 0;
  if (bd.calc_adj) {
    if (lane == 0) {  // :142 chi term of this event, receiver order
      double s = 0.0;
      for (int r = 0; r < ev.nrec; r++) {
        const int k = ev.rec_off + r;
        const double d = bd.syn[k] - bd.tt_obs[k];
        s += bd.weight[k] * (d * d);
      }
      bd.chi_ev[e] = 0.5 * s;
    }
    IterOut ao;
    const V* ta = solve_adjoint<PREC, C>(g, ev.iper, T, ev.nrec, bd.rec_lon + ev.rec_off, bd.rec_lat + ev.rec_off,
                                         bd.adjsrc + ev.rec_off, ws, bd.adj_maxiter, ao, flags);
    aiters = ao.iters;
    mask_store<V, TaT>(g, ev.lon, ev.lat, ta, static_cast<TaT*>(bd.ta_out) + ev.ta_off);  // :125
  } else if (lane == 0) {
    bd.chi_ev[e] = 0.0;
  }
  if (lane == 0) {
    bd.iters_f[e] = fo.iters;
    bd.iters_a[e] = aiters;
    bd.flags[e] = flags;
  }
  __syncthreads();
}

// Persistent: each wave takes the next queue entry until the queue is empty.
// The queue is order[first, first+count), or list[0, *dcount) when list != nullptr
// (the FP64 re-solves of an FP32 batch, written by the preceding launch).
template <int PREC, int C, typename TaT>
__global__ __launch_bounds__(64) void pair_kernel(BatchDev bd, int first, int count, int group, const int* list,
                                                  const int* dcount) {
  using P = Prec<PREC>;
  const Slot ws = make_slot(bd.ws + (size_t)blockIdx.x * bd.ws_slot, P::kSeg, P::kPlain, bd.ws_seg, bd.ws_plain);
  const int n = list ? *dcount : count;
  for (;;) {
    int idx = 0;
    if (threadIdx.x == 0) idx = atomicAdd(bd.counter, 1);
    idx = __shfl(idx, 0);
    if (idx >= n) break;
    run_pair<PREC, C, TaT>(bd, list ? list[idx] : bd.order[first + idx], ws, group);
  }
}

// adj_s[plane][node] = sum over the plane's events (batch order) of Ta_e * svel^-3
// (src/preproc.cpp:28-31), accumulated in FP64.
struct ReduceDev {
  int nplanes;
  const size_t* plane_off;     // output offset of each (grid, period) plane
  const int* plane_n;          // nodes per plane
  const int* ev_off;           // CSR into ev_ids
  const int* ev_ids;
  const double* const* vinv3;  // svel^-3 plane of each (grid, period) plane
  const EventDev* events;
  const void* ta_out;
  double* K;
  size_t total_nodes;
  const size_t* node_prefix;   // prefix sum of plane_n (size nplanes+1)
};

template <typename V>
__global__ void reduce_kernel(ReduceDev rd) {
  const V* ta = static_cast<const V*>(rd.ta_out);
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < rd.total_nodes;
       id += (size_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = rd.nplanes - 1;  // plane containing id
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (rd.node_prefix[mid] <= id) lo = mid; else hi = mid - 1;
    }
    const int pl = lo;
    const size_t n = id - rd.node_prefix[pl];
    const double v3 = rd.vinv3[pl][n];
    double acc = 0.0;
    for (int q = rd.ev_off[pl]; q < rd.ev_off[pl + 1]; q++) {
      const EventDev& ev = rd.events[rd.ev_ids[q]];
      acc += (double)ta[ev.ta_off + n] * v3;
    }
    rd.K[rd.plane_off[pl] + n] = acc;
  }
}

__global__ void chi_kernel(const double* chi_ev, int n, double* chi) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double c = 0.0;
    for (int e = 0; e < n; e++) c += chi_ev[e];
    *chi = c;
  }
}

// ----------------------------------------------------------------------------
// Per-solve kernels (one wave, one solve)
// ----------------------------------------------------------------------------
struct SoloDev {
  GridDev g;
  double src_lon, src_lat;
  const double* T_in;      // backward: forward travel time, column-major
  int nr;
  const double* rlon;
  const double* rlat;
  const double* adj;
  double* out;             // T or Ta, column-major
  char* ws;
  size_t ws_seg, ws_plain;
  int maxiter;
  double* info;            // iters, flags, l1, linf
};

template <int PREC, int C>
__global__ __launch_bounds__(64) void solo_forward(SoloDev sd) {
  using P = Prec<PREC>;
  const Slot ws = make_slot(sd.ws, P::kSeg, P::kPlain, sd.ws_seg, sd.ws_plain);
  int flags = 0;
  const IterOut io = solve_forward<PREC, C>(sd.g, 0, sd.src_lon, sd.src_lat, ws, sd.maxiter, flags);
  for (size_t p = threadIdx.x; p < sd.g.plane; p += kWave) sd.out[p] = ws.pl[0][p];
  if (threadIdx.x == 0) {
    sd.info[0] = io.iters; sd.info[1] = flags; sd.info[2] = io.l1; sd.info[3] = io.linf;
  }
}

template <int PREC, int C>
__global__ __launch_bounds__(64) void solo_backward(SoloDev sd) {
  using P = Prec<PREC>;
  using V = typename P::V;
  const Slot ws = make_slot(sd.ws, P::kSeg, P::kPlain, sd.ws_seg, sd.ws_plain);
  int flags = 0;
  IterOut io;
  const V* ta = solve_adjoint<PREC, C>(sd.g, 0, sd.T_in, sd.nr, sd.rlon, sd.rlat, sd.adj, ws, sd.maxiter, io, flags);
  for (int iy = 0; iy < sd.g.ny; iy++)
    for (int ix = threadIdx.x; ix < sd.g.nx; ix += kWave)
      sd.out[ix + (size_t)iy * sd.g.nx] = (double)ta[seg_index(sd.g, ix, iy)];
  if (threadIdx.x == 0) {
    sd.info[0] = io.iters; sd.info[1] = flags; sd.info[2] = io.l1; sd.info[3] = io.linf;
  }
}

// ----------------------------------------------------------------------------
// Host helpers
// ----------------------------------------------------------------------------
#define SAT_CHECK(x)                                              \
  do {                                                            \
    hipError_t err_ = (x);                                        \
    if (err_ != hipSuccess) {                                     \
      std::fprintf(stderr, "[surfatt_eik] %s:%d %s: %s\n", __FILE__, __LINE__, #x, \
                   hipGetErrorString(err_));                      \
      return (err_ == hipErrorOutOfMemory) ? SAT_ENOMEM : SAT_EDEVICE; \
    }                                                             \
  } while (0)

struct Layout {
  int C, La, P;
};

static Layout layout_for(int nx) {
  Layout L;
  L.C = (nx + kWave - 1) / kWave;
  L.La = (nx + L.C - 1) / L.C;
  L.P = L.La * L.C;
  return L;
}

static bool uniform_axis(const double* v, int n) {
  if (n < 2) return false;
  const double d = v[1] - v[0];
  if (!(d > 0.0) || !std::isfinite(d)) return false;
  for (int i = 1; i < n; i++) {
    const double di = v[i] - v[i - 1];
    if (!(std::fabs(di - d) <= 1e-6 * std::fabs(d))) return false;
  }
  return true;
}

static int device_check() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return SAT_EDEVICE;
  return SAT_OK;
}

template <class T>
static hipError_t dalloc(T** p, size_t n) {
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (n ? n : 1));
}

// Device copy of one grid's geometry + media.
struct GridHost {
  GridDev g{};
  int nx = 0, ny = 0, nper = 0, prec = SAT_PREC_F64;
  double *d_lon = nullptr, *d_lat = nullptr, *d_xx = nullptr, *d_yy = nullptr;
  double *d_model = nullptr;  // 4 * nper * N raw planes (svel, m11, m12, m22) or per-solve inputs
  double *d_fp = nullptr, *d_aA = nullptr, *d_b = nullptr, *d_c = nullptr, *d_v3 = nullptr;
  char* d_seg = nullptr;      // 6 segmented media planes per period (sweep precision)
  char* d_seg64 = nullptr;    // FP32 batches: the FP64 media for the re-solves
  GridDev g64{};              // FP32 batches: g with the FP64 media
  double2* d_trig_yy = nullptr;
  float4* d_rowtab = nullptr;

  void release() {
    void* ps[] = {d_lon, d_lat, d_xx, d_yy, d_model, d_fp, d_aA, d_b, d_c, d_v3, d_seg, d_seg64, d_trig_yy, d_rowtab};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    *this = GridHost();
  }
};

static size_t elem_size(int prec) { return prec == SAT_PREC_F64 ? sizeof(double) : sizeof(float); }

static int grid_setup(GridHost& G, int prec, int nx, int ny, int nper, const double* lon, const double* lat) {
  if (nx < 3 || ny < 3 || nper < 1 || !lon || !lat) return SAT_EINVAL;
  if (!uniform_axis(lon, nx) || !uniform_axis(lat, ny)) return SAT_EINVAL;
  const Layout L = layout_for(nx);
  if (L.C > kMaxC) return SAT_EINVAL;
  G.nx = nx; G.ny = ny; G.nper = nper; G.prec = prec;
  std::vector<double> xx(nx), yy(ny);
  for (int i = 0; i < nx; i++) xx[i] = lon[i] * kDEG2RAD;  // src/eikonal_solver.cpp:41
  for (int j = 0; j < ny; j++) yy[j] = lat[j] * kDEG2RAD;
  const size_t N = (size_t)nx * ny, S = (size_t)L.P * ny;
  SAT_CHECK(dalloc(&G.d_lon, nx));
  SAT_CHECK(dalloc(&G.d_lat, ny));
  SAT_CHECK(dalloc(&G.d_xx, nx));
  SAT_CHECK(dalloc(&G.d_yy, ny));
  SAT_CHECK(dalloc(&G.d_model, 4 * nper * N));
  SAT_CHECK(dalloc(&G.d_fp, nper * N));
  SAT_CHECK(dalloc(&G.d_aA, nper * N));
  SAT_CHECK(dalloc(&G.d_b, nper * N));
  SAT_CHECK(dalloc(&G.d_c, nper * N));
  SAT_CHECK(dalloc(&G.d_v3, nper * N));
  const size_t seg_bytes = nper * S * elem_size(prec);
  SAT_CHECK(dalloc(&G.d_seg, 6 * seg_bytes));
  const size_t seg64_bytes = nper * S * sizeof(double);
  if (prec == SAT_PREC_F32) SAT_CHECK(dalloc(&G.d_seg64, 6 * seg64_bytes));
  SAT_CHECK(hipMemcpy(G.d_lon, lon, sizeof(double) * nx, hipMemcpyHostToDevice));
  SAT_CHECK(hipMemcpy(G.d_lat, lat, sizeof(double) * ny, hipMemcpyHostToDevice));
  SAT_CHECK(hipMemcpy(G.d_xx, xx.data(), sizeof(double) * nx, hipMemcpyHostToDevice));
  SAT_CHECK(hipMemcpy(G.d_yy, yy.data(), sizeof(double) * ny, hipMemcpyHostToDevice));
  GridDev& g = G.g;
  g.nx = nx; g.ny = ny; g.nperiod = nper;
  g.C = L.C; g.La = L.La; g.P = L.P;
  g.seg_plane = S; g.plane = N;
  g.dx = xx[1] - xx[0];  // :46-47
  g.dy = yy[1] - yy[0];
  g.lon = G.d_lon; g.lat = G.d_lat; g.xx = G.d_xx; g.yy = G.d_yy;
  g.f_plain = G.d_fp; g.aA_plain = G.d_aA; g.b_plain = G.d_b; g.c_plain = G.d_c; g.vinv3 = G.d_v3;
  for (int i = 0; i < 6; i++) g.seg[i] = G.d_seg + (size_t)i * seg_bytes;
  SAT_CHECK(dalloc(&G.d_trig_yy, ny));
  SAT_CHECK(dalloc(&G.d_rowtab, ny));
  g.trig_yy = G.d_trig_yy;
  g.rowtab = G.d_rowtab;
  g.iso = 0;
  hipLaunchKernelGGL(row_kernel, dim3((ny + 63) / 64), dim3(64), 0, nullptr, g, G.d_trig_yy, G.d_rowtab);
  SAT_CHECK(hipGetLastError());
  SAT_CHECK(hipDeviceSynchronize());
  G.g64 = g;
  if (G.d_seg64)
    for (int i = 0; i < 6; i++) G.g64.seg[i] = G.d_seg64 + (size_t)i * seg64_bytes;
  return SAT_OK;
}

// Upload four planes (nper each) and rebuild the media on `stream`.
static int grid_media(GridHost& G, int mode, const double* p0, const double* p1, const double* p2,
                      const double* p3, hipStream_t stream) {
  const size_t N = (size_t)G.nx * G.ny, n = (size_t)G.nper * N;
  const double* src[4] = {p0, p1, p2, p3};
  for (int i = 0; i < 4; i++)
    if (!src[i]) return SAT_EINVAL;
  {  // isotropic metric: spha = sphb = 1 and sphc = 0 everywhere
    const double* ma = (mode == 0) ? p1 : p0;   // m11 | spha
    const double* mb = (mode == 0) ? p3 : p1;   // m22 | sphb
    const double* mc = p2;                      // m12 | sphc
    bool iso = true;
    for (size_t k = 0; k < n && iso; k++) iso = ma[k] == 1.0 && mb[k] == 1.0 && mc[k] == 0.0;
    G.g.iso = iso ? 1 : 0;
    G.g64.iso = G.g.iso;
  }
  for (int i = 0; i < 4; i++) {
    SAT_CHECK(hipMemcpyAsync(G.d_model + i * n, src[i], sizeof(double) * n, hipMemcpyHostToDevice, stream));
  }
  const size_t total = (size_t)G.nper * G.ny * G.g.P;
  const int threads = 256;
  const int blocks = (int)std::min<size_t>((total + threads - 1) / threads, 65535);
  for (int pass = 0; pass < (G.d_seg64 ? 2 : 1); pass++) {
    const GridDev& gd = pass ? G.g64 : G.g;
    void* sg[6];
    for (int i = 0; i < 6; i++) sg[i] = const_cast<void*>(gd.seg[i]);
    hipLaunchKernelGGL(media_kernel, dim3(blocks), dim3(threads), 0, stream, gd, mode, pass ? SAT_PREC_F64 : G.prec,
                       G.d_model, G.d_model + n, G.d_model + 2 * n, G.d_model + 3 * n, G.d_fp, G.d_aA, G.d_b,
                       G.d_c, (mode == 0) ? G.d_v3 : nullptr, sg[0], sg[1], sg[2], sg[3], sg[4], sg[5]);
    SAT_CHECK(hipGetLastError());
  }
  return SAT_OK;
}

static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Slot geometry for the largest grid of a precision.
struct SlotGeom {
  size_t seg, plain, slot;
};
static SlotGeom slot_geom(int prec, size_t max_seg_elems, size_t max_plane) {
  SlotGeom s;
  const int nseg = prec == SAT_PREC_F64 ? Prec<SAT_PREC_F64>::kSeg : Prec<SAT_PREC_F32>::kSeg;
  const int npl = prec == SAT_PREC_F64 ? Prec<SAT_PREC_F64>::kPlain : Prec<SAT_PREC_F32>::kPlain;
  s.seg = round_up(max_seg_elems * sizeof(double), 256);  // FP32 path: sigma planes are FP64
  s.plain = round_up(max_plane * sizeof(double), 256);
  s.slot = nseg * s.seg + npl * s.plain;
  return s;
}

}  // namespace sat

// ============================================================================
// C ABI
// ============================================================================
using namespace sat;

namespace {

template <int PREC, int C> struct SoloFwd {
  static void go(dim3 g, hipStream_t s, SoloDev a) { hipLaunchKernelGGL((solo_forward<PREC, C>), g, dim3(kWave), 0, s, a); }
};
template <int PREC, int C> struct SoloBwd {
  static void go(dim3 g, hipStream_t s, SoloDev a) { hipLaunchKernelGGL((solo_backward<PREC, C>), g, dim3(kWave), 0, s, a); }
};
template <int PREC, typename TaT> struct PairKT {
  template <int C> struct K {
    static void go(dim3 g, hipStream_t s, BatchDev a, int first, int count, int group, const int* list,
                   const int* dcount) {
      hipLaunchKernelGGL((pair_kernel<PREC, C, TaT>), g, dim3(kWave), 0, s, a, first, count, group, list, dcount);
    }
  };
};
template <int PREC, int C> struct PairK64 : PairKT<SAT_PREC_F64, double>::K<C> {};
template <int PREC, int C> struct PairK32 : PairKT<SAT_PREC_F32, float>::K<C> {};
template <int PREC, int C> struct PairK64to32 : PairKT<SAT_PREC_F64, float>::K<C> {};

template <template <int, int> class K, int PREC, class... A>
hipError_t dispatch_C1(int C, A... args) {
  switch (C) {
    case 1: K<PREC, 1>::go(args...); break;
    case 2: K<PREC, 2>::go(args...); break;
    case 3: K<PREC, 3>::go(args...); break;
    case 4: K<PREC, 4>::go(args...); break;
    case 5: K<PREC, 5>::go(args...); break;
    case 6: K<PREC, 6>::go(args...); break;
    case 7: K<PREC, 7>::go(args...); break;
    case 8: K<PREC, 8>::go(args...); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <template <int, int> class K, class... A>
hipError_t dispatch(int prec, int C, A... args) {
  return prec == SAT_PREC_F64 ? dispatch_C1<K, SAT_PREC_F64>(C, args...) : dispatch_C1<K, SAT_PREC_F32>(C, args...);
}

int solo_run(bool forward, int precision, int nx, int ny, const double* lon, const double* lat,
             const double* spha, const double* sphb, const double* sphc, const double* fun_or_T, double x0,
             double y0, int nr, const double* rlon, const double* rlat, const double* adj, double* out,
             sat_info* info) {
  if (precision != SAT_PREC_F64 && precision != SAT_PREC_F32) return SAT_EINVAL;
  if (!spha || !sphb || !sphc || !fun_or_T || !out) return SAT_EINVAL;
  if (!forward && nr > 0 && (!rlon || !rlat || !adj)) return SAT_EINVAL;
  if (nr < 0) return SAT_EINVAL;
  int rc = device_check();
  if (rc) return rc;
  GridHost G;
  rc = grid_setup(G, precision, nx, ny, 1, lon, lat);
  if (rc) { G.release(); return rc; }
  const size_t N = (size_t)nx * ny;
  char* d_ws = nullptr;
  double *d_T = nullptr, *d_out = nullptr, *d_r = nullptr, *d_info = nullptr;
  auto cleanup = [&]() {
    void* ps[] = {d_ws, d_T, d_out, d_r, d_info};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    G.release();
  };
#define SOLO_CHECK(x)                                              \
  do {                                                             \
    hipError_t e_ = (x);                                           \
    if (e_ != hipSuccess) {                                        \
      std::fprintf(stderr, "[surfatt_eik] %s: %s\n", #x, hipGetErrorString(e_)); \
      cleanup();                                                   \
      return e_ == hipErrorOutOfMemory ? SAT_ENOMEM : SAT_EDEVICE; \
    }                                                              \
  } while (0)
  hipStream_t stream = nullptr;
  // forward: fun is a media input; backward: the slowness slot is unused (1.0), T separate.
  std::vector<double> ones;
  const double* fin = fun_or_T;
  if (!forward) { ones.assign(N, 1.0); fin = ones.data(); }
  rc = grid_media(G, 1, spha, sphb, sphc, fin, stream);
  if (rc) { cleanup(); return rc; }
  const SlotGeom sg = slot_geom(precision, G.g.seg_plane, N);
  SOLO_CHECK(dalloc(&d_ws, sg.slot));
  SOLO_CHECK(hipMemset(d_ws, 0, sg.slot));
  SOLO_CHECK(dalloc(&d_out, N));
  SOLO_CHECK(dalloc(&d_info, 4));
  SoloDev sd{};
  sd.g = G.g;
  sd.src_lon = x0; sd.src_lat = y0;
  sd.out = d_out; sd.ws = d_ws; sd.ws_seg = sg.seg; sd.ws_plain = sg.plain; sd.info = d_info;
  if (forward) {
    sd.maxiter = kFwdMaxIter;
    SOLO_CHECK((dispatch<SoloFwd>(precision, G.g.C, dim3(1), stream, sd)));
  } else {
    SOLO_CHECK(dalloc(&d_T, N));
    SOLO_CHECK(hipMemcpy(d_T, fun_or_T, sizeof(double) * N, hipMemcpyHostToDevice));
    SOLO_CHECK(dalloc(&d_r, 3 * (size_t)std::max(nr, 1)));
    if (nr > 0) {
      SOLO_CHECK(hipMemcpy(d_r, rlon, sizeof(double) * nr, hipMemcpyHostToDevice));
      SOLO_CHECK(hipMemcpy(d_r + nr, rlat, sizeof(double) * nr, hipMemcpyHostToDevice));
      SOLO_CHECK(hipMemcpy(d_r + 2 * nr, adj, sizeof(double) * nr, hipMemcpyHostToDevice));
    }
    sd.T_in = d_T; sd.nr = nr; sd.rlon = d_r; sd.rlat = d_r + nr; sd.adj = d_r + 2 * nr;
    sd.maxiter = kAdjMaxIter;
    SOLO_CHECK((dispatch<SoloBwd>(precision, G.g.C, dim3(1), stream, sd)));
  }
  SOLO_CHECK(hipDeviceSynchronize());
  SOLO_CHECK(hipMemcpy(out, d_out, sizeof(double) * N, hipMemcpyDeviceToHost));
  double hinfo[4];
  SOLO_CHECK(hipMemcpy(hinfo, d_info, sizeof(hinfo), hipMemcpyDeviceToHost));
  if (info) {
    info->iters = (int)hinfo[0]; info->flags = (int)hinfo[1]; info->l1 = hinfo[2]; info->linf = hinfo[3];
  }
  cleanup();
  return SAT_OK;
#undef SOLO_CHECK
}

}  // namespace

struct sat_batch {
  int precision = SAT_PREC_F64;
  int calc_adj = 1;
  int nevent = 0, nrec = 0;
  std::vector<GridHost> grids;
  std::vector<EventDev> he;        // host copy of the events
  std::vector<EventX> hx;          // host copy of the phase events
  std::vector<int> base_flags;     // per event, set at creation (source on the boundary)
  GridDev* d_grids = nullptr;
  GridDev* d_grids64 = nullptr;    // FP32 batches: grids with FP64 media (re-solves)
  EventDev* d_events = nullptr;
  EventX* d_evx = nullptr;
  int* d_order = nullptr;
  int* d_counter = nullptr;        // re-solve counts | fwd heads | adj heads | re-solve heads (one per group),
                                   // then one head per event for per-solve launches
  double* d_rec = nullptr;         // rec_lon | rec_lat | tt | w | syn | adjsrc
  int* d_rec_ev = nullptr;
  int2* d_rcell = nullptr;
  double* d_rwt = nullptr;
  double* d_chi_ev = nullptr;
  double* d_chi = nullptr;
  int* d_iters = nullptr;          // iters_f | iters_a | flags
  int* d_base_flags = nullptr;
  double* d_conv = nullptr;        // 2 per event
  unsigned char* d_skip = nullptr;
  int* d_fb = nullptr;             // re-solve lists, group g at group_first[g]
  double f32_band = 0.02;
  // per-event planes of the phase pipeline
  size_t seg_total = 0, pl_total = 0, col_total = 0;
  double* d_evs = nullptr;
  double2* d_coltab = nullptr;
  char* d_t0 = nullptr;            // 3 planes (t0, t0x, t0y), sweep precision
  double* d_T0d = nullptr;
  double* d_T = nullptr;
  char* d_co = nullptr;            // 6 (FP64) / 5 (FP32) coefficient planes
  char* d_ta = nullptr;            // masked Ta, sweep precision
  char* d_slots = nullptr;         // phase sweep slots
  size_t slot_plane = 0, slot_bytes = 0;
  int nslots_phase = 0;
  // fused path: per-solve mode and FP64 re-solves
  char* d_ws = nullptr;
  SlotGeom sg{};
  char* d_ws64 = nullptr;
  SlotGeom sg64{};
  int nslots64 = 0;
  double* d_K = nullptr;
  size_t K_total = 0;
  // launch groups: events grouped by (columns per lane C, isotropic grid)
  std::vector<int> group_C, group_iso, group_first, group_count;
  bool groups_dirty = true;
  int max_nx = 0;
  size_t max_seg = 0;
  int nplanes = 0;
  size_t* d_red_sz = nullptr;      // plane_off | node_prefix
  int* d_red_i = nullptr;          // plane_n | ev_off | ev_ids
  const double** d_red_v3 = nullptr;
  ReduceDev rd{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, evA = nullptr;

  void release() {
    for (auto& G : grids) G.release();
    void* ps[] = {d_grids, d_grids64, d_events, d_evx, d_order, d_counter, d_rec, d_rec_ev, d_rcell, d_rwt, d_chi_ev,
                  d_chi, d_iters, d_base_flags, d_conv, d_skip, d_fb, d_evs, d_coltab, d_t0, d_T0d, d_T, d_co, d_ta,
                  d_slots, d_ws, d_ws64, d_K, d_red_sz, d_red_i, (void*)d_red_v3};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    hipEvent_t es[] = {ev0, ev1, ev2, evA};
    for (hipEvent_t e : es)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Sort events into launch groups (same C and isotropy), largest grids first.
int build_groups(sat_batch* b) {
  const int ne = b->nevent;
  std::vector<int> order(ne);
  std::iota(order.begin(), order.end(), 0);
  auto key = [&](int e) {
    const GridDev& g = b->grids[b->he[e].grid].g;
    return std::make_pair(g.C, g.iso);
  };
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    if (key(a) != key(c)) return key(a) < key(c);
    return b->grids[b->he[a].grid].g.plane > b->grids[b->he[c].grid].g.plane;
  });
  b->group_C.clear();
  b->group_iso.clear();
  b->group_first.clear();
  b->group_count.clear();
  for (int i = 0; i < ne;) {
    const auto k = key(order[i]);
    int j = i;
    while (j < ne && key(order[j]) == k) j++;
    b->group_C.push_back(k.first);
    b->group_iso.push_back(k.second);
    b->group_first.push_back(i);
    b->group_count.push_back(j - i);
    i = j;
  }
  if (ne) SAT_CHECK(hipMemcpy(b->d_order, order.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
  std::vector<GridDev> hg(b->grids.size()), hg64(b->grids.size());
  for (size_t gi = 0; gi < b->grids.size(); gi++) {
    hg[gi] = b->grids[gi].g;
    hg64[gi] = b->grids[gi].g64;
  }
  SAT_CHECK(hipMemcpy(b->d_grids, hg.data(), sizeof(GridDev) * hg.size(), hipMemcpyHostToDevice));
  if (b->d_grids64) SAT_CHECK(hipMemcpy(b->d_grids64, hg64.data(), sizeof(GridDev) * hg64.size(), hipMemcpyHostToDevice));
  b->groups_dirty = false;
  return SAT_OK;
}

template <int PREC, int C, bool ISO> struct FwdSweepK {
  static void go(dim3 g, hipStream_t s, PhaseDev a, int first, int count, int group) {
    hipLaunchKernelGGL((fwd_sweep_kernel<PREC, C, ISO>), g, dim3(kWave), 0, s, a, first, count, group);
  }
};
template <int PREC, int C> struct FwdSweepIsoK : FwdSweepK<PREC, C, true> {};
template <int PREC, int C> struct FwdSweepGenK : FwdSweepK<PREC, C, false> {};
template <int PREC, int C> struct AdjSweepK {
  static void go(dim3 g, hipStream_t s, PhaseDev a, int first, int count) {
    hipLaunchKernelGGL((adj_sweep_kernel<PREC, C>), g, dim3(kWave), 0, s, a, first, count);
  }
};

// Host copy of the index boxes, computed with the same IEEE double operations
// as the device (src/eikonal_solver.cpp:119-129 and :528-531).
void event_boxes(const GridHost& G, const std::vector<double>& xx, const std::vector<double>& yy,
                 const double* lon, const double* lat, double elon, double elat, EventX& x, int& flags) {
  const int nx = G.nx, ny = G.ny;
  const double x0 = elon * kDEG2RAD, y0 = elat * kDEG2RAD;
  const double dx = xx[1] - xx[0], dy = yy[1] - yy[0];
  int bx0 = 1 << 30, bx1 = -1, by0 = 1 << 30, by1 = -1;
  for (int ix = 0; ix < nx; ix++)
    if (std::fabs((xx[ix] - x0) / dx) <= 1.0) { bx0 = std::min(bx0, ix); bx1 = std::max(bx1, ix); }
  for (int iy = 0; iy < ny; iy++)
    if (std::fabs((yy[iy] - y0) / dy) <= 1.0) { by0 = std::min(by0, iy); by1 = std::max(by1, iy); }
  x.fb[0] = bx0; x.fb[1] = bx1; x.fb[2] = by0; x.fb[3] = by1;
  flags = 0;
  if (bx0 <= bx1 && by0 <= by1 && (bx0 == 0 || bx1 == nx - 1 || by0 == 0 || by1 == ny - 1))
    flags |= FLAG_SRC_ON_BOUNDARY;
  const double dxd = lon[1] - lon[0], dyd = lat[1] - lat[0];
  x.mb[0] = imax(0, (int)std::floor((elon - 1.5 * dxd - lon[0]) / dxd) + 1);
  x.mb[1] = imin(nx - 1, (int)std::ceil((elon + 1.5 * dxd - lon[0]) / dxd) - 1);
  x.mb[2] = imax(0, (int)std::floor((elat - 1.5 * dyd - lat[0]) / dyd) + 1);
  x.mb[3] = imin(ny - 1, (int)std::ceil((elat + 1.5 * dyd - lat[0]) / dyd) - 1);
}

}  // namespace

extern "C" {

int sat_eik_forward(int precision, int nx, int ny, const double* lon_deg, const double* lat_deg,
                    const double* spha, const double* sphb, const double* sphc, const double* fun,
                    double src_lon_deg, double src_lat_deg, double* T_out, sat_info* info) {
  return solo_run(true, precision, nx, ny, lon_deg, lat_deg, spha, sphb, sphc, fun, src_lon_deg, src_lat_deg, 0,
                  nullptr, nullptr, nullptr, T_out, info);
}

int sat_eik_backward(int precision, int nx, int ny, const double* lon_deg, const double* lat_deg,
                     const double* spha, const double* sphb, const double* sphc, const double* T, int nr,
                     const double* rec_lon_deg, const double* rec_lat_deg, const double* adj, double* Ta_out,
                     sat_info* info) {
  return solo_run(false, precision, nx, ny, lon_deg, lat_deg, spha, sphb, sphc, T, 0.0, 0.0, nr, rec_lon_deg,
                  rec_lat_deg, adj, Ta_out, info);
}

int sat_eik_mask(int nx, int ny, const double* xx, const double* yy, double* Ta, double x0, double y0) {
  // Index box of src/eikonal_solver.cpp:514-536 on a host field: O(box) work, no device round trip.
  if (!xx || !yy || !Ta) return SAT_EINVAL;
  if (nx < 2 || ny < 2) return SAT_OK;
  const double dx = xx[1] - xx[0];
  const double dy = yy[1] - yy[0];
  const int ix_lo = imax(0, (int)std::floor((x0 - 1.5 * dx - xx[0]) / dx) + 1);
  const int ix_hi = imin(nx - 1, (int)std::ceil((x0 + 1.5 * dx - xx[0]) / dx) - 1);
  const int iy_lo = imax(0, (int)std::floor((y0 - 1.5 * dy - yy[0]) / dy) + 1);
  const int iy_hi = imin(ny - 1, (int)std::ceil((y0 + 1.5 * dy - yy[0]) / dy) - 1);
  if (ix_lo <= ix_hi && iy_lo <= iy_hi)
    for (int iy = iy_lo; iy <= iy_hi; iy++)
      for (int ix = ix_lo; ix <= ix_hi; ix++) Ta[ix + (size_t)iy * nx] = 0.0;
  return SAT_OK;
}

int sat_batch_create(const sat_batch_desc* d, sat_batch** out) {
  if (!d || !out) return SAT_EINVAL;
  *out = nullptr;
  if (d->ngrid < 1 || !d->grids || d->nevent < 0 || (d->nevent > 0 && !d->events)) return SAT_EINVAL;
  if (d->precision != SAT_PREC_F64 && d->precision != SAT_PREC_F32) return SAT_EINVAL;
  if (d->nrec_total < 0 || (d->nrec_total > 0 && (!d->rec_lon_deg || !d->rec_lat_deg || !d->tt_obs || !d->weight)))
    return SAT_EINVAL;
  std::vector<int> rec_ev((size_t)std::max(d->nrec_total, 1), -1);
  for (int e = 0; e < d->nevent; e++) {
    const sat_event_desc& ev = d->events[e];
    if (ev.grid < 0 || ev.grid >= d->ngrid) return SAT_EINVAL;
    if (ev.iper < 0 || ev.iper >= d->grids[ev.grid].nperiod) return SAT_EINVAL;
    if (ev.nrec < 0 || ev.rec_off < 0 || (long)ev.rec_off + ev.nrec > d->nrec_total) return SAT_EINVAL;
    if (!std::isfinite(ev.src_lon_deg) || !std::isfinite(ev.src_lat_deg)) return SAT_EINVAL;
    for (int r = 0; r < ev.nrec; r++) {
      if (rec_ev[ev.rec_off + r] != -1) return SAT_EINVAL;  // receivers belong to one event
      rec_ev[ev.rec_off + r] = e;
    }
  }
  int rc = device_check();
  if (rc) return rc;
  sat_batch* b = new (std::nothrow) sat_batch();
  if (!b) return SAT_ENOMEM;
  auto fail = [&](int code) { b->release(); delete b; return code; };
#define B_CHECK(x)                                                            \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      std::fprintf(stderr, "[surfatt_eik] %s: %s\n", #x, hipGetErrorString(e_)); \
      return fail(e_ == hipErrorOutOfMemory ? SAT_ENOMEM : SAT_EDEVICE);      \
    }                                                                         \
  } while (0)
  b->precision = d->precision;
  const bool f32 = b->precision == SAT_PREC_F32;
  b->calc_adj = d->calc_adj ? 1 : 0;
  b->nevent = d->nevent;
  b->nrec = d->nrec_total;
  B_CHECK(hipStreamCreate(&b->stream));
  B_CHECK(hipEventCreate(&b->ev0));
  B_CHECK(hipEventCreate(&b->ev1));
  B_CHECK(hipEventCreate(&b->ev2));
  B_CHECK(hipEventCreate(&b->evA));
  b->grids.resize(d->ngrid);
  size_t max_seg = 0, max_plane = 0;
  int max_nx = 0;
  for (int gi = 0; gi < d->ngrid; gi++) {
    const sat_grid_desc& gd = d->grids[gi];
    rc = grid_setup(b->grids[gi], b->precision, gd.nx, gd.ny, gd.nperiod, gd.lon_deg, gd.lat_deg);
    if (rc) return fail(rc);
    rc = grid_media(b->grids[gi], 0, gd.svel, gd.m11, gd.m12, gd.m22, b->stream);
    if (rc) return fail(rc);
    max_seg = std::max(max_seg, b->grids[gi].g.seg_plane);
    max_plane = std::max(max_plane, b->grids[gi].g.plane);
    max_nx = std::max(max_nx, gd.nx);
  }
  b->max_seg = max_seg;
  b->max_nx = max_nx;
  B_CHECK(dalloc(&b->d_grids, d->ngrid));
  if (f32) B_CHECK(dalloc(&b->d_grids64, d->ngrid));

  // events: per-event plane offsets, index boxes, base flags
  const int ne = d->nevent;
  b->he.resize(ne);
  b->hx.resize(ne);
  b->base_flags.assign(std::max(ne, 1), 0);
  size_t seg_total = 0, pl_total = 0, col_total = 0;
  for (int e = 0; e < ne; e++) {
    const sat_event_desc& ev = d->events[e];
    const sat_grid_desc& gd = d->grids[ev.grid];
    const GridDev& g = b->grids[ev.grid].g;
    b->he[e] = EventDev{ev.grid, ev.iper, ev.src_lon_deg, ev.src_lat_deg, ev.rec_off, ev.nrec, pl_total};
    EventX& x = b->hx[e];
    x.grid = ev.grid; x.iper = ev.iper; x.rec_off = ev.rec_off; x.nrec = ev.nrec;
    x.lon = ev.src_lon_deg; x.lat = ev.src_lat_deg;
    x.seg_off = seg_total; x.pl_off = pl_total; x.col_off = col_total;
    std::vector<double> xx(gd.nx), yy(gd.ny);
    for (int i = 0; i < gd.nx; i++) xx[i] = gd.lon_deg[i] * kDEG2RAD;
    for (int j = 0; j < gd.ny; j++) yy[j] = gd.lat_deg[j] * kDEG2RAD;
    event_boxes(b->grids[ev.grid], xx, yy, gd.lon_deg, gd.lat_deg, ev.src_lon_deg, ev.src_lat_deg, x,
                b->base_flags[e]);
    seg_total += round_up(g.seg_plane, 64);
    pl_total += round_up(g.plane, 64);
    col_total += round_up(g.nx, 64);
  }
  b->seg_total = seg_total;
  b->pl_total = pl_total;
  b->col_total = col_total;
  B_CHECK(dalloc(&b->d_events, ne));
  B_CHECK(dalloc(&b->d_evx, ne));
  if (ne) {
    B_CHECK(hipMemcpy(b->d_events, b->he.data(), sizeof(EventDev) * ne, hipMemcpyHostToDevice));
    B_CHECK(hipMemcpy(b->d_evx, b->hx.data(), sizeof(EventX) * ne, hipMemcpyHostToDevice));
  }
  B_CHECK(dalloc(&b->d_base_flags, std::max(ne, 1)));
  B_CHECK(hipMemcpy(b->d_base_flags, b->base_flags.data(), sizeof(int) * std::max(ne, 1), hipMemcpyHostToDevice));
  B_CHECK(dalloc(&b->d_order, ne));

  // receivers
  const size_t nr = (size_t)d->nrec_total;
  B_CHECK(dalloc(&b->d_rec, 6 * std::max<size_t>(nr, 1)));
  B_CHECK(hipMemset(b->d_rec, 0, sizeof(double) * 6 * std::max<size_t>(nr, 1)));  // receivers of no event: syn 0
  if (nr) {
    B_CHECK(hipMemcpy(b->d_rec, d->rec_lon_deg, sizeof(double) * nr, hipMemcpyHostToDevice));
    B_CHECK(hipMemcpy(b->d_rec + nr, d->rec_lat_deg, sizeof(double) * nr, hipMemcpyHostToDevice));
    B_CHECK(hipMemcpy(b->d_rec + 2 * nr, d->tt_obs, sizeof(double) * nr, hipMemcpyHostToDevice));
    B_CHECK(hipMemcpy(b->d_rec + 3 * nr, d->weight, sizeof(double) * nr, hipMemcpyHostToDevice));
  }
  B_CHECK(dalloc(&b->d_rec_ev, std::max<size_t>(nr, 1)));
  B_CHECK(hipMemcpy(b->d_rec_ev, rec_ev.data(), sizeof(int) * rec_ev.size(), hipMemcpyHostToDevice));
  B_CHECK(dalloc(&b->d_rcell, std::max<size_t>(nr, 1)));
  B_CHECK(dalloc(&b->d_rwt, 3 * std::max<size_t>(nr, 1)));
  B_CHECK(dalloc(&b->d_chi_ev, std::max(ne, 1)));
  B_CHECK(hipMemset(b->d_chi_ev, 0, sizeof(double) * std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_chi, 1));
  B_CHECK(dalloc(&b->d_iters, 3 * (size_t)std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_conv, 2 * (size_t)std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_skip, std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_fb, (size_t)std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_counter, 5 * (size_t)std::max(ne, 1)));  // groups <= events

  // per-event planes of the phase pipeline
  const size_t es = elem_size(b->precision);
  B_CHECK(dalloc(&b->d_evs, 3 * (size_t)std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_coltab, std::max<size_t>(col_total, 1)));
  B_CHECK(dalloc(&b->d_t0, 3 * std::max<size_t>(seg_total, 1) * es));
  if (f32) B_CHECK(dalloc(&b->d_T0d, std::max<size_t>(pl_total, 1)));
  B_CHECK(dalloc(&b->d_T, std::max<size_t>(pl_total, 1)));
  if (b->calc_adj) {
    B_CHECK(dalloc(&b->d_co, (f32 ? 5 : 6) * std::max<size_t>(seg_total, 1) * es));
    B_CHECK(dalloc(&b->d_ta, std::max<size_t>(pl_total, 1) * es));
  }

  // resident waves of the sweep kernels: 3 swept planes each (FP64 sigma/tau sized)
  hipDeviceProp_t prop;
  B_CHECK(hipGetDeviceProperties(&prop, 0));
  const int waves_per_cu = 8;
  b->nslots_phase = std::max(1, std::min(ne, prop.multiProcessorCount * waves_per_cu));
  b->slot_plane = round_up(max_seg * sizeof(double), 256);
  b->slot_bytes = 3 * b->slot_plane;
  B_CHECK(dalloc(&b->d_slots, b->slot_bytes * (size_t)b->nslots_phase));
  B_CHECK(hipMemset(b->d_slots, 0, b->slot_bytes * (size_t)b->nslots_phase));

  // fused path: one slot for per-solve runs; FP32 batches also FP64 slots for re-solves
  b->sg = slot_geom(b->precision, max_seg, max_plane);
  B_CHECK(dalloc(&b->d_ws, b->sg.slot));
  B_CHECK(hipMemset(b->d_ws, 0, b->sg.slot));  // also the all-zero FP32 delta plane
  if (f32) {
    b->nslots64 = std::max(1, std::min(ne, prop.multiProcessorCount));
    b->sg64 = slot_geom(SAT_PREC_F64, max_seg, max_plane);
    B_CHECK(dalloc(&b->d_ws64, b->sg64.slot * (size_t)b->nslots64));
    B_CHECK(hipMemset(b->d_ws64, 0, b->sg64.slot * (size_t)b->nslots64));
  }

  // kernel planes + reduction CSR
  std::vector<size_t> plane_off, node_prefix{0};
  std::vector<const double*> vinv_ptr;
  std::vector<int> plane_n, ev_off{0}, ev_ids;
  size_t K_total = 0;
  for (int gi = 0; gi < d->ngrid; gi++) {
    const GridDev& g = b->grids[gi].g;
    for (int ip = 0; ip < g.nperiod; ip++) {
      plane_off.push_back(K_total);
      vinv_ptr.push_back(b->grids[gi].d_v3 + (size_t)ip * g.plane);
      plane_n.push_back((int)g.plane);
      for (int e = 0; e < ne; e++)
        if (b->he[e].grid == gi && b->he[e].iper == ip) ev_ids.push_back(e);
      ev_off.push_back((int)ev_ids.size());
      K_total += g.plane;
      node_prefix.push_back(K_total);
    }
  }
  b->K_total = K_total;
  b->nplanes = (int)plane_off.size();
  B_CHECK(dalloc(&b->d_K, K_total));
  {
    std::vector<size_t> hs(plane_off);
    hs.insert(hs.end(), node_prefix.begin(), node_prefix.end());
    B_CHECK(dalloc(&b->d_red_sz, hs.size()));
    B_CHECK(hipMemcpy(b->d_red_sz, hs.data(), sizeof(size_t) * hs.size(), hipMemcpyHostToDevice));
    std::vector<int> hi(plane_n);
    hi.insert(hi.end(), ev_off.begin(), ev_off.end());
    hi.insert(hi.end(), ev_ids.begin(), ev_ids.end());
    B_CHECK(dalloc(&b->d_red_i, hi.size()));
    B_CHECK(hipMemcpy(b->d_red_i, hi.data(), sizeof(int) * hi.size(), hipMemcpyHostToDevice));
    B_CHECK(dalloc(&b->d_red_v3, vinv_ptr.size()));
    B_CHECK(hipMemcpy(b->d_red_v3, vinv_ptr.data(), sizeof(const double*) * vinv_ptr.size(), hipMemcpyHostToDevice));
  }
  ReduceDev& rd = b->rd;
  rd.nplanes = b->nplanes;
  rd.plane_off = b->d_red_sz;
  rd.node_prefix = b->d_red_sz + plane_off.size();
  rd.plane_n = b->d_red_i;
  rd.ev_off = b->d_red_i + plane_n.size();
  rd.ev_ids = b->d_red_i + plane_n.size() + ev_off.size();
  rd.vinv3 = b->d_red_v3;
  rd.events = b->d_events;
  rd.ta_out = b->d_ta;
  rd.K = b->d_K;
  rd.total_nodes = K_total;
  B_CHECK(hipStreamSynchronize(b->stream));
  rc = build_groups(b);
  if (rc) return fail(rc);
  *out = b;
  return SAT_OK;
#undef B_CHECK
}

int sat_batch_update_media(sat_batch* b, int grid, const double* svel, const double* m11, const double* m12,
                           const double* m22) {
  if (!b || grid < 0 || grid >= (int)b->grids.size()) return SAT_EINVAL;
  const int iso0 = b->grids[grid].g.iso;
  const int rc = grid_media(b->grids[grid], 0, svel, m11, m12, m22, b->stream);
  if (rc) return rc;
  SAT_CHECK(hipStreamSynchronize(b->stream));  // host planes may be released on return
  if (b->grids[grid].g.iso != iso0) b->groups_dirty = true;
  return SAT_OK;
}

int sat_batch_update_obs(sat_batch* b, const double* tt_obs, const double* weight) {
  if (!b || (b->nrec > 0 && (!tt_obs || !weight))) return SAT_EINVAL;
  const size_t nr = (size_t)b->nrec;
  if (nr) {
    SAT_CHECK(hipMemcpyAsync(b->d_rec + 2 * nr, tt_obs, sizeof(double) * nr, hipMemcpyHostToDevice, b->stream));
    SAT_CHECK(hipMemcpyAsync(b->d_rec + 3 * nr, weight, sizeof(double) * nr, hipMemcpyHostToDevice, b->stream));
    SAT_CHECK(hipStreamSynchronize(b->stream));
  }
  return SAT_OK;
}

int sat_batch_run(sat_batch* b, int mode) {
  if (!b || (mode != SAT_RUN_BATCHED && mode != SAT_RUN_PER_SOLVE)) return SAT_EINVAL;
  if (b->groups_dirty) {
    const int rc = build_groups(b);
    if (rc) return rc;
  }
  const bool f32 = b->precision == SAT_PREC_F32;
  const int ne = b->nevent;
  const size_t nr = (size_t)std::max(b->nrec, 1);
  const size_t ne1 = (size_t)std::max(ne, 1);
  const size_t ng = b->group_C.size();
  int* fb_count = b->d_counter;
  int* fwd_heads = b->d_counter + ng;
  int* adj_heads = b->d_counter + 2 * ng;
  int* res_heads = b->d_counter + 3 * ng;
  int* solo_heads = b->d_counter + 4 * ng;

  BatchDev bd{};
  bd.grids = b->d_grids;
  bd.events = b->d_events;
  bd.order = b->d_order;
  bd.rec_lon = b->d_rec;
  bd.rec_lat = b->d_rec + nr;
  bd.tt_obs = b->d_rec + 2 * nr;
  bd.weight = b->d_rec + 3 * nr;
  bd.syn = b->d_rec + 4 * nr;
  bd.adjsrc = b->d_rec + 5 * nr;
  bd.chi_ev = b->d_chi_ev;
  bd.iters_f = b->d_iters;
  bd.iters_a = b->d_iters + ne1;
  bd.flags = b->d_iters + 2 * ne1;
  bd.ta_out = b->d_ta;
  bd.calc_adj = b->calc_adj;
  bd.fwd_maxiter = kFwdMaxIter;
  bd.adj_maxiter = kAdjMaxIter;
  bd.conv = b->d_conv;
  bd.fb_list = b->d_fb;
  bd.fb_count = fb_count;
  bd.f32_band = b->f32_band;

  SAT_CHECK(hipEventRecord(b->ev0, b->stream));
  SAT_CHECK(hipMemsetAsync(b->d_counter, 0, sizeof(int) * 5 * ne1, b->stream));
  SAT_CHECK(hipMemsetAsync(b->d_skip, 0, ne1, b->stream));
  SAT_CHECK(hipMemsetAsync(b->d_iters, 0, sizeof(int) * 2 * ne1, b->stream));
  SAT_CHECK(hipMemcpyAsync(bd.flags, b->d_base_flags, sizeof(int) * ne1, hipMemcpyDeviceToDevice, b->stream));

  if (mode == SAT_RUN_PER_SOLVE) {
    // one launch per pair of the fused kernel (the plain per-solve path)
    bd.ws = b->d_ws;
    bd.ws_slot = b->sg.slot;
    bd.ws_seg = b->sg.seg;
    bd.ws_plain = b->sg.plain;
    for (size_t gi = 0; gi < ng; gi++)
      for (int i = 0; i < b->group_count[gi]; i++) {
        const int first = b->group_first[gi] + i;
        BatchDev a = bd;
        a.counter = solo_heads + first;
        a.fb_list = bd.fb_list + b->group_first[gi];
        if (!f32)
          SAT_CHECK((dispatch_C1<PairK64, 0>(b->group_C[gi], dim3(1), b->stream, a, first, 1, (int)gi,
                                              (const int*)nullptr, (const int*)nullptr)));
        else
          SAT_CHECK((dispatch_C1<PairK32, 0>(b->group_C[gi], dim3(1), b->stream, a, first, 1, (int)gi,
                                              (const int*)nullptr, (const int*)nullptr)));
      }
  } else {
    PhaseDev pd{};
    pd.grids = b->d_grids;
    pd.ev = b->d_evx;
    pd.nevent = ne;
    pd.evs = b->d_evs;
    pd.coltab = b->d_coltab;
    const size_t es = elem_size(b->precision);
    pd.t0 = b->d_t0;
    pd.t0x = b->d_t0 + b->seg_total * es;
    pd.t0y = b->d_t0 + 2 * b->seg_total * es;
    pd.T0d = b->d_T0d;
    pd.T = b->d_T;
    for (int i = 0; i < 6; i++) pd.co[i] = (b->d_co && i < (f32 ? 5 : 6)) ? b->d_co + i * b->seg_total * es : nullptr;
    pd.ta_out = b->d_ta;
    pd.rec_lon = bd.rec_lon;
    pd.rec_lat = bd.rec_lat;
    pd.tt_obs = bd.tt_obs;
    pd.weight = bd.weight;
    pd.rec_ev = b->d_rec_ev;
    pd.syn = bd.syn;
    pd.adjsrc = bd.adjsrc;
    pd.rcell = b->d_rcell;
    pd.rwt = b->d_rwt;
    pd.chi_ev = b->d_chi_ev;
    pd.iters_f = bd.iters_f;
    pd.iters_a = bd.iters_a;
    pd.flags = bd.flags;
    pd.conv = b->d_conv;
    pd.skip = b->d_skip;
    pd.fb_list = b->d_fb;
    pd.fb_count = fb_count;
    pd.f32_band = b->f32_band;
    pd.slots = b->d_slots;
    pd.slot_bytes = b->slot_bytes;
    pd.slot_plane = b->slot_plane;
    pd.order = b->d_order;
    pd.fwd_maxiter = kFwdMaxIter;
    pd.adj_maxiter = kAdjMaxIter;
    const int max_nodes = (int)b->max_seg;
    if (ne) {
      if (!f32) hipLaunchKernelGGL(trig_kernel<SAT_PREC_F64>, dim3((b->max_nx + 63) / 64, ne), dim3(64), 0, b->stream, pd);
      else hipLaunchKernelGGL(trig_kernel<SAT_PREC_F32>, dim3((b->max_nx + 63) / 64, ne), dim3(64), 0, b->stream, pd);
      SAT_CHECK(hipGetLastError());
      if (!f32) hipLaunchKernelGGL(init_kernel<SAT_PREC_F64>, dim3((max_nodes + 255) / 256, ne), dim3(256), 0, b->stream, pd);
      else hipLaunchKernelGGL(init_kernel<SAT_PREC_F32>, dim3((max_nodes + 255) / 256, ne), dim3(256), 0, b->stream, pd);
      SAT_CHECK(hipGetLastError());
    }
    for (size_t gi = 0; gi < ng; gi++) {
      PhaseDev a = pd;
      a.counter = fwd_heads + gi;
      a.fb_list = pd.fb_list + b->group_first[gi];
      const int nblk = std::min(b->group_count[gi], b->nslots_phase);
      const int C = b->group_C[gi];
      hipError_t err;
      if (!f32) err = dispatch_C1<FwdSweepGenK, SAT_PREC_F64>(C, dim3(nblk), b->stream, a, b->group_first[gi], b->group_count[gi], (int)gi);
      else if (b->group_iso[gi]) err = dispatch_C1<FwdSweepIsoK, SAT_PREC_F32>(C, dim3(nblk), b->stream, a, b->group_first[gi], b->group_count[gi], (int)gi);
      else err = dispatch_C1<FwdSweepGenK, SAT_PREC_F32>(C, dim3(nblk), b->stream, a, b->group_first[gi], b->group_count[gi], (int)gi);
      SAT_CHECK(err);
    }
    if (b->nrec) {
      const int blocks = (int)std::min<size_t>((b->nrec + 255) / 256, 4096);
      hipLaunchKernelGGL(recv_kernel, dim3(blocks), dim3(256), 0, b->stream, pd, b->nrec);
      SAT_CHECK(hipGetLastError());
    }
    if (b->calc_adj && ne) {
      if (!f32) hipLaunchKernelGGL(adj_setup_kernel<SAT_PREC_F64>, dim3((max_nodes + 255) / 256, ne), dim3(256), 0, b->stream, pd);
      else hipLaunchKernelGGL(adj_setup_kernel<SAT_PREC_F32>, dim3((max_nodes + 255) / 256, ne), dim3(256), 0, b->stream, pd);
      SAT_CHECK(hipGetLastError());
      if (!f32) hipLaunchKernelGGL(delta_kernel<SAT_PREC_F64>, dim3((ne + 63) / 64), dim3(64), 0, b->stream, pd);
      else hipLaunchKernelGGL(delta_kernel<SAT_PREC_F32>, dim3((ne + 63) / 64), dim3(64), 0, b->stream, pd);
      SAT_CHECK(hipGetLastError());
      for (size_t gi = 0; gi < ng; gi++) {
        PhaseDev a = pd;
        a.counter = adj_heads + gi;
        const int nblk = std::min(b->group_count[gi], b->nslots_phase);
        hipError_t err;
        if (!f32) err = dispatch_C1<AdjSweepK, SAT_PREC_F64>(b->group_C[gi], dim3(nblk), b->stream, a, b->group_first[gi], b->group_count[gi]);
        else err = dispatch_C1<AdjSweepK, SAT_PREC_F32>(b->group_C[gi], dim3(nblk), b->stream, a, b->group_first[gi], b->group_count[gi]);
        SAT_CHECK(err);
      }
    }
  }
  SAT_CHECK(hipEventRecord(b->evA, b->stream));
  if (f32) {  // FP64 re-solves of the pairs whose FP32 stop decision was within the band
    for (size_t gi = 0; gi < ng; gi++) {
      BatchDev a = bd;
      a.grids = b->d_grids64;
      a.ws = b->d_ws64;
      a.ws_slot = b->sg64.slot;
      a.ws_seg = b->sg64.seg;
      a.ws_plain = b->sg64.plain;
      a.counter = res_heads + gi;
      a.f32_band = 0.0;
      const int* list = bd.fb_list + b->group_first[gi];
      const int nblk = std::min(b->group_count[gi], b->nslots64);
      SAT_CHECK((dispatch_C1<PairK64to32, 0>(b->group_C[gi], dim3(nblk), b->stream, a, 0, 0, (int)gi, list,
                                             (const int*)(fb_count + gi))));
    }
  }
  SAT_CHECK(hipEventRecord(b->ev1, b->stream));
  if (b->calc_adj && b->K_total) {
    const int threads = 256;
    const int blocks = (int)std::min<size_t>((b->K_total + threads - 1) / threads, 65535);
    if (!f32) hipLaunchKernelGGL(reduce_kernel<double>, dim3(blocks), dim3(threads), 0, b->stream, b->rd);
    else hipLaunchKernelGGL(reduce_kernel<float>, dim3(blocks), dim3(threads), 0, b->stream, b->rd);
    SAT_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(chi_kernel, dim3(1), dim3(1), 0, b->stream, b->d_chi_ev, b->calc_adj ? ne : 0, b->d_chi);
  SAT_CHECK(hipGetLastError());
  SAT_CHECK(hipEventRecord(b->ev2, b->stream));
  return SAT_OK;
}

int sat_batch_sync(sat_batch* b) {
  if (!b) return SAT_EINVAL;
  SAT_CHECK(hipStreamSynchronize(b->stream));
  return SAT_OK;
}

int sat_batch_last_ms(sat_batch* b, double* total_ms, double* solve_ms) {
  if (!b) return SAT_EINVAL;
  float t = 0.f, s = 0.f;
  SAT_CHECK(hipEventElapsedTime(&t, b->ev0, b->ev2));
  SAT_CHECK(hipEventElapsedTime(&s, b->ev0, b->evA));
  if (total_ms) *total_ms = t;
  if (solve_ms) *solve_ms = s;
  return SAT_OK;
}

int sat_batch_results(sat_batch* b, double* syn, double* adj_s, double* chi, int* iters_fwd, int* iters_adj,
                      int* flags) {
  if (!b) return SAT_EINVAL;
  SAT_CHECK(hipStreamSynchronize(b->stream));
  const size_t nr = (size_t)std::max(b->nrec, 1);
  const size_t ne = (size_t)std::max(b->nevent, 1);
  if (syn && b->nrec) SAT_CHECK(hipMemcpy(syn, b->d_rec + 4 * nr, sizeof(double) * b->nrec, hipMemcpyDeviceToHost));
  if (adj_s && b->K_total) {
    if (b->calc_adj) SAT_CHECK(hipMemcpy(adj_s, b->d_K, sizeof(double) * b->K_total, hipMemcpyDeviceToHost));
    else std::memset(adj_s, 0, sizeof(double) * b->K_total);
  }
  if (chi) SAT_CHECK(hipMemcpy(chi, b->d_chi, sizeof(double), hipMemcpyDeviceToHost));
  if (iters_fwd && b->nevent) SAT_CHECK(hipMemcpy(iters_fwd, b->d_iters, sizeof(int) * b->nevent, hipMemcpyDeviceToHost));
  if (iters_adj && b->nevent) SAT_CHECK(hipMemcpy(iters_adj, b->d_iters + ne, sizeof(int) * b->nevent, hipMemcpyDeviceToHost));
  if (flags && b->nevent) SAT_CHECK(hipMemcpy(flags, b->d_iters + 2 * ne, sizeof(int) * b->nevent, hipMemcpyDeviceToHost));
  return SAT_OK;
}

int sat_batch_set_f32_band(sat_batch* b, double band) {
  if (!b || !(band >= 0.0 && band < 1.0)) return SAT_EINVAL;
  b->f32_band = band;
  return SAT_OK;
}

int sat_batch_conv(sat_batch* b, double* linf_last, double* linf_prev, int* n_resolved) {
  if (!b) return SAT_EINVAL;
  SAT_CHECK(hipStreamSynchronize(b->stream));
  const int ne = b->nevent;
  if (ne && (linf_last || linf_prev)) {
    std::vector<double> c(2 * (size_t)ne);
    SAT_CHECK(hipMemcpy(c.data(), b->d_conv, sizeof(double) * c.size(), hipMemcpyDeviceToHost));
    for (int e = 0; e < ne; e++) {
      if (linf_last) linf_last[e] = c[2 * (size_t)e];
      if (linf_prev) linf_prev[e] = c[2 * (size_t)e + 1];
    }
  }
  if (n_resolved) {
    *n_resolved = 0;
    if (b->precision == SAT_PREC_F32 && !b->group_C.empty()) {
      std::vector<int> cnt(b->group_C.size());
      SAT_CHECK(hipMemcpy(cnt.data(), b->d_counter, sizeof(int) * cnt.size(), hipMemcpyDeviceToHost));
      for (int c : cnt) *n_resolved += c;
    }
  }
  return SAT_OK;
}

int sat_batch_destroy(sat_batch* b) {
  if (!b) return SAT_OK;
  (void)hipStreamSynchronize(b->stream);
  b->release();
  delete b;
  return SAT_OK;
}

int sat_device_info(char* buf, int buflen) {
  int rc = device_check();
  if (rc) {
    if (buf && buflen > 0) std::snprintf(buf, buflen, "no HIP device");
    return rc;
  }
  hipDeviceProp_t prop;
  SAT_CHECK(hipGetDeviceProperties(&prop, 0));
  if (buf && buflen > 0)
    std::snprintf(buf, buflen, "%s | %s | CUs=%d | build=gfx950", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount);
  return SAT_OK;
}

}  // extern "C"
