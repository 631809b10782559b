// sat_fused.hpp -- one wave solves one (source, period) pair end to end.
//
//   pair_kernel<C>    persistent: one 64-lane wave drains a queue of pairs; per pair
//                     forward solve -> receiver interp2d -> residual -> adjoint
//                     solve -> mask -> masked Ta out (src/preproc.cpp:79-143).
//                     The plain per-solve mode (one launch per pair) and the FP64
//                     re-solves of an FP32 batch.
//   solo_*<C>         the per-solve entry points (FSM_UW_PS_lonlat_2d,
//                     FSM_O1_JSE_lonlat_2d)
// The per-pair workspace is a slot of segmented planes (SegLayout).
#pragma once
#include "../../include/surfatt_eik.h"
#include "sat_crmath.hpp"
#include "sat_device.hpp"
#include "sat_engine.hpp"
#include "sat_policies.hpp"

namespace sat {

template <int PREC> struct Prec;
template <> struct Prec<SAT_PREC_F64> {
  using Fwd = FwdF64;
  using Adj = AdjF64;
  using AdjLds = AdjF64Lds;
  using V = double;
  static constexpr int kSeg = 9;    // T0 T0x T0y u0 u1 u2 | a1p a2m b1p b2m dl d ta0 ta1 ta2
  static constexpr int kPlain = 1;  // T
};
template <> struct Prec<SAT_PREC_F32> {
  using Fwd = FwdF32;
  using Adj = AdjF32;
  using AdjLds = AdjF32Lds;
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
  double* T_out;         // FP64 travel time per event, column-major, at the event's ta_off
                         // (read by the anisotropic kernel terms); may be null
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
  size_t adj_lds_bytes;  // > 0: the launch's dynamic LDS; one-wave adjoints take the LDS tier when they fit
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

template <int PREC>
__device__ inline typename Prec<PREC>::Fwd::Ctx fwd_ctx(const GridDev& g, int iper, const Slot& ws,
                                                         const FixedBox& fb) {
  using V = typename Prec<PREC>::V;
  typename Prec<PREC>::Fwd::Ctx cx;
  const size_t po = (size_t)iper * g.seg_plane;
  cx.g = &g;
  for (int o = 0; o < 2; o++) {  // one segmented plane serves both orientations
    cx.t0[o] = static_cast<const V*>(ws.seg[0]);
    cx.t0x[o] = static_cast<const V*>(ws.seg[1]);
    cx.t0y[o] = static_cast<const V*>(ws.seg[2]);
    cx.a[o] = static_cast<const V*>(g.seg[0]) + po;
    cx.b[o] = static_cast<const V*>(g.seg[1]) + po;
    cx.c[o] = static_cast<const V*>(g.seg[2]) + po;
    cx.f[o] = static_cast<const V*>(g.seg[3]) + po;
    if constexpr (PREC == SAT_PREC_F32) {
      cx.dsx[o] = static_cast<const V*>(g.seg[4]) + po;
      cx.dsy[o] = static_cast<const V*>(g.seg[5]) + po;
    }
  }
  if constexpr (PREC == SAT_PREC_F64) {
    cx.dx = g.dx;
    cx.dy = g.dy;
  } else {
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
  cx.nx = g.nx;
  cx.ny = g.ny;
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

// The LDS exchange of the multi-wave forward and adjoint sweeps of one
// work-group: one buffer serves both (they run one after the other).
template <int PREC, int W>
__device__ inline char* fused_exchange_lds() {
  using P = Prec<PREC>;
  constexpr size_t nf = sizeof(ExchangeLds<typename P::Fwd, W>), na = sizeof(ExchangeLds<typename P::Adj, W>);
  __shared__ __attribute__((aligned(16))) char buf[nf > na ? nf : na];
  return buf;
}
template <int PREC, int W>
__device__ inline double* fused_reduce_lds() {
  __shared__ double red[W];
  return red;
}

// Forward solve (src/eikonal_solver.cpp:22-317) into the plain FP64 plane T = pl[0],
// by a work-group of W waves (W > 1: the lanes' edge values cross through LDS).
template <int PREC, int C, int W = 1>
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
  auto& xch = *reinterpret_cast<ExchangeLds<typename P::Fwd, W>*>(fused_exchange_lds<PREC, W>());
  const U* u = iterate<typename P::Fwd, SegLayout, C, W>(cx, Q[0], Q[1], Q[2], maxiter, kFwdTol, io,
                                                          exchange_ptr<typename P::Fwd, W>(xch),
                                                          fused_reduce_lds<PREC, W>());
  if (!io.converged) flags |= FLAG_NOT_CONVERGED;
  fwd_finish<V, U>(g, u, S[0], ws.pl[1], ws.pl[0]);
  __syncthreads();
  return io;
}

// Adjoint solve (src/eikonal_solver.cpp:323-511) from T (plain FP64); returns Ta plane.
template <int PREC, int C, int W = 1>
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
  auto& xch = *reinterpret_cast<ExchangeLds<typename P::Adj, W>*>(fused_exchange_lds<PREC, W>());
  const V* ta = iterate<typename P::Adj, SegLayout, C, W>(cx, static_cast<V*>(ws.seg[t0]),
                                                           static_cast<V*>(ws.seg[t0 + 1]),
                                                           static_cast<V*>(ws.seg[t0 + 2]), maxiter, kAdjTol, io,
                                                           exchange_ptr<typename P::Adj, W>(xch),
                                                           fused_reduce_lds<PREC, W>());
  if (!io.converged) flags |= FLAG_ADJ_MAXITER;
  return ta;
}

// The LDS tier of the per-pair adjoint: for a grid whose planes fit in the
// work-group's LDS, the adjoint sweeps read and write LDS instead of the slot
// in global memory -- FP32: the five coefficient planes and the three swept
// planes; FP64: the three swept planes (the coefficients stay in the slot).
// The set-up is that of solve_adjoint (into the slot), then copied; the same
// updates in the same order give the same bits.  Plane strides are rounded to
// 4 elements; each swept plane carries the trash row of the idle lanes' stores.
__host__ __device__ inline size_t lds_round4(size_t n) { return (n + 3) / 4 * 4; }
__device__ inline char* fused_dyn_lds() {
  extern __shared__ __attribute__((aligned(16))) char sat_fused_dyn[];
  return sat_fused_dyn;
}
template <int PREC>
__host__ __device__ inline size_t fused_adj_lds_bytes(size_t seg_plane) {
  const size_t sr = lds_round4(seg_plane), st = lds_round4(seg_plane + seg_trash_elems(1));
  return PREC == SAT_PREC_F64 ? 3 * st * sizeof(double) : (5 * sr + 3 * st) * sizeof(float);
}

template <int PREC, int C>
__device__ inline const typename Prec<PREC>::V* solve_adjoint_lds(const GridDev& g, int iper, const double* T,
                                                                   int nr, const double* rlon, const double* rlat,
                                                                   const double* adj, const Slot& ws, int maxiter,
                                                                   IterOut& io, int& flags, char* lds) {
  using P = Prec<PREC>;
  using V = typename P::V;
  if constexpr (PREC == SAT_PREC_F64) adj_setup_f64(g, iper, T, ws);
  else adj_setup_f32(g, iper, T, ws);
  __syncthreads();
  adj_delta<V>(g, nr, rlon, rlat, adj, ws);
  __syncthreads();
  const size_t S = g.seg_plane, sr = lds_round4(S), st = lds_round4(S + seg_trash_elems(1));
  V* L = reinterpret_cast<V*>(lds);
  auto cx = adj_ctx<PREC>(g, ws);
  constexpr int ncoef = (PREC == SAT_PREC_F64) ? 0 : 5;
  V* ta = L + ncoef * sr;  // three swept planes of st elements
  const int t0 = (PREC == SAT_PREC_F64) ? 6 : 5;
  for (size_t q = threadIdx.x; q < S; q += blockDim.x) {
#pragma unroll
    for (int i = 0; i < ncoef; i++) L[i * sr + q] = static_cast<const V*>(ws.seg[i])[q];
    ta[q] = static_cast<const V*>(ws.seg[t0])[q];
  }
  if constexpr (PREC == SAT_PREC_F32) {
    cx.cW = L;
    cx.cE = L + sr;
    cx.cS = L + 2 * sr;
    cx.cN = L + 3 * sr;
    cx.cD = L + 4 * sr;
  }
  __syncthreads();
  const V* res = iterate<typename P::AdjLds, LdsSegLayout, C>(cx, ta, ta + st, ta + 2 * st, maxiter, kAdjTol, io);
  if (!io.converged) flags |= FLAG_ADJ_MAXITER;
  __syncthreads();
  return res;
}

// ----------------------------------------------------------------------------
// One (source, period) pair: src/preproc.cpp:88-143.
// ----------------------------------------------------------------------------
template <int PREC, int C, typename TaT, int W = 1>
__device__ void run_pair(const BatchDev& bd, int e, const Slot& ws, int group) {
  using V = typename Prec<PREC>::V;
  const int lane = threadIdx.x;
  constexpr int NT = kWave * W;
  const EventDev ev = bd.events[e];
  const GridDev& g = bd.grids[ev.grid];
  int flags = 0;
  const IterOut fo = solve_forward<PREC, C, W>(g, ev.iper, ev.lon, ev.lat, ws, bd.fwd_maxiter, flags);
  if (bd.T_out) {  // the event's T where the batched kernels leave it (an FP64 re-solve overwrites it)
    double* To = bd.T_out + ev.ta_off;
    for (size_t p = lane; p < g.plane; p += NT) To[p] = ws.pl[0][p];
  }
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
  for (int r = lane; r < ev.nrec; r += NT) {
    const int k = ev.rec_off + r;
    const double s = interp2d_point(g.lon, g.nx, g.lat, g.ny, T, bd.rec_lon[k], bd.rec_lat[k]);
    bd.syn[k] = s;
    bd.adjsrc[k] = bd.weight[k] * (s - bd.tt_obs[k]);
    outside |= (s != s) ? 1 : 0;
  }
  if (block_or_i(outside)) flags |= FLAG_REC_OUTSIDE;
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
    const V* ta = nullptr;
    if constexpr (W == 1) {
      if (bd.adj_lds_bytes > 0 && fused_adj_lds_bytes<PREC>(g.seg_plane) <= bd.adj_lds_bytes)
        ta = solve_adjoint_lds<PREC, C>(g, ev.iper, T, ev.nrec, bd.rec_lon + ev.rec_off, bd.rec_lat + ev.rec_off,
                                        bd.adjsrc + ev.rec_off, ws, bd.adj_maxiter, ao, flags, fused_dyn_lds());
    }
    if (!ta)
      ta = solve_adjoint<PREC, C, W>(g, ev.iper, T, ev.nrec, bd.rec_lon + ev.rec_off, bd.rec_lat + ev.rec_off,
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
template <int PREC, int C, typename TaT, int W = 1>
__global__ __launch_bounds__(64 * W) void pair_kernel(BatchDev bd, int first, int count, int group,
                                                      const int* list, const int* dcount) {
  using P = Prec<PREC>;
  const Slot ws = make_slot(bd.ws + (size_t)blockIdx.x * bd.ws_slot, P::kSeg, P::kPlain, bd.ws_seg, bd.ws_plain);
  const int n = list ? *dcount : count;
  __shared__ int s_idx;
  for (;;) {
    int idx = 0;
    if constexpr (W == 1) {
      if (threadIdx.x == 0) idx = atomicAdd(bd.counter, 1);
      idx = __shfl(idx, 0);
    } else {
      if (threadIdx.x == 0) s_idx = atomicAdd(bd.counter, 1);
      __syncthreads();
      idx = s_idx;
      __syncthreads();  // every thread has read it before thread 0 takes the next
    }
    if (idx >= n) break;
    run_pair<PREC, C, TaT, W>(bd, list ? list[idx] : bd.order[first + idx], ws, group);
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
  size_t adj_lds_bytes;    // > 0: the launch's dynamic LDS; the adjoint takes the LDS tier when it fits
};

template <int PREC, int C, int W = 1>
__global__ __launch_bounds__(64 * W) void solo_forward(SoloDev sd) {
  using P = Prec<PREC>;
  const Slot ws = make_slot(sd.ws, P::kSeg, P::kPlain, sd.ws_seg, sd.ws_plain);
  int flags = 0;
  const IterOut io = solve_forward<PREC, C, W>(sd.g, 0, sd.src_lon, sd.src_lat, ws, sd.maxiter, flags);
  for (size_t p = threadIdx.x; p < sd.g.plane; p += kWave * W) sd.out[p] = ws.pl[0][p];
  if (threadIdx.x == 0) {
    sd.info[0] = io.iters; sd.info[1] = flags; sd.info[2] = io.l1; sd.info[3] = io.linf;
  }
}

template <int PREC, int C, int W = 1>
__global__ __launch_bounds__(64 * W) void solo_backward(SoloDev sd) {
  using P = Prec<PREC>;
  using V = typename P::V;
  const Slot ws = make_slot(sd.ws, P::kSeg, P::kPlain, sd.ws_seg, sd.ws_plain);
  int flags = 0;
  IterOut io;
  const V* ta = nullptr;
  if constexpr (W == 1) {
    if (sd.adj_lds_bytes > 0 && fused_adj_lds_bytes<PREC>(sd.g.seg_plane) <= sd.adj_lds_bytes)
      ta = solve_adjoint_lds<PREC, C>(sd.g, 0, sd.T_in, sd.nr, sd.rlon, sd.rlat, sd.adj, ws, sd.maxiter, io, flags,
                                      fused_dyn_lds());
  }
  if (!ta)
    ta = solve_adjoint<PREC, C, W>(sd.g, 0, sd.T_in, sd.nr, sd.rlon, sd.rlat, sd.adj, ws, sd.maxiter, io, flags);
  for (int iy = 0; iy < sd.g.ny; iy++)
    for (int ix = threadIdx.x; ix < sd.g.nx; ix += kWave * W)
      sd.out[ix + (size_t)iy * sd.g.nx] = (double)ta[seg_index(sd.g, ix, iy)];
  if (threadIdx.x == 0) {
    sd.info[0] = io.iters; sd.info[1] = flags; sd.info[2] = io.l1; sd.info[3] = io.linf;
  }
}


}  // namespace sat
