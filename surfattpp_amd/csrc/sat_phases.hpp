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
//   adj_setup_kernel   per (event, node): x- and y-face fluxes (src/eikonal_solver.cpp:377-441)
//   delta_kernel       per event: receivers spread in receiver order (:356-375),
//                      chi term (src/preproc.cpp:142)
//   adj_sweep_kernel   persistent, one wave per event: adjoint iterations, mask,
//                      masked Ta (src/eikonal_solver.cpp:471-536)
// Node-parallel work runs at full occupancy; the wave-per-event kernels hold only
// the sweep pipeline in registers.  Every plane a sweep streams is stored
// wavefront-major, one copy per sweep orientation (sat_device.hpp), so each
// step of a sweep reads one contiguous run per plane.
#pragma once
#include <type_traits>

#include "../../include/surfatt_eik.h"
#include "sat_crmath.hpp"
#include "sat_engine.hpp"
#include "sat_policies.hpp"

namespace sat {

struct EventX {
  int grid, iper, rec_off, nrec;
  int fb[4];      // fixed near-source box x0, x1, y0, y1 (src/eikonal_solver.cpp:119-129)
  int mb[4];      // mask box x0, x1, y0, y1 (:528-531)
  double lon, lat;
  size_t w_off;   // element offset of this event's wavefront-major planes
  size_t pl_off;  // element offset of this event's column-major planes
  size_t col_off; // offset into the column trig table
};

struct PhaseDev {
  const GridDev* grids;
  const EventX* ev;
  int ev_base;          // the launch covers events [ev_base, ev_base + nevent) (a chunk)
  int nevent;
  int rec_lo;           // and receivers [rec_lo, rec_lo + nrec) of the launch's receiver pass
  double* evs;          // per event: fun0, sin(y0), cos(y0)
  double2* coltab;      // per (event, column): sin, cos of xx - x0
  // per event, wavefront-major, per orientation; the adjoint planes reuse the
  // forward planes' memory once the forward sweeps are done
  void* t0[2];          // T0, T0x, T0y (sweep precision)
  void* t0x[2];
  void* t0y[2];
  void* fxy[2];         // face fluxes: FP64 (a1, b1); FP32 (a1/dx, b1/dy)
  void* dl[2];          // delta (sweep precision)
  double* T0d;          // FP32 path: FP64 T0, column-major, per event
  double* T;            // travel time, column-major, per event
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
  unsigned char* skip;  // FP32 batches: 0 FP32 result, 1 FP64 re-solve by these kernels,
                        // 2 FP64 re-solve by the per-pair kernel (pool full), kSkipSpec
                        // FP64 solve from the start of the run (speculated; no FP32 sweeps)
  int want_skip;        // the skip value of the events this pass processes (< 0: every event)
  int kden_pass;        // the density-kernel adjoint (src/preproc.cpp:128-136): receiver sources -w,
                        // the field not masked, syn, chi, iteration counts and flags left as they are
  int delta_cap;        // shares per event the receiver spread holds in LDS (launch's dynamic LDS)
  unsigned long long* dmask;  // FP32 batches: per event, dmask_stride words: per row of the adjoint
                              // geometry, ceil(La / 64) words of one bit per lane, set where the
                              // lane's columns hold a receiver share (delta_kernel); null: the adjoint
                              // reads every delta (adj_setup_kernel zeroes the planes)
  int dmask_stride;
  int dmask_lds;        // bytes of dynamic LDS of the FP32 adjoint sweeps (one event's mask)
  int* fb_list;         // FP32 forward: pool re-solves, per launch group
  int* fb_count;
  int* ov_list;         // FP32 forward: re-solves beyond the pool, per launch group
  int* ov_count;
  int* pool_counter;
  int pool_size;
  size_t pool_stride;   // elements of one pool entry of the FP64 planes
  EventX* ev64;         // FP32 forward: the events as the FP64 re-solve sees them
  double f32_band;
  const int* list;      // non-null: events list[0, *list_count) (FP64 re-solve pass)
  const int* list_count;
  int ta_f32;           // masked Ta stored as float (FP64 re-solve of an FP32 batch)
  int scaled_iso;       // FP64 re-solves of an FP32 batch: isotropic grids sweep in scaled
                        // coordinates (gradient planes sqrt(a) T0x, sqrt(b) T0y; FwdF64RIso)
  char* slots;          // per resident wave: 3 swept planes
  size_t slot_bytes, slot_plane;
  const int* order;      // the forward queue of each launch group (event ids)
  const int* order_adj;  // the adjoint queue (the same events, ordered by adjoint rounds)
  int* counter;
  int fwd_maxiter, adj_maxiter;
  int polish;           // SAT_PREC_F32P batch: isotropic grids' T0 gradient planes are FP64, scaled
};

// Rows in flight per lane (register ring depth, sat_engine.hpp): multi-wave
// sweeps, and the single-wave forward and adjoint sweeps (1 = the head/body
// pipeline, fewer registers; >= 2 = the ring, no wait on the prefetches).
#ifndef SAT_RING_DEPTH
#define SAT_RING_DEPTH 4
#endif
constexpr int kRingDepth = SAT_RING_DEPTH;
constexpr int kFwdDepth1 = 1;
constexpr int kAdjDepth1 = 3;
// the one-wave adjoint's rows in flight: the ring of kAdjDepth1 rows, or the
// head/body pipeline for eight columns per lane (the ring's registers would
// leave one wave per SIMD)
template <int C>
constexpr int adj_depth1() { return C >= 8 ? 1 : kAdjDepth1; }

// The multi-wave FP64 sweeps carry the latency-critical re-solves of FP32
// batches, which share the CUs with a throughput kernel on the other stream:
// their waves take the highest instruction-issue priority on their SIMDs
// (s_setprio; the streams themselves keep the default priority).  Multi-wave
// FP32 sweeps are ordinary throughput work and keep the default.
template <int PREC, int W>
__device__ __forceinline__ void latency_critical_priority() {
  if constexpr (PREC == SAT_PREC_F64 && W > 1) __builtin_amdgcn_s_setprio(3);
}

// Rows in flight per lane of the multi-wave sweeps: the register ring of
// kRingDepth rows where it fits, two rows for the FP32 forward at C >= 3
// (its per-row body is ten registers per column), and kRingDepthC1 rows for
// the FP64 sweeps at one column per lane and four waves -- the re-solves,
// whose steps are short enough that four rows in flight leave memory latency
// exposed (at 8 rows the forward kernel holds 256 VGPRs without spilling; two
// columns per lane, or eight waves with their 256-VGPR limit, would spill, so
// they keep kRingDepth).
#ifndef SAT_RING_DEPTH_C1
#define SAT_RING_DEPTH_C1 8
#endif
constexpr int kRingDepthC1 = SAT_RING_DEPTH_C1;
template <int PREC, int C, int W>
constexpr int mw_fwd_depth() {
  if (C >= 5) return 1;  // the wide grids' kMaxC columns per lane: the head/body pipeline
  return (PREC == SAT_PREC_F32 && C >= 3) ? 2 : (PREC == SAT_PREC_F64 && C == 1 && W <= 4) ? kRingDepthC1 : kRingDepth;
}
template <int PREC, int C, int W>
constexpr int mw_adj_depth() {
  if (C >= 5) return 1;
  return (PREC == SAT_PREC_F64 && C == 1 && W <= 4) ? kRingDepthC1 : kRingDepth;
}

template <int PREC> struct PhaseT;
template <> struct PhaseT<SAT_PREC_F64> {
  using V = double;     // T0 planes and adjoint
  using U = double;     // forward swept value (tau)
  using F = double2;    // face fluxes
  using Fwd = FwdF64;
  using FwdIso = FwdF64;
  using FwdIsoP = FwdF64;
  using Adj = AdjF64Face;
};
template <> struct PhaseT<SAT_PREC_F32> {
  using V = float;
  using U = double;     // sigma = tau - 1 in FP64
  using F = float2;
  using Fwd = FwdF32;
  using FwdIso = FwdF32Iso;
  using FwdIsoP = FwdF32IsoP;
  using Adj = AdjF32Face;
};

// Next work-queue entry for the whole work-group (thread 0 pops it).  The
// entry is returned as a wave-uniform value (readfirstlane): the event's
// descriptor and plane pointers derived from it then live in scalar registers,
// which leaves the vector registers to the sweep's row ring.
template <int W>
__device__ inline int pop_work(int* counter, int* s_idx) {
  if constexpr (W == 1) {
    int idx = 0;
    if (threadIdx.x == 0) idx = atomicAdd(counter, 1);
    return __builtin_amdgcn_readfirstlane(__shfl(idx, 0));
  } else {
    __syncthreads();  // the previous entry's readers are done with *s_idx
    if (threadIdx.x == 0) *s_idx = atomicAdd(counter, 1);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*s_idx);
  }
}

// The re-solve mark of a pair solved in FP64 from the start of its run.
constexpr unsigned char kSkipSpec = 3;

// Whether this pass processes event e (its re-solve mark is the pass's).
__device__ __forceinline__ bool pass_event(const PhaseDev& pd, int e) {
  return pd.want_skip < 0 || pd.skip[e] == pd.want_skip;
}

// Event of a node-parallel block: blockIdx.y, or the list entry blockIdx.y; -1 if none.
__device__ inline int block_event(const PhaseDev& pd) {
  if (!pd.list) return pd.ev_base + (int)blockIdx.y;
  return ((int)blockIdx.y < *pd.list_count) ? pd.list[blockIdx.y] : -1;
}

// Decode element q of an orientation-0 wavefront-major plane: column ix, row j;
// false for padding.
__host__ __device__ inline bool wave_decode0(const GridDev& g, int q, int& ix, int& j) {
  const int s = q / g.C, k = q - s * g.C;
  const int dg = s / g.La, l = s - dg * g.La;
  j = dg - l;
  ix = l * g.C + k;
  return j >= 0 && j < g.ny && ix < g.nx;
}

// ---------------------------------------------------------------------------
template <int PREC>
__global__ void trig_kernel(PhaseDev pd) {
  const int e = pd.ev_base + (int)blockIdx.y;
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

// Decode element q of an orientation-o wavefront-major plane: column ix, row j;
// false for padding.
__host__ __device__ inline bool wave_decode(const GridDev& g, int o, int q, int& ix, int& j) {
  const int s = q / g.C, k = q - s * g.C;
  const int dg = s / g.La, l = s - dg * g.La;
  j = (o == 0) ? dg - l : dg + l - (g.La - 1);
  ix = l * g.C + k;
  return j >= 0 && j < g.ny && ix < g.nx;
}

// src/eikonal_solver.cpp:96-117; acos correctly rounded in the FP64 path.
// blockIdx.z = pass: 0 and 1 write the orientation-0 / -1 wavefront-major
// planes (T0, T0x, T0y; thread = plane element, padding skipped), 2 the FP64
// column-major T0 of FP32 sweeps (thread = node): every store coalesced, the
// node's arithmetic repeated per pass (measured: faster than forming T0 once
// and gathering it for the wavefront-major passes).  A thread takes
// kInitPerThread elements (fewer work-groups to dispatch).
constexpr int kInitPerThread = 8;
template <int PREC>
__global__ void init_kernel(PhaseDev pd) {
  using V = typename PhaseT<PREC>::V;
  const int e = block_event(pd);
  if (e < 0) return;
  const EventX ev = pd.ev[e];
  const GridDev& g = pd.grids[ev.grid];
  const int pass = blockIdx.z;
  const double fun0 = pd.evs[3 * (size_t)e], sy0 = pd.evs[3 * (size_t)e + 1], cy0 = pd.evs[3 * (size_t)e + 2];
  const bool scaled = (PREC == SAT_PREC_F32 && pd.polish && g.iso && g.wf64[0]) ||
                      ((PREC == SAT_PREC_F32 || pd.scaled_iso) && g.iso);
  const int n = pass < 2 ? (int)g.wplane : (int)g.plane;
  // kInitPerThread elements per thread, the block's threads side by side
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
    int ix, iy;
    if (pass < 2) {
      if (!wave_decode(g, pass, idx, ix, iy)) continue;
    } else {
      iy = idx / g.nx;
      ix = idx - iy * g.nx;
    }
    const double2 rw = g.trig_yy[iy];
    const double2 cl = pd.coltab[ev.col_off + ix];
    const double syy = rw.x, cyy = rw.y, sdl = cl.x, cdl = cl.y;
    const double tmp = syy * sy0 + cyy * cy0 * cdl;
    const double tmp_c = dmax(-1.0, dmin(1.0, tmp));
    const double ac = (PREC == SAT_PREC_F64) ? cr::acos_cr(tmp_c) : acos(tmp_c);
    const double T0 = ac * kR * fun0;
    if (pass == 2) {  // the FP64 column-major T0 of FP32 sweeps: no gradient
      pd.T0d[ev.pl_off + idx] = T0;
      continue;
    }
    const size_t w = ev.w_off + idx;
    static_cast<V*>(pd.t0[pass])[w] = (V)T0;
    const double sin_d = sqrt(dmax(0.0, 1.0 - tmp_c * tmp_c));
    const bool flat = sin_d < 1e-14;
    const double gx = cyy * cy0 * sdl, gy = -(cyy * sy0 - syy * cy0 * cdl);
    if (PREC == SAT_PREC_F64 && !scaled) {
      // FP64 batches: the reference's expressions, in its operation order
      static_cast<V*>(pd.t0x[pass])[w] = flat ? 0.0 : gx / sin_d * kR * fun0;
      static_cast<V*>(pd.t0y[pass])[w] = flat ? 0.0 : gy / sin_d * kR * fun0;
      continue;
    }
    // FP32 sweeps and FP64 re-solves (the same quantities, one division per
    // node): scaled by sqrt(a) = 1/(R cos y), sqrt(b) = 1/R on isotropic grids,
    // T0x = cy cy0 sdl fun0 R / sin_d and T0y likewise otherwise
    const double r = flat ? 0.0 : (scaled ? fun0 : kR * fun0) / sin_d;
    const double tx = (scaled ? cy0 * sdl : gx) * r, ty = gy * r;
    if (PREC == SAT_PREC_F32 && pd.polish && g.iso && g.wf64[0]) {
      // the polished isotropic FP32 sweeps' scaled gradient, FP64 (FwdF32IsoP)
      static_cast<double*>(pd.t0x[pass])[w] = tx;
      static_cast<double*>(pd.t0y[pass])[w] = ty;
    } else {
      // FP32: the sweeps' gradient (FwdF32Iso scaled, FwdF32); FP64: the re-solves' (FwdF64RIso)
      static_cast<V*>(pd.t0x[pass])[w] = (V)tx;
      static_cast<V*>(pd.t0y[pass])[w] = (V)ty;
    }
  }
}

// Metric modes of the forward sweeps: the general metric, the isotropic metric
// (spha = sphb = 1, sphc = 0), and the isotropic metric with the FP64-polished
// FP32 update (SAT_PREC_F32P batches).
// kMetricExact: FP64 multi-wave sweeps of FP64 batches over grids wider than
// 512 columns, with the reference-order update (the other FP64 multi-wave
// sweeps are the re-solves' reciprocal-based update).
constexpr int kMetricGeneral = 0, kMetricIso = 1, kMetricIsoPolish = 2, kMetricExact = 3;

template <int PREC, int MODE, class Pol>
__device__ inline auto phase_fwd_ctx(const PhaseDev& pd, const GridDev& g, const EventX& ev) {
  constexpr bool ISO = MODE == kMetricIso || MODE == kMetricIsoPolish;
  constexpr bool POL = PREC == SAT_PREC_F32 && MODE == kMetricIsoPolish;
  constexpr bool RISO = std::is_same_v<Pol, FwdF64RIso>;  // FP64 re-solve, isotropic, scaled
  using PT = PhaseT<PREC>;
  using V = typename PT::V;
  typename Pol::Ctx cx;
  const size_t po = (size_t)ev.iper * g.wplane;
  cx.g = &g;
  cx.fb = FixedBox{ev.fb[0], ev.fb[1], ev.fb[2], ev.fb[3]};
  // the polished isotropic FP32 sweeps: FP64 scaled gradient and slowness (FwdF32IsoP)
  using G = std::conditional_t<POL, double, V>;
#pragma unroll
  for (int o = 0; o < 2; o++) {
    cx.t0[o] = static_cast<const V*>(pd.t0[o]) + ev.w_off;
    cx.t0x[o] = static_cast<const G*>(pd.t0x[o]) + ev.w_off;
    cx.t0y[o] = static_cast<const G*>(pd.t0y[o]) + ev.w_off;
    if constexpr (POL) cx.f[o] = g.wf64[o] + po;
    else cx.f[o] = static_cast<const V*>(g.wmed[3][o]) + po;
    if constexpr (!(PREC == SAT_PREC_F32 && ISO) && !RISO) {
      cx.a[o] = static_cast<const V*>(g.wmed[0][o]) + po;
      cx.b[o] = static_cast<const V*>(g.wmed[1][o]) + po;
      cx.c[o] = static_cast<const V*>(g.wmed[2][o]) + po;
    }
    if constexpr (PREC == SAT_PREC_F32 && !ISO) {
      cx.dsx[o] = static_cast<const V*>(g.wmed[4][o]) + po;
      cx.dsy[o] = static_cast<const V*>(g.wmed[5][o]) + po;
    }
  }
  if constexpr (RISO) {
    cx.rowha = g.rowha;
    cx.kb = g.kb;
  } else if constexpr (PREC == SAT_PREC_F64) {
    cx.dx = g.dx;
    cx.dy = g.dy;
    if constexpr (std::is_same_v<Pol, FwdF64R>) {
      cx.idx = 1.0 / g.dx;
      cx.idy = 1.0 / g.dy;
    }
  } else {
    if constexpr (POL) {
      cx.rowha = g.rowha;
      cx.kb = g.kb;
    } else {
      if constexpr (ISO) cx.rowtab = g.rowtab;
      cx.idx = (float)(1.0 / g.dx);
      cx.idy = (float)(1.0 / g.dy);
    }
  }
  return cx;
}

// Waves per SIMD the single-wave FP32 forward sweeps are compiled for: three
// (<= 168 VGPRs) where the row pipeline fits without spills in the step loop --
// with 3 x 1024 resident waves a 3,000-pair batch is resident at once and the
// VALU-bound sweeps of three pairs interleave on each SIMD -- otherwise the
// compiler's own choice.
// The polished update holds the node's FP64 gradient and slowness in the row
// pipeline: two waves per SIMD (no spills) where the plain update fits three.
template <int PREC, int C, int MODE, int W>
constexpr int fwd_waves_per_eu() {
  if (PREC != SAT_PREC_F32 || W != 1) return 1;
  if (MODE == kMetricIsoPolish) return C <= 4 ? 2 : 1;
  return (MODE == kMetricIso ? C <= 4 : C <= 2) ? 3 : 1;
}

// Persistent forward sweeps, one work-group of W waves per event.
template <int PREC, int C, int MODE, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(fwd_waves_per_eu<PREC, C, MODE, W>())))
void fwd_sweep_kernel(PhaseDev pd, int first, int count, int group) {
  using PT = PhaseT<PREC>;
  using U = typename PT::U;
  // multi-wave FP64 launches are the re-solves of FP32 batches (FwdF64R)
  using Pol = std::conditional_t<
      (PREC == SAT_PREC_F64 && W > 1 && MODE != kMetricExact),
      std::conditional_t<MODE == kMetricIso, FwdF64RIso, FwdF64R>,
      std::conditional_t<MODE == kMetricIsoPolish, typename PT::FwdIsoP,
                         std::conditional_t<MODE == kMetricIso, typename PT::FwdIso, typename PT::Fwd>>>;
  const int lane = threadIdx.x;
  const int nthr = kWave * W;
  __shared__ ExchangeLds<Pol, W> xch;
  __shared__ double red[W];
  __shared__ int s_idx;
  latency_critical_priority<PREC, W>();
  char* sb = pd.slots + (size_t)blockIdx.x * pd.slot_bytes;
  U* Q0 = reinterpret_cast<U*>(sb);
  U* Q1 = reinterpret_cast<U*>(sb + pd.slot_plane);
  U* Q2 = reinterpret_cast<U*>(sb + 2 * pd.slot_plane);
  constexpr bool kSigma = PREC == SAT_PREC_F32;
  const int n_work = pd.list ? *pd.list_count : count;
  for (;;) {
    const int idx = pop_work<W>(pd.counter, &s_idx);
    if (idx >= n_work) break;
    const int e = pd.list ? pd.list[idx] : pd.order[first + idx];
    if (!pd.list && pd.skip[e] == kSkipSpec) continue;  // solved in FP64 on the second stream
    const EventX ev = pd.ev[e];
    const GridDev& g = pd.grids[ev.grid];
    const int nx = g.nx, n = (int)g.wplane;
    for (int q = lane; q < n; q += nthr) {  // :119-129 initial tau (sigma), orientation 0
      int ix, iy;
      const bool real = wave_decode0(g, q, ix, iy);
      const bool fixed = real && ix >= ev.fb[0] && ix <= ev.fb[1] && iy >= ev.fb[2] && iy <= ev.fb[3];
      Q0[q] = !real ? U(0) : kSigma ? (fixed ? U(0) : U(9)) : (fixed ? U(1) : U(10));
    }
    __syncthreads();
    const auto cx = uniform(phase_fwd_ctx<PREC, MODE, Pol>(pd, g, ev));
    IterOut io;
    const U* u = iterate<Pol, WaveLayout, C, W, (W > 1) ? mw_fwd_depth<PREC, C, W>() : kFwdDepth1>(cx, Q0, Q1, Q2, pd.fwd_maxiter, kFwdTol,
                                                                          io, exchange_ptr<Pol, W>(xch), red);
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
        unsigned char sk = 0;
        if (amb) {
          const int slot = atomicAdd(pd.pool_counter, 1);
          if (slot < pd.pool_size) {
            pd.ev64[e].w_off = (size_t)slot * pd.pool_stride;
            pd.fb_list[atomicAdd(pd.fb_count + group, 1)] = e;
            sk = 1;
          } else {
            pd.ov_list[atomicAdd(pd.ov_count + group, 1)] = e;
            sk = 2;
          }
        }
        pd.skip[e] = sk;
      }
      if (amb) {
        __syncthreads();
        continue;
      }
    }
    // :317 T = tau .* T0 (FP64) or T0 + sigma*T0 (FP32 sweeps), column-major
    double* T = pd.T + ev.pl_off;
    const typename PT::V* t0 = static_cast<const typename PT::V*>(pd.t0[0]) + ev.w_off;
    for (int iy = 0; iy < g.ny; iy++)
      for (int ix = lane; ix < nx; ix += nthr) {
        const size_t o = wave_index(g, 0, ix, iy);
        const size_t p = ix + (size_t)iy * nx;
        if (kSigma) T[p] = pd.T0d[ev.pl_off + p] + (double)u[o] * pd.T0d[ev.pl_off + p];
        else T[p] = (double)u[o] * (double)t0[o];
      }
    __syncthreads();
  }
}

// recv_kernel (k_phase64.hip): src/preproc.cpp:105-115 per receiver, and the
// receiver's delta cell and weight (src/eikonal_solver.cpp:357-369).
__global__ void recv_kernel(PhaseDev pd, int nrec);

// Per node: x-face flux a1 on the columns 1..nx-1 of the interior rows, y-face
// flux b1 on the rows 1..ny-1 of the interior columns (0 elsewhere), delta = 0.
// FP32 stores (a1/dx, b1/dy) rounded once.
template <int PREC>
__global__ void adj_setup_kernel(PhaseDev pd) {
  using V = typename PhaseT<PREC>::V;
  using F = typename PhaseT<PREC>::F;
  const int e = block_event(pd);
  if (e < 0 || !pass_event(pd, e)) return;
  const EventX ev = pd.ev[e];
  const GridDev& g = pd.grids[ev.grid];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int)g.plane) return;
  const int nx = g.nx, ny = g.ny;
  const int iy = p / nx, ix = p - iy * nx;
  const double* T = pd.T + ev.pl_off;
  const double* a = g.aA_plain + (size_t)ev.iper * g.plane;
  const double* b = g.b_plain + (size_t)ev.iper * g.plane;
  const double* c = g.c_plain + (size_t)ev.iper * g.plane;
  double fx = 0.0, fy = 0.0;
  if (PREC == SAT_PREC_F32 && g.iso) {
    // isotropic metric: a and b depend on the row only and c = 0, so the
    // cross terms vanish; the node's row values stand for its neighbours'
    // (the same (a + a) / 2 and (b + b') / 2), and the a, b, c planes are not read
    const double ar = a[(size_t)iy * nx];
    if (ix >= 1 && iy >= 1 && iy < ny - 1) fx = -(T[p] - T[p - 1]) / g.dx * (ar + ar) / 2.0;
    if (iy >= 1 && ix >= 1 && ix < nx - 1)
      fy = -(T[p] - T[p - nx]) / g.dy * (b[(size_t)iy * nx] + b[(size_t)(iy - 1) * nx]) / 2.0;
  } else {
    if (ix >= 1 && iy >= 1 && iy < ny - 1) fx = face_a1(g, T, a, c, ix, iy);
    if (iy >= 1 && ix >= 1 && ix < nx - 1) fy = face_b1(g, T, b, c, ix, iy);
  }
  F v;
  if constexpr (PREC == SAT_PREC_F64) v = make_double2(fx, fy);
  else v = make_float2((float)(fx / g.dx), (float)(fy / g.dy));
  const bool zero_dl = PREC == SAT_PREC_F64 || !pd.dmask;  // (FP32 with the row mask: delta_kernel)
#pragma unroll
  for (int o = 0; o < 2; o++) {
    const size_t w = ev.w_off + wave_index(g, o, ix, iy);
    static_cast<F*>(pd.fxy[o])[w] = v;
    if (zero_dl) static_cast<V*>(pd.dl[o])[w] = V(0);
  }
}

// src/eikonal_solver.cpp:356-375 and src/preproc.cpp:142, one wave per event.
// The reference adds each receiver's four bilinear shares into delta in
// receiver order (delta starts at 0: adj_setup_kernel); a node shared by
// several receivers gets ((0 + v1) + v2) + ... in that order.  Here the wave
// lists the event's 4*nrec shares (node, value) in that order in LDS, and the
// lane holding a node's first share sums all of that node's shares in order
// and stores the sum: the same bits, no dependent global read-modify-write
// chain.  The scans read kScan shares per batch so their loads overlap.
// FP64: stored in both orientation planes.  FP32: summed in FP64 and rounded
// once into both planes.  The launch sizes the LDS for its largest event
// (delta_cap shares, at most kDeltaShares); an event with more shares takes
// the receiver loop on one lane (the same sums, accumulated in place).  The
// chi term is summed by lane 0 in receiver order.
constexpr int kDeltaShares = 2048;
constexpr int kScan = 16;
__host__ __device__ inline size_t delta_lds_bytes(int cap) {
  return (size_t)(cap + kScan) * (sizeof(double) + sizeof(int));
}

template <int PREC>
__global__ __launch_bounds__(kWave) void delta_kernel(PhaseDev pd) {
  using V = typename PhaseT<PREC>::V;
  extern __shared__ double dsh[];
  const int cap = pd.delta_cap;
  double* s_val = dsh;                                    // [cap + kScan]
  int* s_node = reinterpret_cast<int*>(dsh + cap + kScan);  // [cap + kScan]
  const int e = pd.ev_base + (int)blockIdx.x;
  if (e >= pd.ev_base + pd.nevent || !pass_event(pd, e)) return;
  const int lane = threadIdx.x;
  const EventX ev = pd.ev[e];
  const GridDev& g = pd.grids[ev.grid];
  if (lane == 0 && !pd.kden_pass) {
    double s = 0.0;
    for (int r = 0; r < ev.nrec; r++) {
      const int k = ev.rec_off + r;
      const double d = pd.syn[k] - pd.tt_obs[k];
      s += pd.weight[k] * (d * d);
    }
    pd.chi_ev[e] = 0.5 * s;
  }
  const int k0 = ev.rec_off, ns = 4 * ev.nrec;
  V* dl0 = static_cast<V*>(pd.dl[0]) + ev.w_off;
  V* dl1 = static_cast<V*>(pd.dl[1]) + ev.w_off;
  auto share = [&](int c, int& i, int& j) -> double {  // share m of receiver k, order of :371-374
    const int k = k0 + (c >> 2), m = c & 3;
    const int2 cc = pd.rcell[k];
    const double w = pd.rwt[3 * (size_t)k], r1 = pd.rwt[3 * (size_t)k + 1], r2 = pd.rwt[3 * (size_t)k + 2];
    i = cc.x + (m >> 1);
    j = cc.y + (m & 1);
    return (m == 0) ? w * (1 - r1) * (1 - r2) : (m == 1) ? w * (1 - r1) * r2 : (m == 2) ? w * r1 * (1 - r2) : w * r1 * r2;
  };
  if constexpr (PREC == SAT_PREC_F32) {
    if (pd.dmask) {
      // The row mask: the adjoint reads the delta planes only in the lane
      // segments (C columns of one row) marked here, which are zeroed first in
      // both orientations; elsewhere the planes hold the forward's values.
      unsigned long long* msk = pd.dmask + (size_t)e * pd.dmask_stride;
      const int mw = (g.La + kWave - 1) / kWave;
      for (int i = lane; i < g.ny * mw; i += kWave) msk[i] = 0ull;
      for (int c = lane; c < ns; c += kWave) {
        int i, j;
        (void)share(c, i, j);
        const int l = i / g.C;
        for (int k = 0; k < g.C; k++) {
          const int ix = l * g.C + k;
          dl0[wave_index(g, 0, ix, j)] = V(0);
          dl1[wave_index(g, 1, ix, j)] = V(0);
        }
      }
      __threadfence_block();  // the zeroes are written before the values and the marks
      __syncthreads();
      for (int c = lane; c < ns; c += kWave) {
        int i, j;
        (void)share(c, i, j);
        const int l = i / g.C;
        atomicOr(msk + (size_t)j * mw + (l >> 6), 1ull << (l & 63));
      }
    }
  }
  if (ns > cap) {
    if (lane != 0) return;
    double* acc = pd.T0d + ev.pl_off;  // FP32: an FP64 accumulator (T0d is no longer needed once T is formed)
    for (int pass = 0; pass < ((PREC == SAT_PREC_F64) ? 2 : 3); pass++) {
      for (int c = 0; c < ns; c++) {
        int i, j;
        const double v = share(c, i, j);
        const size_t w0 = wave_index(g, 0, i, j), w1 = wave_index(g, 1, i, j);
        const size_t pp = i + (size_t)j * g.nx;
        if constexpr (PREC == SAT_PREC_F64) {
          if (pass == 0) dl0[w0] += v;
          else dl1[w1] = dl0[w0];
        } else {
          if (pass == 0) {
            acc[pp] = 0.0;
          } else if (pass == 1) {
            acc[pp] += v;
          } else {
            const float f = (float)acc[pp];
            dl0[w0] = f;
            dl1[w1] = f;
          }
        }
      }
    }
    return;
  }
  for (int c = lane; c < ns + kScan; c += kWave) {
    if (c < ns) {
      int i, j;
      s_val[c] = share(c, i, j);
      s_node[c] = i + j * g.nx;
    } else {
      s_val[c] = 0.0;
      s_node[c] = -1;  // scan padding: matches no node
    }
  }
  __syncthreads();
  for (int c = lane; c < ns; c += kWave) {
    const int n = s_node[c];
    bool first = true;
    for (int q0 = 0; q0 < c && first; q0 += kScan) {
      bool hit = false;
#pragma unroll
      for (int u = 0; u < kScan; u++) hit |= (q0 + u < c) & (s_node[q0 + u] == n);
      first = !hit;
    }
    if (!first) continue;
    double sum = 0.0;
    for (int q0 = c; q0 < ns; q0 += kScan) {  // reads up to ns - 1 + kScan - 1: the padding
      int nd[kScan];
      double vv[kScan];
#pragma unroll
      for (int u = 0; u < kScan; u++) {
        nd[u] = s_node[q0 + u];
        vv[u] = s_val[q0 + u];
      }
#pragma unroll
      for (int u = 0; u < kScan; u++)
        if (nd[u] == n) sum += vv[u];  // in share order
    }
    const int j = n / g.nx, i = n - j * g.nx;
    const size_t w0 = wave_index(g, 0, i, j), w1 = wave_index(g, 1, i, j);
    if constexpr (PREC == SAT_PREC_F64) {
      dl0[w0] = sum;
      dl1[w1] = sum;
    } else {
      const float f = (float)sum;
      dl0[w0] = f;
      dl1[w1] = f;
    }
  }
}

// Occupancy of the one-wave FP32 adjoint sweeps on eight columns per lane:
// three waves per SIMD, so that a batch of up to three pairs per SIMD is
// resident at once (the memory-bound sweeps gain more from waves in flight than
// from the registers beyond 168).
template <int PREC, int C, int W>
constexpr int adj_waves_per_eu() { return (PREC == SAT_PREC_F32 && W == 1 && C == 8) ? 3 : 1; }

// Dynamic LDS of an adjoint sweep launch: FP32, one event's delta row mask.
template <int PREC>
inline size_t adj_dyn_lds(const PhaseDev& pd) {
  return PREC == SAT_PREC_F32 ? (size_t)pd.dmask_lds : 0;
}

// Persistent adjoint sweeps, one work-group of W waves per event; mask and store.
template <int PREC, int C, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(adj_waves_per_eu<PREC, C, W>())))
void adj_sweep_kernel(PhaseDev pd, int first, int count) {
  using PT = PhaseT<PREC>;
  using V = typename PT::V;
  using F = typename PT::F;
  using Pol = typename PT::Adj;
  const int lane = threadIdx.x;
  const int nthr = kWave * W;
  __shared__ ExchangeLds<Pol, W> xch;
  __shared__ double red[W];
  __shared__ int s_idx;
  extern __shared__ unsigned long long s_mask[];  // FP32: the event's delta row mask (pd.dmask_lds bytes)
  latency_critical_priority<PREC, W>();
  char* sb = pd.slots + (size_t)blockIdx.x * pd.slot_bytes;
  V* Q0 = reinterpret_cast<V*>(sb);
  V* Q1 = reinterpret_cast<V*>(sb + pd.slot_plane);
  V* Q2 = reinterpret_cast<V*>(sb + 2 * pd.slot_plane);
  const int n_work = pd.list ? *pd.list_count : count;
  for (;;) {
    const int idx = pop_work<W>(pd.counter, &s_idx);
    if (idx >= n_work) break;
    const int e = pd.list ? pd.list[idx] : pd.order_adj[first + idx];
    if (!pass_event(pd, e)) continue;
    const EventX ev = pd.ev[e];
    const GridDev& g = pd.grids[ev.grid];
    const int nx = g.nx, ny = g.ny, n = (int)g.wplane;
    for (int q = lane; q < n; q += nthr) {  // :446-449 boundary 0, interior 100, orientation 0
      int ix, iy;
      const bool real = wave_decode0(g, q, ix, iy);
      Q0[q] = (real && ix >= 1 && ix < nx - 1 && iy >= 1 && iy < ny - 1) ? V(100) : V(0);
    }
    // FP32: the row mask in LDS (without one: a single all-lanes word for every row)
    const int mw = (PREC == SAT_PREC_F32 && pd.dmask) ? (g.La + kWave - 1) / kWave : 0;
    if constexpr (PREC == SAT_PREC_F32) {
      if (mw) {
        const unsigned long long* msk = pd.dmask + (size_t)e * pd.dmask_stride;
        for (int i = lane; i < ny * mw; i += nthr) s_mask[i] = msk[i];
      } else if (lane == 0) {
        s_mask[0] = ~0ull;
      }
    }
    __syncthreads();
    typename Pol::Ctx cx;
    cx.g = &g;
    cx.nx = g.nx;
    cx.ny = g.ny;
#pragma unroll
    for (int o = 0; o < 2; o++) {
      cx.fxy[o] = static_cast<const F*>(pd.fxy[o]) + ev.w_off;
      cx.dl[o] = static_cast<const V*>(pd.dl[o]) + ev.w_off;
    }
    if constexpr (PREC == SAT_PREC_F64) {
      cx.dx = g.dx;
      cx.dy = g.dy;
    } else {
      cx.rmask = (typename Pol::MaskLds)s_mask;
      cx.mw = mw;
      cx.dl_bytes = (int)(g.wplane * sizeof(V));
    }
    IterOut io;
    const V* ta = iterate<Pol, WaveLayout, C, W, (W > 1) ? mw_adj_depth<PREC, C, W>() : adj_depth1<C>()>(uniform(cx), Q0, Q1, Q2, pd.adj_maxiter, kAdjTol,
                                                                           io, exchange_ptr<Pol, W>(xch), red);
    if (lane == 0 && !pd.kden_pass) {
      pd.iters_a[e] = io.iters;
      if (!io.converged) atomicOr(pd.flags + e, FLAG_ADJ_MAXITER);
    }
    // :514-536 mask, column-major (the density field is not masked: the
    // reference masks adj_field a second time instead, src/preproc.cpp:136)
    const bool masked = !pd.kden_pass;
    for (int iy = 0; iy < ny; iy++)
      for (int ix = lane; ix < nx; ix += nthr) {
        const bool m = masked && ix >= ev.mb[0] && ix <= ev.mb[1] && iy >= ev.mb[2] && iy <= ev.mb[3];
        const V v = m ? V(0) : ta[wave_index(g, 0, ix, iy)];
        const size_t o = ev.pl_off + ix + (size_t)iy * nx;
        if (PREC == SAT_PREC_F64 && pd.ta_f32) static_cast<float*>(pd.ta_out)[o] = (float)v;
        else static_cast<V*>(pd.ta_out)[o] = v;
      }
    __syncthreads();
  }
}

}  // namespace sat
