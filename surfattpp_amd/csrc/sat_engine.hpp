// sat_engine.hpp -- the lane-skewed Gauss-Seidel sweep shared by every solve.
//
// One 64-lane wave sweeps one grid.  Lane l owns columns [l*C, l*C+C) of
// every row.  In a sweep with directions (XD, YD) lane l handles sweep-row r
// at step t = r + lag(l), lag = l for XD > 0 and La-1-l for XD < 0, and visits
// its C columns in XD order.
//
// A node reads its 4 axis neighbours.  The neighbours that precede it in the
// reference's x-outer / y-inner order (src/eikonal_solver.cpp:149-154,
// :474-478) are exactly those already updated here: the previous column in
// the lane (this step), the previous row (this lane, last step) and, for the
// first column, the adjacent lane's last column (one step earlier; it arrives
// by a cross-lane shuffle of that lane's previous row).  The following
// neighbours still hold the previous sweep's values: the next column, the
// next row (prefetched) and, for the last column, the other adjacent lane's
// first column (shuffled from that lane's prefetched row).  Each sweep reads
// `tin` and writes `tout`, so every update sees the same values as the
// sequential loop and the result is identical iterate for iterate.
//
// Register pipeline.  A row is split into a head (swept value u and the
// neighbour auxiliary aux) and a body (the node's own coefficients).  At step
// t the lane computes row r from cur (head + body of row r) and the head of
// row r+1, while the body of row r+1 and the head of row r+2 are in flight.
//
// Storage layouts (Lay):
//   SegLayout   rows of pitch P = La*C, lane l's part at j*P + l*C, in every
//               sweep (per-solve kernels: one plane serves all orientations).
//   WaveLayout  wavefront-major (sat_device.hpp), one copy per orientation of
//               every streamed plane: a step's loads are one contiguous run.
//               The swept planes change orientation where the next sweep's
//               does: sweep 1 (+x,+y) writes orientation 1 for sweep 2
//               (+x,-y), sweep 3 (-x,+y) writes orientation 0 for sweep 4
//               (-x,-y), whose output starts the next round in orientation 0.
//
// Waves per grid (W).  W = 1: one wave sweeps the grid and the adjacent
// lanes' values cross by ds_bpermute.  W > 1: a work-group of W waves sweeps
// one grid, lane l of 64W (C = ceil(nx / 64W) columns each); the values cross
// through LDS, double-buffered by step parity with one barrier per step (a
// lane reads slot t&1 at step t; it is written again at step t+2, after the
// barrier of step t+1).  Fewer columns per lane shorten a grid's latency.
#pragma once
#include <type_traits>

#include "sat_device.hpp"

namespace sat {

// Every swept plane has kTrashElems elements past its end: under SegLayout the
// stores of lanes outside the grid at a step go there, so that every load and
// store of a step is issued unconditionally (conditional memory operations
// make the compiler's wait counts conservative, which would wait for the
// prefetches issued in the same step).  WaveLayout drops those stores instead
// (store_row).
constexpr int kTrashElems = kWave * kMaxC;

// The geometry a sweep needs, read from the grid descriptor once before the
// step loop (inside the loop the compiler could not keep it in registers: the
// stores to the swept plane might alias the descriptor, and every reload would
// wait for all loads in flight).
struct SweepGeom {
  int nx, ny, La, P, trash0;
};

// kStepRun: every lane's row of a step lies in one contiguous run of the
// plane, and a row outside the grid (a lane before or after its part of the
// sweep) addresses the plane's padding inside that run.  Such loads are issued
// unclamped (for the inactive lanes at the address of the last active lane):
// they touch only the cache lines the active lanes read anyway.  A clamped row
// instead sends each idle lane to its own distant line, every step.
struct SegLayout {
  static constexpr bool kStepRun = false;
  static constexpr bool kLds = false;
  __device__ static __forceinline__ SweepGeom geom(const GridDev& g) {
    return SweepGeom{g.nx, g.ny, g.La, g.P, (int)g.seg_plane};
  }
  template <int C, int O>
  __host__ __device__ static __forceinline__ int base(const SweepGeom& s, int j, int lane) {
    return j * s.P + lane * C;
  }
  template <int C>
  __host__ __device__ static __forceinline__ int trash(const SweepGeom& s, int lane) { return s.trash0 + lane * C; }
};
// SegLayout planes held in LDS (the per-pair adjoint of grids that fit, sat_fused.hpp):
// the same addressing; the swept planes' stores go to LDS.
struct LdsSegLayout : SegLayout {
  static constexpr bool kLds = true;
};
struct WaveLayout {
  static constexpr bool kStepRun = true;
  static constexpr bool kLds = false;
  __device__ static __forceinline__ SweepGeom geom(const GridDev& g) {
    return SweepGeom{g.nx, g.ny, g.La, g.P, (int)g.wplane};
  }
  template <int C, int O>
  __host__ __device__ static __forceinline__ int base(const SweepGeom& s, int j, int lane) {
    const int q = (O == 0) ? j + lane : j - lane + s.La - 1;
    return (q * s.La + lane) * C;
  }
  template <int C>
  __host__ __device__ static __forceinline__ int trash(const SweepGeom& s, int lane) { return s.trash0 + lane * C; }
};

// Addressing of a sweep with directions (XD, YD), shared by the sweep loops
// below and by the host-side check of every address they issue
// (tests/cpp/layout_check.cpp).  Lane l handles sweep-row r at step r + lag(l).
template <int XD, int YD>
struct SweepAddr {
  int lane, La, ny, lag, nsteps, lane_c, lag_c;
  __host__ __device__ __forceinline__ SweepAddr(const SweepGeom& g, int lane_) {
    lane = lane_;
    La = g.La;
    ny = g.ny;
    lag = (XD > 0) ? lane : (La - 1 - lane);
    nsteps = ny + La - 1;
    // loads are issued for every lane and step; rows and lanes outside the grid
    // read an in-bounds location whose values are not used (see kStepRun)
    lane_c = imin(lane, La - 1);
    lag_c = (XD > 0) ? lane_c : (La - 1 - lane_c);
  }
  __host__ __device__ __forceinline__ int rowj(int r) const { return (YD > 0) ? r : (ny - 1 - r); }
  __host__ __device__ __forceinline__ bool in_range(int r) const { return lane < La && r >= 0 && r < ny; }
  // the clamped grid row of sweep-row r (row tables, neighbour presence)
  __host__ __device__ __forceinline__ int row_clamped(int r) const { return rowj(imin(imax(r, 0), ny - 1)); }
  // element offset of the C values the lane loads for sweep-row r
  template <class Lay, int C>
  __host__ __device__ __forceinline__ int load_base(const SweepGeom& g, int r) const {
    constexpr int O = (XD * YD > 0) ? 0 : 1;
    if constexpr (Lay::kStepRun) {
      const int ta = imin(r + lag, nsteps - 1);  // the step this row belongs to
      return Lay::template base<C, O>(g, rowj(ta - lag_c), lane_c);
    } else {
      return Lay::template base<C, O>(g, row_clamped(r), lane_c);
    }
  }
};

// Policy P provides:
//   using V;  using A;  static constexpr bool kAux;
//   template <int C> struct Head { V u[C]; A aux[C]; };
//   template <int C> struct Body { ...; V st[C]; };      (st: iterate at the round start)
//   template <int C, int O> static void load_head(const Ctx&, Head<C>&, int base, int j, const V* tin)
//   template <int C, int O, bool NORM> static void load_body(const Ctx&, Body<C>&, int base, int j, const V* tst)
//                                    (st is read from tst only in the sweep that forms the norm)
//   static bool fixed(const Ctx&, int ix, int j)
//   template <int C> static V update(const Ctx&, const Head<C>&, const Body<C>&, int k, int j,
//                                    V uW, A aW, bool hW, V uE, A aE, bool hE,
//                                    V uS, A aS, bool hS, V uN, A aN, bool hN)
//   template <int C> static double norm(const Head<C>&, const Body<C>&, int k, V newv)
// O is the orientation of the sweep (0: XD*YD > 0, 1: XD*YD < 0).
// Work-group barrier that orders LDS only: the published values cross in LDS,
// and a full __syncthreads() would also wait for every global load in flight
// (the prefetched rows), stalling each step for a memory round trip.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Per-lane values published to the neighbouring lanes each step (W > 1): the
// previous-row edge value (pu, pa) and the next-row edge value (nu, na), one
// array per value and step parity.  Each value goes out in its own LDS store:
// a combined 128-bit store would need a computed value and a just-loaded one in
// adjacent registers, and the copy that puts them there waits for the load.
template <class Pol, int W> struct Exchange {
  typename Pol::V pu[2][kWave * W], nu[2][kWave * W];
  typename Pol::A pa[2][kWave * W], na[2][kWave * W];
};
// The exchange a kernel declares in LDS (none for single-wave sweeps, which
// cross lanes by ds_bpermute) and the pointer the sweeps take.
template <class Pol, int W>
using ExchangeLds = std::conditional_t<(W > 1), Exchange<Pol, W>, char>;
template <class Pol, int W>
__device__ __forceinline__ Exchange<Pol, W>* exchange_ptr(ExchangeLds<Pol, W>& x) {
  if constexpr (W > 1) return &x;
  else return nullptr;
}

// A lane's C new values of one step.  Lanes with no live row (outside their
// part of the sweep, or beyond the active lanes) store nothing: under
// WaveLayout their raw buffer stores carry an offset past the plane and are
// dropped; under SegLayout they go to the plane's trash row.  Either way the
// store instruction is unconditional (a branch around it would make the
// compiler wait for the prefetches issued in the same step).
// store_offset: the first store's byte offset in the plane's raw buffer
// (WaveLayout; kDropOffset, past the plane's num_records, for a lane with no
// live row) or element offset (SegLayout; the trash row for no live row); the
// lane's C values go to consecutive elements from there.
template <class Lay, int C, int OOUT, class V>
__host__ __device__ __forceinline__ int store_offset(const SweepGeom& g, bool live, int j, int lane) {
  if constexpr (Lay::kStepRun) return live ? Lay::template base<C, OOUT>(g, j, lane) * (int)sizeof(V) : kDropOffset;
  else return live ? Lay::template base<C, OOUT>(g, j, lane) : Lay::template trash<C>(g, lane);
}
template <class Lay, int C, int OOUT, class V>
__device__ __forceinline__ void store_row(const SweepGeom& g, V* __restrict__ tout, bool live, int j, int lane,
                                          const V (&newv)[C]) {
  const int off = store_offset<Lay, C, OOUT, V>(g, live, j, lane);
  if constexpr (Lay::kLds) {
    auto* p = (__attribute__((address_space(3))) V*)tout;
#pragma unroll
    for (int k = 0; k < C; k++) p[off + k] = newv[k];
  } else if constexpr (Lay::kStepRun) {
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(tout, g.trash0 * (int)sizeof(V));
#pragma unroll
    for (int k = 0; k < C; k++) bst(rs, off + k * (int)sizeof(V), newv[k]);
  } else {
#pragma unroll
    for (int k = 0; k < C; k++) gst(tout + off + k, newv[k]);
  }
}

template <class Pol, class Lay, int C, int XD, int YD, bool NORM, int OOUT, int W = 1>
__device__ __forceinline__ void sweep(const typename Pol::Ctx& cx, const typename Pol::V* __restrict__ tin,
                                      typename Pol::V* __restrict__ tout,
                                      const typename Pol::V* __restrict__ tst, double& lmax, double& lsum,
                                      Exchange<Pol, W>* __restrict__ xch = nullptr) {
  using V = typename Pol::V;
  using A = typename Pol::A;
  using Head = typename Pol::template Head<C>;
  using Body = typename Pol::template Body<C>;
  constexpr int O = (XD * YD > 0) ? 0 : 1;
  const SweepGeom g = uniform(Lay::geom(*cx.g));
  const int lane = threadIdx.x;
  const int nx = g.nx;
  const SweepAddr<XD, YD> sa(g, lane);
  const int lag = sa.lag, nsteps = sa.nsteps;
  auto load_head = [&](Head& H, int r) {
    Pol::template load_head<C, O>(cx, H, sa.template load_base<Lay, C>(g, r), sa.row_clamped(r), tin);
  };
  auto load_body = [&](Body& B, int r) {
    Pol::template load_body<C, O, NORM>(cx, B, sa.template load_base<Lay, C>(g, r), sa.row_clamped(r), tst);
  };

  Head cur_h, nxt_h, nn_h;
  Body cur_b, nxt_b;
  V prev_u[C];
  A prev_a[C];
#pragma unroll
  for (int k = 0; k < C; k++) { prev_u[k] = V(0); prev_a[k] = A{}; }
  load_head(cur_h, -lag);
  load_body(cur_b, -lag);
  load_head(nxt_h, 1 - lag);

  for (int t = 0; t < nsteps; ++t) {
    const int r = t - lag;
    load_body(nxt_b, r + 1);
    load_head(nn_h, r + 2);
    V nw_u, od_u;
    A nw_a{}, od_a{};
    if constexpr (W == 1) {
      nw_u = shfl_any((XD > 0) ? prev_u[C - 1] : prev_u[0], lane - XD);
      od_u = shfl_any((XD > 0) ? nxt_h.u[0] : nxt_h.u[C - 1], lane + XD);
      if constexpr (Pol::kAux) {
        nw_a = shfl_any((XD > 0) ? prev_a[C - 1] : prev_a[0], lane - XD);
        od_a = shfl_any((XD > 0) ? nxt_h.aux[0] : nxt_h.aux[C - 1], lane + XD);
      }
    } else {
      const int par = t & 1;
      xch->pu[par][lane] = (XD > 0) ? prev_u[C - 1] : prev_u[0];
      xch->nu[par][lane] = (XD > 0) ? nxt_h.u[0] : nxt_h.u[C - 1];
      if constexpr (Pol::kAux) {
        xch->pa[par][lane] = (XD > 0) ? prev_a[C - 1] : prev_a[0];
        xch->na[par][lane] = (XD > 0) ? nxt_h.aux[0] : nxt_h.aux[C - 1];
      }
      lds_barrier();
      // lanes at the ends read a clamped slot; those values are never used
      // (the node is on the grid edge or outside the active lanes)
      const int src_w = imin(imax(lane - XD, 0), kWave * W - 1);
      const int src_o = imin(imax(lane + XD, 0), kWave * W - 1);
      nw_u = xch->pu[par][src_w];
      od_u = xch->nu[par][src_o];
      if constexpr (Pol::kAux) {
        nw_a = xch->pa[par][src_w];
        od_a = xch->na[par][src_o];
      }
    }
    {
      // every lane evaluates and stores at every step: rows outside the grid
      // compute on clamped inputs and store to the plane's trash row
      const bool live = sa.in_range(r);
      const int j = sa.row_clamped(r);
      const bool hS = j > 0, hN = j < g.ny - 1;
      V newv[C];
#pragma unroll
      for (int kk = 0; kk < C; kk++) {
        const int k = (XD > 0) ? kk : (C - 1 - kk);
        const int ix = lane * C + k;
        V uW, uE;
        A aW{}, aE{};
        if (XD > 0) {
          if (k > 0) { uW = newv[k - 1]; aW = cur_h.aux[k - 1]; } else { uW = nw_u; aW = nw_a; }
          if (k < C - 1) { uE = cur_h.u[k + 1]; aE = cur_h.aux[k + 1]; } else { uE = od_u; aE = od_a; }
        } else {
          if (k < C - 1) { uE = newv[k + 1]; aE = cur_h.aux[k + 1]; } else { uE = nw_u; aE = nw_a; }
          if (k > 0) { uW = cur_h.u[k - 1]; aW = cur_h.aux[k - 1]; } else { uW = od_u; aW = od_a; }
        }
        V uS, uN;
        A aS, aN;
        if (YD > 0) { uS = prev_u[k]; aS = prev_a[k]; uN = nxt_h.u[k]; aN = nxt_h.aux[k]; }
        else { uN = prev_u[k]; aN = prev_a[k]; uS = nxt_h.u[k]; aS = nxt_h.aux[k]; }
        // padding columns and fixed nodes keep their value
        const bool valid = live && ix < nx;
        const V upd = Pol::template update<C>(cx, cur_h, cur_b, k, j, uW, aW, ix > 0, uE, aE, ix < nx - 1, uS, aS,
                                              hS, uN, aN, hN);
        newv[k] = (!valid || Pol::fixed(cx, ix, j)) ? cur_h.u[k] : upd;
        if (NORM) {
          const double d = valid ? Pol::template norm<C>(cur_h, cur_b, k, newv[k]) : 0.0;
          lsum += d;
          lmax = dmax(lmax, d);
        }
      }
      store_row<Lay, C, OOUT>(g, tout, live, j, lane, newv);
#pragma unroll
      for (int k = 0; k < C; k++) { prev_u[k] = newv[k]; prev_a[k] = cur_h.aux[k]; }
    }
    cur_h = nxt_h;
    cur_b = nxt_b;
    nxt_h = nn_h;
  }
}

// The same sweep with a ring of D+1 rows in registers (D >= 2): at step t the
// lane computes row r from ring slot t % (D+1), reads row r+1 from the next
// slot, and issues the loads of row r + D into the slot row r-1 left, so each
// row's loads have D-1 to D steps to arrive.  The step loop is unrolled by
// D+1 (slot indices are compile-time); the steps past the end are no-ops for
// every lane (r >= ny), which keeps the per-step barrier uniform.  Used where
// few columns per lane leave registers for the ring (W > 1).
template <class Pol, class Lay, int C, int XD, int YD, bool NORM, int OOUT, int W, int D>
__device__ __forceinline__ void sweep_ring(const typename Pol::Ctx& cx, const typename Pol::V* __restrict__ tin,
                                           typename Pol::V* __restrict__ tout,
                                           const typename Pol::V* __restrict__ tst, double& lmax, double& lsum,
                                           Exchange<Pol, W>* __restrict__ xch) {
  static_assert(D >= 2, "the ring needs two rows in flight");
  using V = typename Pol::V;
  using A = typename Pol::A;
  using Head = typename Pol::template Head<C>;
  using Body = typename Pol::template Body<C>;
  constexpr int O = (XD * YD > 0) ? 0 : 1;
  constexpr int R = D + 1;
  const SweepGeom g = uniform(Lay::geom(*cx.g));
  const int lane = threadIdx.x;
  const int nx = g.nx;
  const SweepAddr<XD, YD> sa(g, lane);
  const int lag = sa.lag, nsteps = sa.nsteps;
  auto load = [&](Head& H, Body& B, int r) {  // unconditional, in bounds (see kTrashElems, kStepRun)
    const int j = sa.row_clamped(r);
    const int base = sa.template load_base<Lay, C>(g, r);
    Pol::template load_head<C, O>(cx, H, base, j, tin);
    Pol::template load_body<C, O, NORM>(cx, B, base, j, tst);
  };

  Head H[R];
  Body B[R];
  V prev_u[C];
  A prev_a[C];
#pragma unroll
  for (int k = 0; k < C; k++) { prev_u[k] = V(0); prev_a[k] = A{}; }
#pragma unroll
  for (int s = 0; s < D; s++) load(H[s], B[s], s - lag);
  // The ring's first rows are in registers before the loop starts.  Without this
  // wait the compiler merges, at the loop head, the loads still in flight on
  // entry with those of the back edge and waits for nearly every load issued
  // so far at each loop head, which takes the prefetch depth back to one row;
  // with it the waits in the loop keep D - 1 rows in flight.
  wait_vmem();

  for (int t0 = 0; t0 < nsteps; t0 += R) {
#pragma unroll
    for (int s = 0; s < R; s++) {
      const int t = t0 + s;
      const int r = t - lag;
      Head& cur_h = H[s];
      Body& cur_b = B[s];
      Head& nxt_h = H[(s + 1) % R];
      load(H[(s + D) % R], B[(s + D) % R], r + D);
      V nw_u, od_u;
      A nw_a{}, od_a{};
      if constexpr (W == 1) {
        nw_u = shfl_any((XD > 0) ? prev_u[C - 1] : prev_u[0], lane - XD);
        od_u = shfl_any((XD > 0) ? nxt_h.u[0] : nxt_h.u[C - 1], lane + XD);
        if constexpr (Pol::kAux) {
          nw_a = shfl_any((XD > 0) ? prev_a[C - 1] : prev_a[0], lane - XD);
          od_a = shfl_any((XD > 0) ? nxt_h.aux[0] : nxt_h.aux[C - 1], lane + XD);
        }
      } else {
        const int par = t & 1;
        xch->pu[par][lane] = (XD > 0) ? prev_u[C - 1] : prev_u[0];
        xch->nu[par][lane] = (XD > 0) ? nxt_h.u[0] : nxt_h.u[C - 1];
        if constexpr (Pol::kAux) {
          xch->pa[par][lane] = (XD > 0) ? prev_a[C - 1] : prev_a[0];
          xch->na[par][lane] = (XD > 0) ? nxt_h.aux[0] : nxt_h.aux[C - 1];
        }
        lds_barrier();
        const int src_w = imin(imax(lane - XD, 0), kWave * W - 1);
        const int src_o = imin(imax(lane + XD, 0), kWave * W - 1);
        nw_u = xch->pu[par][src_w];
        od_u = xch->nu[par][src_o];
        if constexpr (Pol::kAux) {
          nw_a = xch->pa[par][src_w];
          od_a = xch->na[par][src_o];
        }
      }
      {
        const bool live = sa.in_range(r);
        const int j = sa.row_clamped(r);
        const bool hS = j > 0, hN = j < g.ny - 1;
        V newv[C];
#pragma unroll
        for (int kk = 0; kk < C; kk++) {
          const int k = (XD > 0) ? kk : (C - 1 - kk);
          const int ix = lane * C + k;
          V uW, uE;
          A aW{}, aE{};
          if (XD > 0) {
            if (k > 0) { uW = newv[k - 1]; aW = cur_h.aux[k - 1]; } else { uW = nw_u; aW = nw_a; }
            if (k < C - 1) { uE = cur_h.u[k + 1]; aE = cur_h.aux[k + 1]; } else { uE = od_u; aE = od_a; }
          } else {
            if (k < C - 1) { uE = newv[k + 1]; aE = cur_h.aux[k + 1]; } else { uE = nw_u; aE = nw_a; }
            if (k > 0) { uW = cur_h.u[k - 1]; aW = cur_h.aux[k - 1]; } else { uW = od_u; aW = od_a; }
          }
          V uS, uN;
          A aS, aN;
          if (YD > 0) { uS = prev_u[k]; aS = prev_a[k]; uN = nxt_h.u[k]; aN = nxt_h.aux[k]; }
          else { uN = prev_u[k]; aN = prev_a[k]; uS = nxt_h.u[k]; aS = nxt_h.aux[k]; }
          const bool valid = live && ix < nx;
          const V upd = Pol::template update<C>(cx, cur_h, cur_b, k, j, uW, aW, ix > 0, uE, aE, ix < nx - 1, uS,
                                                aS, hS, uN, aN, hN);
          newv[k] = (!valid || Pol::fixed(cx, ix, j)) ? cur_h.u[k] : upd;
          if (NORM) {
            const double d = valid ? Pol::template norm<C>(cur_h, cur_b, k, newv[k]) : 0.0;
            lsum += d;
            lmax = dmax(lmax, d);
          }
        }
        store_row<Lay, C, OOUT>(g, tout, live, j, lane, newv);
#pragma unroll
        for (int k = 0; k < C; k++) { prev_u[k] = newv[k]; prev_a[k] = cur_h.aux[k]; }
      }
    }
  }
}

// One sweep: the head/body pipeline (D = 1) or the register ring (D >= 2).
template <class Pol, class Lay, int C, int XD, int YD, bool NORM, int OOUT, int W, int D>
__device__ __forceinline__ void sweep_any(const typename Pol::Ctx& cx, const typename Pol::V* __restrict__ tin,
                                          typename Pol::V* __restrict__ tout,
                                          const typename Pol::V* __restrict__ tst, double& lmax, double& lsum,
                                          Exchange<Pol, W>* __restrict__ xch) {
  if constexpr (D == 1) sweep<Pol, Lay, C, XD, YD, NORM, OOUT, W>(cx, tin, tout, tst, lmax, lsum, xch);
  else sweep_ring<Pol, Lay, C, XD, YD, NORM, OOUT, W, D>(cx, tin, tout, tst, lmax, lsum, xch);
}

struct IterOut {
  int iters;
  bool converged;
  double l1, linf;
  double linf_prev;  // Linf of the round before the last (-1 if none)
};

// Iterate 4-sweep rounds until L1 < tol and Linf < tol (src/eikonal_solver.cpp:146-314,
// :471-509).  Three rotating planes: B0 holds the iterate at the start of a round
// (orientation 0 under WaveLayout).  Returns the plane holding the result
// (orientation 0).
// Work-group reductions in a fixed order (W > 1: per wave, then across waves).
template <int W>
__device__ inline double block_max(double v, double* red) {
  v = wave_max(v);
  if constexpr (W == 1) {
    return v;
  } else {
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < W; w++) r = dmax(r, red[w]);
    return r;
  }
}
template <int W>
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  if constexpr (W == 1) {
    return v;
  } else {
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < W; w++) r += red[w];
    return r;
  }
}

template <class Pol, class Lay, int C, int W = 1, int D = 1>
__device__ typename Pol::V* iterate(const typename Pol::Ctx& cx, typename Pol::V* b0, typename Pol::V* b1,
                                    typename Pol::V* b2, int maxiter, double tol, IterOut& io,
                                    Exchange<Pol, W>* xch = nullptr, double* red = nullptr) {
  using V = typename Pol::V;
  V* B0 = b0;
  V* B1 = b1;
  V* B2 = b2;
  V* res = B0;
  io.iters = maxiter;
  io.converged = false;
  io.linf = -1.0;
  io.linf_prev = -1.0;
  const double npts = (double)(cx.g->nx * cx.g->ny);
  for (int it = 0; it < maxiter; ++it) {
    double lmax = 0.0, lsum = 0.0;
    sweep_any<Pol, Lay, C, 1, 1, false, 1, W, D>(cx, B0, B1, nullptr, lmax, lsum, xch);
    __syncthreads();
    sweep_any<Pol, Lay, C, 1, -1, false, 1, W, D>(cx, B1, B2, nullptr, lmax, lsum, xch);
    __syncthreads();
    sweep_any<Pol, Lay, C, -1, 1, false, 0, W, D>(cx, B2, B1, nullptr, lmax, lsum, xch);
    __syncthreads();
    sweep_any<Pol, Lay, C, -1, -1, true, 0, W, D>(cx, B1, B2, B0, lmax, lsum, xch);
    __syncthreads();
    const double Linf = block_max<W>(lmax, red);
    const double L1 = block_sum<W>(lsum, red) / npts;
    io.l1 = L1;
    io.linf_prev = io.linf;
    io.linf = Linf;
    res = B2;
    if (L1 < tol && Linf < tol) {
      io.iters = it + 1;
      io.converged = true;
      return res;
    }
    V* n0 = B2;
    V* n1 = B0;
    V* n2 = B1;
    B0 = n0; B1 = n1; B2 = n2;
  }
  return res;
}

}  // namespace sat
