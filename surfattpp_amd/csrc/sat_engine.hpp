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
#pragma once
#include "sat_device.hpp"

namespace sat {

struct SegLayout {
  template <int C, int O>
  __device__ static __forceinline__ int base(const GridDev& g, int j, int lane) { return j * g.P + lane * C; }
};
struct WaveLayout {
  template <int C, int O>
  __device__ static __forceinline__ int base(const GridDev& g, int j, int lane) {
    return wave_base<C, O>(g, j, lane);
  }
};

// Policy P provides:
//   using V;  using A;  static constexpr bool kAux;
//   template <int C> struct Head { V u[C]; A aux[C]; };
//   template <int C> struct Body { ...; V st[C]; };      (st: iterate at the round start)
//   template <int C, int O> static void load_head(const Ctx&, Head<C>&, int base, int j, const V* tin)
//   template <int C, int O> static void load_body(const Ctx&, Body<C>&, int base, int j, const V* tst)
//   static bool fixed(const Ctx&, int ix, int j)
//   template <int C> static V update(const Ctx&, const Head<C>&, const Body<C>&, int k, int j,
//                                    V uW, A aW, bool hW, V uE, A aE, bool hE,
//                                    V uS, A aS, bool hS, V uN, A aN, bool hN)
//   template <int C> static double norm(const Head<C>&, const Body<C>&, int k, V newv)
// O is the orientation of the sweep (0: XD*YD > 0, 1: XD*YD < 0).
template <class Pol, class Lay, int C, int XD, int YD, bool NORM, int OOUT>
__device__ __forceinline__ void sweep(const typename Pol::Ctx& cx, const typename Pol::V* __restrict__ tin,
                                      typename Pol::V* __restrict__ tout,
                                      const typename Pol::V* __restrict__ tst, double& lmax, double& lsum) {
  using V = typename Pol::V;
  using A = typename Pol::A;
  using Head = typename Pol::template Head<C>;
  using Body = typename Pol::template Body<C>;
  constexpr int O = (XD * YD > 0) ? 0 : 1;
  const GridDev& g = *cx.g;
  const int lane = threadIdx.x;
  const int La = g.La, nx = g.nx, ny = g.ny;
  const bool lane_ok = lane < La;
  const int lag = (XD > 0) ? lane : (La - 1 - lane);
  const int nsteps = ny + La - 1;
  auto rowj = [&](int r) { return (YD > 0) ? r : (ny - 1 - r); };
  auto in_range = [&](int r) { return lane_ok && r >= 0 && r < ny; };
  auto load_head = [&](Head& H, int r) {
    if (in_range(r)) {
      const int j = rowj(r);
      Pol::template load_head<C, O>(cx, H, Lay::template base<C, O>(g, j, lane), j, tin);
    }
  };
  auto load_body = [&](Body& B, int r) {
    if (in_range(r)) {
      const int j = rowj(r);
      Pol::template load_body<C, O>(cx, B, Lay::template base<C, O>(g, j, lane), j, NORM ? tst : nullptr);
    }
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
    const V nw_u = shfl_any((XD > 0) ? prev_u[C - 1] : prev_u[0], lane - XD);
    const V od_u = shfl_any((XD > 0) ? nxt_h.u[0] : nxt_h.u[C - 1], lane + XD);
    A nw_a{}, od_a{};
    if constexpr (Pol::kAux) {
      nw_a = shfl_any((XD > 0) ? prev_a[C - 1] : prev_a[0], lane - XD);
      od_a = shfl_any((XD > 0) ? nxt_h.aux[0] : nxt_h.aux[C - 1], lane + XD);
    }
    if (in_range(r)) {
      const int j = rowj(r);
      const bool hS = j > 0, hN = j < ny - 1;
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
        // every lane evaluates the update; padding columns and fixed nodes keep their value
        const bool valid = ix < nx;
        const V upd = Pol::template update<C>(cx, cur_h, cur_b, k, j, uW, aW, ix > 0, uE, aE, ix < nx - 1, uS, aS,
                                              hS, uN, aN, hN);
        newv[k] = (!valid || Pol::fixed(cx, ix, j)) ? cur_h.u[k] : upd;
        if (NORM) {
          const double d = valid ? Pol::template norm<C>(cur_h, cur_b, k, newv[k]) : 0.0;
          lsum += d;
          lmax = dmax(lmax, d);
        }
      }
      const int base = Lay::template base<C, OOUT>(g, j, lane);
#pragma unroll
      for (int k = 0; k < C; k++) tout[base + k] = newv[k];
#pragma unroll
      for (int k = 0; k < C; k++) { prev_u[k] = newv[k]; prev_a[k] = cur_h.aux[k]; }
    }
    cur_h = nxt_h;
    cur_b = nxt_b;
    nxt_h = nn_h;
  }
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
template <class Pol, class Lay, int C>
__device__ typename Pol::V* iterate(const typename Pol::Ctx& cx, typename Pol::V* b0, typename Pol::V* b1,
                                    typename Pol::V* b2, int maxiter, double tol, IterOut& io) {
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
    sweep<Pol, Lay, C, 1, 1, false, 1>(cx, B0, B1, nullptr, lmax, lsum);
    __syncthreads();
    sweep<Pol, Lay, C, 1, -1, false, 1>(cx, B1, B2, nullptr, lmax, lsum);
    __syncthreads();
    sweep<Pol, Lay, C, -1, 1, false, 0>(cx, B2, B1, nullptr, lmax, lsum);
    __syncthreads();
    sweep<Pol, Lay, C, -1, -1, true, 0>(cx, B1, B2, B0, lmax, lsum);
    __syncthreads();
    const double Linf = wave_max(lmax);
    const double L1 = wave_sum(lsum) / npts;
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
