// sat_policies.hpp -- node updates and per-solve set-up for the two precisions.
//
//   FwdF64 / AdjF64   the reference's FP64 arithmetic in its operation order
//                     (src/eikonal_solver.cpp); kernels within 1e-12 of the oracle.
//   FwdF32 / AdjF32   FP32 sweeps.  The forward node update runs in FP32 on the
//                     factored perturbation sigma = tau - 1: each quadratic is solved
//                     for the increment w about the smaller upwind neighbour, which
//                     removes the O((dist/dx)^2) cancellation of the reference's
//                     coefficients (a*px^2*tau^2 ~ 1e3 against f^2 ~ 0.1 at 1 km
//                     cells) that FP32 cannot carry.  sigma itself is kept in FP64
//                     (sigma_new = sigma_ref + w), so the per-update rounding is that
//                     of w (~1e-3) rather than of sigma (~0.1): FP32 storage let it
//                     accumulate along rays to ~1e-5 in the kernels.  The analytic
//                     factor T0, the travel time T = T0 + sigma*T0, the adjoint
//                     coefficients, the convergence norm and the kernel
//                     accumulation are FP64.
#pragma once
#include "sat_crmath.hpp"
#include "sat_device.hpp"
#include "sat_engine.hpp"

namespace sat {

struct FixedBox {  // near-source nodes held at tau = 1 (src/eikonal_solver.cpp:119-129)
  int x0, x1, y0, y1;
};

// Per-solve workspace: segmented planes in sweep precision, plain FP64 planes.
struct Slot {
  void* seg[9];
  double* pl[4];
};

// ============================================================================
// FP64, reference operation order.  Plane pointers are per sweep orientation
// (the per-solve kernels set both to the same segmented plane).
// ============================================================================
struct FwdF64 {
  using V = double;
  using A = double;
  static constexpr bool kAux = true;
  struct Ctx {
    const GridDev* g;
    const double *t0[2], *t0x[2], *t0y[2], *a[2], *b[2], *c[2], *f[2];
    FixedBox fb;
    double dx, dy;
  };
  template <int C> struct Head { double u[C], aux[C]; };
  template <int C> struct Body { double t0x[C], t0y[C], a[C], b[C], c[C], f[C], st[C]; };
  template <int C, int O>
  __device__ static void load_head(const Ctx& cx, Head<C>& H, int base, int, const double* tin) {
#pragma unroll
    for (int k = 0; k < C; k++) { H.u[k] = gld(tin, base + k); H.aux[k] = gld(cx.t0[O], base + k); }
  }
  template <int C, int O, bool NORM>
  __device__ static void load_body(const Ctx& cx, Body<C>& B, int base, int, const double* tst) {
#pragma unroll
    for (int k = 0; k < C; k++) {
      B.t0x[k] = gld(cx.t0x[O], base + k);
      B.t0y[k] = gld(cx.t0y[O], base + k);
      B.a[k] = gld(cx.a[O], base + k);
      B.b[k] = gld(cx.b[O], base + k);
      B.c[k] = gld(cx.c[O], base + k);
      B.f[k] = gld(cx.f[O], base + k);
      B.st[k] = NORM ? gld(tst, base + k) : 0.0;
    }
  }
  __device__ static bool fixed(const Ctx& cx, int ix, int j) {
    return j >= cx.fb.y0 && j <= cx.fb.y1 && ix >= cx.fb.x0 && ix <= cx.fb.x1;
  }
  template <int C>
  __device__ static double update(const Ctx& cx, const Head<C>& h, const Body<C>& r, int k, int, double uW,
                                  double aW, bool hW, double uE, double aE, bool hE, double uS, double aS, bool hS,
                                  double uN, double aN, bool hN) {
    return fwd_update_f64(h.u[k], h.aux[k], r.t0x[k], r.t0y[k], r.a[k], r.b[k], r.c[k], r.f[k], cx.dx, cx.dy, hW,
                          uW, aW, hE, uE, aE, hS, uS, aS, hN, uN, aN);
  }
  template <int C>
  __device__ static double norm(const Head<C>& h, const Body<C>& r, int k, double v) {
    return fabs(v - r.st[k]) * h.aux[k];  // :295
  }
};

// ============================================================================
// FP64 re-solves of FP32 batch pairs.  A re-solve must reproduce the reference's
// stop decision, which needs FP64 iterates, not the reference's exact bits: its
// iterates differ from the reference's by rounding (~1e-16 relative), which moves
// a convergence norm of ~1e-4 by ~1e-12 relative.  One work-group sweeps one pair
// (a single wave per SIMD), so each step's latency is the chain of dependent
// FP64 operations of one node; the IEEE divide and square-root sequences hold
// VCC and run one after another.  This update shares one reciprocal per
// denominator and uses the hardware estimates refined in FMA, so the
// independent candidates interleave.
// ============================================================================

// 1/x: hardware estimate and two Newton-Raphson steps (about 1 ulp; x normal and
// finite -- otherwise the result is not finite and the candidate is rejected).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
// sqrt(x) for x > 0: reciprocal square-root estimate, one Goldschmidt step and
// two residual corrections (about 1 ulp).
__device__ __forceinline__ double sqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

// min(old, smallest accepted candidate): a balanced tree of minima over the 16
// candidates (rejected ones as +inf), then against the old value.  Accepted
// candidates are never NaN (every acceptance test compares them), so this is
// the value of the reference's scan (:281-283) in 5 dependent steps instead of
// 32; the tree's minima are single v_min_f64 (no NaN reaches them), the last
// step keeps std::min's rule (a NaN old value stays, as in the scan).
template <int N>
__device__ __forceinline__ double min_accepted(double old, const double (&cand)[N], const bool (&acc)[N]) {
  static_assert((N & (N - 1)) == 0, "a power of two");
  double m[N];
#pragma unroll
  for (int i = 0; i < N; i++) m[i] = acc[i] ? cand[i] : __builtin_inf();
#pragma unroll
  for (int s = N / 2; s > 0; s >>= 1)
#pragma unroll
    for (int i = 0; i < s; i++) m[i] = __builtin_fmin(m[i], m[i + s]);
  return dmin(old, m[0]);
}

// The candidates and acceptance tests of fwd_update_f64 (src/eikonal_solver.cpp:155-283)
// with h = t0/dx and k = t0/dy formed as t0 * (1/dx), one reciprocal per
// quadratic (1/(2A)), per one-point denominator and of det, and FMA contraction.
__device__ inline double fwd_update_f64r(double tau_self, double t0, double t0x, double t0y, double a, double b,
                                         double c, double f, double idx, double idy, bool hW, double tW,
                                         double T0W, bool hE, double tE, double T0E, bool hS, double tS,
                                         double T0S, bool hN, double tN, double T0N) {
#pragma clang fp contract(fast)
  const double h = t0 * idx, k = t0 * idy;
  const double det = a * b - c * c;
  const double f2 = f * f;
  double cand[16];
  bool acc[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const double sx = (q < 2) ? -1.0 : 1.0;
    const double sy = (q & 1) ? 1.0 : -1.0;
    const bool ok0 = ((q < 2) ? hW : hE) & ((q & 1) ? hN : hS);
    const double tX = (q < 2) ? tW : tE;
    const double tY = (q & 1) ? tN : tS;
    const double px = t0x - sx * h;
    const double qx = sx * h * tX;
    const double py = t0y - sy * k;
    const double qy = sy * k * tY;
    const double eq_a = a * px * px + b * py * py - 2.0 * c * px * py;
    const double eq_b = 2.0 * (a * px * qx + b * py * qy - c * (px * qy + py * qx));
    const double eq_c = a * qx * qx + b * qy * qy - 2.0 * c * qx * qy - f2;
    const double Delta = eq_b * eq_b - 4.0 * eq_a * eq_c;
    const bool ok = ok0 & !(Delta < 0.0 || fabs(eq_a) < 1e-30);
    const double sq = (Delta > 0.0) ? sqrt_nr(Delta) : 0.0;
    const double r2a = rcp_nr(2.0 * eq_a);
#pragma unroll
    for (int sgn = 0; sgn < 2; sgn++) {
      const double tmp_tau = (-eq_b + (sgn == 0 ? sq : -sq)) * r2a;
      const double Tx = px * tmp_tau + qx;
      const double Ty = py * tmp_tau + qy;
      cand[2 * q + sgn] = tmp_tau;
      acc[2 * q + sgn] = ok & (-sx * (a * Tx - c * Ty) >= 0.0) & (-sy * (b * Ty - c * Tx) >= 0.0);
    }
  }
  const double idet = rcp_nr(det);
  const double dis_x = sqrt_nr(f2 * b * idet);
  const double dis_y = sqrt_nr(f2 * a * idet);
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const bool hd = (d == 0) ? hW : (d == 1) ? hE : (d == 2) ? hS : hN;
    const double tD = (d == 0) ? tW : (d == 1) ? tE : (d == 2) ? tS : tN;
    const double T0D = (d == 0) ? T0W : (d == 1) ? T0E : (d == 2) ? T0S : T0N;
    const double g = (d < 2) ? t0x : t0y;
    const double hh = (d < 2) ? h : k;
    const double p = (d == 0 || d == 2) ? g + hh : g - hh;
    const double qd = (d == 0 || d == 2) ? -hh * tD : hh * tD;
    const double dis = (d < 2) ? dis_x : dis_y;
    const double rp = rcp_nr(p);
    const bool ok = hd & (det > 0.0);
#pragma unroll
    for (int sgn = 0; sgn < 2; sgn++) {
      const double tmp_tau = ((sgn == 0 ? dis : -dis) - qd) * rp;
      cand[8 + 2 * d + sgn] = tmp_tau;
      acc[8 + 2 * d + sgn] = ok & (tmp_tau * t0 >= tD * T0D) & (tmp_tau > tD * 0.5);
    }
  }
  return min_accepted(tau_self, cand, acc);
}

struct FwdF64R : FwdF64 {
  struct Ctx : FwdF64::Ctx {
    double idx, idy;  // 1/dx, 1/dy
  };
  template <int C>
  __device__ static double update(const Ctx& cx, const Head<C>& h, const Body<C>& r, int k, int, double uW,
                                  double aW, bool hW, double uE, double aE, bool hE, double uS, double aS, bool hS,
                                  double uN, double aN, bool hN) {
    return fwd_update_f64r(h.u[k], h.aux[k], r.t0x[k], r.t0y[k], r.a[k], r.b[k], r.c[k], r.f[k], cx.idx, cx.idy,
                           hW, uW, aW, hE, uE, aE, hS, uS, aS, hN, uN, aN);
  }
};

// The re-solve update of an isotropic metric (spha = sphb = 1, sphc = 0) in the
// scaled coordinates of fwd_update_f32s: gradient planes sqrt(a) T0x and
// sqrt(b) T0y, h = t0 sqrt(a)/dx, k = t0 sqrt(b)/dy.  Every quantity of
// fwd_update_f64r scales by sqrt(a) or sqrt(b) (> 0), so each candidate and
// each acceptance test is the same as there: eq_a, eq_b, eq_c keep their
// values (without the cross terms, c = 0), the causality tests keep their
// signs, and the one-point distance sqrt(f^2 b/det) is f / sqrt(a), i.e. f
// after scaling (det = a b > 0).  About half the FP64 operations of the
// general update, and no reciprocal or square root of det.
__device__ inline double fwd_update_f64r_iso(double tau_self, double t0, double gx, double gy, double f, double ha,
                                             double kb, bool hW, double tW, double T0W, bool hE, double tE,
                                             double T0E, bool hS, double tS, double T0S, bool hN, double tN,
                                             double T0N) {
#pragma clang fp contract(fast)
  const double h = t0 * ha, k = t0 * kb;
  const double f2 = f * f;
  double cand[16];
  bool acc[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const double sx = (q < 2) ? -1.0 : 1.0;
    const double sy = (q & 1) ? 1.0 : -1.0;
    const bool ok0 = ((q < 2) ? hW : hE) & ((q & 1) ? hN : hS);
    const double tX = (q < 2) ? tW : tE;
    const double tY = (q & 1) ? tN : tS;
    const double px = gx - sx * h;
    const double qx = sx * h * tX;
    const double py = gy - sy * k;
    const double qy = sy * k * tY;
    const double eq_a = px * px + py * py;
    const double eq_b = 2.0 * (px * qx + py * qy);
    const double eq_c = qx * qx + qy * qy - f2;
    const double Delta = eq_b * eq_b - 4.0 * eq_a * eq_c;
    const bool ok = ok0 & !(Delta < 0.0 || fabs(eq_a) < 1e-30);
    const double sq = (Delta > 0.0) ? sqrt_nr(Delta) : 0.0;
    const double r2a = rcp_nr(2.0 * eq_a);
#pragma unroll
    for (int sgn = 0; sgn < 2; sgn++) {
      const double tmp_tau = (-eq_b + (sgn == 0 ? sq : -sq)) * r2a;
      const double Tx = px * tmp_tau + qx;
      const double Ty = py * tmp_tau + qy;
      cand[2 * q + sgn] = tmp_tau;
      acc[2 * q + sgn] = ok & (-sx * Tx >= 0.0) & (-sy * Ty >= 0.0);
    }
  }
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const bool hd = (d == 0) ? hW : (d == 1) ? hE : (d == 2) ? hS : hN;
    const double tD = (d == 0) ? tW : (d == 1) ? tE : (d == 2) ? tS : tN;
    const double T0D = (d == 0) ? T0W : (d == 1) ? T0E : (d == 2) ? T0S : T0N;
    const double g = (d < 2) ? gx : gy;
    const double hh = (d < 2) ? h : k;
    const double p = (d == 0 || d == 2) ? g + hh : g - hh;
    const double qd = (d == 0 || d == 2) ? -hh * tD : hh * tD;
    const double rp = rcp_nr(p);
#pragma unroll
    for (int sgn = 0; sgn < 2; sgn++) {
      const double tmp_tau = ((sgn == 0 ? f : -f) - qd) * rp;
      cand[8 + 2 * d + sgn] = tmp_tau;
      acc[8 + 2 * d + sgn] = hd & (tmp_tau * t0 >= tD * T0D) & (tmp_tau > tD * 0.5);
    }
  }
  return min_accepted(tau_self, cand, acc);
}

// FP64 re-solves of isotropic grids: scaled gradient planes, the slowness per
// node and sqrt(a)/dx per row (GridDev::rowha).
struct FwdF64RIso {
  using V = double;
  using A = double;
  static constexpr bool kAux = true;
  struct Ctx {
    const GridDev* g;
    const double *t0[2], *t0x[2], *t0y[2], *f[2];
    const double* rowha;
    double kb;
    FixedBox fb;
  };
  template <int C> struct Head { double u[C], aux[C]; };
  template <int C> struct Body { double t0x[C], t0y[C], f[C], st[C]; double ha; };
  template <int C, int O>
  __device__ static void load_head(const Ctx& cx, Head<C>& H, int base, int, const double* tin) {
#pragma unroll
    for (int k = 0; k < C; k++) { H.u[k] = gld(tin, base + k); H.aux[k] = gld(cx.t0[O], base + k); }
  }
  template <int C, int O, bool NORM>
  __device__ static void load_body(const Ctx& cx, Body<C>& B, int base, int j, const double* tst) {
#pragma unroll
    for (int k = 0; k < C; k++) {
      B.t0x[k] = gld(cx.t0x[O], base + k);
      B.t0y[k] = gld(cx.t0y[O], base + k);
      B.f[k] = gld(cx.f[O], base + k);
      B.st[k] = NORM ? gld(tst, base + k) : 0.0;
    }
    B.ha = gld(cx.rowha, (unsigned)j);
  }
  __device__ static bool fixed(const Ctx& cx, int ix, int j) {
    return j >= cx.fb.y0 && j <= cx.fb.y1 && ix >= cx.fb.x0 && ix <= cx.fb.x1;
  }
  template <int C>
  __device__ static double update(const Ctx& cx, const Head<C>& h, const Body<C>& r, int k, int, double uW,
                                  double aW, bool hW, double uE, double aE, bool hE, double uS, double aS, bool hS,
                                  double uN, double aN, bool hN) {
    return fwd_update_f64r_iso(h.u[k], h.aux[k], r.t0x[k], r.t0y[k], r.f[k], r.ha, cx.kb, hW, uW, aW, hE, uE, aE,
                               hS, uS, aS, hN, uN, aN);
  }
  template <int C>
  __device__ static double norm(const Head<C>& h, const Body<C>& r, int k, double v) {
    return fabs(v - r.st[k]) * h.aux[k];  // :295
  }
};

// Per-solve adjoint: precomputed a1p, a2m, b1p, b2m, dl, d (one segmented plane each).
struct AdjF64 {
  using V = double;
  using A = double;
  static constexpr bool kAux = false;
  struct Ctx {
    const GridDev* g;
    int nx, ny;
    const double *a1p, *a2m, *b1p, *b2m, *dl, *d;
    double dx, dy;
  };
  template <int C> struct Head { double u[C], aux[C]; };
  template <int C> struct Body { double a1p[C], a2m[C], b1p[C], b2m[C], dl[C], d[C], st[C]; };
  template <int C, int O>
  __device__ static void load_head(const Ctx&, Head<C>& H, int base, int, const double* tin) {
#pragma unroll
    for (int k = 0; k < C; k++) { H.u[k] = gld(tin, base + k); H.aux[k] = 0.0; }
  }
  template <int C, int O, bool NORM>
  __device__ static void load_body(const Ctx& cx, Body<C>& B, int base, int, const double* tst) {
#pragma unroll
    for (int k = 0; k < C; k++) {
      B.a1p[k] = gld(cx.a1p, base + k);
      B.a2m[k] = gld(cx.a2m, base + k);
      B.b1p[k] = gld(cx.b1p, base + k);
      B.b2m[k] = gld(cx.b2m, base + k);
      B.dl[k] = gld(cx.dl, base + k);
      B.d[k] = gld(cx.d, base + k);
      B.st[k] = NORM ? gld(tst, base + k) : 0.0;
    }
  }
  __device__ static bool fixed(const Ctx& cx, int ix, int j) {
    return ix == 0 || ix >= cx.nx - 1 || j == 0 || j == cx.ny - 1;
  }
  template <int C>
  __device__ static double update(const Ctx& cx, const Head<C>&, const Body<C>& r, int k, int, double uW, double,
                                  bool, double uE, double, bool, double uS, double, bool, double uN, double, bool) {
    return adj_update_f64(r.a1p[k], r.a2m[k], r.b1p[k], r.b2m[k], r.dl[k], r.d[k], uW, uE, uS, uN, cx.dx, cx.dy);
  }
  template <int C>
  __device__ static double norm(const Head<C>& h, const Body<C>& r, int k, double v) {
    return fabs(v - r.st[k]);  // :501
  }
};

// Batched adjoint, face form.  a2 at node (ix, iy) is the same expression, in
// the same operation order, as a1 at (ix+1, iy), and b2 at (ix, iy) as b1 at
// (ix, iy+1) (src/eikonal_solver.cpp:407-441), so one x-face and one y-face
// flux per node, F = (a1, b1), carry all four: the node reads its own F and,
// through the neighbour auxiliaries, a2 = F_E.x and b2 = F_N.y.  The splits,
// d (:479-481) and the update (:483-490) are evaluated per visit in the
// reference's order, so the result is the same bits as the precomputed form.
struct AdjF64Face {
  using V = double;
  using A = double2;
  static constexpr bool kAux = true;
  struct Ctx {
    const GridDev* g;
    int nx, ny;
    const double2* fxy[2];  // (a1, b1) per node
    const double* dl[2];    // delta
    double dx, dy;
  };
  template <int C> struct Head { double u[C]; double2 aux[C]; };
  template <int C> struct Body { double dl[C], st[C]; };
  template <int C, int O>
  __device__ static void load_head(const Ctx& cx, Head<C>& H, int base, int, const double* tin) {
#pragma unroll
    for (int k = 0; k < C; k++) { H.u[k] = gld(tin, base + k); H.aux[k] = gld(cx.fxy[O], base + k); }
  }
  template <int C, int O, bool NORM>
  __device__ static void load_body(const Ctx& cx, Body<C>& B, int base, int, const double* tst) {
#pragma unroll
    for (int k = 0; k < C; k++) {
      B.dl[k] = gld(cx.dl[O], base + k);
      B.st[k] = NORM ? gld(tst, base + k) : 0.0;
    }
  }
  __device__ static bool fixed(const Ctx& cx, int ix, int j) {
    return ix == 0 || ix >= cx.nx - 1 || j == 0 || j == cx.ny - 1;
  }
  template <int C>
  __device__ static double update(const Ctx& cx, const Head<C>& h, const Body<C>& r, int k, int, double uW, double2,
                                  bool, double uE, double2 aE, bool, double uS, double2, bool, double uN, double2 aN,
                                  bool) {
    const double a1 = h.aux[k].x, b1 = h.aux[k].y, a2 = aE.x, b2 = aN.y;
    const double a1m = (a1 - fabs(a1)) / 2.0, a1p = (a1 + fabs(a1)) / 2.0;
    const double a2m = (a2 - fabs(a2)) / 2.0, a2p = (a2 + fabs(a2)) / 2.0;
    const double b1m = (b1 - fabs(b1)) / 2.0, b1p = (b1 + fabs(b1)) / 2.0;
    const double b2m = (b2 - fabs(b2)) / 2.0, b2p = (b2 + fabs(b2)) / 2.0;
    const double d = (a2p - a1m) / cx.dx + (b2p - b1m) / cx.dy;
    return adj_update_f64(a1p, a2m, b1p, b2m, r.dl[k], d, uW, uE, uS, uN, cx.dx, cx.dy);
  }
  template <int C>
  __device__ static double norm(const Head<C>&, const Body<C>& r, int k, double v) {
    return fabs(v - r.st[k]);  // :501
  }
};

// ============================================================================
// FP32 sweeps on sigma = tau - 1
// ============================================================================

// One forward node in increment form.  Same candidate set and acceptance tests
// as src/eikonal_solver.cpp:173-283, written for sigma relative to the smallest
// present neighbour value u_ref (so every difference below is the variation of
// sigma over about a cell, exact to FP32 relative precision):
//   dX = sigma_X - u_ref, tau_ref = 1 + u_ref,
//   two-point (sx, sy): dr = min(dX, dY), tau_r = tau_ref + dr, tau = tau_r + w,
//     Tx = Tx0 + px*w,  Tx0 = t0x*tau_r + sx*h*(dX - dr),  px = t0x - sx*h
//     (y alike), a*Tx^2 + b*Ty^2 - 2c*Tx*Ty = f^2 solved for w with the stable
//     quadratic formula (w0 = s/A, w1 = C/s); accepted if -sx*(a*Tx - c*Ty) >= 0
//     and -sy*(b*Ty - c*Tx) >= 0; candidate dr + w.
//   one-point (W): tau_W = tau_ref + dW, w = (+-dis - t0x*tau_W)/(t0x + h), accepted
//     if (tau_W + w)*T0 >= tau_W*T0_W and tau_W + w > tau_W/2; candidate dW + w.
// Candidates are compared with sigma_self - u_ref; the smallest, if below it, is
// the new value u_ref + candidate formed in FP64 (the reference keeps the
// smaller of the old and the new tau).  Values only ever decrease, so the
// rounding of a large change would persist: measuring from u_ref keeps the
// rounded quantities small.
//
// An absent neighbour is passed as NaN, and the tests are stated so that NaN
// rejects, which leaves no presence masks in the instruction stream:
//  - the NaN drops out of u_ref (fmin ignores NaN) and turns every candidate
//    that uses the neighbour into NaN, as does a negative discriminant (sqrt),
//    |A| < 1e-30 or det <= 0 (dis is NaN then); fminf ignores NaN;
//  - two-point causality is max(sx*(a*Tx - c*Ty), sy*(b*Ty - c*Tx)) <= 0, and
//    with c == 0, a, b > 0 (ISO) max(sx*Tx, sy*Ty) <= 0; a NaN root makes both
//    operands NaN: rejected;
//  - the second root C/s needs no s != 0 guard: s == 0 only with C == 0, where
//    C/s is NaN and the first root carries the same value;
//  - one-point: both tests are lower bounds on w (t0, tau_W > 0):
//    w >= tau_W*max(T0_W/t0 - 1, -1/2).
// ISO: the isotropic metric (c == 0, a and b constant along a row, both > 0).
// dsx = f*sqrt(b/det), dsy = f*sqrt(a/det) (f/sqrt(a), f/sqrt(b) when ISO).
template <bool ISO>
__device__ __forceinline__ double fwd_update_f32(double u_self, float t0, float t0x, float t0y, float a, float b,
                                                 float c, float f2, float dsx, float dsy, float h, float kk,
                                                 double uW, float T0W, double uE, float T0E, double uS, float T0S,
                                                 double uN, float T0N) {
#pragma clang fp contract(fast)
  const float kInf = __builtin_huge_valf();
  const float kNaNf = __builtin_nanf("");
  const double u_ref = fmin(fmin(uW, uE), fmin(uS, uN));
  const float ts = (float)(1.0 + u_ref);
  const float dW = (float)(uW - u_ref), dE = (float)(uE - u_ref);
  const float dS = (float)(uS - u_ref), dN = (float)(uN - u_ref);
  // px = t0x - sx*h, py = t0y - sy*k for s = -1 (m) and +1 (p)
  const float pxm = t0x + h, pxp = t0x - h, pym = t0y + kk, pyp = t0y - kk;
  const float apm = a * pxm, app = a * pxp, bpm = b * pym, bpp = b * pyp;
  const float Axm = apm * pxm, Axp = app * pxp, Bym = bpm * pym, Byp = bpp * pyp;
  float best = kInf;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const bool xm = q < 2, ym = !(q & 1);
    const float sxh = xm ? -h : h, syk = ym ? -kk : kk;
    const float dX = xm ? dW : dE, dY = ym ? dS : dN;
    const float px = xm ? pxm : pxp, py = ym ? pym : pyp;
    const float apx = xm ? apm : app, bpy = ym ? bpm : bpp;
    const float dr = fminf(dX, dY);
    const float tr = ts + dr;
    const float Tx0 = t0x * tr + sxh * (dX - dr);
    const float Ty0 = t0y * tr + syk * (dY - dr);
    float Aq, Bh, Cq;
    if constexpr (ISO) {
      Aq = (xm ? Axm : Axp) + (ym ? Bym : Byp);
      Bh = apx * Tx0 + bpy * Ty0;
      Cq = Tx0 * (a * Tx0) + Ty0 * (b * Ty0) - f2;
    } else {
      Aq = (xm ? Axm : Axp) + (ym ? Bym : Byp) - 2.0f * c * px * py;
      Bh = apx * Tx0 + bpy * Ty0 - c * (px * Ty0 + py * Tx0);
      Cq = Tx0 * (a * Tx0) + Ty0 * (b * Ty0) - 2.0f * c * Tx0 * Ty0 - f2;
    }
    float D = Bh * Bh - Aq * Cq;
    D = (fabsf(Aq) >= 1e-30f) ? D : kNaNf;
    const float sq = __builtin_amdgcn_sqrtf(D);
    const float sv = -(Bh + __builtin_copysignf(sq, Bh));
    const float w0 = sv * __builtin_amdgcn_rcpf(Aq);
    const float w1 = Cq * __builtin_amdgcn_rcpf(sv);
    float cw[2];
#pragma unroll
    for (int g = 0; g < 2; g++) {
      const float w = g ? w1 : w0;
      const float Tx = Tx0 + px * w;
      const float Ty = Ty0 + py * w;
      const float cxv = ISO ? Tx : a * Tx - c * Ty;
      const float cyv = ISO ? Ty : b * Ty - c * Tx;
      const float m = fmaxf(xm ? -cxv : cxv, ym ? -cyv : cyv);
      cw[g] = (m <= 0.0f) ? w : kInf;
    }
    best = fminf(best, dr + fminf(cw[0], cw[1]));
  }
  const float rt0 = __builtin_amdgcn_rcpf(t0);
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const float dq = (d == 0) ? dW : (d == 1) ? dE : (d == 2) ? dS : dN;
    const float T0n = (d == 0) ? T0W : (d == 1) ? T0E : (d == 2) ? T0S : T0N;
    const float td = (d < 2) ? t0x : t0y;
    const float p = (d == 0) ? pxm : (d == 1) ? pxp : (d == 2) ? pym : pyp;
    const float dis = (d < 2) ? dsx : dsy;
    const float rp = __builtin_amdgcn_rcpf(p);
    const float tnd = ts + dq;
    const float Td0 = td * tnd;
    const float wp = (dis - Td0) * rp;
    const float wm = (-dis - Td0) * rp;
    const float thw = tnd * fmaxf(T0n * rt0 - 1.0f, -0.5f);
    const float cp = (wp >= thw) ? wp : kInf;
    const float cm = (wm >= thw) ? wm : kInf;
    best = fminf(best, dq + fminf(cp, cm));
  }
  const double v = u_ref + (double)best;  // +inf when no candidate was accepted
  return ((best < (float)(u_self - u_ref)) & (v < u_self)) ? v : u_self;
}

struct FwdF32 {
  using V = double;  // sigma storage
  using A = float;   // T0 of a neighbour
  static constexpr bool kAux = true;
  struct Ctx {
    const GridDev* g;
    const float *t0[2], *t0x[2], *t0y[2], *a[2], *b[2], *c[2], *f[2], *dsx[2], *dsy[2];
    FixedBox fb;
    float idx, idy;  // 1/dx, 1/dy (radians)
  };
  template <int C> struct Head { double u[C]; float aux[C]; };
  template <int C> struct Body { double st[C]; float t0x[C], t0y[C], a[C], b[C], c[C], f[C], dsx[C], dsy[C]; };
  template <int C, int O>
  __device__ static void load_head(const Ctx& cx, Head<C>& H, int base, int, const double* tin) {
#pragma unroll
    for (int k = 0; k < C; k++) { H.u[k] = gld(tin, base + k); H.aux[k] = gld(cx.t0[O], base + k); }
  }
  template <int C, int O, bool NORM>
  __device__ static void load_body(const Ctx& cx, Body<C>& B, int base, int, const double* tst) {
#pragma unroll
    for (int k = 0; k < C; k++) {
      B.t0x[k] = gld(cx.t0x[O], base + k);
      B.t0y[k] = gld(cx.t0y[O], base + k);
      B.a[k] = gld(cx.a[O], base + k);
      B.b[k] = gld(cx.b[O], base + k);
      B.c[k] = gld(cx.c[O], base + k);
      B.f[k] = gld(cx.f[O], base + k);
      B.dsx[k] = gld(cx.dsx[O], base + k);
      B.dsy[k] = gld(cx.dsy[O], base + k);
      B.st[k] = NORM ? gld(tst, base + k) : 0.0;
    }
  }
  __device__ static bool fixed(const Ctx& cx, int ix, int j) {
    return j >= cx.fb.y0 && j <= cx.fb.y1 && ix >= cx.fb.x0 && ix <= cx.fb.x1;
  }
  template <int C>
  __device__ static double update(const Ctx& cx, const Head<C>& hd, const Body<C>& r, int k, int, double uW,
                                  float aW, bool hW, double uE, float aE, bool hE, double uS, float aS, bool hS,
                                  double uN, float aN, bool hN) {
    const float t0 = hd.aux[k];
    const double kNaN = __builtin_nan("");
    return fwd_update_f32<false>(hd.u[k], t0, r.t0x[k], r.t0y[k], r.a[k], r.b[k], r.c[k], r.f[k] * r.f[k], r.dsx[k],
                                 r.dsy[k], t0 * cx.idx, t0 * cx.idy, hW ? uW : kNaN, aW, hE ? uE : kNaN, aE,
                                 hS ? uS : kNaN, aS, hN ? uN : kNaN, aN);
  }
  template <int C>
  __device__ static double norm(const Head<C>& h, const Body<C>& r, int k, double v) {
    return fabs(v - r.st[k]) * (double)h.aux[k];
  }
};

// The FP32 update of an isotropic metric (spha = sphb = 1, sphc = 0) in the
// scaled coordinates X = sqrt(a) Tx, Y = sqrt(b) Ty, where the node equation
// a Tx^2 + b Ty^2 = f^2 reads X^2 + Y^2 = f^2.  The gradient planes hold
// sqrt(a) T0x and sqrt(b) T0y, h = t0 sqrt(a)/dx and k = t0 sqrt(b)/dy (per
// row), and every candidate of fwd_update_f32<true> keeps its value and its
// acceptance test: px, py, Tx0, Ty0 and the roots' Tx, Ty all scale by sqrt(a)
// or sqrt(b) (> 0, so no causality sign changes), the quadratic's
// coefficients become Aq = px^2 + py^2 (= the reference's eq_a),
// Bh = px X0 + py Y0 and Cq = X0^2 + Y0^2 - f^2, and the one-point distances
// sqrt(f^2 b/det), sqrt(f^2 a/det) become f.  Aq needs no test against the
// reference's 1e-30: it is a sum of squares, zero only where px = py = 0,
// and there both roots give Tx = Ty = NaN, which no causality test accepts.
__device__ __forceinline__ double fwd_update_f32s(double u_self, float t0, float t0x, float t0y, float f, float h,
                                                  float kk, double uW, float T0W, double uE, float T0E, double uS,
                                                  float T0S, double uN, float T0N) {
#pragma clang fp contract(fast)
  const float kInf = __builtin_huge_valf();
  const double u_ref = fmin(fmin(uW, uE), fmin(uS, uN));
  const float ts = (float)(1.0 + u_ref);
  const float dW = (float)(uW - u_ref), dE = (float)(uE - u_ref);
  const float dS = (float)(uS - u_ref), dN = (float)(uN - u_ref);
  const float f2 = f * f;
  const float pxm = t0x + h, pxp = t0x - h, pym = t0y + kk, pyp = t0y - kk;
  const float Axm = pxm * pxm, Axp = pxp * pxp, Bym = pym * pym, Byp = pyp * pyp;
  float best = kInf;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const bool xm = q < 2, ym = !(q & 1);
    const float sxh = xm ? -h : h, syk = ym ? -kk : kk;
    const float dX = xm ? dW : dE, dY = ym ? dS : dN;
    const float px = xm ? pxm : pxp, py = ym ? pym : pyp;
    const float dr = fminf(dX, dY);
    const float tr = ts + dr;
    const float X0 = t0x * tr + sxh * (dX - dr);
    const float Y0 = t0y * tr + syk * (dY - dr);
    const float Aq = (xm ? Axm : Axp) + (ym ? Bym : Byp);
    const float Bh = px * X0 + py * Y0;
    const float Cq = X0 * X0 + Y0 * Y0 - f2;
    const float D = Bh * Bh - Aq * Cq;
    const float sq = __builtin_amdgcn_sqrtf(D);  // NaN for D < 0: no candidate
    const float sv = -(Bh + __builtin_copysignf(sq, Bh));
    const float w0 = sv * __builtin_amdgcn_rcpf(Aq);
    const float w1 = Cq * __builtin_amdgcn_rcpf(sv);
    float cw[2];
#pragma unroll
    for (int g = 0; g < 2; g++) {
      const float w = g ? w1 : w0;
      const float X = X0 + px * w;
      const float Y = Y0 + py * w;
      const float m = fmaxf(xm ? -X : X, ym ? -Y : Y);
      cw[g] = (m <= 0.0f) ? w : kInf;
    }
    best = fminf(best, dr + fminf(cw[0], cw[1]));
  }
  const float rt0 = __builtin_amdgcn_rcpf(t0);
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const float dq = (d == 0) ? dW : (d == 1) ? dE : (d == 2) ? dS : dN;
    const float T0n = (d == 0) ? T0W : (d == 1) ? T0E : (d == 2) ? T0S : T0N;
    const float td = (d < 2) ? t0x : t0y;
    const float p = (d == 0) ? pxm : (d == 1) ? pxp : (d == 2) ? pym : pyp;
    const float rp = __builtin_amdgcn_rcpf(p);
    const float tnd = ts + dq;
    const float Td0 = td * tnd;
    const float wp = (f - Td0) * rp;
    const float wm = (-f - Td0) * rp;
    const float thw = tnd * fmaxf(T0n * rt0 - 1.0f, -0.5f);
    const float cp = (wp >= thw) ? wp : kInf;
    const float cm = (wm >= thw) ? wm : kInf;
    best = fminf(best, dq + fminf(cp, cm));
  }
  const double v = u_ref + (double)best;  // +inf when no candidate was accepted
  return ((best < (float)(u_self - u_ref)) & (v < u_self)) ? v : u_self;
}

// Isotropic metric (spha = sphb = 1, sphc = 0 on every node): scaled gradient
// planes, the slowness per node and {sqrt(a)/dx, sqrt(b)/dy} per row.
struct FwdF32Iso {
  using V = double;
  using A = float;
  static constexpr bool kAux = true;
  struct Ctx {
    const GridDev* g;
    const float *t0[2], *t0x[2], *t0y[2], *f[2];
    const float2* rowtab;
    FixedBox fb;
    float idx, idy;
  };
  template <int C> struct Head { double u[C]; float aux[C]; };
  template <int C> struct Body { double st[C]; float t0x[C], t0y[C], f[C]; float2 row; };
  template <int C, int O>
  __device__ static void load_head(const Ctx& cx, Head<C>& H, int base, int, const double* tin) {
#pragma unroll
    for (int k = 0; k < C; k++) { H.u[k] = gld(tin, base + k); H.aux[k] = gld(cx.t0[O], base + k); }
  }
  template <int C, int O, bool NORM>
  __device__ static void load_body(const Ctx& cx, Body<C>& B, int base, int j, const double* tst) {
#pragma unroll
    for (int k = 0; k < C; k++) {
      B.t0x[k] = gld(cx.t0x[O], base + k);
      B.t0y[k] = gld(cx.t0y[O], base + k);
      B.f[k] = gld(cx.f[O], base + k);
      B.st[k] = NORM ? gld(tst, base + k) : 0.0;
    }
    B.row = gld(cx.rowtab, (unsigned)j);
  }
  __device__ static bool fixed(const Ctx& cx, int ix, int j) {
    return j >= cx.fb.y0 && j <= cx.fb.y1 && ix >= cx.fb.x0 && ix <= cx.fb.x1;
  }
  template <int C>
  __device__ static double update(const Ctx& cx, const Head<C>& hd, const Body<C>& r, int k, int, double uW,
                                  float aW, bool hW, double uE, float aE, bool hE, double uS, float aS, bool hS,
                                  double uN, float aN, bool hN) {
    const float t0 = hd.aux[k];
    const double kNaN = __builtin_nan("");
    return fwd_update_f32s(hd.u[k], t0, r.t0x[k], r.t0y[k], r.f[k], t0 * r.row.x, t0 * r.row.y, hW ? uW : kNaN,
                           aW, hE ? uE : kNaN, aE, hS ? uS : kNaN, aS, hN ? uN : kNaN, aN);
  }
  template <int C>
  __device__ static double norm(const Head<C>& h, const Body<C>& r, int k, double v) {
    return fabs(v - r.st[k]) * (double)h.aux[k];
  }
};

// Isotropic metric (spha = sphb = 1, sphc = 0 on every node), FP32 candidate
// search with an FP64 polish of the winner.
//
// Coordinates.  With X = sqrt(a) Tx, Y = sqrt(b) Ty the node equation
// a Tx^2 + b Ty^2 = f^2 is X^2 + Y^2 = f^2, and every candidate and
// acceptance test of fwd_update_f32<true> is the same after scaling: the T0
// gradient planes hold sqrt(a) T0x and sqrt(b) T0y, h = t0 sqrt(a)/dx (per row)
// and k = t0 sqrt(b)/dy, the one-point distances become f, and the causality
// signs are unchanged (sqrt(a), sqrt(b) > 0).
//
// Polish.  Rounding the per-node inputs (T0 gradient, slowness) to FP32
// perturbs each node's equation by ~1e-7 of f^2 and moves its root by ~1e-7
// of a cell's travel time; along rays these add up to T noise whose
// differences, taken by the adjoint's face fluxes, reached 1e-5 of the
// kernels on the 401x481 workload.  So the search finds the winning candidate
// and its increment w in FP32, and then the winner's residual is evaluated in
// FP64 from the FP64 inputs and neighbour values, and one Newton step
// w' = w - r(w)/r'(w) (the correction itself in FP32) gives the FP64 root to
// ~1e-14 relative.  The new value is taken when it is below the old one in
// FP64 (also where the FP32 search saw no decrease: those tiny decreases
// move the convergence norm).  The node's own T0 enters through h only and
// stays FP32.
// A candidate tagged with its group in the low 3 bits of its FP32 encoding:
// the running minimum then carries the winner's group (a change of at most 7
// ulp in the compared value; the polish recomputes the value anyway).
__device__ __forceinline__ float tag_group(float c, unsigned g) {
  return __builtin_bit_cast(float, (__builtin_bit_cast(unsigned, c) & ~7u) | g);
}

__device__ __forceinline__ double fwd_update_f32p(double u_self, float t0, double t0xd, double t0yd, double fd,
                                                  double had, double kbd, double uW, float T0W, double uE, float T0E,
                                                  double uS, float T0S, double uN, float T0N) {
#pragma clang fp contract(fast)
  const float kInf = __FLT_MAX__;  // "rejected" (see tag_group)
  const float kNaNf = __builtin_nanf("");
  const double u_ref = fmin(fmin(uW, uE), fmin(uS, uN));
  const float ts = (float)(1.0 + u_ref);
  const float dW = (float)(uW - u_ref), dE = (float)(uE - u_ref);
  const float dS = (float)(uS - u_ref), dN = (float)(uN - u_ref);
  const float t0x = (float)t0xd, t0y = (float)t0yd, f = (float)fd;
  const float h = t0 * (float)had, kk = t0 * (float)kbd;
  const float f2 = f * f;
  const float pxm = t0x + h, pxp = t0x - h, pym = t0y + kk, pyp = t0y - kk;
  const float Axm = pxm * pxm, Axp = pxp * pxp, Bym = pym * pym, Byp = pyp * pyp;
  // groups: 0..3 two-point quadrants (bit 1: x neighbour E, bit 0: y neighbour N),
  // 4..7 one-point W, E, S, N.  A NaN candidate (an absent neighbour, a
  // negative discriminant) stays a quiet NaN when tagged; the minimum ignores it.
  float best = kInf;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const bool xm = q < 2, ym = !(q & 1);
    const float sxh = xm ? -h : h, syk = ym ? -kk : kk;
    const float dX = xm ? dW : dE, dY = ym ? dS : dN;
    const float px = xm ? pxm : pxp, py = ym ? pym : pyp;
    const float dr = fminf(dX, dY);
    const float tr = ts + dr;
    const float X0 = t0x * tr + sxh * (dX - dr);
    const float Y0 = t0y * tr + syk * (dY - dr);
    const float Aq = (xm ? Axm : Axp) + (ym ? Bym : Byp);
    const float Bh = px * X0 + py * Y0;
    const float Cq = X0 * X0 + Y0 * Y0 - f2;
    float D = Bh * Bh - Aq * Cq;
    D = (fabsf(Aq) >= 1e-30f) ? D : kNaNf;
    const float sq = __builtin_amdgcn_sqrtf(D);
    const float sv = -(Bh + __builtin_copysignf(sq, Bh));
    const float w0 = sv * __builtin_amdgcn_rcpf(Aq);
    const float w1 = Cq * __builtin_amdgcn_rcpf(sv);
    float cw[2];
#pragma unroll
    for (int g = 0; g < 2; g++) {
      const float w = g ? w1 : w0;
      const float X = X0 + px * w;
      const float Y = Y0 + py * w;
      const float m = fmaxf(xm ? -X : X, ym ? -Y : Y);
      cw[g] = (m <= 0.0f) ? w : kInf;
    }
    best = fminf(best, tag_group(dr + fminf(cw[0], cw[1]), (unsigned)q));
  }
  const float rt0 = __builtin_amdgcn_rcpf(t0);
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const float dq = (d == 0) ? dW : (d == 1) ? dE : (d == 2) ? dS : dN;
    const float T0n = (d == 0) ? T0W : (d == 1) ? T0E : (d == 2) ? T0S : T0N;
    const float td = (d < 2) ? t0x : t0y;
    const float p = (d == 0) ? pxm : (d == 1) ? pxp : (d == 2) ? pym : pyp;
    const float rp = __builtin_amdgcn_rcpf(p);
    const float tnd = ts + dq;
    const float Td0 = td * tnd;
    const float wp = (f - Td0) * rp;
    const float wm = (-f - Td0) * rp;
    const float thw = tnd * fmaxf(T0n * rt0 - 1.0f, -0.5f);
    const float cp = (wp >= thw) ? wp : kInf;
    const float cm = (wm >= thw) ? wm : kInf;
    best = fminf(best, tag_group(dq + fminf(cp, cm), 4u + (unsigned)d));
  }
  // FP64 polish of the winner, one residual form for every group: the x and
  // y terms of the two-point quadratic X^2 + Y^2 = f^2, the one-point
  // equations being the same with the other term dropped (X = +-f).  When no
  // candidate was accepted best is ~FLT_MAX and so is v: the old value stays.
  const unsigned g = __builtin_bit_cast(unsigned, best) & 7u;
  const bool two = g < 4u;
  const bool useW = two ? (g < 2u) : (g == 4u);
  const bool useS = two ? !(g & 1u) : (g == 6u);
  const bool hasX = two || g < 6u;
  const bool hasY = two || g >= 6u;
  const double t0d = (double)t0;
  const double h64 = t0d * had, k64 = t0d * kbd;
  const double ux = useW ? uW : uE, uy = useS ? uS : uN;
  const double dX = ux - u_ref, dY = uy - u_ref;
  const double dr = two ? fmin(dX, dY) : (hasX ? dX : dY);
  const double tr = 1.0 + u_ref + dr;
  const double sxh = useW ? -h64 : h64, syk = useS ? -k64 : k64;
  const double px = t0xd - sxh, py = t0yd - syk;
  const double w = (double)best - dr;
  const double X = hasX ? t0xd * tr + sxh * (dX - dr) + px * w : 0.0;
  const double Y = hasY ? t0yd * tr + syk * (dY - dr) + py * w : 0.0;
  const double r = X * X + Y * Y - fd * fd;
  const double rd = 2.0 * (X * px + Y * py);
  const double v = u_ref + dr + (w - (double)((float)r * __builtin_amdgcn_rcpf((float)rd)));
  return (v < u_self) ? v : u_self;
}

struct FwdF32IsoP {
  using V = double;  // sigma storage
  using A = float;   // T0 of a neighbour
  static constexpr bool kAux = true;
  struct Ctx {
    const GridDev* g;
    const float* t0[2];
    const double *t0x[2], *t0y[2], *f[2];  // sqrt(a) T0x, sqrt(b) T0y, slowness (FP64)
    const double* rowha;                    // sqrt(a)/dx per row
    double kb;                              // sqrt(b)/dy
    FixedBox fb;
  };
  template <int C> struct Head { double u[C]; float aux[C]; };
  template <int C> struct Body { double st[C], t0x[C], t0y[C], f[C]; double ha; };
  template <int C, int O>
  __device__ static void load_head(const Ctx& cx, Head<C>& H, int base, int, const double* tin) {
#pragma unroll
    for (int k = 0; k < C; k++) { H.u[k] = gld(tin, base + k); H.aux[k] = gld(cx.t0[O], base + k); }
  }
  template <int C, int O, bool NORM>
  __device__ static void load_body(const Ctx& cx, Body<C>& B, int base, int j, const double* tst) {
#pragma unroll
    for (int k = 0; k < C; k++) {
      B.t0x[k] = gld(cx.t0x[O], base + k);
      B.t0y[k] = gld(cx.t0y[O], base + k);
      B.f[k] = gld(cx.f[O], base + k);
      B.st[k] = NORM ? gld(tst, base + k) : 0.0;
    }
    B.ha = gld(cx.rowha, j);
  }
  __device__ static bool fixed(const Ctx& cx, int ix, int j) {
    return j >= cx.fb.y0 && j <= cx.fb.y1 && ix >= cx.fb.x0 && ix <= cx.fb.x1;
  }
  template <int C>
  __device__ static double update(const Ctx& cx, const Head<C>& hd, const Body<C>& r, int k, int, double uW,
                                  float aW, bool hW, double uE, float aE, bool hE, double uS, float aS, bool hS,
                                  double uN, float aN, bool hN) {
    const double kNaN = __builtin_nan("");
    return fwd_update_f32p(hd.u[k], hd.aux[k], r.t0x[k], r.t0y[k], r.f[k], r.ha, cx.kb, hW ? uW : kNaN, aW,
                           hE ? uE : kNaN, aE, hS ? uS : kNaN, aS, hN ? uN : kNaN, aN);
  }
  template <int C>
  __device__ static double norm(const Head<C>& h, const Body<C>& r, int k, double v) {
    return fabs(v - r.st[k]) * (double)h.aux[k];
  }
};

// Per-solve FP32 adjoint: coefficients divided through by d (one segmented plane each).
struct AdjF32 {
  using V = float;
  using A = float;
  static constexpr bool kAux = false;
  struct Ctx {
    const GridDev* g;
    int nx, ny;
    const float *cW, *cE, *cS, *cN, *cD;
  };
  template <int C> struct Head { float u[C], aux[C]; };
  template <int C> struct Body { float cW[C], cE[C], cS[C], cN[C], cD[C], st[C]; };
  template <int C, int O>
  __device__ static void load_head(const Ctx&, Head<C>& H, int base, int, const float* tin) {
#pragma unroll
    for (int k = 0; k < C; k++) { H.u[k] = gld(tin, base + k); H.aux[k] = 0.0f; }
  }
  template <int C, int O, bool NORM>
  __device__ static void load_body(const Ctx& cx, Body<C>& B, int base, int, const float* tst) {
#pragma unroll
    for (int k = 0; k < C; k++) {
      B.cW[k] = gld(cx.cW, base + k);
      B.cE[k] = gld(cx.cE, base + k);
      B.cS[k] = gld(cx.cS, base + k);
      B.cN[k] = gld(cx.cN, base + k);
      B.cD[k] = gld(cx.cD, base + k);
      B.st[k] = NORM ? gld(tst, base + k) : 0.0f;
    }
  }
  __device__ static bool fixed(const Ctx& cx, int ix, int j) {
    return ix == 0 || ix >= cx.nx - 1 || j == 0 || j == cx.ny - 1;
  }
  template <int C>
  __device__ static float update(const Ctx&, const Head<C>&, const Body<C>& r, int k, int, float uW, float, bool,
                                 float uE, float, bool, float uS, float, bool, float uN, float, bool) {
#pragma clang fp contract(fast)
    return r.cD[k] + r.cW[k] * uW + r.cE[k] * uE + r.cS[k] * uS + r.cN[k] * uN;
  }
  template <int C>
  __device__ static double norm(const Head<C>& h, const Body<C>& r, int k, float v) {
    return fabs((double)v - (double)r.st[k]);
  }
};

// The per-pair adjoints with their swept planes in LDS (LdsSegLayout): FP32
// with the coefficient planes in LDS too, FP64 with the coefficients read from
// the slot in global memory.  The same updates; only the loads differ.
struct AdjF32Lds : AdjF32 {
  template <int C, int O>
  __device__ static void load_head(const Ctx&, Head<C>& H, int base, int, const float* tin) {
#pragma unroll
    for (int k = 0; k < C; k++) { H.u[k] = lld(tin, base + k); H.aux[k] = 0.0f; }
  }
  template <int C, int O, bool NORM>
  __device__ static void load_body(const Ctx& cx, Body<C>& B, int base, int, const float* tst) {
#pragma unroll
    for (int k = 0; k < C; k++) {
      B.cW[k] = lld(cx.cW, base + k);
      B.cE[k] = lld(cx.cE, base + k);
      B.cS[k] = lld(cx.cS, base + k);
      B.cN[k] = lld(cx.cN, base + k);
      B.cD[k] = lld(cx.cD, base + k);
      B.st[k] = NORM ? lld(tst, base + k) : 0.0f;
    }
  }
};
struct AdjF64Lds : AdjF64 {
  template <int C, int O>
  __device__ static void load_head(const Ctx&, Head<C>& H, int base, int, const double* tin) {
#pragma unroll
    for (int k = 0; k < C; k++) { H.u[k] = lld(tin, base + k); H.aux[k] = 0.0; }
  }
  template <int C, int O, bool NORM>
  __device__ static void load_body(const Ctx& cx, Body<C>& B, int base, int, const double* tst) {
#pragma unroll
    for (int k = 0; k < C; k++) {
      B.a1p[k] = gld(cx.a1p, base + k);
      B.a2m[k] = gld(cx.a2m, base + k);
      B.b1p[k] = gld(cx.b1p, base + k);
      B.b2m[k] = gld(cx.b2m, base + k);
      B.dl[k] = gld(cx.dl, base + k);
      B.d[k] = gld(cx.d, base + k);
      B.st[k] = NORM ? lld(tst, base + k) : 0.0;
    }
  }
};

// Batched FP32 adjoint, face form (see AdjF64Face).  The face fluxes are stored
// pre-divided by dx and dy: g = (a1/dx, b1/dy), so d = max(gE.x,0) - min(g.x,0)
// + max(gN.y,0) - min(g.y,0) and Ta = (delta + max(g.x,0)*Ta_W - min(gE.x,0)*Ta_E
// + max(g.y,0)*Ta_S - min(gN.y,0)*Ta_N) / d.
// The receiver term delta is zero except in the cells of the event's receivers.
// With a row mask (rmask: per row, mw words of one bit per lane, in LDS) a lane
// reads its C delta values from the plane only in the rows where its columns
// hold a receiver share (delta_kernel zeroes and marks those segments); its
// other rows' raw buffer loads carry an offset past the plane, which the
// hardware answers with zeros and no memory access: the sweeps stream 16
// instead of 20 bytes per node.
struct AdjF32Face {
  using V = float;
  using A = float2;
  static constexpr bool kAux = true;
  using MaskLds = const __attribute__((address_space(3))) unsigned long long*;
  struct Ctx {
    const GridDev* g;
    int nx, ny;
    const float2* fxy[2];
    const float* dl[2];
    MaskLds rmask;
    int mw;
    int dl_bytes;  // one delta plane of the event
  };
  template <int C> struct Head { float u[C]; float2 aux[C]; };
  template <int C> struct Body { float dl[C], st[C]; };
  template <int C, int O>
  __device__ static void load_head(const Ctx& cx, Head<C>& H, int base, int, const float* tin) {
#pragma unroll
    for (int k = 0; k < C; k++) { H.u[k] = gld(tin, base + k); H.aux[k] = gld(cx.fxy[O], base + k); }
  }
  template <int C, int O, bool NORM>
  __device__ static void load_body(const Ctx& cx, Body<C>& B, int base, int j, const float* tst) {
    const int lane = threadIdx.x;
    // mw == 0: one word of all lanes for every row (no mask)
    const unsigned long long m = cx.rmask[cx.mw ? j * cx.mw + imin(lane >> 6, cx.mw - 1) : 0];
    const int off = ((m >> (lane & 63)) & 1ull) ? base * (int)sizeof(float) : kDropOffset;
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(const_cast<float*>(cx.dl[O]), cx.dl_bytes);
    if constexpr (C % 4 == 0) {
      typedef float v4 __attribute__((ext_vector_type(4)));
#pragma unroll
      for (int k = 0; k < C; k += 4) {
        const v4 v = __builtin_bit_cast(v4, __builtin_amdgcn_raw_buffer_load_b128(rs, off + k * 4, 0, 0));
        B.dl[k] = v.x;
        B.dl[k + 1] = v.y;
        B.dl[k + 2] = v.z;
        B.dl[k + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < C; k++)
        B.dl[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off + k * 4, 0, 0));
    }
#pragma unroll
    for (int k = 0; k < C; k++) B.st[k] = NORM ? gld(tst, base + k) : 0.0f;
  }
  __device__ static bool fixed(const Ctx& cx, int ix, int j) {
    return ix == 0 || ix >= cx.nx - 1 || j == 0 || j == cx.ny - 1;
  }
  template <int C>
  __device__ static float update(const Ctx&, const Head<C>& h, const Body<C>& r, int k, int, float uW, float2, bool,
                                 float uE, float2 aE, bool, float uS, float2, bool, float uN, float2 aN, bool) {
#pragma clang fp contract(fast)
    const float2 s = h.aux[k];
    const float a1p = fmaxf(s.x, 0.0f), a1m = fminf(s.x, 0.0f);
    const float a2p = fmaxf(aE.x, 0.0f), a2m = fminf(aE.x, 0.0f);
    const float b1p = fmaxf(s.y, 0.0f), b1m = fminf(s.y, 0.0f);
    const float b2p = fmaxf(aN.y, 0.0f), b2m = fminf(aN.y, 0.0f);
    const float d = (a2p - a1m) + (b2p - b1m);
    if (fabsf(d) < (float)kAdjEps) return 0.0f;
    const float e = r.dl[k] + (uW * a1p - uE * a2m) + (uS * b1p - uN * b2m);
    return e * __builtin_amdgcn_rcpf(d);
  }
  template <int C>
  __device__ static double norm(const Head<C>&, const Body<C>& r, int k, float v) {
    return fabs((double)v - (double)r.st[k]);
  }
};

// ============================================================================
// Per-solve set-up and finish
// ============================================================================

// src/eikonal_solver.cpp:67-131.  Analytic factor in FP64 for every node of the
// grid; stores it in the sweep precision (t0, t0x, t0y segmented), the FP64 T0
// (plain, FP32 path only) and the initial sweep value (tau or sigma).
template <typename V, typename U>
__device__ inline FixedBox fwd_init(const GridDev& g, int iper, double src_lon, double src_lat, V* t0s, V* t0xs,
                                    V* t0ys, U* u0, double* T0d, int& flags) {
  const int lane = threadIdx.x;
  const int nx = g.nx, ny = g.ny, P = g.P;
  const double x0 = src_lon * kDEG2RAD;
  const double y0 = src_lat * kDEG2RAD;
  const double dx = g.dx, dy = g.dy;
  int idx0 = (int)floor((x0 - g.xx[0]) / dx);
  int idy0 = (int)floor((y0 - g.yy[0]) / dy);
  idx0 = imax(0, imin(idx0, nx - 2));
  idy0 = imax(0, imin(idy0, ny - 2));
  const double r1 = dmin(1.0, (x0 - g.xx[idx0]) / dx);
  const double r2 = dmin(1.0, (y0 - g.yy[idy0]) / dy);
  const double fun0 = bilinear_cm(g.f_plain + (size_t)iper * g.plane, nx, idx0, idy0, r1, r2);

  // The fixed set {|(xx-x0)/dx| <= 1} x {|(yy-y0)/dy| <= 1} is an index box
  // (the test is monotone in the coordinate).
  const int nt = blockDim.x;  // 64 per wave of the work-group
  int bx0 = 1 << 30, bx1 = -1, by0 = 1 << 30, by1 = -1;
  for (int ix = lane; ix < nx; ix += nt)
    if (fabs((g.xx[ix] - x0) / dx) <= 1.0) { bx0 = imin(bx0, ix); bx1 = imax(bx1, ix); }
  for (int iy = lane; iy < ny; iy += nt)
    if (fabs((g.yy[iy] - y0) / dy) <= 1.0) { by0 = imin(by0, iy); by1 = imax(by1, iy); }
  FixedBox fb;
  fb.x0 = block_min_i(bx0);
  fb.x1 = block_max_i(bx1);
  fb.y0 = block_min_i(by0);
  fb.y1 = block_max_i(by1);
  if (fb.x0 <= fb.x1 && fb.y0 <= fb.y1 && (fb.x0 == 0 || fb.x1 == nx - 1 || fb.y0 == 0 || fb.y1 == ny - 1))
    flags |= FLAG_SRC_ON_BOUNDARY;

  constexpr bool kSigma = sizeof(V) == 4;
  // Correctly rounded sin/cos/acos (sat_crmath.hpp): the same bits as any
  // correctly rounded CPU evaluation.  sin/cos of the longitude offsets are
  // evaluated once per column (a thread owns at most kMaxCols columns: the
  // segmented pitch is at most kMaxC columns per lane of the work-group).
  constexpr int kMaxCols = kMaxC;
  double sdl[kMaxCols], cdl[kMaxCols];
#pragma unroll
  for (int m = 0; m < kMaxCols; m++) {
    const int q = lane + nt * m;
    sdl[m] = 0.0;
    cdl[m] = 1.0;
    if (q < nx) cr::sincos_cr(g.xx[q] - x0, sdl[m], cdl[m]);
  }
  double sy0, cy0;
  cr::sincos_cr(y0, sy0, cy0);
  for (int iy = 0; iy < ny; iy++) {
    double syy, cyy;
    cr::sincos_cr(g.yy[iy], syy, cyy);
#pragma unroll
    for (int m = 0; m < kMaxCols; m++) {
      const int q = lane + nt * m;
      if (q >= P) continue;
      const size_t o = (size_t)iy * P + q;
      if (q >= nx) {
        t0s[o] = V(0); t0xs[o] = V(0); t0ys[o] = V(0); u0[o] = kSigma ? U(9) : U(10);
        continue;
      }
      const double tmp = syy * sy0 + cyy * cy0 * cdl[m];
      const double tmp_c = dmax(-1.0, dmin(1.0, tmp));
      const double T0 = cr::acos_cr(tmp_c) * kR * fun0;
      const double sin_d = sqrt(dmax(0.0, 1.0 - tmp_c * tmp_c));
      double T0x, T0y;
      if (sin_d < 1e-14) {
        T0x = 0.0;
        T0y = 0.0;
      } else {
        T0x = cyy * cy0 * sdl[m] / sin_d * kR * fun0;
        T0y = -(cyy * sy0 - syy * cy0 * cdl[m]) / sin_d * kR * fun0;
      }
      t0s[o] = (V)T0;
      t0xs[o] = (V)T0x;
      t0ys[o] = (V)T0y;
      if (kSigma) T0d[q + (size_t)iy * nx] = T0;
      const bool fixed = (q >= fb.x0 && q <= fb.x1 && iy >= fb.y0 && iy <= fb.y1);
      u0[o] = kSigma ? (fixed ? U(0) : U(9)) : (fixed ? U(1) : U(10));
    }
  }
  return fb;
}

// :317 T = tau .* T0 (FP64) or T = T0 + sigma*T0 (FP32 sweeps), column-major.
template <typename V, typename U>
__device__ inline void fwd_finish(const GridDev& g, const U* u, const V* t0s, const double* T0d, double* T) {
  const int nx = g.nx, ny = g.ny;
  for (int iy = 0; iy < ny; iy++)
    for (int ix = threadIdx.x; ix < nx; ix += blockDim.x) {
      const int o = seg_index(g, ix, iy);
      const size_t p = ix + (size_t)iy * nx;
      if (sizeof(V) == 8) T[p] = (double)u[o] * (double)t0s[o];
      else T[p] = T0d[p] + (double)u[o] * T0d[p];
    }
}

// src/eikonal_solver.cpp:407-441 face fluxes at interior node (ix, iy).
struct Faces {
  double a1m, a1p, a2m, a2p, b1m, b1p, b2m, b2p;
};
__device__ inline Faces faces_at(const GridDev& g, const double* __restrict__ T, const double* __restrict__ a,
                                 const double* __restrict__ b, const double* __restrict__ c, int ix, int iy) {
  const int nx = g.nx;
  const double dx = g.dx, dy = g.dy;
#define TT(i, j) T[(i) + (size_t)(j) * nx]
#define AA(i, j) a[(i) + (size_t)(j) * nx]
#define BB(i, j) b[(i) + (size_t)(j) * nx]
#define CC(i, j) c[(i) + (size_t)(j) * nx]
  Faces F;
  const double a1 = -(TT(ix, iy) - TT(ix - 1, iy)) / dx * (AA(ix, iy) + AA(ix - 1, iy)) / 2.0
      + (TT(ix, iy + 1) - TT(ix, iy - 1) + TT(ix - 1, iy + 1) - TT(ix - 1, iy - 1)) / (4.0 * dy)
        * (CC(ix, iy) + CC(ix - 1, iy)) / 2.0;
  F.a1m = (a1 - fabs(a1)) / 2.0;
  F.a1p = (a1 + fabs(a1)) / 2.0;
  const double a2 = -(TT(ix + 1, iy) - TT(ix, iy)) / dx * (AA(ix + 1, iy) + AA(ix, iy)) / 2.0
      + (TT(ix + 1, iy + 1) - TT(ix + 1, iy - 1) + TT(ix, iy + 1) - TT(ix, iy - 1)) / (4.0 * dy)
        * (CC(ix + 1, iy) + CC(ix, iy)) / 2.0;
  F.a2m = (a2 - fabs(a2)) / 2.0;
  F.a2p = (a2 + fabs(a2)) / 2.0;
  const double b1 = -(TT(ix, iy) - TT(ix, iy - 1)) / dy * (BB(ix, iy) + BB(ix, iy - 1)) / 2.0
      + (TT(ix + 1, iy) - TT(ix - 1, iy) + TT(ix + 1, iy - 1) - TT(ix - 1, iy - 1)) / (4.0 * dx)
        * (CC(ix, iy) + CC(ix, iy - 1)) / 2.0;
  F.b1m = (b1 - fabs(b1)) / 2.0;
  F.b1p = (b1 + fabs(b1)) / 2.0;
  const double b2 = -(TT(ix, iy + 1) - TT(ix, iy)) / dy * (BB(ix, iy + 1) + BB(ix, iy)) / 2.0
      + (TT(ix + 1, iy + 1) - TT(ix - 1, iy + 1) + TT(ix + 1, iy) - TT(ix - 1, iy)) / (4.0 * dx)
        * (CC(ix, iy + 1) + CC(ix, iy)) / 2.0;
  F.b2m = (b2 - fabs(b2)) / 2.0;
  F.b2p = (b2 + fabs(b2)) / 2.0;
#undef TT
#undef AA
#undef BB
#undef CC
  return F;
}

// The x-face flux a1 at (ix, iy) and the y-face flux b1 (src/eikonal_solver.cpp:409-413,
// :423-427), operation for operation; a1 at (ix+1, iy) is a2 at (ix, iy) and b1 at
// (ix, iy+1) is b2 at (ix, iy) (:416-420, :430-434).
__device__ inline double face_a1(const GridDev& g, const double* __restrict__ T, const double* __restrict__ a,
                                 const double* __restrict__ c, int ix, int iy) {
  const int nx = g.nx;
#define TT(i, j) T[(i) + (size_t)(j) * nx]
#define AA(i, j) a[(i) + (size_t)(j) * nx]
#define CC(i, j) c[(i) + (size_t)(j) * nx]
  return -(TT(ix, iy) - TT(ix - 1, iy)) / g.dx * (AA(ix, iy) + AA(ix - 1, iy)) / 2.0
      + (TT(ix, iy + 1) - TT(ix, iy - 1) + TT(ix - 1, iy + 1) - TT(ix - 1, iy - 1)) / (4.0 * g.dy)
        * (CC(ix, iy) + CC(ix - 1, iy)) / 2.0;
#undef TT
#undef AA
#undef CC
}
__device__ inline double face_b1(const GridDev& g, const double* __restrict__ T, const double* __restrict__ b,
                                 const double* __restrict__ c, int ix, int iy) {
  const int nx = g.nx;
#define TT(i, j) T[(i) + (size_t)(j) * nx]
#define BB(i, j) b[(i) + (size_t)(j) * nx]
#define CC(i, j) c[(i) + (size_t)(j) * nx]
  return -(TT(ix, iy) - TT(ix, iy - 1)) / g.dy * (BB(ix, iy) + BB(ix, iy - 1)) / 2.0
      + (TT(ix + 1, iy) - TT(ix - 1, iy) + TT(ix + 1, iy - 1) - TT(ix - 1, iy - 1)) / (4.0 * g.dx)
        * (CC(ix, iy) + CC(ix, iy - 1)) / 2.0;
#undef TT
#undef BB
#undef CC
}

// FP64 adjoint set-up: a1p, a2m, b1p, b2m, dl = 0, d, ta0 (segmented, seg[0..6]).
__device__ inline void adj_setup_f64(const GridDev& g, int iper, const double* __restrict__ T, const Slot& ws) {
  const int nx = g.nx, ny = g.ny, P = g.P;
  const double* a = g.aA_plain + (size_t)iper * g.plane;
  const double* b = g.b_plain + (size_t)iper * g.plane;
  const double* c = g.c_plain + (size_t)iper * g.plane;
  double* S[7];
#pragma unroll
  for (int i = 0; i < 7; i++) S[i] = static_cast<double*>(ws.seg[i]);
  for (int iy = 0; iy < ny; iy++)
    for (int q = threadIdx.x; q < P; q += blockDim.x) {
      const size_t o = (size_t)iy * P + q;
      double v[7] = {0, 0, 0, 0, 0, 0, 0};
      if (q >= 1 && q < nx - 1 && iy >= 1 && iy < ny - 1) {
        const Faces F = faces_at(g, T, a, b, c, q, iy);
        v[0] = F.a1p; v[1] = F.a2m; v[2] = F.b1p; v[3] = F.b2m;
        v[5] = (F.a2p - F.a1m) / g.dx + (F.b2p - F.b1m) / g.dy;  // :479-481
        v[6] = 100.0;                                            // :446-449
      }
#pragma unroll
      for (int i = 0; i < 7; i++) S[i][o] = v[i];
    }
}

// FP32 adjoint set-up: cW, cE, cS, cN (segmented seg[0..3], float), cD = 0 (seg[4]),
// ta0 (seg[5]); the FP64 diagonal d into pl[2].  Ta = cD + cW*Ta_W + cE*Ta_E +
// cS*Ta_S + cN*Ta_N is src/eikonal_solver.cpp:483-490 divided through by d.
__device__ inline void adj_setup_f32(const GridDev& g, int iper, const double* __restrict__ T, const Slot& ws) {
  const int nx = g.nx, ny = g.ny, P = g.P;
  const double* a = g.aA_plain + (size_t)iper * g.plane;
  const double* b = g.b_plain + (size_t)iper * g.plane;
  const double* c = g.c_plain + (size_t)iper * g.plane;
  float* S[6];
#pragma unroll
  for (int i = 0; i < 6; i++) S[i] = static_cast<float*>(ws.seg[i]);
  double* dpl = ws.pl[2];
  for (int iy = 0; iy < ny; iy++)
    for (int q = threadIdx.x; q < P; q += blockDim.x) {
      const size_t o = (size_t)iy * P + q;
      float v[6] = {0, 0, 0, 0, 0, 0};
      if (q >= 1 && q < nx - 1 && iy >= 1 && iy < ny - 1) {
        const Faces F = faces_at(g, T, a, b, c, q, iy);
        const double d = (F.a2p - F.a1m) / g.dx + (F.b2p - F.b1m) / g.dy;
        dpl[q + (size_t)iy * nx] = d;
        if (!(fabs(d) < kAdjEps)) {
          const double rdx = 1.0 / (g.dx * d), rdy = 1.0 / (g.dy * d);
          v[0] = (float)(F.a1p * rdx);
          v[1] = (float)(-F.a2m * rdx);
          v[2] = (float)(F.b1p * rdy);
          v[3] = (float)(-F.b2m * rdy);
        }
        v[5] = 100.0f;
      }
#pragma unroll
      for (int i = 0; i < 6; i++) S[i][o] = v[i];
    }
}

// src/eikonal_solver.cpp:356-375 receivers spread onto delta, in receiver order,
// by lane 0.  FP64: into the dense dl plane (seg[4]).  FP32: into the persistent
// all-zero FP64 plane pl[3], then cD = dl/d at the touched nodes, then pl[3] is
// cleared again.
template <typename V>
__device__ inline void adj_delta(const GridDev& g, int nr, const double* rlon, const double* rlat, const double* adj,
                                 const Slot& ws) {
  if (threadIdx.x != 0) return;
  const int nx = g.nx, ny = g.ny;
  const double dx = g.dx, dy = g.dy;
  constexpr bool k64 = sizeof(V) == 8;
  double* dl = k64 ? static_cast<double*>(ws.seg[4]) : ws.pl[3];
  auto at = [&](int i, int j) -> size_t { return k64 ? (size_t)seg_index(g, i, j) : (size_t)i + (size_t)j * nx; };
  for (int pass = 0; pass < (k64 ? 1 : 3); pass++) {
    for (int ir = 0; ir < nr; ir++) {
      const double xrec = rlon[ir] * kDEG2RAD;
      const double yrec = rlat[ir] * kDEG2RAD;
      int idx0 = (int)floor((xrec - g.xx[0]) / dx);
      int idy0 = (int)floor((yrec - g.yy[0]) / dy);
      idx0 = imax(0, imin(idx0, nx - 2));
      idy0 = imax(0, imin(idy0, ny - 2));
      if (pass == 0) {
        const double r1 = dmin(1.0, (xrec - g.xx[idx0]) / dx);
        const double r2 = dmin(1.0, (yrec - g.yy[idy0]) / dy);
        const double area = dx * kR * cr::cos_cr(yrec) * dy * kR;
        const double w = adj[ir] / area;
        dl[at(idx0, idy0)] += w * (1 - r1) * (1 - r2);
        dl[at(idx0, idy0 + 1)] += w * (1 - r1) * r2;
        dl[at(idx0 + 1, idy0)] += w * r1 * (1 - r2);
        dl[at(idx0 + 1, idy0 + 1)] += w * r1 * r2;
      } else {
        for (int m = 0; m < 4; m++) {
          const int i = idx0 + (m >> 1), j = idy0 + (m & 1);
          const size_t p = (size_t)i + (size_t)j * nx;
          if (pass == 1) {
            const bool interior = i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1;
            const double d = ws.pl[2][p];
            static_cast<float*>(ws.seg[4])[seg_index(g, i, j)] =
                (interior && !(fabs(d) < kAdjEps)) ? (float)(dl[p] / d) : 0.0f;
          } else {
            dl[p] = 0.0;
          }
        }
      }
    }
  }
}

// src/eikonal_solver.cpp:514-536 in degrees; writes the masked Ta column-major.
template <typename V, typename O>
__device__ inline void mask_store(const GridDev& g, double src_lon, double src_lat, const V* __restrict__ ta,
                                  O* __restrict__ out) {
  const int nx = g.nx, ny = g.ny;
  const double dxd = g.lon[1] - g.lon[0];
  const double dyd = g.lat[1] - g.lat[0];
  const int ix_lo = imax(0, (int)floor((src_lon - 1.5 * dxd - g.lon[0]) / dxd) + 1);
  const int ix_hi = imin(nx - 1, (int)ceil((src_lon + 1.5 * dxd - g.lon[0]) / dxd) - 1);
  const int iy_lo = imax(0, (int)floor((src_lat - 1.5 * dyd - g.lat[0]) / dyd) + 1);
  const int iy_hi = imin(ny - 1, (int)ceil((src_lat + 1.5 * dyd - g.lat[0]) / dyd) - 1);
  for (int iy = 0; iy < ny; iy++)
    for (int ix = threadIdx.x; ix < nx; ix += blockDim.x) {
      const bool m = ix >= ix_lo && ix <= ix_hi && iy >= iy_lo && iy <= iy_hi;
      out[ix + (size_t)iy * nx] = m ? O(0) : (O)ta[seg_index(g, ix, iy)];
    }
}

}  // namespace sat
