// sat_surfdisp.hip -- surface-wave dispersion of layered columns on the GPU:
// surfdisp96 (src/surfker/surfdisp96.f, Computer Programs in Seismology,
// Herrmann & Russell, with its SurfATT modifications) for every column of a
// 3-D model, as SurfGrid::fwdsurf does column by column on the host
// (src/surf_grid.cpp:194-245 -> surfker::build_disp_req, src/surfdisp.cpp:60-139,
// and surfker::surfdisp, :428-462).
//
// One thread per column.  A column's periods run in order because each root
// search starts from the previous period's root (surfdisp96.f:281-305); the
// root search brackets a sign change of the period equation in steps of dc and
// refines it by interval halving / Neville iteration (getsol :405-491, nevill
// :583-695).  The period equation is the Haskell-Thomson SH propagator (Love,
// dltar1 :741-799) or Dunkin's compound P-SV matrix (Rayleigh, dltar4
// :805-895 with var :901-1008, normc :1014-1041, dnka :1047-1094), FP64 with
// exponent tracking.  The model is REAL*4 as in the reference: thickness,
// velocities and density are rounded to float before the solve (src/surfdisp.cpp:
// 434-444), the earth-flattening transform (sphere :497-576) and the starting
// value (gtsolh :379-397) compute in float, and the returned velocity is the
// root rounded to float (:327).  Every float/double conversion and every
// operation order of the Fortran is kept; the library is compiled without FMA
// contraction.  A column whose fundamental mode has no root gets flag 1 and
// zeros from that period on (the reference prints a message and stops the
// program, :343-372).
//
// Work per column: ~kmax * (10-30 period-equation evaluations) * nlayer layer
// steps of ~150 FP64 operations and one or two sin/cos or exp per wave type.
// Compute-bound FP64; no dense contraction, so no MFMA.  Layer arrays live in
// global memory laid out [field][layer][column] so a wave's loads of one layer
// are contiguous.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../include/surfatt_eik.h"

namespace {

constexpr int kMaxLayers = 512;
constexpr int kMaxPeriods = 1024;

struct Col {
  // layer fields of this column, stride ncol between layers
  float *d, *a, *b, *rho, *rtp, *dtp, *btp;
  size_t s;
  int mmax;
  __device__ float& D(int i) const { return d[(size_t)i * s]; }  // 0-based layer index
  __device__ float& A(int i) const { return a[(size_t)i * s]; }
  __device__ float& B(int i) const { return b[(size_t)i * s]; }
  __device__ float& R(int i) const { return rho[(size_t)i * s]; }
  __device__ float& RT(int i) const { return rtp[(size_t)i * s]; }
  __device__ float& DT(int i) const { return dtp[(size_t)i * s]; }
  __device__ float& BT(int i) const { return btp[(size_t)i * s]; }
};

__device__ inline double dsign1(double x) { return copysign(1.0, x); }  // Fortran dsign(1d0, x)

// var (surfdisp96.f:901-1008): eigenfunction terms of one layer; only the
// outputs the callers use are returned.
struct VarOut {
  double w, cosp, a0, cpcq, cpy, cpz, cqw, cqx, xy, xz, wy, wz;
};
__device__ inline VarOut var(double p, double q, double ra, double rb, double wvno, double xka, double xkb,
                             double dpth) {
  VarOut o;
  double pex = 0.0, sex = 0.0;
  double x = 0.0, cosp = 0.0, w = 0.0, sinp;
  if (wvno < xka) {
    sinp = sin(p);
    w = sinp / ra;
    x = -(ra * sinp);
    cosp = cos(p);
  } else if (wvno == xka) {
    cosp = 1.0;
    w = dpth;
    x = 0.0;
  } else if (wvno > xka) {
    pex = p;
    double fac = 0.0;
    if (p < 16.0) fac = exp(-2.0 * p);
    cosp = (1.0 + fac) * 0.5;
    sinp = (1.0 - fac) * 0.5;
    w = sinp / ra;
    x = ra * sinp;
  }
  double y = 0.0, z = 0.0, cosq = 0.0, sinq;
  if (wvno < xkb) {
    sinq = sin(q);
    y = sinq / rb;
    z = -(rb * sinq);
    cosq = cos(q);
  } else if (wvno == xkb) {
    cosq = 1.0;
    y = dpth;
    z = 0.0;
  } else if (wvno > xkb) {
    sex = q;
    double fac = 0.0;
    if (q < 16.0) fac = exp(-2.0 * q);
    cosq = (1.0 + fac) * 0.5;
    sinq = (1.0 - fac) * 0.5;
    y = sinq / rb;
    z = rb * sinq;
  }
  const double exa = pex + sex;
  o.a0 = 0.0;
  if (exa < 60.0) o.a0 = exp(-exa);
  o.cpcq = cosp * cosq;
  o.cpy = cosp * y;
  o.cpz = cosp * z;
  o.cqw = cosq * w;
  o.cqx = cosq * x;
  o.xy = x * y;
  o.xz = x * z;
  o.wy = w * y;
  o.wz = w * z;
  o.w = w;
  o.cosp = cosp;
  return o;
}

// dltar4 (surfdisp96.f:805-895): Rayleigh period equation, Dunkin's compound
// matrices from the half-space up, normalised per layer (normc).
__device__ double dltar4(double wvno, double omga, const Col& m, int llw) {
  double omega = omga;
  if (omega < 1.0e-4) omega = 1.0e-4;
  const double wvno2 = wvno * wvno;
  const int mm = m.mmax - 1;  // 0-based half-space
  double xka = omega / (double)m.A(mm);
  double xkb = omega / (double)m.B(mm);
  double wvnop = wvno + xka, wvnom = fabs(wvno - xka);
  double ra = sqrt(wvnop * wvnom);
  wvnop = wvno + xkb;
  wvnom = fabs(wvno - xkb);
  double rb = sqrt(wvnop * wvnom);
  double t = (double)m.B(mm) / omega;
  double gammk = 2.0 * t * t;
  double gam = gammk * wvno2;
  double gamm1 = gam - 1.0;
  double rho1 = (double)m.R(mm);
  double e[5];
  e[0] = rho1 * rho1 * (gamm1 * gamm1 - gam * gammk * ra * rb);
  e[1] = -(rho1 * ra);
  e[2] = rho1 * (gamm1 - gammk * ra * rb);
  e[3] = rho1 * rb;
  e[4] = wvno2 - ra * rb;
  // the next layer's values are loaded one layer ahead, so their loads are in
  // flight while the current layer's matrix is formed
  float nA = m.A(mm - 1), nB = m.B(mm - 1), nD = m.D(mm - 1), nR = m.R(mm - 1);
  for (int l = mm - 1; l >= llw - 1; l--) {
    const float lA = nA, lB = nB, lD = nD, lR = nR;
    if (l >= 1) {
      nA = m.A(l - 1);
      nB = m.B(l - 1);
      nD = m.D(l - 1);
      nR = m.R(l - 1);
    }
    xka = omega / (double)lA;
    xkb = omega / (double)lB;
    t = (double)lB / omega;
    gammk = 2.0 * t * t;
    gam = gammk * wvno2;
    wvnop = wvno + xka;
    wvnom = fabs(wvno - xka);
    ra = sqrt(wvnop * wvnom);
    wvnop = wvno + xkb;
    wvnom = fabs(wvno - xkb);
    rb = sqrt(wvnop * wvnom);
    const double dpth = (double)lD;
    rho1 = (double)lR;
    const double p = ra * dpth, q = rb * dpth;
    const VarOut v = var(p, q, ra, rb, wvno, xka, xkb, dpth);
    // dnka (surfdisp96.f:1047-1094)
    const double one = 1.0, two = 2.0;
    gamm1 = gam - one;
    const double twgm1 = gam + gamm1;
    const double gmgmk = gam * gammk;
    const double gmgm1 = gam * gamm1;
    const double gm1sq = gamm1 * gamm1;
    const double rho2 = rho1 * rho1;
    const double a0pq = v.a0 - v.cpcq;
    double ca[5][5];  // ca[j][i] = Fortran ca(j+1, i+1)
    ca[0][0] = v.cpcq - two * gmgm1 * a0pq - gmgmk * v.xz - wvno2 * gm1sq * v.wy;
    ca[0][1] = (wvno2 * v.cpy - v.cqx) / rho1;
    ca[0][2] = -((twgm1 * a0pq + gammk * v.xz + wvno2 * gamm1 * v.wy) / rho1);
    ca[0][3] = (v.cpz - wvno2 * v.cqw) / rho1;
    ca[0][4] = -((two * wvno2 * a0pq + v.xz + wvno2 * wvno2 * v.wy) / rho2);
    ca[1][0] = (gmgmk * v.cpz - gm1sq * v.cqw) * rho1;
    ca[1][1] = v.cpcq;
    ca[1][2] = gammk * v.cpz - gamm1 * v.cqw;
    ca[1][3] = -v.wz;
    ca[1][4] = ca[0][3];
    ca[3][0] = (gm1sq * v.cpy - gmgmk * v.cqx) * rho1;
    ca[3][1] = -v.xy;
    ca[3][2] = gamm1 * v.cpy - gammk * v.cqx;
    ca[3][3] = ca[1][1];
    ca[3][4] = ca[0][1];
    ca[4][0] = -((two * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * v.xz + gm1sq * gm1sq * v.wy) * rho2);
    ca[4][1] = ca[3][0];
    ca[4][2] = -((gammk * gamm1 * twgm1 * a0pq + gam * gammk * gammk * v.xz + gamm1 * gm1sq * v.wy) * rho1);
    ca[4][3] = ca[1][0];
    ca[4][4] = ca[0][0];
    const double tt = -(two * wvno2);
    ca[2][0] = tt * ca[4][2];
    ca[2][1] = tt * ca[3][2];
    ca[2][2] = v.a0 + two * (v.cpcq - ca[0][0]);
    ca[2][3] = tt * ca[1][2];
    ca[2][4] = tt * ca[0][2];
    double ee[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      double cr = 0.0;
#pragma unroll
      for (int j = 0; j < 5; j++) cr = cr + e[j] * ca[j][i];
      ee[i] = cr;
    }
    // normc (surfdisp96.f:1014-1041)
    double t1 = 0.0;
#pragma unroll
    for (int i = 0; i < 5; i++)
      if (fabs(ee[i]) > t1) t1 = fabs(ee[i]);
    if (t1 < 1.0e-40) t1 = 1.0;
#pragma unroll
    for (int i = 0; i < 5; i++) e[i] = ee[i] / t1;
  }
  if (llw != 1) {  // water layer on top
    xka = omega / (double)m.A(0);
    wvnop = wvno + xka;
    wvnom = fabs(wvno - xka);
    ra = sqrt(wvnop * wvnom);
    const double dpth = (double)m.D(0);
    rho1 = (double)m.R(0);
    const double p = ra * dpth;
    const double znul = 1.0e-05;
    const VarOut v = var(p, znul, ra, znul, wvno, xka, znul, dpth);
    const double w0 = -(rho1 * v.w);
    return v.cosp * e[0] + w0 * e[1];
  }
  return e[0];
}

// dltar1 (surfdisp96.f:741-799): Love period equation.
__device__ double dltar1(double wvno, double omega, const Col& m, int llw) {
  const int mm = m.mmax - 1;
  double beta1 = (double)m.B(mm);
  double rho1 = (double)m.R(mm);
  double xkb = omega / beta1;
  double wvnop = wvno + xkb, wvnom = fabs(wvno - xkb);
  double rb = sqrt(wvnop * wvnom);
  double e1 = rho1 * rb;
  double e2 = 1.0 / (beta1 * beta1);
  float nB = m.B(mm - 1), nD = m.D(mm - 1), nR = m.R(mm - 1);  // one layer ahead
  for (int l = mm - 1; l >= llw - 1; l--) {
    const float lB = nB, lD = nD, lR = nR;
    if (l >= 1) {
      nB = m.B(l - 1);
      nD = m.D(l - 1);
      nR = m.R(l - 1);
    }
    beta1 = (double)lB;
    rho1 = (double)lR;
    const double xmu = rho1 * beta1 * beta1;
    xkb = omega / beta1;
    wvnop = wvno + xkb;
    wvnom = fabs(wvno - xkb);
    rb = sqrt(wvnop * wvnom);
    const double q = (double)lD * rb;
    double cosq, y, z, sinq;
    if (wvno < xkb) {
      sinq = sin(q);
      y = sinq / rb;
      z = -(rb * sinq);
      cosq = cos(q);
    } else if (wvno == xkb) {
      cosq = 1.0;
      y = (double)lD;
      z = 0.0;
    } else {
      double fac = 0.0;
      if (q < 16.0) fac = exp(-2.0 * q);
      cosq = (1.0 + fac) * 0.5;
      sinq = (1.0 - fac) * 0.5;
      y = sinq / rb;
      z = rb * sinq;
    }
    const double e10 = e1 * cosq + e2 * xmu * z;
    const double e20 = e1 * y / xmu + e2 * cosq;
    double xnor = fabs(e10);
    const double ynor = fabs(e20);
    if (ynor > xnor) xnor = ynor;
    if (xnor < 1.0e-40) xnor = 1.0;
    e1 = e10 / xnor;
    e2 = e20 / xnor;
  }
  return e1;
}

__device__ inline double dltar(double wvno, double omega, int kk, const Col& m, int llw) {
  return kk == 1 ? dltar1(wvno, omega, m, llw) : dltar4(wvno, omega, m, llw);
}

// half (surfdisp96.f:697-712)
__device__ inline void half(double c1, double c2, double& c3, double& del3, double omega, int ifunc, const Col& m,
                            int llw) {
  c3 = 0.5 * (c1 + c2);
  const double wvno = omega / c3;
  del3 = dltar(wvno, omega, ifunc, m, llw);
}

// nevill (surfdisp96.f:583-695)
__device__ double nevill(double t, double c1, double c2, double del1, double del2, int ifunc, const Col& m, int llw,
                         double twopi) {
  double x[20], y[20];
  const double omega = twopi / t;
  double c3, del3;
  half(c1, c2, c3, del3, omega, ifunc, m, llw);
  int nev = 1, nctrl = 1, mm = 1;
  const double s_001 = (double)0.01f;  // REAL*4 literal 0.01 in a double expression
  while (true) {
    nctrl = nctrl + 1;
    if (nctrl >= 100) break;
    if (c3 < fmin(c1, c2) || c3 > fmax(c1, c2)) {
      nev = 0;
      half(c1, c2, c3, del3, omega, ifunc, m, llw);
    }
    const double s13 = del1 - del3;
    const double s32 = del3 - del2;
    if (dsign1(del3) * dsign1(del1) < 0.0) {
      c2 = c3;
      del2 = del3;
    } else {
      c1 = c3;
      del1 = del3;
    }
    if (fabs(c1 - c2) <= 1.0e-6 * c1) break;
    if (dsign1(s13) != dsign1(s32)) nev = 0;
    const double ss1 = fabs(del1);
    const double s1 = s_001 * ss1;
    const double ss2 = fabs(del2);
    const double s2 = s_001 * ss2;
    if (s1 > ss2 || s2 > ss1 || nev == 0) {
      half(c1, c2, c3, del3, omega, ifunc, m, llw);
      nev = 1;
      mm = 1;
    } else {
      if (nev == 2) {
        x[mm] = c3;  // x(m+1)
        y[mm] = del3;
      } else {
        x[0] = c1;
        y[0] = del1;
        x[1] = c2;
        y[1] = del2;
        mm = 1;
      }
      bool bad = false;
      for (int kk = 1; kk <= mm; kk++) {
        const int j = mm - kk + 1;  // Fortran index
        const double denom = y[mm] - y[j - 1];
        if (fabs(denom) < 1.0e-10 * fabs(y[mm])) {
          bad = true;
          break;
        }
        x[j - 1] = (-(y[j - 1] * x[j]) + y[mm] * x[j - 1]) / denom;
      }
      if (!bad) {
        c3 = x[0];
        const double wvno = omega / c3;
        del3 = dltar(wvno, omega, ifunc, m, llw);
        nev = 2;
        mm = mm + 1;
        if (mm > 10) mm = 10;
      } else {
        half(c1, c2, c3, del3, omega, ifunc, m, llw);
        nev = 1;
        mm = 1;
      }
    }
  }
  return c3;
}

// getsol (surfdisp96.f:399-491), for G lanes that solve one column together.
// del1st is the period-equation value saved at the first period of a mode (a
// SAVEd common in the reference).
//
// The bracketing walk steps c by dc until the period equation changes sign;
// the points it visits depend only on c, dc, clow and the direction, not on
// the equation's values.  So the G lanes evaluate the next G points at once
// (lane g the g-th), exchange the values, and replay the walk over them in
// order: the first sign change, or the first bound left, ends it exactly where
// the one-point-at-a-time walk would, with the same points and values; values
// past that point are discarded.  G = 1 is the reference's walk.  Neville /
// halving refinement is sequential and all G lanes run it redundantly.
//
// max_steps (twice the walk across [cm, betmx + dc], set by the caller) only
// stops a walk that inputs outside the solver's domain would make endless,
// reporting no root.
template <int G>
__device__ int getsol(double t1, double& c1, double clow, double dc, double cm, float betmx, int ifunc, int ifirst,
                      double& del1st, const Col& m, int llw, int max_steps, int g) {
  const double twopi = 2.0 * 3.141592653589793;
  double omega = twopi / t1;
  double wvno = omega / c1;
  double del1 = dltar(wvno, omega, ifunc, m, llw);
  if (ifirst == 1) del1st = del1;
  const double plmn = dsign1(del1st) * dsign1(del1);
  int idir;
  if (ifirst == 1)
    idir = +1;
  else if (plmn >= 0.0)
    idir = +1;
  else
    idir = -1;
  int steps = 0;
  for (;;) {
    // the next G points of the walk from (c1, idir), with the clow resets
    double pts[G];
    {
      double c = c1;
      int dir = idir;
#pragma unroll
      for (int i = 0; i < G; i++) {
        double c2 = dir > 0 ? c + dc : c - dc;
        if (c2 <= clow) {
          dir = +1;
          c = clow;
          c2 = c + dc;  // clow + dc > clow: the walk goes on upward
        }
        pts[i] = c2;
        c = c2;
      }
    }
    double mine = pts[0];
#pragma unroll
    for (int i = 1; i < G; i++)
      if (g == i) mine = pts[i];
    omega = twopi / t1;
    wvno = omega / mine;
    const double dmine = dltar(wvno, omega, ifunc, m, llw);
    double dels[G];
#pragma unroll
    for (int i = 0; i < G; i++) dels[i] = (G == 1) ? dmine : __shfl(dmine, i, G);
    // replay the walk over the G points
    double c = c1;
    int dir = idir;
#pragma unroll
    for (int i = 0; i < G; i++) {
      if (++steps > max_steps) return -1;
      double c2 = dir > 0 ? c + dc : c - dc;
      if (c2 <= clow) {
        dir = +1;
        c = clow;
        if (++steps > max_steps) return -1;
        c2 = c + dc;
      }
      const double del2 = dels[i];
      if (dsign1(del1) != dsign1(del2)) {
        const double cn = nevill(t1, c, c2, del1, del2, ifunc, m, llw, twopi);
        c1 = cn;
        if (c1 > (double)betmx) return -1;
        return 1;
      }
      c = c2;
      del1 = del2;
      if (c < cm) {
        c1 = c;
        return -1;
      }
      if (c >= ((double)betmx + dc)) {
        c1 = c;
        return -1;
      }
    }
    c1 = c;
    idir = dir;
  }
}

// gtsolh (surfdisp96.f:379-397): REAL*4 Newton iterations for the half-space
// Rayleigh velocity.
__device__ float gtsolh(float a, float b) {
  float c = 0.95f * b;
  for (int i = 0; i < 5; i++) {
    const float gamma = b / a;
    const float kappa = c / b;
    const float k2 = kappa * kappa;
    const float gk = gamma * kappa;
    const float gk2 = gk * gk;
    const float fac1 = sqrtf(1.0f - gk2);
    const float fac2 = sqrtf(1.0f - k2);
    const float fr = (2.0f - k2) * (2.0f - k2) - 4.0f * fac1 * fac2;
    float frp = -(4.0f * (2.0f - k2) * kappa) + 4.0f * fac2 * gamma * gamma * kappa / fac1 + 4.0f * fac1 * kappa / fac2;
    frp = frp / b;
    c = c - fr / frp;
  }
  return c;
}

struct DispDev {
  int ncol, mmax, kmax, iflsph, iwave, mode, igr;
  const double* t;   // [kmax]
  float* lay;        // [7][mmax][ncol]: d a b rho rtp dtp btp
  double* c;         // [kmax][ncol]
  double* cb;        // [kmax][ncol]
  double* cg;        // [ncol][kmax]
  int* flags;        // [ncol]
  float* minvs;      // [ncol] minval of the input vs
  double* cstart;    // [ncol] starting phase velocity cc (surfdisp96.f:217-236)
  float* betmx;      // [ncol] largest shear velocity after flattening
  int* llw;          // [ncol] 1, or 2 with a water layer on top
};

__device__ inline Col column(const DispDev& g, int col) {
  const size_t s = (size_t)g.ncol, lm = (size_t)g.mmax * s;
  return Col{g.lay + col, g.lay + lm + col, g.lay + 2 * lm + col, g.lay + 3 * lm + col, g.lay + 4 * lm + col,
             g.lay + 5 * lm + col, g.lay + 6 * lm + col, s, g.mmax};
}

// surfdisp96.f:98-236 for one column: earth flattening (sphere with iflag 0,
// then the density mapping of the requested wave type), the extremal shear
// velocities and the starting value.  Writes the transformed layers back and
// the column's cc, betmx and llw.
__global__ void prep_kernel(DispDev g) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= g.ncol) return;
  const Col m = column(g, col);
  const int mmax = g.mmax;
  int llw = 1;
  if (m.B(0) <= 0.0f) llw = 2;
  float dhalf = 0.0f;
  if (g.iflsph == 1) {  // sphere(0, 0)
    const double ar = 6370.0;
    double dr = 0.0, r0 = ar;
    m.D(mmax - 1) = 1.0f;
    for (int i = 0; i < mmax; i++) {
      m.DT(i) = m.D(i);
      m.RT(i) = m.R(i);
    }
    for (int i = 0; i < mmax; i++) {
      dr = dr + (double)m.D(i);
      const double r1 = ar - dr;
      const double z0 = ar * log(ar / r0);
      const double z1 = ar * log(ar / r1);
      m.D(i) = (float)(z1 - z0);
      const double tmp = (ar + ar) / (r0 + r1);
      m.A(i) = (float)((double)m.A(i) * tmp);
      m.B(i) = (float)((double)m.B(i) * tmp);
      m.BT(i) = (float)tmp;
      r0 = r1;
    }
    dhalf = m.D(mmax - 1);
    m.D(mmax - 1) = 0.0f;
  }
  float betmx = -1.e20f, betmn = 1.e20f;
  int jmn = 0, jsol = 1;
  for (int i = 0; i < mmax; i++) {
    if (m.B(i) > 0.01f && m.B(i) < betmn) {
      betmn = m.B(i);
      jmn = i;
      jsol = 1;
    } else if (m.B(i) <= 0.01f && m.A(i) < betmn) {
      betmn = m.A(i);
      jmn = i;
      jsol = 0;
    }
    if (m.B(i) > betmx) betmx = m.B(i);
  }
  const int ifunc = g.iwave;  // 1 Love, 2 Rayleigh: the one wave type requested
  if (g.iflsph == 1) {        // sphere(ifunc, 1): density mapping
    m.D(mmax - 1) = dhalf;
    for (int i = 0; i < mmax; i++) {
      if (ifunc == 1) {
        const float bt = m.BT(i);
        const float b2 = bt * bt, b4 = b2 * b2;
        m.R(i) = m.RT(i) * (1.0f / (b4 * bt));
      } else {
        m.R(i) = m.RT(i) * (float)pow((double)m.BT(i), (double)-2.275f);
      }
    }
    m.D(mmax - 1) = 0.0f;
  }
  float cc1;
  if (jsol == 0)
    cc1 = betmn;
  else
    cc1 = gtsolh(m.A(jmn), m.B(jmn));
  cc1 = 0.95f * cc1;
  cc1 = 0.90f * cc1;
  g.cstart[col] = (double)cc1;
  g.betmx[col] = betmx;
  g.llw[col] = llw;
}

// surfdisp96.f:237-374 for one column, G lanes per column (getsol).  The lanes
// of a column hold the same state and write the same values.
template <int G>
__global__ __launch_bounds__(64) void surfdisp_kernel(DispDev g) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int col = tid / G, lane = tid % G;
  if (col >= g.ncol) return;  // whole groups: G divides the block size
  const size_t s = (size_t)g.ncol;
  const Col m = column(g, col);
  const int kmax = g.kmax;
  double* cgc = g.cg + (size_t)col * kmax;
  auto C = [&](int k) -> double& { return g.c[(size_t)k * s + col]; };   // 0-based period
  auto CB = [&](int k) -> double& { return g.cb[(size_t)k * s + col]; };
  int flag = 0;
  const float sone0 = 1.500f;
  const float vsmin = g.minvs[col];
  const float ddc0 = vsmin < 1.0f ? 0.001f : (vsmin < 2.0f ? 0.002f : 0.005f);
  const float h0 = 0.005f;
  const int llw = g.llw[col];
  const float betmx = g.betmx[col];
  const double one = 1.0e-2;
  const int ifunc = g.iwave;
  const float ddc = ddc0;
  float sone = sone0;
  const float h = h0;
  if (sone < 0.01f) sone = 2.0f;
  const double onea = (double)sone;
  const double cc = g.cstart[col];
  double dc = (double)ddc;
  dc = fabs(dc);
  double c1 = cc;
  const double cm = cc;
  for (int i = 0; i < kmax; i++) {
    CB(i) = 0.0;
    C(i) = 0.0;
  }
  double del1st = 0.0;
  const double span = ((double)betmx + dc - cm) / dc;
  const int max_steps = (span > 0.0 && span < 1.0e7) ? 2 * (int)span + 64 : 64;
  int ift = 999;
  for (int iq = 1; iq <= g.mode; iq++) {
    const int is = 0, ie = kmax - 1;
    int kfail = -1;
    for (int k = is; k <= ie; k++) {
      if (k + 1 >= ift) {
        kfail = k;
        break;
      }
      double t1 = g.t[k];
      float t1a, t1b = 0.0f;
      if (g.igr > 0) {
        t1a = (float)(t1 / (double)(1.0f + h));
        t1b = (float)(t1 / (double)(1.0f - h));
        t1 = (double)t1a;
      } else {
        t1a = (float)t1;
      }
      double clow;
      int ifirst;
      if (k == is && iq == 1) {
        c1 = cc;
        clow = cc;
        ifirst = 1;
      } else if (k == is && iq > 1) {
        c1 = C(is) + one * dc;
        clow = c1;
        ifirst = 1;
      } else if (k > is && iq > 1) {
        ifirst = 0;
        clow = C(k) + one * dc;
        c1 = C(k - 1);
        if (c1 < clow) c1 = clow;
      } else {
        ifirst = 0;
        c1 = C(k - 1) - onea * dc;
        clow = cm;
      }
      int iret = getsol<G>(t1, c1, clow, dc, cm, betmx, ifunc, ifirst, del1st, m, llw, max_steps, lane);
      if (iret == -1) {
        kfail = k;
        break;
      }
      C(k) = c1;
      if (g.igr > 0) {
        t1 = (double)t1b;
        ifirst = 0;
        clow = CB(k) + one * dc;
        c1 = c1 - onea * dc;
        iret = getsol<G>(t1, c1, clow, dc, cm, betmx, ifunc, ifirst, del1st, m, llw, max_steps, lane);
        if (iret == -1) c1 = C(k);
        CB(k) = c1;
      } else {
        c1 = 0.0;
      }
      const float cc0 = (float)C(k);
      const float ccb = (float)c1;
      if (g.igr == 0) {
        cgc[k] = (double)cc0;
      } else {
        const float gvel = (1.0f / t1a - 1.0f / t1b) / (1.0f / (t1a * cc0) - 1.0f / (t1b * ccb));
        cgc[k] = (double)gvel;
      }
    }
    if (kfail >= 0) {
      if (iq == 1) flag = 1;  // no root in the fundamental mode: the reference stops here
      ift = kfail + 1;
      for (int i = kfail; i <= ie; i++) cgc[i] = 0.0;
      if (iq == 1) break;
    }
  }
  if (g.flags) g.flags[col] = flag;
}

// Lanes per column: the most (up to 8) that keep columns x lanes within one
// wave per SIMD.  More lanes shorten a column's walk but add redundant
// refinement work, so they pay only while each wave has its SIMD to itself
// (measured: 3,233 columns run twice as fast with 8 lanes as with 1; 40,401
// columns -- 0.6 waves per SIMD with 1 lane -- 16 % slower with 2 and 26 %
// with 4).
int lanes_per_column(int ncol) {
  const long target = 1024L * 64;  // 256 CUs x 4 SIMDs x one wave of 64 lanes
  if ((long)ncol * 8 <= target) return 8;
  if ((long)ncol * 4 <= target) return 4;
  if ((long)ncol * 2 <= target) return 2;
  return 1;
}

// SURFATT_SURFDISP_LANES=1|2|4|8 fixes the lanes per column (the results do
// not depend on it; tests compare the settings).
int solve_lanes(int ncol) {
  if (const char* e = std::getenv("SURFATT_SURFDISP_LANES")) {
    const int v = std::atoi(e);
    if (v == 1 || v == 2 || v == 4 || v == 8) return v;
  }
  return lanes_per_column(ncol);
}

hipError_t launch_solve(const DispDev& g, int G) {
  const long threads = (long)g.ncol * G;
  const dim3 grid((unsigned)((threads + 63) / 64)), block(64);
  switch (G) {
    case 8: hipLaunchKernelGGL(surfdisp_kernel<8>, grid, block, 0, nullptr, g); break;
    case 4: hipLaunchKernelGGL(surfdisp_kernel<4>, grid, block, 0, nullptr, g); break;
    case 2: hipLaunchKernelGGL(surfdisp_kernel<2>, grid, block, 0, nullptr, g); break;
    default: hipLaunchKernelGGL(surfdisp_kernel<1>, grid, block, 0, nullptr, g); break;
  }
  return hipGetLastError();
}

// refine_request (src/surfdisp.cpp:60-111) on a column of the 3-D model: nz
// nodes -> nz + 1 layers (the last node's layer repeated, then a half-space
// of zero thickness), Brocher vp and density where they are not given
// (include/utils.h:185-187, :223-225), rounded to float (:434-444).
struct LayerDev {
  int nx, ny, nz, mmax;
  const double* z;     // [nz]
  const double* vs3d;  // [nx][ny][nz]
  const double* vp3d;  // may be null
  const double* rho3d;
  float* lay;
  float* minvs;
  int colmajor_xy;  // 1: columns ordered ix + iy*nx (the batch's planes); 0: ix*ny + iy
};

__global__ void layers_kernel(LayerDev g) {
  const int ncol = g.nx * g.ny;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  int ix, iy;
  if (g.colmajor_xy) {
    iy = col / g.nx;
    ix = col - iy * g.nx;
  } else {
    ix = col / g.ny;
    iy = col - ix * g.ny;
  }
  const size_t base = ((size_t)ix * g.ny + iy) * g.nz;
  const size_t s = ncol, lm = (size_t)g.mmax * s;
  float* d = g.lay;
  float* a = g.lay + lm;
  float* b = g.lay + 2 * lm;
  float* r = g.lay + 3 * lm;
  float vmin = 3.4e38f;
  double prev_thk = 0.0;
  for (int k = 0; k < g.nz; k++) {
    const double vs = g.vs3d[base + k];
    double vp, rho;
    if (g.vp3d) {
      vp = g.vp3d[base + k];
      rho = g.rho3d[base + k];
    } else {
      vp = 0.9409 + 2.0947 * vs - 0.8206 * vs * vs + 0.2683 * vs * vs * vs - 0.0251 * vs * vs * vs * vs;
      rho = 1.6612 * vp - 0.4721 * vp * vp + 0.0671 * vp * vp * vp - 0.0043 * vp * vp * vp * vp +
            0.000106 * vp * vp * vp * vp * vp;
    }
    double thk;
    if (k == g.nz - 1)
      thk = (g.nz > 1) ? prev_thk : 0.0;
    else
      thk = g.z[k + 1] - g.z[k];
    prev_thk = thk;
    d[(size_t)k * s + col] = (float)thk;
    a[(size_t)k * s + col] = (float)vp;
    b[(size_t)k * s + col] = (float)vs;
    r[(size_t)k * s + col] = (float)rho;
    vmin = fminf(vmin, (float)vs);
  }
  const int h = g.nz;  // half-space: the last node's properties, zero thickness
  d[(size_t)h * s + col] = 0.0f;
  a[(size_t)h * s + col] = a[(size_t)(h - 1) * s + col];
  b[(size_t)h * s + col] = b[(size_t)(h - 1) * s + col];
  r[(size_t)h * s + col] = r[(size_t)(h - 1) * s + col];
  g.minvs[col] = vmin;
}

// explicit layers [ncol][nlayer] -> the [field][layer][col] layout
__global__ void copy_layers_kernel(int ncol, int mmax, const float* thk, const float* vp, const float* vs,
                                   const float* rho, float* lay, float* minvs) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const size_t s = ncol, lm = (size_t)mmax * s;
  float vmin = 3.4e38f;
  for (int i = 0; i < mmax; i++) {
    const size_t src = (size_t)col * mmax + i;
    lay[(size_t)i * s + col] = thk[src];
    lay[lm + (size_t)i * s + col] = vp[src];
    lay[2 * lm + (size_t)i * s + col] = vs[src];
    lay[3 * lm + (size_t)i * s + col] = rho[src];
    vmin = fminf(vmin, vs[src]);
  }
  minvs[col] = vmin;
}

#define D_CHECK(x)                                                                 \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "[surfatt_eik] %s: %s\n", #x, hipGetErrorString(e_));  \
      rc = (e_ == hipErrorOutOfMemory) ? SAT_ENOMEM : SAT_EDEVICE;                 \
      goto done;                                                                   \
    }                                                                              \
  } while (0)

int check_common(int mmax, int kmax, int iflsph, int iwave, int mode, int igr, const double* t) {
  if (mmax < 2 || mmax > kMaxLayers || kmax < 1 || kmax > kMaxPeriods) return SAT_EINVAL;
  if ((iflsph != 0 && iflsph != 1) || (iwave != 1 && iwave != 2) || mode < 1 || igr < 0 || !t) return SAT_EINVAL;
  for (int k = 0; k < kmax; k++)
    if (!(t[k] > 0.0) || std::isinf(t[k])) return SAT_EINVAL;
  return SAT_OK;
}

// Layer values the solver's domain needs: finite, thickness >= 0, vp > 0,
// vs >= 0 (0: a water layer, allowed on top only), density > 0.
bool layer_ok(double thk, double vp, double vs, double rho, bool top) {
  if (!std::isfinite(thk) || !std::isfinite(vp) || !std::isfinite(vs) || !std::isfinite(rho)) return false;
  if (thk < 0.0 || vp <= 0.0 || vs < 0.0 || rho <= 0.0) return false;
  if (vs <= 0.01 && !top) return false;
  return true;
}

}  // namespace

extern "C" int sat_surfdisp96(int ncol, int nlayer, const float* thk, const float* vp, const float* vs,
                              const float* rho, int iflsph, int iwave, int mode, int igr, int kmax, const double* t,
                              double* cg, int* flags, double* kernel_ms) {
  if (ncol < 1 || !thk || !vp || !vs || !rho || !cg) return SAT_EINVAL;
  int rc = check_common(nlayer, kmax, iflsph, iwave, mode, igr, t);
  if (rc != SAT_OK) return rc;
  for (size_t i = 0; i < (size_t)ncol * nlayer; i++)
    if (!layer_ok(thk[i], vp[i], vs[i], rho[i], i % nlayer == 0)) return SAT_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SAT_EDEVICE;
  const size_t lm = (size_t)nlayer * ncol, nin = lm;
  char* buf = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DispDev g{};
  // one allocation: t | c | cb | cg | flags | lay | minvs | the four input arrays
  const size_t bytes = sizeof(double) * ((size_t)kmax + 3 * (size_t)kmax * ncol + ncol) + sizeof(int) * 2 * ncol +
                       sizeof(float) * (7 * lm + 2 * (size_t)ncol + 4 * nin);
  D_CHECK(hipMalloc(&buf, bytes));
  {
    double* dt = (double*)buf;
    double* dc = dt + kmax;
    double* dcb = dc + (size_t)kmax * ncol;
    double* dcg = dcb + (size_t)kmax * ncol;
    double* cst = dcg + (size_t)kmax * ncol;
    int* dfl = (int*)(cst + ncol);
    int* llw = dfl + ncol;
    float* lay = (float*)(llw + ncol);
    float* minvs = lay + 7 * lm;
    float* bmx = minvs + ncol;
    float* in = bmx + ncol;
    g = DispDev{ncol, nlayer, kmax, iflsph, iwave, mode, igr, dt, lay, dc, dcb, dcg, dfl, minvs, cst, bmx, llw};
    D_CHECK(hipMemcpy(dt, t, sizeof(double) * kmax, hipMemcpyHostToDevice));
    D_CHECK(hipMemcpy(in, thk, sizeof(float) * nin, hipMemcpyHostToDevice));
    D_CHECK(hipMemcpy(in + nin, vp, sizeof(float) * nin, hipMemcpyHostToDevice));
    D_CHECK(hipMemcpy(in + 2 * nin, vs, sizeof(float) * nin, hipMemcpyHostToDevice));
    D_CHECK(hipMemcpy(in + 3 * nin, rho, sizeof(float) * nin, hipMemcpyHostToDevice));
    D_CHECK(hipMemset(dcg, 0, sizeof(double) * (size_t)kmax * ncol));
    hipLaunchKernelGGL(copy_layers_kernel, dim3((ncol + 255) / 256), dim3(256), 0, nullptr, ncol, nlayer, in,
                       in + nin, in + 2 * nin, in + 3 * nin, lay, minvs);
    D_CHECK(hipGetLastError());
  }
  D_CHECK(hipEventCreate(&e0));
  D_CHECK(hipEventCreate(&e1));
  D_CHECK(hipEventRecord(e0, nullptr));
  hipLaunchKernelGGL(prep_kernel, dim3((ncol + 63) / 64), dim3(64), 0, nullptr, g);
  D_CHECK(hipGetLastError());
  D_CHECK(launch_solve(g, solve_lanes(ncol)));
  D_CHECK(hipEventRecord(e1, nullptr));
  D_CHECK(hipEventSynchronize(e1));
  if (kernel_ms) {
    float ms = 0.f;
    D_CHECK(hipEventElapsedTime(&ms, e0, e1));
    *kernel_ms = ms;
  }
  D_CHECK(hipMemcpy(cg, g.cg, sizeof(double) * (size_t)kmax * ncol, hipMemcpyDeviceToHost));
  if (flags) D_CHECK(hipMemcpy(flags, g.flags, sizeof(int) * ncol, hipMemcpyDeviceToHost));
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (buf) (void)hipFree(buf);
  return rc;
}

extern "C" int sat_fwdsurf(int nx, int ny, int nz, const double* zgrids, const double* vs3d, const double* vp3d,
                           const double* rho3d, int nper, const double* periods, int iflsph, int iwave, int mode,
                           int igr, double* svel, int* flags, double* kernel_ms) {
  if (nx < 1 || ny < 1 || nz < 2 || !zgrids || !vs3d || !svel) return SAT_EINVAL;
  if ((vp3d == nullptr) != (rho3d == nullptr)) return SAT_EINVAL;
  const int mmax = nz + 1;
  int rc = check_common(mmax, nper, iflsph, iwave, mode, igr, periods);
  if (rc != SAT_OK) return rc;
  for (int k = 0; k < nz; k++)
    if (!std::isfinite(zgrids[k])) return SAT_EINVAL;
  for (size_t i = 0; i < (size_t)nx * ny * nz; i++) {
    const bool top = i % nz == 0;
    if (!std::isfinite(vs3d[i]) || vs3d[i] < 0.0 || (vs3d[i] <= 0.01 && !top)) return SAT_EINVAL;
    if (vp3d && !layer_ok(1.0, vp3d[i], vs3d[i], rho3d[i], top)) return SAT_EINVAL;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SAT_EDEVICE;
  const int ncol = nx * ny;
  const size_t n3 = (size_t)ncol * nz, lm = (size_t)mmax * ncol;
  char* buf = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DispDev g{};
  LayerDev L{};
  const int nmod = vp3d ? 3 : 1;
  const size_t bytes = sizeof(double) * ((size_t)nper + 3 * (size_t)nper * ncol + ncol + nz + nmod * n3) +
                       sizeof(int) * 2 * ncol + sizeof(float) * (7 * lm + 2 * (size_t)ncol);
  D_CHECK(hipMalloc(&buf, bytes));
  {
    double* dt = (double*)buf;
    double* dc = dt + nper;
    double* dcb = dc + (size_t)nper * ncol;
    double* dcg = dcb + (size_t)nper * ncol;
    double* cst = dcg + (size_t)nper * ncol;
    double* dz = cst + ncol;
    double* dvs = dz + nz;
    double* dvp = vp3d ? dvs + n3 : nullptr;
    double* drho = vp3d ? dvs + 2 * n3 : nullptr;
    int* dfl = (int*)(dvs + nmod * n3);
    int* llw = dfl + ncol;
    float* lay = (float*)(llw + ncol);
    float* minvs = lay + 7 * lm;
    float* bmx = minvs + ncol;
    g = DispDev{ncol, mmax, nper, iflsph, iwave, mode, igr, dt, lay, dc, dcb, dcg, dfl, minvs, cst, bmx, llw};
    L = LayerDev{nx, ny, nz, mmax, dz, dvs, dvp, drho, lay, minvs, 1};
    D_CHECK(hipMemcpy(dt, periods, sizeof(double) * nper, hipMemcpyHostToDevice));
    D_CHECK(hipMemcpy(dz, zgrids, sizeof(double) * nz, hipMemcpyHostToDevice));
    D_CHECK(hipMemcpy(dvs, vs3d, sizeof(double) * n3, hipMemcpyHostToDevice));
    if (vp3d) {
      D_CHECK(hipMemcpy(dvp, vp3d, sizeof(double) * n3, hipMemcpyHostToDevice));
      D_CHECK(hipMemcpy(drho, rho3d, sizeof(double) * n3, hipMemcpyHostToDevice));
    }
    D_CHECK(hipMemset(dcg, 0, sizeof(double) * (size_t)nper * ncol));
  }
  D_CHECK(hipEventCreate(&e0));
  D_CHECK(hipEventCreate(&e1));
  D_CHECK(hipEventRecord(e0, nullptr));
  hipLaunchKernelGGL(layers_kernel, dim3((ncol + 255) / 256), dim3(256), 0, nullptr, L);
  D_CHECK(hipGetLastError());
  hipLaunchKernelGGL(prep_kernel, dim3((ncol + 63) / 64), dim3(64), 0, nullptr, g);
  D_CHECK(hipGetLastError());
  D_CHECK(launch_solve(g, solve_lanes(ncol)));
  D_CHECK(hipEventRecord(e1, nullptr));
  D_CHECK(hipEventSynchronize(e1));
  if (kernel_ms) {
    float ms = 0.f;
    D_CHECK(hipEventElapsedTime(&ms, e0, e1));
    *kernel_ms = ms;
  }
  {
    // cg is [col][nper] with col = ix + iy*nx: the period-major planes [nper][ny][nx]
    double* host = new (std::nothrow) double[(size_t)nper * ncol];
    if (!host) {
      rc = SAT_ENOMEM;
      goto done;
    }
    hipError_t e = hipMemcpy(host, g.cg, sizeof(double) * (size_t)nper * ncol, hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
      for (int c = 0; c < ncol; c++)
        for (int p = 0; p < nper; p++) svel[(size_t)p * ncol + c] = host[(size_t)c * nper + p];
    }
    delete[] host;
    if (e != hipSuccess) {
      rc = SAT_EDEVICE;
      goto done;
    }
  }
  if (flags) D_CHECK(hipMemcpy(flags, g.flags, sizeof(int) * ncol, hipMemcpyDeviceToHost));
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (buf) (void)hipFree(buf);
  return rc;
}
