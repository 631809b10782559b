// sat_device.hpp -- device-side building blocks of the gfx950 eikonal path.
//
// Storage on the device ("segmented rows"): a wave of 64 lanes sweeps one
// grid; lane l owns the C consecutive columns [l*C, l*C+C) of every row.
// A plane of an nx-by-ny grid is stored row after row with pitch P = La*C
// (La = ceil(nx/C) active lanes), so lane l's part of row iy starts at
// iy*P + l*C.  Columns ix >= nx in the last lane are padding.
//
// Gauss-Seidel order.  The reference visits, per sweep, x outer and y inner
// (src/eikonal_solver.cpp:149-154, :474-478) and every node update reads only
// its 4 axis neighbours, so node (ix,iy) depends on exactly the neighbours
// that precede it in that order.  Lane l processes row r of the sweep at step
// t = r + lag(l), lag = l for +x sweeps and La-1-l for -x sweeps: the x-side
// neighbour of its first column was finished by the adjacent lane one step
// earlier (passed by a cross-lane shuffle), the y-side neighbour by itself.
// Every node sees the same neighbour values as in the sequential loop, so the
// result is identical iterate for iterate.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace sat {

constexpr double kPI = 3.14159265358979323846264338327950288;  // include/config.h:54
constexpr double kR = 6371.0;                                   // include/config.h:55
constexpr double kDEG2RAD = kPI / 180.0;                        // include/config.h:68
constexpr int kFwdMaxIter = 2000;                               // src/eikonal_solver.cpp:15
constexpr double kFwdTol = 1e-4;                                // :16
constexpr int kAdjMaxIter = 1000;                               // :339
constexpr double kAdjTol = 1e-9;                                // :338
constexpr double kAdjEps = 1e-14;                               // :337
constexpr int kWave = 64;
constexpr int kMaxC = 8;  // columns per lane of a one-wave sweep: nx <= 512

// Wider grids (512 < nx <= 4096) sweep with kMaxC columns per lane on a
// work-group of W = 2, 4 or 8 waves per pair (the lanes exchange their edge
// values through LDS), in every path: per-solve, per-pair, batched FP64 and
// FP32 and the FP64 re-solves.  wide_waves: that W, 1 for nx <= 512, 0 above
// kMaxWideW * 512 columns (refused).
constexpr int kMaxWideW = 8;
__host__ __device__ constexpr int wide_waves(int nx) {
  const int c1 = (nx + kWave - 1) / kWave;
  if (c1 <= kMaxC) return 1;
  for (int w = 2; w <= kMaxWideW; w *= 2)
    if (c1 <= kMaxC * w) return w;
  return 0;
}
// Elements past the end of a segmented swept plane: the stores of lanes with
// no row at a step go there (one row of W waves' lanes, kMaxC columns each).
__host__ __device__ constexpr int seg_trash_elems(int W) { return kWave * W * kMaxC; }

constexpr int FLAG_NOT_CONVERGED = 1;
constexpr int FLAG_SRC_ON_BOUNDARY = 2;
constexpr int FLAG_ADJ_MAXITER = 4;
constexpr int FLAG_REC_OUTSIDE = 8;

__host__ __device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min
__host__ __device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max
__host__ __device__ inline int imin(int a, int b) { return (b < a) ? b : a; }
__host__ __device__ inline int imax(int a, int b) { return (a < b) ? b : a; }

// Geometry and per-period media of one grid, as seen by the kernels.
struct GridDev {
  int nx, ny, nperiod;
  int C, La, P;            // columns per lane, active lanes, segmented row pitch
  size_t seg_plane;        // P*ny
  size_t plane;            // nx*ny
  double dx, dy;           // radians
  const double* lon;       // degrees [nx]
  const double* lat;       // degrees [ny]
  const double* xx;        // radians [nx]   (lon * DEG2RAD, eikonal_solver.cpp:41)
  const double* yy;        // radians [ny]
  // per-period media, [nperiod] planes each
  const double* f_plain;   // slowness 1/svel, column-major (bilinear at source)
  const double* aA_plain;  // adjoint metric a  (spha/(R2*cos*cos), :385)
  const double* b_plain;   // b = sphb/R2
  const double* c_plain;   // c = sphc/(R2*cos)
  const double* vinv3;     // svel^-3 (preproc.cpp:29)
  // forward media, segmented, sweep precision: FP64 a, b, c, f; FP32 a, b, c, f^2,
  // dis_x = sqrt(f^2 b/det), dis_y = sqrt(f^2 a/det) (-1 where det <= 0)
  const void* seg[6];
  const double2* trig_yy;  // per row: sin(yy), cos(yy), correctly rounded
  int iso;                 // spha = sphb = 1, sphc = 0 on every node of every period
  // Isotropic FP32 batched sweeps solve in scaled coordinates X = sqrt(a) Tx,
  // Y = sqrt(b) Ty (a = 1/(R cos(lat))^2, b = 1/R^2): per row
  // rowha = sqrt(a)/dx = 1/(R cos(lat) dx), kb = sqrt(b)/dy = 1/(R dy), FP64,
  // and rowtab = {sqrt(a)/dx, sqrt(b)/dy} rounded to FP32.
  const float2* rowtab;
  const double* rowha;
  double kb;
  // the forward media again in wavefront-major order (batched sweeps), one copy
  // per orientation, [nperiod] planes of wplane elements each
  size_t wplane;           // (ny + La - 1) * La * C
  const void* wmed[6][2];
  const double* wf64[2];   // FP32 batches: the slowness again in FP64, wavefront-major per orientation
};

// Geometry of a grid swept by W waves of 64 lanes: C = ceil(ceil(nx / 64) / W)
// columns per lane, La = ceil(nx / C) active lanes, and the size of one
// wavefront-major plane (below), (ny + La - 1) * La * C elements.
struct WaveGeom {
  int C, La;
  size_t wplane;
};
// Two pairs per wave (FP32 isotropic forward sweeps): each half of the wave's
// lanes sweeps one pair, at most kPair2MaxC columns per lane.
constexpr int kHalfWave = 32;
constexpr int kPair2MaxC = 7;

__host__ __device__ inline WaveGeom wave_geom(int nx, int ny, int W) {
  WaveGeom w;
  w.C = ((nx + kWave - 1) / kWave + W - 1) / W;
  w.La = (nx + w.C - 1) / w.C;
  w.wplane = (size_t)(ny + w.La - 1) * w.La * w.C;
  return w;
}

__device__ inline int seg_index(const GridDev& g, int ix, int iy) {
  return iy * g.P + ix;  // lane l, column k -> l*C + k == ix
}

// ----------------------------------------------------------------------------
// Wavefront-major storage.  In a sweep with directions (XD, YD) lane l handles
// row j at step t with j + l constant (orientation 0: XD*YD > 0) or j - l
// constant (orientation 1: XD*YD < 0) across the wave.  Storing a plane ordered
// by that diagonal, lane-minor, makes every load and store of a step one
// contiguous run of La*C elements.  Diagonal index q = j + l (o = 0) or
// j - l + La - 1 (o = 1), q in [0, ny + La - 1); element (q*La + l)*C + k.
// Entries with j outside [0, ny) or column l*C + k >= nx are padding.
// ----------------------------------------------------------------------------
template <int C, int O>
__device__ __forceinline__ int wave_base(const GridDev& g, int j, int lane) {
  const int q = (O == 0) ? j + lane : j - lane + g.La - 1;
  return (q * g.La + lane) * C;
}
__host__ __device__ inline size_t wave_index(int C, int La, int o, int ix, int iy) {
  const int l = ix / C, k = ix - l * C;
  const int q = (o == 0) ? iy + l : iy - l + La - 1;
  return ((size_t)q * La + l) * C + k;
}
__device__ inline size_t wave_index(const GridDev& g, int o, int ix, int iy) {
  return wave_index(g.C, g.La, o, ix, iy);
}

// Load through a global-memory pointer.  A pointer the compiler cannot prove
// global (one read from device memory, such as a GridDev plane) is loaded with a
// flat instruction, which also counts on the LDS counter: each wait for the
// step's cross-lane exchange would then also wait for the prefetched rows.
template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
// HIP's vector types are classes whose copy goes through a generic reference:
// load them as the compiler's own vector types.
__device__ __forceinline__ float2 gld(const float2* p) {
  typedef float v2 __attribute__((ext_vector_type(2)));
  const v2 v = *(const __attribute__((address_space(1))) v2*)p;
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 gld(const float4* p) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = *(const __attribute__((address_space(1))) v4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ double2 gld(const double2* p) {
  typedef double v2 __attribute__((ext_vector_type(2)));
  const v2 v = *(const __attribute__((address_space(1))) v2*)p;
  return make_double2(v.x, v.y);
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}

// Raw buffer stores over one plane: a store whose byte offset lies at or beyond
// the plane's size is dropped by the hardware (no memory access), so lanes with
// nothing to store issue the same unconditional instruction with kDropOffset.
// Word 3 of the descriptor: the gfx9 raw-buffer format (32-bit data format,
// range-checked against num_records in bytes).
constexpr int kBufferWord3 = 0x00020000;
constexpr int kDropOffset = 0x40000000;
// The descriptor lives in scalar registers: the plane pointer is made
// wave-uniform explicitly (it is the same on every lane; a pointer the compiler
// keeps in vector registers would turn each store into a waterfall loop).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(void* p, int bytes) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  void* pu = (void*)(((unsigned long long)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(pu, (short)0, __builtin_amdgcn_readfirstlane(bytes), kBufferWord3);
}
typedef unsigned int sat_u32x2 __attribute__((__vector_size__(8)));
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, int off, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(sat_u32x2, v), r, off, 0, 0);
}
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, int off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}
// Element i of a plane, addressed as a wave-uniform base plus a 32-bit byte
// offset (planes are < 4 GiB): the load takes the base from scalar registers
// and needs no 64-bit vector address.  (Stores keep pointer arithmetic: with a
// 32-bit offset that may wrap, consecutive stores are not merged.)
template <class T>
__device__ __forceinline__ T gld(const T* base, unsigned i) {
  const unsigned off = i * (unsigned)sizeof(T);
  return gld(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off));
}

// Element i of a plane held in LDS (a generic pointer into the work-group's
// shared memory).
template <class T>
__device__ __forceinline__ T lld(const T* p, unsigned i) {
  return ((const __attribute__((address_space(3))) T*)p)[i];
}

// A wave-uniform copy of v (readfirstlane of every 32-bit word): for descriptors
// read from device memory that the compiler cannot prove unchanged by the
// kernel's own stores, so it loads them into vector registers.  Scalar copies
// leave the vector registers to the sweeps' row rings.  Every lane must hold
// the same v.
template <class T>
__device__ __forceinline__ T uniform(const T& v) {
  static_assert(sizeof(T) % 4 == 0, "32-bit words");
  int w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); i++) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

// s_waitcnt vmcnt(0): wait for every vector memory operation of the wave.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Cross-lane reads of the swept value and its auxiliary (ds_bpermute).
__device__ __forceinline__ float shfl_any(float v, int src) { return __shfl(v, src); }
__device__ __forceinline__ double shfl_any(double v, int src) { return __shfl(v, src); }
__device__ __forceinline__ float2 shfl_any(float2 v, int src) {
  return make_float2(__shfl(v.x, src), __shfl(v.y, src));
}
__device__ __forceinline__ double2 shfl_any(double2 v, int src) {
  return make_double2(__shfl(v.x, src), __shfl(v.y, src));
}

// ----------------------------------------------------------------------------
// Wave-level helpers (blockDim.x == 64, one wave per work-group).
// ----------------------------------------------------------------------------
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o));
  return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = imin(v, __shfl_xor(v, o));
  return v;
}
__device__ inline int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = imax(v, __shfl_xor(v, o));
  return v;
}
__device__ inline int wave_or_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

// The same over a work-group of blockDim.x / 64 <= kMaxWideW waves (every
// thread of the block calls them; one wave: no LDS, no barrier).
__device__ inline int block_reduce_i(int v, int op) {  // op: 0 min, 1 max, 2 or
  v = op == 0 ? wave_min_i(v) : op == 1 ? wave_max_i(v) : wave_or_i(v);
  if (blockDim.x == kWave) return v;
  __shared__ int buf[kMaxWideW];
  __syncthreads();  // an earlier call's readers are done with buf
  if ((threadIdx.x & (kWave - 1)) == 0) buf[threadIdx.x / kWave] = v;
  __syncthreads();
  int r = buf[0];
  for (int w = 1; w < (int)(blockDim.x / kWave); w++)
    r = op == 0 ? imin(r, buf[w]) : op == 1 ? imax(r, buf[w]) : (r | buf[w]);
  return r;
}
__device__ inline int block_min_i(int v) { return block_reduce_i(v, 0); }
__device__ inline int block_max_i(int v) { return block_reduce_i(v, 1); }
__device__ inline int block_or_i(int v) { return block_reduce_i(v, 2); }

// include/utils.h:28-36 bilinear at fractional index (idx0+r1, idy0+r2), column-major.
__device__ inline double bilinear_cm(const double* M, int nx, int idx0, int idy0, double r1, double r2) {
  return (1 - r1) * (1 - r2) * M[idx0 + idy0 * nx]
       + (1 - r1) * r2 * M[idx0 + (idy0 + 1) * nx]
       + r1 * (1 - r2) * M[idx0 + 1 + idy0 * nx]
       + r1 * r2 * M[idx0 + 1 + (idy0 + 1) * nx];
}

// std::lower_bound over an ascending array.
__device__ inline int lower_bound_d(const double* v, int n, double q) {
  int lo = 0, cnt = n;
  while (cnt > 0) {
    const int step = cnt / 2;
    const int it = lo + step;
    if (v[it] < q) { lo = it + 1; cnt -= step + 1; } else cnt = step;
  }
  return lo;
}

// include/utils.h:358-409 interp2d at one query point (NaN outside the grid).
__device__ inline double interp2d_point(const double* xg, int nx, const double* yg, int ny,
                                        const double* z, double qx, double qy) {
  if (qx < xg[0] || qx > xg[nx - 1] || qy < yg[0] || qy > yg[ny - 1]) return __builtin_nan("");
  int i = lower_bound_d(xg, nx, qx);
  if (i == nx) --i;
  if (i > 0 && xg[i] > qx) --i;
  i = imax(0, imin(i, nx - 2));
  int j = lower_bound_d(yg, ny, qy);
  if (j == ny) --j;
  if (j > 0 && yg[j] > qy) --j;
  j = imax(0, imin(j, ny - 2));
  const double x1 = xg[i], x2 = xg[i + 1];
  const double y1 = yg[j], y2 = yg[j + 1];
  const double tx = (qx - x1) / (x2 - x1);
  const double ty = (qy - y1) / (y2 - y1);
  return (1 - tx) * (1 - ty) * z[i + j * nx]
       + (tx) * (1 - ty) * z[i + 1 + j * nx]
       + (1 - tx) * (ty) * z[i + (j + 1) * nx]
       + (tx) * (ty) * z[i + 1 + (j + 1) * nx];
}

// ----------------------------------------------------------------------------
// FP64 node updates, in the reference's operation order (compiled with
// -ffp-contract=off, IEEE division and square root).
// ----------------------------------------------------------------------------

// One forward node (src/eikonal_solver.cpp:155-283).  W/E/S/N = (ix-1,iy),
// (ix+1,iy), (ix,iy-1), (ix,iy+1); h* says the neighbour exists; t* its tau,
// T0* its factor T0.
//
// Straight-line: every candidate of the reference is evaluated with the
// reference's operations and operand order, whether or not the reference
// would reach it, and an acceptance mask (the reference's skip and acceptance
// conditions) decides whether it enters the running minimum.  The minimum is
// taken in the reference's candidate order, so the result is the same bits;
// without branches the compiler interleaves the independent candidates, which
// hides the latency of the FP64 divide and square-root sequences.  A rejected
// candidate may be NaN or infinite (square root of a negative discriminant,
// division by a vanishing coefficient); it never reaches the minimum.
__device__ inline double fwd_update_f64(double tau_self, double t0, double t0x, double t0y,
                                        double a, double b, double c, double f,
                                        double dx, double dy,
                                        bool hW, double tW, double T0W, bool hE, double tE, double T0E,
                                        bool hS, double tS, double T0S, bool hN, double tN, double T0N) {
  const double det = a * b - c * c;
  double cand[16];
  bool acc[16];
  // :173-219 two-point quadrants (sx, sy) = (-1,-1), (-1,+1), (+1,-1), (+1,+1)
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int sx = (k < 2) ? -1 : 1;
    const int sy = (k & 1) ? 1 : -1;
    const bool hx = (sx < 0) ? hW : hE;
    const bool hy = (sy < 0) ? hS : hN;
    const double tX = (sx < 0) ? tW : tE;
    const double tY = (sy < 0) ? tS : tN;
    const double px = t0x - sx * (t0 / dx);
    const double qx = sx * (t0 / dx) * tX;
    const double py = t0y - sy * (t0 / dy);
    const double qy = sy * (t0 / dy) * tY;
    const double eq_a = a * px * px + b * py * py - 2.0 * c * px * py;
    const double eq_b = 2.0 * (a * px * qx + b * py * qy - c * (px * qy + py * qx));
    const double eq_c = a * qx * qx + b * qy * qy - 2.0 * c * qx * qy - f * f;
    const double Delta = eq_b * eq_b - 4.0 * eq_a * eq_c;
    const bool ok = hx & hy & !(Delta < 0.0 || fabs(eq_a) < 1e-30);
    const double sq = sqrt(Delta);
    const double two_a = 2.0 * eq_a;
#pragma unroll
    for (int sgn = 0; sgn < 2; sgn++) {
      const double tmp_tau = (-eq_b + (sgn == 0 ? sq : -sq)) / two_a;
      const double Tx = px * tmp_tau + qx;
      const double Ty = py * tmp_tau + qy;
      const double char_x = a * Tx - c * Ty;
      const double char_y = b * Ty - c * Tx;
      cand[2 * k + sgn] = tmp_tau;
      acc[2 * k + sgn] = ok & (-sx * char_x >= 0.0) & (-sy * char_y >= 0.0);
    }
  }
  // :224-279 one-point stencils W, E (dis from b), S, N (dis from a)
  const double dis_x = sqrt(f * f * b / det);
  const double dis_y = sqrt(f * f * a / det);
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const bool hd = (d == 0) ? hW : (d == 1) ? hE : (d == 2) ? hS : hN;
    const double tD = (d == 0) ? tW : (d == 1) ? tE : (d == 2) ? tS : tN;
    const double T0D = (d == 0) ? T0W : (d == 1) ? T0E : (d == 2) ? T0S : T0N;
    const double g = (d < 2) ? t0x : t0y;
    const double h = (d < 2) ? dx : dy;
    // W, S: p = g + t0/h, q = -(t0/h)*t;  E, N: p = g - t0/h, q = (t0/h)*t
    const double p = (d == 0 || d == 2) ? g + t0 / h : g - t0 / h;
    const double q = (d == 0 || d == 2) ? -(t0 / h) * tD : (t0 / h) * tD;
    const double dis = (d < 2) ? dis_x : dis_y;
    const bool ok = hd & (det > 0.0);
#pragma unroll
    for (int sgn = 0; sgn < 2; sgn++) {
      const double tmp_tau = ((sgn == 0 ? dis : -dis) - q) / p;
      cand[8 + 2 * d + sgn] = tmp_tau;
      acc[8 + 2 * d + sgn] = ok & (tmp_tau * t0 >= tD * T0D) & (tmp_tau > tD * 0.5);
    }
  }
  double tn = tau_self;
#pragma unroll
  for (int i = 0; i < 16; i++) tn = acc[i] ? dmin(tn, cand[i]) : tn;
  return tn;
}

// One adjoint node (src/eikonal_solver.cpp:479-490).
__device__ inline double adj_update_f64(double a1p, double a2m, double b1p, double b2m, double dlt, double d,
                                        double tW, double tE, double tS, double tN, double dx, double dy) {
  if (fabs(d) < kAdjEps) return 0.0;
  const double e = (tW * a1p - tE * a2m) / dx + (tS * b1p - tN * b2m) / dy;
  return (dlt + e) / d;
}

}  // namespace sat
