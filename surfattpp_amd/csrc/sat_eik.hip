// sat_eik.hip -- kernels and C ABI of the gfx950 eikonal forward + adjoint path.
//
// Kernels here (the sweep kernels are in sat_fused.hpp / sat_phases.hpp,
// instantiated by k_*.hip)
//   media_kernel      per-period metric/slowness planes (src/eikonal_solver.cpp:55-65,
//                     :380-389; src/preproc.cpp:97-103 and :29), segmented and
//                     wavefront-major
//   row_kernel        per-row latitude trig and the isotropic FP32 row metric
//   reduce_kernel     per-period kernel planes, adj_s[iper] = sum_e Ta_e * svel^-3 in
//                     event order (src/preproc.cpp:28-31)
//   chi_kernel        chi = sum_e 0.5*sum_r w r^2 in event order (src/preproc.cpp:142)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <type_traits>
#include <tuple>
#include <vector>

#include "../../include/surfatt_eik.h"
#include "sat_launch.hpp"

namespace sat {

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
    void* sp[6] = {s0, s1, s2, s3, s4, s5};
    const bool wave = q < nx && g.wmed[0][0] != nullptr;
    size_t wi[2] = {0, 0};
    if (wave)
      for (int o = 0; o < 2; o++) wi[o] = (size_t)ip * g.wplane + wave_index(g, o, q, iy);
    if (wave && g.wf64[0])
      for (int o = 0; o < 2; o++) const_cast<double*>(g.wf64[o])[wi[o]] = vf;
    if (prec == SAT_PREC_F64) {
      const double v[4] = {va, vb, vc, vf};
      for (int i = 0; i < 4; i++) {
        if (s0) static_cast<double*>(sp[i])[so] = v[i];
        if (wave)
          for (int o = 0; o < 2; o++) static_cast<double*>(const_cast<void*>(g.wmed[i][o]))[wi[o]] = v[i];
      }
    } else {
      // one-point distances; NaN where det <= 0 (the FP32 update drops NaN candidates)
      const double det = va * vb - vc * vc;
      const float v[6] = {(float)va, (float)vb, (float)vc, (float)vf,
                          (det > 0.0) ? (float)sqrt(vf * vf * vb / det) : __builtin_nanf(""),
                          (det > 0.0) ? (float)sqrt(vf * vf * va / det) : __builtin_nanf("")};
      for (int i = 0; i < 6; i++) {
        if (s0) static_cast<float*>(sp[i])[so] = v[i];
        if (wave)
          for (int o = 0; o < 2; o++) static_cast<float*>(const_cast<void*>(g.wmed[i][o]))[wi[o]] = v[i];
      }
    }
  }
}

// Isotropic metric test on the device: *flag |= 1 where a node has
// spha != 1, sphb != 1 or sphc != 0 (the planes may have come from host or
// device memory).
__global__ void iso_check_kernel(const double* ma, const double* mb, const double* mc, size_t n, int* flag) {
  int bad = 0;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    bad |= !(ma[k] == 1.0 && mb[k] == 1.0 && mc[k] == 0.0);
  if (__any(bad) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(flag, 1);
}

// Per row: correctly rounded sin/cos of the latitude; isotropic grids also the
// row metric's scale sqrt(a)/dx in FP64 and {sqrt(a)/dx, sqrt(b)/dy} in FP32
// (src/eikonal_solver.cpp:55-65 with spha = sphb = 1, sphc = 0: a = 1/(R cos)^2,
// b = 1/R^2).
__global__ void row_kernel(GridDev g, double2* trig_yy, float2* rowtab, double* rowha) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  if (iy >= g.ny) return;
  double s, c;
  cr::sincos_cr(g.yy[iy], s, c);
  trig_yy[iy] = make_double2(s, c);
  if (rowha) rowha[iy] = 1.0 / (kR * c * g.dx);  // sqrt(a)/dx, a = 1/(R cos)^2
  if (rowtab) rowtab[iy] = make_float2((float)(1.0 / (kR * c * g.dx)), (float)(1.0 / (kR * g.dy)));
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
  const size_t* ev_taoff;      // parallel to ev_ids: the event's Ta plane offset
  int ev_lo, ev_hi;            // events of this pass (a chunk of the batch order)
  int accumulate;              // add to K (a later chunk) instead of overwriting it
};

// First position in ev_ids[q0, q1) (ascending) holding an event >= e.
__device__ __forceinline__ int first_event_at_or_after(const int* ev_ids, int q0, int q1, int e) {
  while (q0 < q1) {
    const int m = (q0 + q1) >> 1;
    if (ev_ids[m] < e) q0 = m + 1;
    else q1 = m;
  }
  return q0;
}

// Per node of every (grid, period) plane: K += Ta * svel^-3 over the plane's
// events of this pass, in batch order (src/preproc.cpp:28-31).  The events of
// a plane are listed in ascending batch order, so a chunk's events are one run
// of the list.  The Ta offsets come from a table parallel to the list, and the
// loads of kReduceUnroll events are issued before they are summed (the sum
// keeps the event order): one dependent load per event instead of a chain of
// three, and several in flight per thread.
constexpr int kReduceUnroll = 8;
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
    // the chunks are consecutive ranges of the batch order: continuing from the
    // previous chunk's partial sum keeps the reference's summation order
    double acc = rd.accumulate ? rd.K[rd.plane_off[pl] + n] : 0.0;
    int q0 = rd.ev_off[pl], q1 = rd.ev_off[pl + 1];
    q0 = first_event_at_or_after(rd.ev_ids, q0, q1, rd.ev_lo);
    q1 = first_event_at_or_after(rd.ev_ids, q0, q1, rd.ev_hi);
    int q = q0;
    for (; q + kReduceUnroll <= q1; q += kReduceUnroll) {
      size_t off[kReduceUnroll];
      V v[kReduceUnroll];
#pragma unroll
      for (int u = 0; u < kReduceUnroll; u++) off[u] = rd.ev_taoff[q + u];
#pragma unroll
      for (int u = 0; u < kReduceUnroll; u++) v[u] = ta[off[u] + n];
#pragma unroll
      for (int u = 0; u < kReduceUnroll; u++) acc += (double)v[u] * v3;
    }
    for (; q < q1; q++) acc += (double)ta[rd.ev_taoff[q] + n] * v3;
    rd.K[rd.plane_off[pl] + n] = acc;
  }
}

// Anisotropic kernel terms of accumulate_kernels (src/preproc.cpp:37-65) for
// the (grid, period) planes of one grid: per node, over the plane's events in
// batch order, the travel-time gradient (gradient_2_geo, include/utils.h:
// 595-649: central differences inside, one-sided at the edges, dx = dlon_i *
// R * DEG2RAD * cos(lat_j), dy = dlat_j * R * DEG2RAD) and
//   adj_xi  += Ta * (Tx^2 (c^2 - a^2) + 2 TxTy (ac - bc) + Ty^2 (b^2 - c^2)),
//   adj_eta += Ta * (2 ac Tx^2 - 2 (ab + c^2) TxTy + 2 bc Ty^2),
// a, b, c the topography metric (SurfGrid a, b, c), in the reference's
// operation order, FP64.
struct AnisoDev {
  GridDev g;
  int nper, pl0;                // planes pl0 .. pl0 + nper - 1 of the reduction CSR
  const int* ev_off;
  const int* ev_ids;
  const EventDev* events;
  const double* T;              // per event, column-major, at events[e].ta_off
  const void* ta;               // masked Ta, sweep precision, same offsets
  const double* abc;            // a, b, c: [3][nper][ny][nx]
  double* xi;                   // [nper][ny][nx]
  double* eta;
  int ev_lo, ev_hi;             // the events of this pass (a chunk of the batch order)
  int accumulate;               // add to xi, eta (a later chunk) instead of overwriting them
};

template <typename V>
__global__ void aniso_kernel(AnisoDev ad) {
  const GridDev& g = ad.g;
  const int nx = g.nx, ny = g.ny;
  const size_t N = (size_t)nx * ny, total = N * ad.nper;
  const double km = kR * kDEG2RAD;
  const V* ta = static_cast<const V*>(ad.ta);
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total; id += (size_t)gridDim.x * blockDim.x) {
    const int p = (int)(id / N);
    const size_t n = id - (size_t)p * N;
    const int iy = (int)(n / nx), ix = (int)(n - (size_t)iy * nx);
    const double cos_lat = g.trig_yy[iy].y;
    auto dxk = [&](int i) { return ((g.lon[i] - g.lon[i - 1]) * km) * cos_lat; };
    auto dyk = [&](int j) { return (g.lat[j] - g.lat[j - 1]) * km; };
    const double a = ad.abc[id], b = ad.abc[total + id], c = ad.abc[2 * total + id];
    const double a2 = a * a, b2 = b * b, c2 = c * c, ac = a * c, bc = b * c, ab = a * b;
    // the stencil of gradient_2_geo at this node, the same for every event:
    // central differences inside, one-sided at the edges
    const size_t oxp = n + (ix < nx - 1 ? 1 : 0), oxm = n - (ix > 0 ? 1 : 0);
    const size_t oyp = n + (iy < ny - 1 ? (size_t)nx : 0), oym = n - (iy > 0 ? (size_t)nx : 0);
    const double dxd = (ix == 0) ? dxk(1) : (ix == nx - 1) ? dxk(nx - 1) : 2.0 * dxk(ix);
    const double dyd = (iy == 0) ? dyk(1) : (iy == ny - 1) ? dyk(ny - 1) : 2.0 * dyk(iy);
    // a later chunk continues the previous chunks' sums: the batch order is kept
    double xi = ad.accumulate ? ad.xi[id] : 0.0, eta = ad.accumulate ? ad.eta[id] : 0.0;
    const int pl = ad.pl0 + p;
    const int qe = ad.ev_off[pl + 1];
    // kEv events' loads in flight at once; the sums still run in batch order
    constexpr int kEv = 4;
    int q0 = ad.ev_off[pl];
    int q1 = qe;
    while (q0 < q1 && ad.ev_ids[q0] < ad.ev_lo) q0++;  // the plane's events in batch order: the chunk's run
    while (q1 > q0 && ad.ev_ids[q1 - 1] >= ad.ev_hi) q1--;
    for (int q = q0; q < q1; q += kEv) {
      double tx[kEv], ty[kEv], adj[kEv];
#pragma unroll
      for (int k = 0; k < kEv; k++) {
        if (q + k < q1) {
          const EventDev& ev = ad.events[ad.ev_ids[q + k]];
          const double* f = ad.T + ev.ta_off;
          tx[k] = (f[oxp] - f[oxm]) / dxd;
          ty[k] = (f[oyp] - f[oym]) / dyd;
          adj[k] = (double)ta[ev.ta_off + n];
        }
      }
#pragma unroll
      for (int k = 0; k < kEv; k++) {
        if (q + k < q1) {
          const double tx2 = tx[k] * tx[k], ty2 = ty[k] * ty[k], txty = tx[k] * ty[k];
          xi = xi + adj[k] * (tx2 * (-a2 + c2) + 2.0 * txty * (ac - bc) + ty2 * (-c2 + b2));
          eta = eta + adj[k] * (2.0 * ac * tx2 - 2.0 * (ab + c2) * txty + 2.0 * bc * ty2);
        }
      }
    }
    ad.xi[id] = xi;
    ad.eta[id] = eta;
  }
}

// Start-of-run state: work-queue heads and list lengths, re-solve marks and
// iteration counts zeroed, flags from the creation-time flags.  A kernel, not
// memset/memcpy nodes, so that the run can be captured in a graph and replayed.
__global__ void reset_kernel(int* counters, size_t n_counters, unsigned char* skip, int* iters, int* flags,
                             const int* base_flags, size_t ne) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_counters + ne;
       i += (size_t)gridDim.x * blockDim.x) {
    if (i < n_counters) {
      counters[i] = 0;
    } else {
      const size_t e = i - n_counters;
      skip[e] = 0;
      iters[e] = 0;
      iters[ne + e] = 0;
      flags[e] = base_flags[e];
    }
  }
}

__global__ void zero_ints_kernel(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Longest pairs first.  Block g reorders launch group g's queue,
// order[rng[g].x, rng[g].x + rng[g].y), by the events' rounds in the previous
// run (key[e]: the forward or the adjoint iteration count, 0 before the first
// run), most rounds first, equal rounds keeping their previous order (a stable
// counting sort).  The persistent sweep kernels pop their queue in this order:
// the pairs that took the most rounds start first, one per SIMD, and the
// shorter ones fill in behind them (longest-processing-time-first), so the
// SIMDs' shares of sweep steps come out even.  No result depends on it: every
// pair is solved from its own start values wherever it runs, and the kernel
// and chi sums run in batch order.
constexpr int kOrderThreads = 1024;
constexpr int kOrderBuckets = 16;  // 15 and more rounds share the first bucket
__global__ __launch_bounds__(kOrderThreads) void order_kernel(int* order, int* tmp, const int* key,
                                                              const int2* rng) {
  constexpr int kWaves = kOrderThreads / kWave;
  __shared__ int hist[kOrderBuckets], start[kOrderBuckets];
  __shared__ int wcnt[kWaves][kOrderBuckets];
  const int first = rng[blockIdx.x].x, n = rng[blockIdx.x].y;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  auto bucket = [&](int e) { return kOrderBuckets - 1 - imin(imax(key[e], 0), kOrderBuckets - 1); };
  if (tid < kOrderBuckets) hist[tid] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += kOrderThreads) {
    const int e = order[first + i];
    tmp[first + i] = e;
    atomicAdd(&hist[bucket(e)], 1);
  }
  __syncthreads();
  if (tid == 0) {
    int s = 0;
    for (int c = 0; c < kOrderBuckets; c++) {
      start[c] = s;
      s += hist[c];
    }
  }
  __syncthreads();
  for (int base = 0; base < n; base += kOrderThreads) {
    const int i = base + tid;
    const int e = i < n ? tmp[first + i] : -1;
    const int c = e >= 0 ? bucket(e) : kOrderBuckets;
    int rank = 0;
    for (int b = 0; b < kOrderBuckets; b++) {  // rank among this wave's lanes of bucket b
      const unsigned long long m = __ballot(c == b);
      if (lane == 0) wcnt[wv][b] = __popcll(m);
      if (c == b) rank = __popcll(m & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    if (e >= 0) {
      int pos = start[c] + rank;
      for (int w = 0; w < wv; w++) pos += wcnt[w][c];
      order[first + pos] = e;
    }
    __syncthreads();
    if (tid < kOrderBuckets) {
      int s = 0;
      for (int w = 0; w < kWaves; w++) s += wcnt[w][tid];
      start[tid] += s;
    }
    __syncthreads();
  }
}

// Speculative FP64 re-solves.  An FP32 batch re-solves in FP64 the pairs whose
// stop decision lies within the band of the tolerance, so that every forward
// iteration count is the FP64 one; the decision is known only when the pair's
// FP32 sweeps end, so those re-solves wait for the whole FP32 forward launch
// and then run alone at the end of the run.  Consecutive runs of an inversion
// differ by one model update, so a pair that needed a re-solve in a recent run
// is likely to need one again: such pairs are solved in FP64 from the start of
// the run, on the second stream beside the FP32 forward sweeps, and the FP32
// pipeline passes over them (mark kSkipSpec).  Their result is their FP64
// solution, which the band rule would have given them or which is closer to
// the reference than the FP32 one; the band rule still re-solves the pairs it
// finds among the others.

// Runs a prediction lasts: refreshed while the pair is re-solved by the band
// rule, or speculated and its FP64 stop decision is within the band itself.
constexpr unsigned char kSpecAge = 8;
// Counters of the speculated solves: the three list lengths (spec_kernel), the
// forward queue heads of the two update kinds, the adjoint queue head.
constexpr int kSpecCounters = 8;

// Before a run's reset: each pair's prediction age from the previous run's
// re-solve marks (and, for speculated pairs, their FP64 convergence norms).
__global__ void spec_age_kernel(unsigned char* age, const unsigned char* skip, const double* conv, double band,
                                int ne) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const unsigned char s = skip[e];
  unsigned char a = age[e];
  if (s == 1 || s == 2) {
    a = kSpecAge;
  } else if (s == kSkipSpec) {
    const double l = conv[2 * (size_t)e], lp = conv[2 * (size_t)e + 1];
    const bool near = (l >= (1.0 - band) * kFwdTol && l < kFwdTol) || (lp >= 0.0 && lp < (1.0 + band) * kFwdTol);
    a = near ? kSpecAge : (a > 0 ? (unsigned char)(a - 1) : (unsigned char)0);
  } else {
    a = 0;
  }
  age[e] = a;
}

// After the reset: the speculated pairs, in launch-group and queue order, each
// with a slot of the FP64 pool until the pool is full: list[0, count[0]) holds
// all of them, list[ne, ne + count[1]) those of the groups whose re-solves take
// the isotropic update, list[2 ne, 2 ne + count[2]) the others; the pool
// counter then starts the band rule's re-solves past them.  One block: the
// choice of pairs does not depend on timing.
__global__ __launch_bounds__(kOrderThreads) void spec_kernel(const unsigned char* age, unsigned char* skip,
                                                             const int* order, const int2* rng, const int* group_iso,
                                                             int ng, int ne, int* list, int* count,
                                                             int* pool_counter, int pool_size, size_t pool_stride,
                                                             EventX* ev64) {
  constexpr int kWaves = kOrderThreads / kWave;
  __shared__ int wcnt[kWaves];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  int total = 0, n_iso = 0, n_gen = 0;
  for (int g = 0; g < ng; g++) {
    const int first = rng[g].x, n = rng[g].y;
    const bool iso = group_iso[g] != 0;
    for (int base = 0; base < n; base += kOrderThreads) {
      const int i = base + tid;
      const int e = i < n ? order[first + i] : -1;
      const bool p = e >= 0 && age[e] > 0;
      const unsigned long long m = __ballot(p);
      if (lane == 0) wcnt[wv] = __popcll(m);
      __syncthreads();
      int rank = __popcll(m & ((1ull << lane) - 1ull)), tot = 0;
      for (int w = 0; w < kWaves; w++) {
        if (w < wv) rank += wcnt[w];
        tot += wcnt[w];
      }
      const int slot = total + rank;
      if (p && slot < pool_size) {
        list[slot] = e;
        list[(iso ? ne + n_iso : 2 * ne + n_gen) + rank] = e;
        ev64[e].w_off = (size_t)slot * pool_stride;
        skip[e] = kSkipSpec;
      }
      const int kept = imax(0, imin(tot, pool_size - total));
      total += kept;
      if (iso) n_iso += kept;
      else n_gen += kept;
      __syncthreads();  // wcnt is rewritten by the next pass
    }
  }
  if (tid == 0) {
    count[0] = total;
    count[1] = n_iso;
    count[2] = n_gen;
    *pool_counter = total;
  }
}

// One block of kChiThreads: each pass stages kChiThreads per-event terms in LDS
// with one coalesced load, and thread 0 adds them in event order (the
// reference's sequential sum; the loads of a pass are all in flight at once).
constexpr int kChiThreads = 256;
__global__ __launch_bounds__(kChiThreads) void chi_kernel(const double* chi_ev, int n, double* chi) {
  __shared__ double v[kChiThreads];
  double c = 0.0;
  for (int e0 = 0; e0 < n; e0 += kChiThreads) {
    const int e = e0 + (int)threadIdx.x;
    v[threadIdx.x] = e < n ? chi_ev[e] : 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
      const int m = n - e0 < kChiThreads ? n - e0 : kChiThreads;
      for (int i = 0; i < m; i++) c += v[i];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *chi = c;
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
  int W;  // waves per pair of the one-pair kernels (wide_waves; 0: too wide)
};

// Launch groups: one per (columns per lane C, isotropic metric).
constexpr int kMaxGroups = 2 * kMaxC;
// Columns per lane of the FP32 adjoint of one-wave grids (GridHost::gadj):
// 8 x 4 bytes, one 32-byte sector per lane and row.
constexpr int kAdjC = 8;

// Up to 512 columns: one wave, C = ceil(nx / 64) columns per lane.  Wider:
// kMaxC columns per lane on W = wide_waves(nx) waves.
static Layout layout_for(int nx) {
  Layout L;
  L.W = wide_waves(nx);
  L.C = L.W > 1 ? kMaxC : (nx + kWave - 1) / kWave;
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

// Dynamic LDS a one-wave per-pair adjoint may hold (the LDS tier, sat_fused.hpp):
// the device's LDS per work-group (read once per process) less 4 KiB for the
// kernels' static LDS.  SURFATT_ADJ_LDS=0 turns the tier off (read per call).
static size_t adj_lds_budget() {
  if (const char* v = std::getenv("SURFATT_ADJ_LDS"))
    if (std::atoi(v) == 0) return 0;
  static const size_t budget = [] {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return (size_t)0;
    const size_t lds = (size_t)prop.sharedMemPerBlock;
    return lds > 4096 ? lds - 4096 : (size_t)0;
  }();
  return budget;
}
// The dynamic LDS of a one-wave launch whose adjoints run on a grid of this
// segmented plane size: the tier's need when it fits the budget, else 0.
static size_t adj_lds_bytes(int prec, size_t seg_plane, int W) {
  if (W != 1) return 0;
  const size_t need = prec == SAT_PREC_F64 ? fused_adj_lds_bytes<SAT_PREC_F64>(seg_plane)
                                           : fused_adj_lds_bytes<SAT_PREC_F32>(seg_plane);
  return need <= adj_lds_budget() ? need : 0;
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
  float2* d_rowtab = nullptr;
  double* d_rowha = nullptr;
  int* d_flag = nullptr;      // iso_check_kernel's result
  double* d_wf64 = nullptr;   // FP32 batches: FP64 slowness, wavefront-major, per orientation
  char* d_wmed = nullptr;     // batched sweeps: wavefront-major media, per orientation
  char* d_wmedmw = nullptr;   // FP32 batches: FP64 wavefront-major media of the re-solve geometry
  GridDev gmw{};              // FP32 batches: FP64 media, resolve_waves() waves per grid (re-solves)
  GridDev gseg{};             // g with one wave per pair (per-pair kernels; equals g when W == 1)
  int W = 1;                  // waves per pair of the batched sweeps
  int Wseg = 1;               // waves per pair of the one-pair kernels (> 1: a grid wider than 512 columns)
  // FP32 batches, one wave per pair, at most 32 lanes of <= kPair2MaxC columns:
  // the forward geometry of the two-pairs-per-wave sweeps (isotropic metric)
  bool p2 = false;
  GridDev gp2{};
  char* d_wmedp2 = nullptr;   // its wavefront-major media
  // FP32 batches, one wave per pair, fewer than kAdjC columns per lane: the
  // adjoint's geometry, kAdjC columns per lane (adjoint_geometry)
  bool adj8 = false;
  GridDev gadj{};

  // wavefront-major elements of one event's work plane (every geometry its
  // forward and adjoint planes may take)
  size_t work_wplane() const {
    size_t w = g.wplane;
    if (p2) w = std::max(w, gp2.wplane);
    if (adj8) w = std::max(w, gadj.wplane);
    return w;
  }
  // the geometry of the adjoint phase (face fluxes, delta, sweeps, Ta)
  const GridDev& gadjoint() const { return adj8 ? gadj : g; }
  bool p2_on = false;          // the batch's groups run this grid's pairs two per wave
  // the forward geometry of the batched sweeps for the current grouping
  const GridDev& gfwd() const { return p2_on ? gp2 : g; }

  void release() {
    void* ps[] = {d_lon, d_lat, d_xx, d_yy, d_model, d_fp, d_aA, d_b, d_c, d_v3, d_seg, d_seg64, d_trig_yy, d_rowtab,
                  d_rowha, d_wf64, d_wmed, d_wmedmw, d_flag, d_wmedp2};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    *this = GridHost();
  }
};

static size_t elem_size(int prec) { return prec == SAT_PREC_F64 ? sizeof(double) : sizeof(float); }

// wave: also the wavefront-major media of the batched sweeps, laid out for W
// waves per pair (W * 64 lanes, ceil(ceil(nx / 64) / W) columns per lane).
static int grid_setup(GridHost& G, int prec, int nx, int ny, int nper, const double* lon, const double* lat,
                      bool wave, int W = 1, bool polish = false, int wr = 0) {
  if (nx < 3 || ny < 3 || nper < 1 || !lon || !lat) return SAT_EINVAL;
  if (!uniform_axis(lon, nx) || !uniform_axis(lat, ny)) return SAT_EINVAL;
  const Layout L = layout_for(nx);
  if (L.W == 0) return SAT_EINVAL;  // wider than kMaxWideW * 512 columns
  G.nx = nx; G.ny = ny; G.nper = nper; G.prec = prec;
  G.Wseg = L.W;
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
  SAT_CHECK(dalloc(&G.d_rowha, ny));
  SAT_CHECK(dalloc(&G.d_flag, 1));
  g.trig_yy = G.d_trig_yy;
  g.rowtab = G.d_rowtab;
  g.rowha = G.d_rowha;
  g.kb = 1.0 / (kR * g.dy);  // sqrt(b)/dy, b = 1/R^2
  g.iso = 0;
  hipLaunchKernelGGL(row_kernel, dim3((ny + 63) / 64), dim3(64), 0, nullptr, g, G.d_trig_yy, G.d_rowtab, G.d_rowha);
  SAT_CHECK(hipGetLastError());
  SAT_CHECK(hipDeviceSynchronize());
  g.wplane = (size_t)(ny + L.La - 1) * L.La * L.C;
  for (int i = 0; i < 6; i++) g.wmed[i][0] = g.wmed[i][1] = nullptr;
  G.g64 = g;
  if (G.d_seg64)
    for (int i = 0; i < 6; i++) G.g64.seg[i] = G.d_seg64 + (size_t)i * seg64_bytes;
  G.gseg = g;  // one wave per pair: the segmented layout of the per-pair kernels
  G.W = W;
  if (wave && W > 1 && G.Wseg == 1) {  // the batched sweeps' wave geometry (the segmented pitch P stays)
    const WaveGeom wg = wave_geom(nx, ny, W);
    g.C = wg.C;
    g.La = wg.La;
    g.wplane = wg.wplane;
  }
  if (wave) {  // padding entries stay zero
    const int nmed = prec == SAT_PREC_F64 ? 4 : 6;
    const size_t pb = (size_t)nper * g.wplane * elem_size(prec);
    SAT_CHECK(dalloc(&G.d_wmed, (size_t)nmed * 2 * pb));
    SAT_CHECK(hipMemset(G.d_wmed, 0, (size_t)nmed * 2 * pb));
    for (int i = 0; i < nmed; i++)
      for (int o = 0; o < 2; o++) g.wmed[i][o] = G.d_wmed + (size_t)(2 * i + o) * pb;
    if (prec == SAT_PREC_F32 && polish && G.Wseg == 1) {  // (wider grids take the plain FP32 sweeps)
      // the FP64 slowness of the polished isotropic sweeps (g only: the FP64
      // copies g64 and gmw were taken above and keep wf64 null)
      const size_t pf = (size_t)nper * g.wplane;
      SAT_CHECK(dalloc(&G.d_wf64, 2 * pf));
      SAT_CHECK(hipMemset(G.d_wf64, 0, 2 * pf * sizeof(double)));
      g.wf64[0] = G.d_wf64;
      g.wf64[1] = G.d_wf64 + pf;
    }
    if (prec == SAT_PREC_F32) {
      // re-solve geometry: Wr = resolve_waves(ceil(nx / 64)) waves per grid, or
      // the batch's common wr, C = ceil(nx / (64 Wr))
      GridDev& m = G.gmw;
      m = G.g64;
      const WaveGeom wg = wave_geom(nx, ny, wr > 0 ? wr : resolve_waves((nx + kWave - 1) / kWave));
      m.C = wg.C;
      m.La = wg.La;
      m.P = m.La * m.C;
      m.wplane = wg.wplane;
      if (G.Wseg > 1) {  // wide grids: the batch's own geometry, Wseg waves of kMaxC columns
        m.C = L.C;
        m.La = L.La;
        m.P = L.P;
        m.wplane = (size_t)(ny + L.La - 1) * L.La * L.C;
      }
      for (int i = 0; i < 6; i++) m.seg[i] = nullptr;
      const size_t pbm = (size_t)nper * m.wplane * sizeof(double);
      SAT_CHECK(dalloc(&G.d_wmedmw, (size_t)4 * 2 * pbm));
      SAT_CHECK(hipMemset(G.d_wmedmw, 0, (size_t)4 * 2 * pbm));
      for (int i = 0; i < 4; i++)
        for (int o = 0; o < 2; o++) m.wmed[i][o] = G.d_wmedmw + (size_t)(2 * i + o) * pbm;
    }
    const int c2 = (nx + kHalfWave - 1) / kHalfWave;
    if (prec == SAT_PREC_F32 && !polish && W == 1 && G.Wseg == 1 && c2 <= kPair2MaxC) {
      // two pairs per wave: 32 lanes of c2 columns each
      GridDev& q = G.gp2;
      q = g;
      q.C = c2;
      q.La = (nx + c2 - 1) / c2;
      q.wplane = (size_t)(ny + q.La - 1) * q.La * q.C;
      for (int i = 0; i < 6; i++) q.seg[i] = nullptr;
      q.wf64[0] = q.wf64[1] = nullptr;
      const size_t pb2 = (size_t)nper * q.wplane * sizeof(float);
      SAT_CHECK(dalloc(&G.d_wmedp2, (size_t)6 * 2 * pb2));
      SAT_CHECK(hipMemset(G.d_wmedp2, 0, (size_t)6 * 2 * pb2));
      for (int i = 0; i < 6; i++)
        for (int o = 0; o < 2; o++) q.wmed[i][o] = G.d_wmedp2 + (size_t)(2 * i + o) * pb2;
      G.p2 = true;
    }
    if (prec == SAT_PREC_F32 && W == 1 && G.Wseg == 1 && g.C < kAdjC) {
      // The FP32 adjoint with kAdjC columns per lane.  Its sweeps store 4-byte
      // values, and the two sweeps of a round that change the planes'
      // orientation (1 and 3) store each lane's C values into a different run;
      // pieces of fewer than 32 bytes are written back as partial 32-byte
      // sectors, each costing a whole sector (C = 4: twice the bytes written,
      // measured with the TCC write counters).  Eight FP32 columns per lane
      // make every piece a whole sector; nx <= 448 keeps one wave.
      GridDev& a = G.gadj;
      a = g;
      a.C = kAdjC;
      a.La = (nx + kAdjC - 1) / kAdjC;
      a.P = a.La * a.C;
      a.wplane = (size_t)(ny + a.La - 1) * a.La * a.C;
      for (int i = 0; i < 6; i++) a.wmed[i][0] = a.wmed[i][1] = nullptr;
      a.wf64[0] = a.wf64[1] = nullptr;
      G.adj8 = true;
    }
  }
  return SAT_OK;
}

// Upload four planes (nper each) and rebuild the media on `stream`.
static int grid_media(GridHost& G, int mode, const double* p0, const double* p1, const double* p2,
                      const double* p3, hipStream_t stream) {
  const size_t N = (size_t)G.nx * G.ny, n = (size_t)G.nper * N;
  const double* src[4] = {p0, p1, p2, p3};
  for (int i = 0; i < 4; i++)
    if (!src[i]) return SAT_EINVAL;
  // the planes come from host or device memory (hipMemcpyDefault tells them apart)
  for (int i = 0; i < 4; i++) {
    SAT_CHECK(hipMemcpyAsync(G.d_model + i * n, src[i], sizeof(double) * n, hipMemcpyDefault, stream));
  }
  {  // isotropic metric: spha = sphb = 1 and sphc = 0 everywhere
    const double* ma = G.d_model + (size_t)((mode == 0) ? 1 : 0) * n;  // m11 | spha
    const double* mb = G.d_model + (size_t)((mode == 0) ? 3 : 1) * n;  // m22 | sphb
    const double* mc = G.d_model + 2 * n;                             // m12 | sphc
    SAT_CHECK(hipMemsetAsync(G.d_flag, 0, sizeof(int), stream));
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(iso_check_kernel, dim3(blocks), dim3(256), 0, stream, ma, mb, mc, n, G.d_flag);
    SAT_CHECK(hipGetLastError());
    int flag = 0;
    SAT_CHECK(hipMemcpyAsync(&flag, G.d_flag, sizeof(int), hipMemcpyDeviceToHost, stream));
    SAT_CHECK(hipStreamSynchronize(stream));
    G.g.iso = flag ? 0 : 1;
    G.g64.iso = G.g.iso;
    G.gmw.iso = G.g.iso;
    G.gseg.iso = G.g.iso;
    G.gp2.iso = G.g.iso;
    G.gadj.iso = G.g.iso;
  }
  const size_t total = (size_t)G.nper * G.ny * G.g.P;
  const int threads = 256;
  const int blocks = (int)std::min<size_t>((total + threads - 1) / threads, 65535);
  // passes: the sweep-precision media (g), the FP64 media of FP32 batches
  // (g64), their re-solve geometry (gmw), the two-pairs-per-wave geometry (gp2)
  const GridDev* passes[4];
  int npass = 0;
  passes[npass++] = &G.g;
  if (G.d_seg64) passes[npass++] = &G.g64;
  if (G.d_wmedmw) passes[npass++] = &G.gmw;
  if (G.p2) passes[npass++] = &G.gp2;
  for (int pass = 0; pass < npass; pass++) {
    const GridDev& gd = *passes[pass];
    void* sg[6];
    for (int i = 0; i < 6; i++) sg[i] = const_cast<void*>(gd.seg[i]);  // null: no segmented planes
    const int mprec = (pass == 0 || &gd == &G.gp2) ? G.prec : SAT_PREC_F64;
    hipLaunchKernelGGL(media_kernel, dim3(blocks), dim3(threads), 0, stream, gd, mode, mprec,
                       G.d_model, G.d_model + n, G.d_model + 2 * n, G.d_model + 3 * n, G.d_fp, G.d_aA, G.d_b,
                       G.d_c, (mode == 0) ? G.d_v3 : nullptr, sg[0], sg[1], sg[2], sg[3], sg[4], sg[5]);
    SAT_CHECK(hipGetLastError());
  }
  return SAT_OK;
}

static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Elements between the two halves' planes of a two-pairs-per-wave sweep slot.
static size_t pair2_hstride(const GridDev& g) { return round_up(g.wplane, 64); }

// Slot geometry for the largest grid of a precision.
struct SlotGeom {
  size_t seg, plain, slot;
};
static SlotGeom slot_geom(int prec, size_t max_seg_elems, size_t max_plane, int W = 1) {
  SlotGeom s;
  const int nseg = prec == SAT_PREC_F64 ? Prec<SAT_PREC_F64>::kSeg : Prec<SAT_PREC_F32>::kSeg;
  const int npl = prec == SAT_PREC_F64 ? Prec<SAT_PREC_F64>::kPlain : Prec<SAT_PREC_F32>::kPlain;
  // FP32 path: sigma planes are FP64; every plane carries the sweeps' trash row
  s.seg = round_up((max_seg_elems + seg_trash_elems(W)) * sizeof(double), 256);
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

int solo_run(bool forward, int precision, int nx, int ny, const double* lon, const double* lat,
             const double* spha, const double* sphb, const double* sphc, const double* fun_or_T, double x0,
             double y0, int nr, const double* rlon, const double* rlat, const double* adj, double* out,
             sat_info* info) {
  if (precision == SAT_PREC_F32P) precision = SAT_PREC_F32;  // the per-solve kernels have no polished update
  if (precision != SAT_PREC_F64 && precision != SAT_PREC_F32) return SAT_EINVAL;
  if (!spha || !sphb || !sphc || !fun_or_T || !out) return SAT_EINVAL;
  if (!forward && nr > 0 && (!rlon || !rlat || !adj)) return SAT_EINVAL;
  if (nr < 0) return SAT_EINVAL;
  int rc = device_check();
  if (rc) return rc;
  GridHost G;
  rc = grid_setup(G, precision, nx, ny, 1, lon, lat, false);
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
  const SlotGeom sg = slot_geom(precision, G.g.seg_plane, N, G.Wseg);
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
    SOLO_CHECK(precision == SAT_PREC_F64 ? launch_solo<SAT_PREC_F64>(1, G.g.C, G.Wseg, stream, sd)
                                         : launch_solo<SAT_PREC_F32>(1, G.g.C, G.Wseg, stream, sd));
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
    sd.adj_lds_bytes = adj_lds_bytes(precision, G.g.seg_plane, G.Wseg);
    SOLO_CHECK(precision == SAT_PREC_F64 ? launch_solo<SAT_PREC_F64>(0, G.g.C, G.Wseg, stream, sd)
                                         : launch_solo<SAT_PREC_F32>(0, G.g.C, G.Wseg, stream, sd));
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

// ----------------------------------------------------------------------------
// Allocation-free per-solve forward (sat_eik_forward_noalloc): the same
// kernels as sat_eik_forward, with every plane they need carved out of one
// file-scope device array instead of allocated per call.  Each GPU holds its
// own copy of the array for the life of the process; every call on a GPU
// shares it, so calls that overlap on one GPU race.
// ----------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(24) << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// The array's device address on the current device: hipGetSymbolAddress once
// per device (the host-side name of the array is not a device address).
int scratch_address(void** out) {
  static std::mutex mu;
  static void* addr[kNoAllocMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices) return SAT_EDEVICE;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      std::fprintf(stderr, "[surfatt_eik] hipGetSymbolAddress(scratch): %s\n", hipGetErrorString(e));
      return SAT_EDEVICE;
    }
    addr[dev] = p;
  }
  *out = addr[dev];
  return SAT_OK;
}

}  // namespace

namespace sat {
// Axes in radians (src/eikonal_solver.cpp:41-47) and the allocating entry's
// uniform-axis test (uniform_axis) on device-resident axes: meta[0] = 1 if
// both axes are uniform and ascending, meta[1] = dx, meta[2] = dy.
__global__ void noalloc_axes_kernel(int nx, int ny, const double* lon, const double* lat, double* xx, double* yy,
                                    double* meta) {
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const double dlon = lon[1] - lon[0], dlat = lat[1] - lat[0];
  if (threadIdx.x == 0 && (!(dlon > 0.0) || !isfinite(dlon) || !(dlat > 0.0) || !isfinite(dlat))) bad = 1;
  for (int i = threadIdx.x; i < nx; i += blockDim.x) {
    xx[i] = lon[i] * kDEG2RAD;
    if (i > 0 && !(fabs((lon[i] - lon[i - 1]) - dlon) <= 1e-6 * fabs(dlon))) atomicOr(&bad, 1);
  }
  for (int j = threadIdx.x; j < ny; j += blockDim.x) {
    yy[j] = lat[j] * kDEG2RAD;
    if (j > 0 && !(fabs((lat[j] - lat[j - 1]) - dlat) <= 1e-6 * fabs(dlat))) atomicOr(&bad, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    meta[0] = bad ? 0.0 : 1.0;
    meta[1] = lon[1] * kDEG2RAD - lon[0] * kDEG2RAD;  // xx[1] - xx[0]
    meta[2] = lat[1] * kDEG2RAD - lat[0] * kDEG2RAD;
  }
}

// Zero n 8-byte words (a kernel, not a memset node).
__global__ void fill_zero_kernel(unsigned long long* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}
}  // namespace sat

namespace {

int forward_noalloc(int precision, int nx, int ny, const double* lon, const double* lat, const double* spha,
                    const double* sphb, const double* sphc, const double* fun, double x0, double y0, double* T_out,
                    sat_info* info) {
  if (precision == SAT_PREC_F32P) precision = SAT_PREC_F32;  // the per-solve kernels have no polished update
  if (precision != SAT_PREC_F64 && precision != SAT_PREC_F32) return SAT_EINVAL;
  if (!lon || !lat || !spha || !sphb || !sphc || !fun || !T_out) return SAT_EINVAL;
  if (nx < 3 || ny < 3 || !std::isfinite(x0) || !std::isfinite(y0)) return SAT_EINVAL;
  const Layout L = layout_for(nx);
  if (L.W == 0) return SAT_EINVAL;
  int rc = device_check();
  if (rc) return rc;
  // the scratch array's layout for this grid
  const size_t N = (size_t)nx * ny, S = (size_t)L.P * ny, es = elem_size(precision);
  const SlotGeom sg = slot_geom(precision, S, N, L.W);
  size_t off = 0;
  auto take = [&](size_t bytes) { const size_t o = off; off += round_up(bytes, 256); return o; };
  const size_t o_xx = take(nx * sizeof(double)), o_yy = take(ny * sizeof(double));
  const size_t o_meta = take(4 * sizeof(double)), o_info = take(4 * sizeof(double));
  const size_t o_trig = take(ny * sizeof(double2)), o_row = take(ny * sizeof(float2));
  size_t o_pl[4], o_seg[6];
  for (int i = 0; i < 4; i++) o_pl[i] = take(N * sizeof(double));
  for (int i = 0; i < 6; i++) o_seg[i] = take(S * es);
  const size_t o_slot = take(sg.slot);
  if (off > kNoAllocScratchBytes) {
    std::fprintf(stderr,
                 "[surfatt_eik] sat_eik_forward_noalloc: a %dx%d grid needs %zu bytes of scratch; the static "
                 "scratch holds %zu (use sat_eik_forward)\n",
                 nx, ny, off, kNoAllocScratchBytes);
    return SAT_EINVAL;
  }
  void* base = nullptr;
  rc = scratch_address(&base);
  if (rc) return rc;
  char* sp = static_cast<char*>(base);
  auto at = [&](size_t o) { return sp + o; };
  double* meta = reinterpret_cast<double*>(at(o_meta));
  hipStream_t stream = nullptr;
  hipLaunchKernelGGL(noalloc_axes_kernel, dim3(1), dim3(256), 0, stream, nx, ny, lon, lat,
                     reinterpret_cast<double*>(at(o_xx)), reinterpret_cast<double*>(at(o_yy)), meta);
  SAT_CHECK(hipGetLastError());
  double hm[3];
  SAT_CHECK(hipMemcpy(hm, meta, sizeof(hm), hipMemcpyDeviceToHost));
  if (hm[0] != 1.0) return SAT_EINVAL;  // non-uniform or descending axis
  GridDev g{};
  g.nx = nx; g.ny = ny; g.nperiod = 1;
  g.C = L.C; g.La = L.La; g.P = L.P;
  g.seg_plane = S; g.plane = N;
  g.dx = hm[1]; g.dy = hm[2];
  g.lon = lon; g.lat = lat;
  g.xx = reinterpret_cast<const double*>(at(o_xx));
  g.yy = reinterpret_cast<const double*>(at(o_yy));
  double* pl[4];
  for (int i = 0; i < 4; i++) pl[i] = reinterpret_cast<double*>(at(o_pl[i]));
  g.f_plain = pl[0]; g.aA_plain = pl[1]; g.b_plain = pl[2]; g.c_plain = pl[3]; g.vinv3 = nullptr;
  for (int i = 0; i < 6; i++) g.seg[i] = at(o_seg[i]);
  g.trig_yy = reinterpret_cast<const double2*>(at(o_trig));
  g.rowtab = reinterpret_cast<const float2*>(at(o_row));
  g.iso = 0;
  g.wplane = (size_t)(ny + L.La - 1) * L.La * L.C;
  for (int i = 0; i < 6; i++) g.wmed[i][0] = g.wmed[i][1] = nullptr;
  hipLaunchKernelGGL(row_kernel, dim3((ny + 63) / 64), dim3(64), 0, stream, g, reinterpret_cast<double2*>(at(o_trig)),
                     reinterpret_cast<float2*>(at(o_row)), nullptr);
  SAT_CHECK(hipGetLastError());
  {  // media mode 1: (spha, sphb, sphc, fun) as FSM_UW_PS_lonlat_2d takes them
    const int blocks = (int)std::min<size_t>((S + 255) / 256, 65535);
    hipLaunchKernelGGL(media_kernel, dim3(blocks), dim3(256), 0, stream, g, 1, precision, spha, sphb, sphc, fun, pl[0],
                       pl[1], pl[2], pl[3], nullptr, at(o_seg[0]), at(o_seg[1]), at(o_seg[2]), at(o_seg[3]),
                       at(o_seg[4]), at(o_seg[5]));
    SAT_CHECK(hipGetLastError());
  }
  {  // the slot starts at zero, as the allocating entry's does (the previous call's planes are stale)
    const size_t words = sg.slot / 8;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)std::min<size_t>((words + 255) / 256, 4096)), dim3(256), 0,
                       stream, reinterpret_cast<unsigned long long*>(at(o_slot)), words);
    SAT_CHECK(hipGetLastError());
  }
  SoloDev sd{};
  sd.g = g;
  sd.src_lon = x0; sd.src_lat = y0;
  sd.out = T_out; sd.ws = at(o_slot); sd.ws_seg = sg.seg; sd.ws_plain = sg.plain;
  sd.info = reinterpret_cast<double*>(at(o_info));
  sd.maxiter = kFwdMaxIter;
  SAT_CHECK(precision == SAT_PREC_F64 ? launch_solo<SAT_PREC_F64>(1, g.C, L.W, stream, sd)
                                      : launch_solo<SAT_PREC_F32>(1, g.C, L.W, stream, sd));
  double hinfo[4];
  SAT_CHECK(hipMemcpy(hinfo, sd.info, sizeof(hinfo), hipMemcpyDeviceToHost));  // waits for the solve
  if (info) {
    info->iters = (int)hinfo[0]; info->flags = (int)hinfo[1]; info->l1 = hinfo[2]; info->linf = hinfo[3];
  }
  return SAT_OK;
}

}  // namespace

struct sat_batch {
  int precision = SAT_PREC_F64;    // SAT_PREC_F64 or SAT_PREC_F32 (also for SAT_PREC_F32P batches)
  bool polish = false;             // SAT_PREC_F32P: isotropic grids sweep with the FP64-polished update
  int calc_adj = 1;
  int nevent = 0, nrec = 0;
  std::vector<GridHost> grids;
  std::vector<EventDev> he;        // host copy of the events
  std::vector<EventX> hx;          // host copy of the phase events
  std::vector<int> base_flags;     // per event, set at creation (source on the boundary)
  GridDev* d_grids = nullptr;
  GridDev* d_grids_fwd = nullptr;  // the forward geometry of each grid (GridHost::gfwd)
  GridDev* d_grids_seg = nullptr;  // one-wave geometry of the per-pair kernels
  GridDev* d_grids_adj = nullptr;  // the adjoint phase's geometry (GridHost::gadjoint)
  GridDev* d_grids64 = nullptr;    // FP32 batches: grids with FP64 media (per-pair re-solves)
  GridDev* d_grids64mw = nullptr;  // FP32 batches: FP64 media, re-solve geometry (batched re-solves)
  EventDev* d_events = nullptr;
  EventX* d_evx = nullptr;
  int* d_order = nullptr;          // per launch group: the forward queue (event ids)
  int* d_order_adj = nullptr;      // and the adjoint queue, each longest pairs first (order_kernel)
  int* d_order_tmp = nullptr;
  int2* d_group_rng = nullptr;     // per launch group: first, count (positions in the queues)
  bool order_by_rounds = true;     // SAT_OPT_ORDER (SURFATT_ORDER=0 at creation: queues in batch order)
  int opt_pair2 = -1;              // SAT_OPT_PAIR2: -1 the library's rule (SURFATT_PAIR2 overrides), 0, 1
  bool opt_adj_cols = true;        // SAT_OPT_ADJ_COLS: the FP32 adjoint of one-wave grids on kAdjC columns
  bool opt_spec = true;            // SAT_OPT_SPECULATE: FP64 from the start for recently re-solved pairs
  unsigned char* d_age = nullptr;  // per event: runs its speculation lasts (spec_age_kernel)
  int* d_spec = nullptr;           // speculated pairs: all | isotropic re-solve update | general (spec_kernel)
  int* d_group_iso = nullptr;      // per launch group: 1 if its re-solves take the isotropic update
  int wr_all = 0;                  // FP32 batches: waves per pair of every grid's re-solves (0: per grid)
  hipEvent_t spec_ev = nullptr;    // the column tables are ready: the speculated solves may start
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
  int* d_fb = nullptr;             // re-solve lists (FP64 phase kernels), group g at group_first[g]
  int* d_ov = nullptr;             // re-solves beyond the pool (per-pair kernel), same layout
  EventX* d_evx64 = nullptr;       // FP32 batches: events with the re-solve pool offsets
  char* d_work64 = nullptr;        // FP32 batches: FP64 re-solve pool (pool_size entries)
  char* d_slots64 = nullptr;       // FP32 batches: sweep slots of the re-solve kernels
  int pool_size = 0, nslots64p = 0;
  size_t pool_stride = 0, slot_plane64 = 0, slot_bytes64 = 0;
  double f32_band = 0.02;
  // per-event planes of the phase pipeline
  size_t w_total = 0, pl_total = 0, col_total = 0;
  double* d_evs = nullptr;
  double2* d_coltab = nullptr;
  char* d_work = nullptr;          // wavefront-major: t0, t0x, t0y x 2 orientations, then the
                                   // adjoint's face fluxes and delta x 2 in the same bytes
  double* d_T0d = nullptr;
  unsigned long long* d_dmask = nullptr;  // FP32: per event, the delta row mask (PhaseDev::dmask)
  int dmask_stride = 0;
  double* d_T = nullptr;
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
  // group_C: columns per lane of one wave per pair (per-pair kernels, re-solves);
  // group_W, group_wC: waves per pair and columns per lane of the batched sweeps
  std::vector<int> group_C, group_iso, group_first, group_count, group_W, group_wC;
  std::vector<int> group_p2;       // > 0: the two-pairs-per-wave forward sweeps' columns per lane (one grid per group)
  std::vector<size_t> group_p2h;   // and their half-plane stride (pair2_hstride)
  std::vector<int> group_wide;     // 1: a grid wider than 512 columns (W waves of kMaxC columns, every kernel)
  std::vector<int> group_adj8;     // 1: the FP32 adjoint sweeps kAdjC columns per lane on one wave (gadj)
  // per group: sweep slots [group_slot0, group_slot0 + group_nblk); launch order
  // of the groups (most work first)
  std::vector<int> group_slot0, group_nblk, launch_order;
  std::vector<double> group_work;  // sweep steps x columns per lane over the group's pairs
  // per group: dynamic LDS of its one-wave per-pair launches (the adjoint's LDS
  // tier; 0 where no grid of the group fits), batch precision and FP64
  std::vector<size_t> group_lds, group_lds64;
  // Chunks: consecutive ranges of the batch order whose per-event device
  // workspace fits the cap, run one after the other (one chunk when uncapped):
  // events, receiver span, groups [chunk_g0[c], chunk_g0[c+1]).
  std::vector<int> chunk_first, chunk_count, chunk_rec_lo, chunk_rec_n, chunk_g0;
  size_t ws_cap = 0, ws_event_max = 0;
  bool groups_dirty = true;
  int max_nx = 0;
  size_t max_seg = 0, max_plane = 0, max_wplane = 0;
  int nplanes = 0;
  size_t* d_red_sz = nullptr;      // plane_off | node_prefix
  int* d_red_i = nullptr;          // plane_n | ev_off | ev_ids
  const double** d_red_v3 = nullptr;
  ReduceDev rd{};
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;   // FP64 re-solves, concurrent with the FP32 adjoint phase
  // Launch groups of different grid shapes run concurrently: group k of the
  // launch order on `stream` (k % (kSide+1) == 0) or side[k % (kSide+1) - 1].
  // The FP64 re-solves fan out from stream2 the same way (side2).
  static constexpr int kSide = 3;
  hipStream_t side[kSide] = {}, side2[kSide] = {};
  hipEvent_t fork_ev = nullptr, fork2_ev = nullptr;
  hipEvent_t join_ev[kSide] = {}, join2_ev[kSide] = {};
  // time marks: start, init done, forward sweeps done, adjoint set-up done,
  // adjoint sweeps done, re-solves start / end (stream2), end
  static constexpr int kMarks = 8;
  hipEvent_t mk[kMarks] = {};
  // SAT_RUN_GRAPH: the captured batched run and its executable instance
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool last_graph = false;
  bool ran = false;                // a run has been queued: every event's T and Ta are on the device
  // accumulate_kernels' optional terms, formed inside every run (chunks included):
  // the anisotropic terms of the grids with a registered metric and the density kernel
  std::vector<double*> d_aniso;    // per grid: a, b, c, then xi, eta ([nper][ny][nx] each), or null
  std::vector<char> aniso_ran;     // per grid: the last run formed its terms
  bool kden = false, kden_ran = false;
  double* d_Kden = nullptr;        // [K_total] density accumulators (kden_s_local)

  void release() {
    for (auto& G : grids) G.release();
    for (double* p : d_aniso)
      if (p) (void)hipFree(p);
    d_aniso.clear();
    if (d_Kden) (void)hipFree(d_Kden);
    void* ps[] = {d_grids, d_grids_fwd, d_grids_seg, d_grids_adj, d_grids64, d_grids64mw, d_events, d_evx, d_order, d_order_adj, d_order_tmp, d_group_rng, d_counter, d_rec, d_rec_ev, d_rcell, d_rwt, d_chi_ev,
                  d_chi, d_iters, d_base_flags, d_conv, d_skip, d_fb, d_ov, d_evx64, d_work64, d_slots64, d_evs, d_coltab, d_work, d_T0d, d_T, d_ta,
                  d_slots, d_ws, d_ws64, d_K, d_red_sz, d_red_i, (void*)d_red_v3, d_age, d_spec, d_group_iso, d_dmask};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : mk)
      if (e) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (fork2_ev) (void)hipEventDestroy(fork2_ev);
    if (spec_ev) (void)hipEventDestroy(spec_ev);
    for (int i = 0; i < kSide; i++) {
      if (join_ev[i]) (void)hipEventDestroy(join_ev[i]);
      if (join2_ev[i]) (void)hipEventDestroy(join2_ev[i]);
      if (side[i]) (void)hipStreamDestroy(side[i]);
      if (side2[i]) (void)hipStreamDestroy(side2[i]);
    }
    if (stream) (void)hipStreamDestroy(stream);
    if (stream2) (void)hipStreamDestroy(stream2);
  }
};

namespace {

// Sort events into launch groups (same C and isotropy), largest grids first.
int build_groups(sat_batch* b) {
  const int ne = b->nevent;
  std::vector<int> order(ne);
  std::iota(order.begin(), order.end(), 0);
  // Two pairs per wave for the grids whose one-pair sweeps take one column per
  // lane (nx <= 64): there the lane-skew ramp is La - 1 of ny + La - 1 steps
  // and halving La pays (measured on the MI355X, 3,000 pairs: 61x71 forward
  // sweeps 5.9 -> 5.1 ms); with more columns per lane the halves' longer
  // per-step chains on fewer resident waves lose more than the ramp gains
  // (101x101 10.3 -> 11.0 ms, 201x201 33.5 -> 41.1 ms).  SURFATT_PAIR2=0 / 1
  // switches it off / on wherever a grid allows it (tests, tuning).
  std::vector<char> use_p2(b->grids.size(), 0);
  {
    const char* env = std::getenv("SURFATT_PAIR2");
    const int force = b->opt_pair2 >= 0 ? b->opt_pair2 : (env ? std::atoi(env) : -1);
    for (size_t gi = 0; gi < b->grids.size(); gi++) {
      const GridHost& G = b->grids[gi];
      if (!G.p2 || !G.g.iso) continue;
      use_p2[gi] = force >= 0 ? (force > 0) : (G.g.C == 1);
    }
  }
  auto key = [&](int e) {  // metric mode: kMetricGeneral, kMetricIso, kMetricIsoPolish
    const GridHost& G = b->grids[b->he[e].grid];
    // two pairs per wave: a group holds one grid (the two halves share its geometry)
    const int p2grid = use_p2[b->he[e].grid] ? b->he[e].grid : -1;
    int mode;
    if (b->precision == SAT_PREC_F64) mode = G.Wseg > 1 ? kMetricExact : (G.g.iso ? kMetricIso : kMetricGeneral);
    else mode = G.g.iso ? ((b->polish && G.g.wf64[0]) ? kMetricIsoPolish : kMetricIso) : kMetricGeneral;
    return std::make_tuple(G.W, G.gseg.C, mode, p2grid, G.Wseg > 1 ? 1 : 0);
  };
  b->group_W.clear();
  b->group_wC.clear();
  b->group_C.clear();
  b->group_iso.clear();
  b->group_p2.clear();
  b->group_p2h.clear();
  b->group_wide.clear();
  b->group_adj8.clear();
  b->group_first.clear();
  b->group_count.clear();
  b->group_slot0.clear();
  b->group_nblk.clear();
  b->group_lds.clear();
  b->group_lds64.clear();
  b->launch_order.clear();
  b->group_work.clear();
  b->chunk_g0.assign(1, 0);
  const int nchunk = (int)b->chunk_first.size();
  for (int c = 0; c < nchunk; c++) {
    // within a chunk: events sorted by launch group, largest grids first
    const int c0 = b->chunk_first[c], cn = b->chunk_count[c];
    std::stable_sort(order.begin() + c0, order.begin() + c0 + cn, [&](int a, int d) {
      if (key(a) != key(d)) return key(a) < key(d);
      return b->grids[b->he[a].grid].g.plane > b->grids[b->he[d].grid].g.plane;
    });
    const size_t g_lo = b->group_C.size();
    for (int i = c0; i < c0 + cn;) {
      const auto k = key(order[i]);
      int j = i;
      while (j < c0 + cn && key(order[j]) == k) j++;
      const bool p2 = std::get<3>(k) >= 0;
      b->group_W.push_back(std::get<0>(k));
      b->group_C.push_back(std::get<1>(k));
      // (a wide grid: kMaxC columns per lane on every one of its waves)
      const bool wide = std::get<4>(k) != 0;
      b->group_wC.push_back(wide ? kMaxC : (std::get<1>(k) + std::get<0>(k) - 1) / std::get<0>(k));
      b->group_wide.push_back(wide ? 1 : 0);
      b->group_adj8.push_back(b->opt_adj_cols && b->grids[b->he[order[i]].grid].adj8 ? 1 : 0);
      b->group_iso.push_back(std::get<2>(k));
      b->group_p2.push_back(p2 ? b->grids[std::get<3>(k)].gp2.C : 0);
      b->group_p2h.push_back(p2 ? pair2_hstride(b->grids[std::get<3>(k)].gp2) : 0);
      b->group_first.push_back(i);
      b->group_count.push_back(j - i);
      i = j;
    }
    const size_t g_hi = b->group_C.size();
    if (g_hi - g_lo > (size_t)kMaxGroups) return SAT_EINVAL;
    // Each group's persistent sweep kernels get a share of the sweep slots in
    // proportion to its events (rounded up: at most nslots_phase + kMaxGroups);
    // a chunk's groups run concurrently, the one with the most sweep steps first.
    std::vector<double> work(g_hi - g_lo, 0.0);
    int slot = 0;
    for (size_t gi = g_lo; gi < g_hi; gi++) {
      const long share = ((long)b->nslots_phase * b->group_count[gi] + cn - 1) / std::max(cn, 1);
      // sweep slots (waves) of the group: one per pair in flight (a two-pairs
      // group's forward sweeps use half of them, two pairs each)
      b->group_nblk.push_back((int)std::max(1L, std::min<long>(b->group_count[gi], share)));
      b->group_slot0.push_back(slot);
      slot += b->group_nblk.back();
      size_t lds = 0, lds64 = 0;
      for (int k = 0; k < b->group_count[gi]; k++) {
        const GridHost& G = b->grids[b->he[order[b->group_first[gi] + k]].grid];
        const GridDev& g = G.g;
        work[gi - g_lo] += (double)(g.ny + g.La - 1) * g.C;
        const int w = b->group_wide[gi] ? b->group_W[gi] : 1;  // the per-pair kernels' waves
        lds = std::max(lds, adj_lds_bytes(b->precision, std::max(G.gseg.seg_plane, g.seg_plane), w));
        lds64 = std::max(lds64, adj_lds_bytes(SAT_PREC_F64, std::max(G.g64.seg_plane, g.seg_plane), w));
      }
      b->group_lds.push_back(lds);
      b->group_lds64.push_back(lds64);
    }
    b->group_work.insert(b->group_work.end(), work.begin(), work.end());
    std::vector<int> lo(g_hi - g_lo);
    std::iota(lo.begin(), lo.end(), (int)g_lo);
    std::stable_sort(lo.begin(), lo.end(), [&](int a, int d) { return work[a - g_lo] > work[d - g_lo]; });
    b->launch_order.insert(b->launch_order.end(), lo.begin(), lo.end());
    b->chunk_g0.push_back((int)g_hi);
  }
  if (ne) {
    SAT_CHECK(hipMemcpy(b->d_order, order.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
    SAT_CHECK(hipMemcpy(b->d_order_adj, order.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
    std::vector<int2> rng(b->group_first.size());
    for (size_t gi = 0; gi < rng.size(); gi++) rng[gi] = make_int2(b->group_first[gi], b->group_count[gi]);
    if (!rng.empty()) SAT_CHECK(hipMemcpy(b->d_group_rng, rng.data(), sizeof(int2) * rng.size(), hipMemcpyHostToDevice));
    std::vector<int> giso(b->group_iso.size());
    for (size_t gi = 0; gi < giso.size(); gi++) giso[gi] = b->group_iso[gi] != kMetricGeneral ? 1 : 0;
    if (!giso.empty()) SAT_CHECK(hipMemcpy(b->d_group_iso, giso.data(), sizeof(int) * giso.size(), hipMemcpyHostToDevice));
  }
  std::vector<GridDev> hg(b->grids.size()), hg64(b->grids.size()), hgmw(b->grids.size()), hgs(b->grids.size());
  std::vector<GridDev> hgf(b->grids.size()), hga(b->grids.size());
  for (size_t gi = 0; gi < b->grids.size(); gi++) {
    b->grids[gi].p2_on = use_p2[gi] != 0;
    hg[gi] = b->grids[gi].g;
    hgf[gi] = b->grids[gi].gfwd();
    hga[gi] = b->opt_adj_cols ? b->grids[gi].gadjoint() : b->grids[gi].g;
    hg64[gi] = b->grids[gi].g64;
    hgmw[gi] = b->grids[gi].gmw;
    hgs[gi] = b->grids[gi].gseg;
    hgs[gi].iso = hg[gi].iso;
  }
  SAT_CHECK(hipMemcpy(b->d_grids, hg.data(), sizeof(GridDev) * hg.size(), hipMemcpyHostToDevice));
  SAT_CHECK(hipMemcpy(b->d_grids_fwd, hgf.data(), sizeof(GridDev) * hgf.size(), hipMemcpyHostToDevice));
  SAT_CHECK(hipMemcpy(b->d_grids_adj, hga.data(), sizeof(GridDev) * hga.size(), hipMemcpyHostToDevice));
  SAT_CHECK(hipMemcpy(b->d_grids_seg, hgs.data(), sizeof(GridDev) * hgs.size(), hipMemcpyHostToDevice));
  if (b->d_grids64) SAT_CHECK(hipMemcpy(b->d_grids64, hg64.data(), sizeof(GridDev) * hg64.size(), hipMemcpyHostToDevice));
  if (b->d_grids64mw)
    SAT_CHECK(hipMemcpy(b->d_grids64mw, hgmw.data(), sizeof(GridDev) * hgmw.size(), hipMemcpyHostToDevice));
  b->groups_dirty = false;
  return SAT_OK;
}

// Waves per pair of each grid's batched sweeps.  FP64 batches: one (the
// reference's operation order is kept by the one-wave kernels).  FP32 batches:
// the descriptor's value (1, 2 or 4), or for 0 the library's choice: start at
// one wave per pair and, while the batch's waves stay within three per SIMD
// (the residency the FP32 forward sweeps are compiled for), double the waves
// of the grid whose pairs take the most sweep steps
// ((ny + La - 1) * C per sweep), so that batches too small to fill the device
// split their longest pairs over more lanes.
int choose_waves(const sat_batch_desc& d, bool f32, bool polish, int cus, std::vector<int>& w) {
  const int ng = d.ngrid;
  w.assign(ng, 1);
  std::vector<long> nev(ng, 0);
  for (int e = 0; e < d.nevent; e++) nev[d.events[e].grid]++;
  std::vector<bool> autow(ng, false);
  for (int gi = 0; gi < ng; gi++) {
    const int v = d.grids[gi].waves;
    const int ww = wide_waves(d.grids[gi].nx);
    if (ww == 0) return SAT_EINVAL;
    if (ww > 1) {  // wider than 512 columns: wide_waves waves of kMaxC columns, in every precision
      if (v != 0 && v != ww) return SAT_EINVAL;
      w[gi] = ww;
      continue;
    }
    if (v != 0 && v != 1 && v != 2 && v != 4) return SAT_EINVAL;
    if (v > 1 && (!f32 || polish)) return SAT_EINVAL;  // FP64 and polished sweeps: one wave per pair
    if (v == 0) autow[gi] = f32 && !polish;
    else w[gi] = v;
  }
  auto steps = [&](int gi, int W) {  // sweep steps x columns of one pair
    const int C = ((d.grids[gi].nx + kWave - 1) / kWave + W - 1) / W;
    const int La = (d.grids[gi].nx + C - 1) / C;
    return (double)(d.grids[gi].ny + La - 1) * C;
  };
  const long cap = 3L * 4 * cus;  // three waves per SIMD
  long total = 0;
  for (int gi = 0; gi < ng; gi++) total += nev[gi] * w[gi];
  for (;;) {
    int best = -1;
    for (int gi = 0; gi < ng; gi++) {
      if (!autow[gi] || nev[gi] == 0 || w[gi] >= 4 || d.grids[gi].nx <= kWave * w[gi]) continue;
      if (total + nev[gi] * w[gi] > cap) continue;
      if (best < 0 || steps(gi, w[gi]) > steps(best, w[best])) best = gi;
    }
    if (best < 0) break;
    total += nev[best] * w[best];
    w[best] *= 2;
  }
  return SAT_OK;
}

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

namespace sat {
int combine_impl(int nx, int ny, int nz, int nper, int mode, const double* adj_s, const double* adj_dev,
                 const double* sen_vs, const double* sen_vp, const double* sen_rho, const double* vs3d,
                 const double* vp3d, double* ker_vs, double* ker_vp, double* ker_rho, double* kernel_ms);
}  // namespace sat

extern "C" {

int sat_eik_forward(int precision, int nx, int ny, const double* lon_deg, const double* lat_deg,
                    const double* spha, const double* sphb, const double* sphc, const double* fun,
                    double src_lon_deg, double src_lat_deg, double* T_out, sat_info* info) {
  return solo_run(true, precision, nx, ny, lon_deg, lat_deg, spha, sphb, sphc, fun, src_lon_deg, src_lat_deg, 0,
                  nullptr, nullptr, nullptr, T_out, info);
}

int sat_eik_forward_noalloc(int precision, int nx, int ny, const double* lon_deg, const double* lat_deg,
                            const double* spha, const double* sphb, const double* sphc, const double* fun,
                            double src_lon_deg, double src_lat_deg, double* T_out, sat_info* info) {
  return forward_noalloc(precision, nx, ny, lon_deg, lat_deg, spha, sphb, sphc, fun, src_lon_deg, src_lat_deg, T_out,
                         info);
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
  if (d->precision != SAT_PREC_F64 && d->precision != SAT_PREC_F32 && d->precision != SAT_PREC_F32P)
    return SAT_EINVAL;
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
  b->polish = d->precision == SAT_PREC_F32P;
  b->precision = b->polish ? SAT_PREC_F32 : d->precision;
  const bool f32 = b->precision == SAT_PREC_F32;
  b->calc_adj = d->calc_adj ? 1 : 0;
  b->nevent = d->nevent;
  b->nrec = d->nrec_total;
  B_CHECK(hipStreamCreate(&b->stream));
  B_CHECK(hipStreamCreate(&b->stream2));
  for (int i = 0; i < sat_batch::kMarks; i++) B_CHECK(hipEventCreate(&b->mk[i]));
  B_CHECK(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming));
  B_CHECK(hipEventCreateWithFlags(&b->fork2_ev, hipEventDisableTiming));
  B_CHECK(hipEventCreateWithFlags(&b->spec_ev, hipEventDisableTiming));
  for (int i = 0; i < sat_batch::kSide; i++) {
    B_CHECK(hipStreamCreate(&b->side[i]));
    B_CHECK(hipStreamCreate(&b->side2[i]));
    B_CHECK(hipEventCreateWithFlags(&b->join_ev[i], hipEventDisableTiming));
    B_CHECK(hipEventCreateWithFlags(&b->join2_ev[i], hipEventDisableTiming));
  }
  b->grids.resize(d->ngrid);
  b->d_aniso.assign(d->ngrid, nullptr);
  b->aniso_ran.assign(d->ngrid, 0);
  std::vector<int> waves;
  {
    hipDeviceProp_t prop;
    B_CHECK(hipGetDeviceProperties(&prop, 0));
    rc = choose_waves(*d, f32, b->polish, prop.multiProcessorCount, waves);
    if (rc) return fail(rc);
  }
  // FP32 batches whose grids (up to 512 columns) would re-solve on different
  // geometries re-solve all of them on kMultiWaveWide waves of one column per
  // lane, so that one launch per phase covers every speculated pair
  if (f32) {
    int w0 = 0, c0 = 0;
    for (int gi = 0; gi < d->ngrid; gi++) {
      const int c1 = (d->grids[gi].nx + kWave - 1) / kWave;
      if (c1 > kMaxC) continue;  // a wide grid keeps its own geometry
      const int w = resolve_waves(c1), c = (c1 + w - 1) / w;
      if (w0 && (w != w0 || c != c0)) b->wr_all = kMultiWaveWide;
      w0 = w;
      c0 = c;
    }
  }
  size_t max_seg = 0, max_plane = 0, max_wplane = 0;
  int max_nx = 0;
  for (int gi = 0; gi < d->ngrid; gi++) {
    const sat_grid_desc& gd = d->grids[gi];
    rc = grid_setup(b->grids[gi], b->precision, gd.nx, gd.ny, gd.nperiod, gd.lon_deg, gd.lat_deg, true, waves[gi],
                    b->polish, b->wr_all);
    if (rc) return fail(rc);
    rc = grid_media(b->grids[gi], 0, gd.svel, gd.m11, gd.m12, gd.m22, b->stream);
    if (rc) return fail(rc);
    max_seg = std::max(max_seg, b->grids[gi].g.seg_plane);
    max_plane = std::max(max_plane, b->grids[gi].g.plane);
    max_wplane = std::max(max_wplane, b->grids[gi].work_wplane());
    max_nx = std::max(max_nx, gd.nx);
  }
  b->max_seg = max_seg;
  b->max_plane = max_plane;
  b->max_wplane = max_wplane;
  b->max_nx = max_nx;
  B_CHECK(dalloc(&b->d_grids, d->ngrid));
  B_CHECK(dalloc(&b->d_grids_fwd, d->ngrid));
  B_CHECK(dalloc(&b->d_grids_seg, d->ngrid));
  B_CHECK(dalloc(&b->d_grids_adj, d->ngrid));
  if (f32) B_CHECK(dalloc(&b->d_grids64, d->ngrid));
  if (f32) B_CHECK(dalloc(&b->d_grids64mw, d->ngrid));

  // events: per-event plane offsets, index boxes, base flags
  const int ne = d->nevent;
  b->he.resize(ne);
  b->hx.resize(ne);
  b->base_flags.assign(std::max(ne, 1), 0);
  // Per-event workspace, in bytes: wavefront-major work planes, FP64 T (and
  // the FP64 T0 of FP32 sweeps), the masked Ta, the column trig table.
  // (the work planes hold 8-byte elements in polished batches: their
  // isotropic sweeps read the T0 gradient in FP64)
  const size_t es = elem_size(b->precision);
  const size_t wes = b->polish ? sizeof(double) : es;
  auto ev_bytes = [&](const GridHost& G) {
    const GridDev& g = G.g;
    const size_t w = round_up(G.work_wplane(), 64), pl = round_up(g.plane, 64), col = round_up(g.nx, 64);
    return 6 * w * wes + (f32 ? 2 : 1) * pl * sizeof(double) + (b->calc_adj ? pl * es : 0) + col * sizeof(double2);
  };
  size_t ws_all = 0;
  for (int e = 0; e < ne; e++) {
    const size_t eb = ev_bytes(b->grids[d->events[e].grid]);
    ws_all += eb;
    b->ws_event_max = std::max(b->ws_event_max, eb);
  }
  // The cap: the caller's, or -- when the whole batch would not fit in three
  // quarters of the free device memory -- half of the free memory.
  b->ws_cap = d->workspace_cap;
  if (b->ws_cap == 0) {
    size_t mfree = 0, mtotal = 0;
    B_CHECK(hipMemGetInfo(&mfree, &mtotal));
    if (ws_all > mfree / 4 * 3) b->ws_cap = mfree / 2;
  }
  if (b->ws_cap != 0 && b->ws_cap < b->ws_event_max) return fail(SAT_ENOMEM);
  size_t w_total = 0, pl_total = 0, col_total = 0;   // maxima over the chunks
  size_t w_c = 0, pl_c = 0, col_c = 0, bytes_c = 0;  // the current chunk
  for (int e = 0; e < ne; e++) {
    const sat_event_desc& ev = d->events[e];
    const sat_grid_desc& gd = d->grids[ev.grid];
    const GridDev& g = b->grids[ev.grid].g;
    const size_t eb = ev_bytes(b->grids[ev.grid]);
    if (e == 0 || (b->ws_cap != 0 && bytes_c + eb > b->ws_cap)) {  // a new chunk
      b->chunk_first.push_back(e);
      b->chunk_count.push_back(0);
      w_c = pl_c = col_c = bytes_c = 0;
    }
    b->chunk_count.back()++;
    bytes_c += eb;
    // offsets within the chunk's workspace
    b->he[e] = EventDev{ev.grid, ev.iper, ev.src_lon_deg, ev.src_lat_deg, ev.rec_off, ev.nrec, pl_c};
    EventX& x = b->hx[e];
    x.grid = ev.grid; x.iper = ev.iper; x.rec_off = ev.rec_off; x.nrec = ev.nrec;
    x.lon = ev.src_lon_deg; x.lat = ev.src_lat_deg;
    x.w_off = w_c; x.pl_off = pl_c; x.col_off = col_c;
    std::vector<double> xx(gd.nx), yy(gd.ny);
    for (int i = 0; i < gd.nx; i++) xx[i] = gd.lon_deg[i] * kDEG2RAD;
    for (int j = 0; j < gd.ny; j++) yy[j] = gd.lat_deg[j] * kDEG2RAD;
    event_boxes(b->grids[ev.grid], xx, yy, gd.lon_deg, gd.lat_deg, ev.src_lon_deg, ev.src_lat_deg, x,
                b->base_flags[e]);
    w_c += round_up(b->grids[ev.grid].work_wplane(), 64);
    pl_c += round_up(g.plane, 64);
    col_c += round_up(g.nx, 64);
    w_total = std::max(w_total, w_c);
    pl_total = std::max(pl_total, pl_c);
    col_total = std::max(col_total, col_c);
  }
  if (ne == 0) {
    b->chunk_first.push_back(0);
    b->chunk_count.push_back(0);
  }
  // each chunk's receiver span
  for (size_t c = 0; c < b->chunk_first.size(); c++) {
    int lo = INT_MAX, hi = 0;
    for (int e = b->chunk_first[c]; e < b->chunk_first[c] + b->chunk_count[c]; e++)
      if (d->events[e].nrec > 0) {
        lo = std::min(lo, d->events[e].rec_off);
        hi = std::max(hi, d->events[e].rec_off + d->events[e].nrec);
      }
    b->chunk_rec_lo.push_back(lo == INT_MAX ? 0 : lo);
    b->chunk_rec_n.push_back(lo == INT_MAX ? 0 : hi - lo);
  }
  b->w_total = w_total;
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
  B_CHECK(dalloc(&b->d_order_adj, ne));
  B_CHECK(dalloc(&b->d_order_tmp, ne));
  B_CHECK(dalloc(&b->d_group_rng, std::max(ne, 1)));  // a group holds at least one event
  if (const char* v = std::getenv("SURFATT_ORDER")) b->order_by_rounds = std::atoi(v) != 0;

  // receivers
  const size_t nr = (size_t)d->nrec_total;
  B_CHECK(dalloc(&b->d_rec, 6 * std::max<size_t>(nr, 1)));
  B_CHECK(hipMemset(b->d_rec, 0, sizeof(double) * 6 * std::max<size_t>(nr, 1)));  // receivers of no event: syn 0
  if (nr) {
    B_CHECK(hipMemcpy(b->d_rec, d->rec_lon_deg, sizeof(double) * nr, hipMemcpyDefault));
    B_CHECK(hipMemcpy(b->d_rec + nr, d->rec_lat_deg, sizeof(double) * nr, hipMemcpyDefault));
    B_CHECK(hipMemcpy(b->d_rec + 2 * nr, d->tt_obs, sizeof(double) * nr, hipMemcpyDefault));
    B_CHECK(hipMemcpy(b->d_rec + 3 * nr, d->weight, sizeof(double) * nr, hipMemcpyDefault));
  }
  B_CHECK(dalloc(&b->d_rec_ev, std::max<size_t>(nr, 1)));
  B_CHECK(hipMemcpy(b->d_rec_ev, rec_ev.data(), sizeof(int) * rec_ev.size(), hipMemcpyHostToDevice));
  B_CHECK(dalloc(&b->d_rcell, std::max<size_t>(nr, 1)));
  B_CHECK(dalloc(&b->d_rwt, 3 * std::max<size_t>(nr, 1)));
  B_CHECK(dalloc(&b->d_chi_ev, std::max(ne, 1)));
  B_CHECK(hipMemset(b->d_chi_ev, 0, sizeof(double) * std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_chi, 1));
  B_CHECK(dalloc(&b->d_iters, 3 * (size_t)std::max(ne, 1)));
  // no rounds before the first run: the first reorder keeps the grouped batch order
  B_CHECK(hipMemset(b->d_iters, 0, sizeof(int) * 3 * (size_t)std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_conv, 2 * (size_t)std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_skip, std::max(ne, 1)));
  B_CHECK(hipMemset(b->d_skip, 0, std::max(ne, 1)));  // no re-solves before the first run
  B_CHECK(dalloc(&b->d_age, std::max(ne, 1)));
  B_CHECK(hipMemset(b->d_age, 0, std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_fb, (size_t)std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_ov, (size_t)std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_spec, 3 * (size_t)std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_group_iso, (size_t)std::max(ne, 1)));  // groups <= events
  // groups <= events; then one re-solve pool counter per chunk, then the
  // speculated pairs' counts and queue heads
  B_CHECK(dalloc(&b->d_counter, 8 * (size_t)std::max(ne, 1) + 2 + b->chunk_first.size() + kSpecCounters));

  // per-event planes of the phase pipeline (sized for the largest chunk)
  B_CHECK(dalloc(&b->d_evs, 3 * (size_t)std::max(ne, 1)));
  B_CHECK(dalloc(&b->d_coltab, std::max<size_t>(col_total, 1)));
  // 6 forward planes of wes bytes, or 2 face-flux planes of 2*es and 2 delta planes
  const size_t work_bytes = 6 * std::max<size_t>(w_total, 1) * wes;
  B_CHECK(dalloc(&b->d_work, work_bytes));
  B_CHECK(hipMemset(b->d_work, 0, work_bytes));
  if (f32) B_CHECK(dalloc(&b->d_T0d, std::max<size_t>(pl_total, 1)));
  if (f32) {
    // the delta row masks: per event, ny rows of one bit per lane of either
    // adjoint geometry; the adjoint sweeps hold one in LDS (beyond kMaskLds
    // bytes the planes are read whole instead; SURFATT_MASK_LDS lowers the
    // limit, for tests of that path)
    size_t kMaskLds = 48 << 10;
    if (const char* v = std::getenv("SURFATT_MASK_LDS")) kMaskLds = std::min<size_t>(kMaskLds, std::strtoul(v, nullptr, 10));
    size_t stride = 1;
    for (const GridHost& G : b->grids) {
      const int la = std::max(G.g.La, G.adj8 ? G.gadj.La : 0);
      stride = std::max(stride, (size_t)G.g.ny * (size_t)((la + kWave - 1) / kWave));
    }
    if (stride * sizeof(unsigned long long) <= kMaskLds) {
      b->dmask_stride = (int)stride;
      B_CHECK(dalloc(&b->d_dmask, stride * (size_t)std::max(ne, 1)));
    }
  }
  B_CHECK(dalloc(&b->d_T, std::max<size_t>(pl_total, 1)));
  if (b->calc_adj) {
    B_CHECK(dalloc(&b->d_ta, std::max<size_t>(pl_total, 1) * es));
  }

  // resident waves of the sweep kernels: 3 swept planes each (FP64 sigma/tau sized)
  hipDeviceProp_t prop;
  B_CHECK(hipGetDeviceProperties(&prop, 0));
  // SURFATT_SWEEP_WAVES_PER_CU (1..64): fewer sweep slots, so that each slot
  // takes several pairs from the work queue in turn (tuning and tests)
  int waves_per_cu = 16;
  if (const char* v = std::getenv("SURFATT_SWEEP_WAVES_PER_CU")) {
    const int w = std::atoi(v);
    if (w >= 1 && w <= 64) waves_per_cu = w;
  }
  const int max_chunk = *std::max_element(b->chunk_count.begin(), b->chunk_count.end());
  b->nslots_phase = std::max(1, std::min(max_chunk, prop.multiProcessorCount * waves_per_cu));
  // + the sweeps' trash row; two-pairs-per-wave sweeps keep two halves per plane
  size_t slot_elems = max_wplane + kTrashElems;
  for (const GridHost& G : b->grids)
    if (G.p2) slot_elems = std::max(slot_elems, 2 * pair2_hstride(G.gp2));
  b->slot_plane = round_up(slot_elems * sizeof(double), 256);
  b->slot_bytes = 3 * b->slot_plane;
  // concurrent launch groups share the slots; each group's share is rounded up
  const size_t nslots_alloc = (size_t)b->nslots_phase + kMaxGroups;
  B_CHECK(dalloc(&b->d_slots, b->slot_bytes * nslots_alloc));
  B_CHECK(hipMemset(b->d_slots, 0, b->slot_bytes * nslots_alloc));

  // FP32 batches: a pool of FP64 phase-kernel planes for the re-solves
  if (f32) {
    size_t max_wplane_mw = 0;
    for (const GridHost& G : b->grids) max_wplane_mw = std::max(max_wplane_mw, G.gmw.wplane);
    b->slot_plane64 = round_up((max_wplane_mw + kTrashElems) * sizeof(double), 256);
    b->slot_bytes64 = 3 * b->slot_plane64;
    b->pool_size = std::max(1, std::min(ne, std::max(64, (ne + 15) / 16)));
    b->pool_stride = round_up(max_wplane_mw, 64);
    const size_t pool_bytes = 6 * b->pool_stride * sizeof(double) * (size_t)b->pool_size;
    B_CHECK(dalloc(&b->d_work64, pool_bytes));
    B_CHECK(hipMemset(b->d_work64, 0, pool_bytes));
    B_CHECK(dalloc(&b->d_evx64, std::max(ne, 1)));
    if (ne) B_CHECK(hipMemcpy(b->d_evx64, b->hx.data(), sizeof(EventX) * ne, hipMemcpyHostToDevice));
    b->nslots64p = std::min(b->pool_size, 2 * prop.multiProcessorCount);
    // concurrent re-solve groups: each group's share rounded up
    const size_t nslots64_alloc = (size_t)b->nslots64p + kMaxGroups;
    B_CHECK(dalloc(&b->d_slots64, b->slot_bytes64 * nslots64_alloc));
    B_CHECK(hipMemset(b->d_slots64, 0, b->slot_bytes64 * nslots64_alloc));
  }

  // fused path: one slot for per-solve runs; FP32 batches also FP64 slots for re-solves
  int max_wseg = 1;
  for (const GridHost& G : b->grids) max_wseg = std::max(max_wseg, G.Wseg);
  b->sg = slot_geom(b->precision, max_seg, max_plane, max_wseg);
  B_CHECK(dalloc(&b->d_ws, b->sg.slot));
  B_CHECK(hipMemset(b->d_ws, 0, b->sg.slot));  // also the all-zero FP32 delta plane
  if (f32) {
    b->nslots64 = std::max(1, std::min(ne, 64));
    b->sg64 = slot_geom(SAT_PREC_F64, max_seg, max_plane, max_wseg);
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
    for (int e : ev_ids) hs.push_back(b->he[e].ta_off);
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
  rd.ev_taoff = b->d_red_sz + plane_off.size() + node_prefix.size();
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
    SAT_CHECK(hipMemcpyAsync(b->d_rec + 2 * nr, tt_obs, sizeof(double) * nr, hipMemcpyDefault, b->stream));
    SAT_CHECK(hipMemcpyAsync(b->d_rec + 3 * nr, weight, sizeof(double) * nr, hipMemcpyDefault, b->stream));
    SAT_CHECK(hipStreamSynchronize(b->stream));
  }
  return SAT_OK;
}

namespace {

// Queue one run of the pipeline on the batch's streams.  capturing: the calls
// are being recorded into a graph; the timing-only marks are left out (the
// events that order the streams stay).
int enqueue_run(sat_batch* b, int mode, bool capturing) {
  const bool f32 = b->precision == SAT_PREC_F32;
  const int ne = b->nevent;
  const size_t nr = (size_t)std::max(b->nrec, 1);
  const size_t ne1 = (size_t)std::max(ne, 1);
  const size_t ng = b->group_C.size();
  // work-queue heads and list lengths, zeroed at the start of every run
  int* fb_count = b->d_counter;               // pool re-solves per group
  int* fwd_heads = b->d_counter + ng;
  int* adj_heads = b->d_counter + 2 * ng;
  int* resf_heads = b->d_counter + 3 * ng;    // re-solve forward sweeps
  int* resa_heads = b->d_counter + 4 * ng;    // re-solve adjoint sweeps
  int* ov_count = b->d_counter + 5 * ng;      // re-solves beyond the pool per group
  int* ov_heads = b->d_counter + 6 * ng;
  int* solo_heads = b->d_counter + 7 * ng + 1;
  int* pool_counters = b->d_counter + 8 * ne1 + 2;  // one per chunk
  const size_t nch = b->chunk_first.size();
  int* spec_count = b->d_counter + 8 * ne1 + 2 + nch;  // spec_kernel's list lengths, then queue heads
  const size_t n_counters = 8 * ne1 + 2 + nch + kSpecCounters;
  // FP64 solves from the start of the run for the pairs re-solved recently
  // (spec_kernel): one-chunk batched or graph runs of FP32 batches with a band,
  // no grid wider than 512 columns
  bool any_wide = false, spec_mode[2] = {false, false};  // isotropic / general re-solve update
  for (size_t gi = 0; gi < ng; gi++) {
    any_wide |= b->group_wide[gi] != 0;
    spec_mode[b->group_iso[gi] != kMetricGeneral ? 0 : 1] = true;
  }
  const bool spec = f32 && b->opt_spec && b->f32_band > 0.0 && mode != SAT_RUN_PER_SOLVE && nch == 1 && ng && ne &&
                    !any_wide;

  BatchDev bd{};
  bd.grids = b->d_grids_seg;
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
  bd.T_out = b->d_T;
  bd.calc_adj = b->calc_adj;
  bd.fwd_maxiter = kFwdMaxIter;
  bd.adj_maxiter = kAdjMaxIter;
  bd.conv = b->d_conv;
  bd.fb_list = b->d_fb;
  bd.fb_count = fb_count;
  bd.f32_band = b->f32_band;

  PhaseDev pd{};
  // the receivers and the adjoint phase (face fluxes, delta, sweeps, Ta) in the
  // adjoint's geometry (GridHost::gadjoint)
  pd.grids = b->d_grids_adj;
  pd.ev = b->d_evx;
  pd.ev_base = 0;
  pd.nevent = ne;
  pd.rec_lo = 0;
  pd.evs = b->d_evs;
  pd.coltab = b->d_coltab;
  // bytes of one work plane (8-byte elements in polished batches)
  const size_t wb = b->w_total * (b->polish ? sizeof(double) : elem_size(b->precision));
  for (int o = 0; o < 2; o++) {
    pd.t0[o] = b->d_work + (0 + o) * wb;
    pd.t0x[o] = b->d_work + (2 + o) * wb;
    pd.t0y[o] = b->d_work + (4 + o) * wb;
    pd.fxy[o] = b->d_work + 2 * o * wb;
    pd.dl[o] = b->d_work + (4 + o) * wb;
  }
  pd.T0d = b->d_T0d;
  pd.T = b->d_T;
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
  pd.want_skip = 0;
  {
    // the receiver spread's LDS: the largest event's shares, at most kDeltaShares
    int max_nrec = 0;
    for (const EventX& x : b->hx) max_nrec = std::max(max_nrec, x.nrec);
    pd.delta_cap = std::min(kDeltaShares, std::max(4 * max_nrec, kScan));
  }
  pd.dmask = b->d_dmask;
  pd.dmask_stride = b->dmask_stride;
  pd.dmask_lds = f32 ? (int)(b->d_dmask ? b->dmask_stride * sizeof(unsigned long long) : sizeof(unsigned long long)) : 0;
  pd.ov_list = b->d_ov;
  pd.ov_count = ov_count;
  pd.pool_counter = pool_counters;
  pd.pool_size = b->pool_size;
  pd.pool_stride = b->pool_stride;
  pd.ev64 = b->d_evx64;
  pd.list = nullptr;
  pd.list_count = nullptr;
  pd.ta_f32 = 0;
  pd.fb_list = b->d_fb;
  pd.fb_count = fb_count;
  pd.f32_band = b->f32_band;
  pd.slots = b->d_slots;
  pd.slot_bytes = b->slot_bytes;
  pd.slot_plane = b->slot_plane;
  pd.order = b->d_order;
  pd.order_adj = b->d_order_adj;
  pd.fwd_maxiter = kFwdMaxIter;
  pd.adj_maxiter = kAdjMaxIter;
  pd.polish = b->polish ? 1 : 0;
  const unsigned node_bx = (unsigned)((b->max_plane + 255) / 256);
  // init_kernel: kInitPerThread wavefront-major elements (either geometry) or nodes per thread
  size_t init_n = b->max_plane;
  for (const GridHost& G : b->grids) init_n = std::max({init_n, G.g.wplane, G.gmw.wplane});
  const unsigned init_bx = (unsigned)((init_n + 256 * kInitPerThread - 1) / (256 * kInitPerThread));
  hipEvent_t* mk = b->mk;
  // waves per pair and columns per lane of the re-solve geometry:
  // Wr = resolve_waves(C), ceil(nx / (64 Wr)) = ceil(C / Wr)
  // (wide grids: the batch's own W waves of kMaxC columns)
  auto wmw = [&](size_t gi) {
    return b->group_wide[gi] ? b->group_W[gi] : (b->wr_all ? b->wr_all : resolve_waves(b->group_C[gi]));
  };
  auto cmw = [&](size_t gi) { return b->group_wide[gi] ? kMaxC : (b->group_C[gi] + wmw(gi) - 1) / wmw(gi); };
  // waves of the one-pair kernels (pair_kernel) of a group
  auto wseg = [&](size_t gi) { return b->group_wide[gi] ? b->group_W[gi] : 1; };
  // columns per lane and waves of a group's batched adjoint sweeps
  auto adj_cols = [&](size_t gi) { return b->group_adj8[gi] ? kAdjC : b->group_wC[gi]; };
  auto adj_waves = [&](size_t gi) { return b->group_adj8[gi] ? 1 : b->group_W[gi]; };

  auto mark = [&](int i, hipStream_t st) -> hipError_t {
    return capturing ? hipSuccess : hipEventRecord(mk[i], st);
  };
  // the queues, longest pairs of the previous run first (its rounds are still in
  // iters_f / iters_a until reset_kernel clears them)
  if (b->order_by_rounds && ng && ne) {
    hipLaunchKernelGGL(order_kernel, dim3((unsigned)ng), dim3(kOrderThreads), 0, b->stream, b->d_order,
                       b->d_order_tmp, bd.iters_f, b->d_group_rng);
    hipLaunchKernelGGL(order_kernel, dim3((unsigned)ng), dim3(kOrderThreads), 0, b->stream, b->d_order_adj,
                       b->d_order_tmp, bd.iters_a, b->d_group_rng);
    SAT_CHECK(hipGetLastError());
  }
  // the pairs' speculation ages from the previous run's re-solve marks, which
  // reset_kernel clears
  if (f32 && b->f32_band > 0.0 && ne) {
    hipLaunchKernelGGL(spec_age_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, b->stream, b->d_age,
                       b->d_skip, b->d_conv, b->f32_band, ne);
    SAT_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(reset_kernel, dim3((unsigned)std::min<size_t>((n_counters + ne1 + 255) / 256, 4096)), dim3(256),
                     0, b->stream, b->d_counter, n_counters, b->d_skip, b->d_iters, bd.flags, b->d_base_flags, ne1);
  SAT_CHECK(hipGetLastError());
  if (spec) {
    hipLaunchKernelGGL(spec_kernel, dim3(1), dim3(kOrderThreads), 0, b->stream, b->d_age, b->d_skip, b->d_order,
                       b->d_group_rng, b->d_group_iso, (int)ng, (int)ne1, b->d_spec, spec_count, pool_counters,
                       b->pool_size, b->pool_stride, b->d_evx64);
    SAT_CHECK(hipGetLastError());
  }

  // The chunks run one after the other through the same workspace (one chunk
  // unless the batch exceeds the workspace cap).  The phase marks describe the
  // last chunk; mk[0] -> mk[7] spans the whole run.
  const int nchunk = (int)b->chunk_first.size();
  for (int c = 0; c < nchunk; c++) {
    const int cn = b->chunk_count[c];
    const size_t g_lo = (size_t)b->chunk_g0[c], g_hi = (size_t)b->chunk_g0[c + 1];
    const int nrec_c = b->chunk_rec_n[c];
    PhaseDev pc = pd;
    pc.ev_base = b->chunk_first[c];
    pc.nevent = cn;
    pc.rec_lo = b->chunk_rec_lo[c];
    pc.pool_counter = pool_counters + c;
    const dim3 node_grid(node_bx, (unsigned)std::max(cn, 1));
    const int rec_blocks = (int)std::max(1, std::min((nrec_c + 255) / 256, 4096));

    // Launch fn(gi, stream) for every group of the chunk, concurrently: from the
    // base stream (main or stream2) over its side streams, joined back at the end.
    // A stream's kernels run one after the other (so do those of streams that
    // share a hardware queue): each group goes to the stream with the least work
    // so far, in launch order (largest first).
    auto fan_out_on = [&](hipStream_t base, hipStream_t* sides, hipEvent_t fork, hipEvent_t* joins, int max_streams,
                          auto&& fn) -> int {
      const size_t ngc = g_hi - g_lo;
      const int nused = (int)std::min<size_t>(ngc, (size_t)max_streams);
      if (nused > 1) {
        SAT_CHECK(hipEventRecord(fork, base));
        for (int k = 1; k < nused; k++) SAT_CHECK(hipStreamWaitEvent(sides[k - 1], fork, 0));
      }
      double load[sat_batch::kSide + 1] = {};
      for (size_t k = 0; k < ngc; k++) {
        const int gi = b->launch_order[g_lo + k];
        int sk = 0;
        for (int j = 1; j < nused; j++)
          if (load[j] < load[sk]) sk = j;
        load[sk] += b->group_work[gi];
        const int rc = fn(gi, sk == 0 ? base : sides[sk - 1]);
        if (rc) return rc;
      }
      if (nused > 1) {
        for (int k = 1; k < nused; k++) {
          SAT_CHECK(hipEventRecord(joins[k - 1], sides[k - 1]));
          SAT_CHECK(hipStreamWaitEvent(base, joins[k - 1], 0));
        }
      }
      return SAT_OK;
    };
    // the main launch groups keep one hardware queue free for the speculated
    // solves on stream2 (streams are spread over the device's queues in creation
    // order: stream, stream2, then the side streams)
    const int main_streams = spec ? sat_batch::kSide : sat_batch::kSide + 1;
    auto fan_out = [&](auto&& fn) -> int {
      return fan_out_on(b->stream, b->side, b->fork_ev, b->join_ev, main_streams, fn);
    };
    auto fan_out2 = [&](auto&& fn) -> int {
      return fan_out_on(b->stream2, b->side2, b->fork2_ev, b->join2_ev, sat_batch::kSide + 1, fn);
    };

    // The FP64 phase kernels on stream2 over pool lists (group g's list at
    // lists + group_first[g], its length counts[g]; queue heads fheads, aheads):
    // init, forward sweeps, receivers, adjoint set-up, delta and adjoint sweeps
    // of the events whose re-solve mark is `want`.
    auto pool_phase = [&](int want) {
      PhaseDev q = pc;
      q.grids = b->d_grids64mw;
      q.ev = b->d_evx64;
      const size_t wb = b->pool_stride * (size_t)b->pool_size * sizeof(double);
      for (int o = 0; o < 2; o++) {
        q.t0[o] = b->d_work64 + (0 + o) * wb;
        q.t0x[o] = b->d_work64 + (2 + o) * wb;
        q.t0y[o] = b->d_work64 + (4 + o) * wb;
        q.fxy[o] = b->d_work64 + 2 * o * wb;
        q.dl[o] = b->d_work64 + (4 + o) * wb;
      }
      q.slots = b->d_slots64;
      q.slot_bytes = b->slot_bytes64;
      q.slot_plane = b->slot_plane64;
      q.want_skip = want;
      q.ta_f32 = 1;
      q.scaled_iso = 1;
      return q;
    };
    auto pool_pipeline = [&](int* lists, int* counts, int* fheads, int* aheads, int want) -> int {
      hipStream_t s2 = b->stream2;
      const PhaseDev q = pool_phase(want);
      // the groups' re-solves run concurrently; each takes a share of the
      // re-solve slots in proportion to its events, as the main sweeps do
      auto res_slots = [&](int gi, PhaseDev& a, int ub) {
        const int share = std::max(1, (int)(((long)b->nslots64p * b->group_count[gi] + cn - 1) / cn));
        int base = 0;
        for (size_t k = g_lo; k < (size_t)gi; k++)
          base += std::max(1, (int)(((long)b->nslots64p * b->group_count[k] + cn - 1) / cn));
        a.slots = b->d_slots64 + (size_t)base * b->slot_bytes64;
        return dim3(std::min(ub, share));
      };
      int rc = fan_out2([&](int gi, hipStream_t st) -> int {
        const int ub = std::min(b->group_count[gi], b->pool_size);
        PhaseDev a = q;
        a.list = lists + b->group_first[gi];
        a.list_count = counts + gi;
        SAT_CHECK(launch_init<SAT_PREC_F64>(dim3(init_bx, ub, 2), st, a));
        a.counter = fheads + gi;
        const dim3 grid = res_slots(gi, a, ub);
        // isotropic grids: the scaled update (init_kernel wrote scaled gradient planes)
        const int rmode = b->group_iso[gi] != kMetricGeneral ? kMetricIso : kMetricGeneral;
        SAT_CHECK(launch_fwd_sweep<SAT_PREC_F64>(rmode, cmw(gi), wmw(gi), grid, st, a, 0, 0, gi));
        return SAT_OK;
      });
      if (rc) return rc;
      if (nrec_c) SAT_CHECK(launch_recv(dim3(rec_blocks), s2, q, nrec_c));
      if (b->calc_adj) {
        rc = fan_out2([&](int gi, hipStream_t st) -> int {
          const int ub = std::min(b->group_count[gi], b->pool_size);
          PhaseDev a = q;
          a.list = lists + b->group_first[gi];
          a.list_count = counts + gi;
          SAT_CHECK(launch_adj_setup<SAT_PREC_F64>(dim3(node_bx, ub), st, a));
          return SAT_OK;
        });
        if (rc) return rc;
        SAT_CHECK(launch_delta<SAT_PREC_F64>(dim3((unsigned)cn), s2, q));  // one wave per event
        rc = fan_out2([&](int gi, hipStream_t st) -> int {
          const int ub = std::min(b->group_count[gi], b->pool_size);
          PhaseDev a = q;
          a.list = lists + b->group_first[gi];
          a.list_count = counts + gi;
          a.counter = aheads + gi;
          const dim3 grid = res_slots(gi, a, ub);
          SAT_CHECK(launch_adj_sweep<SAT_PREC_F64>(cmw(gi), wmw(gi), grid, st, a, 0, 0));
          return SAT_OK;
        });
        if (rc) return rc;
      }
      return SAT_OK;
    };

    // The speculated pairs' FP64 solves: one chain on stream2, each phase one
    // launch over all of them (the batch's grids share one re-solve geometry):
    // init, forward sweeps (one launch per update kind), receivers, adjoint
    // set-up, delta, adjoint sweeps.
    auto spec_pipeline = [&]() -> int {
      hipStream_t s2 = b->stream2;
      const PhaseDev q = pool_phase(kSkipSpec);
      size_t g0 = g_lo;
      while (b->group_wide[g0]) g0++;  // (no wide group when speculating)
      const int Wr = wmw(g0), Cr = cmw(g0);
      const int ub = b->pool_size;
      const dim3 sgrid((unsigned)std::min(ub, b->nslots64p));
      PhaseDev a = q;
      a.list = b->d_spec;
      a.list_count = spec_count;
      SAT_CHECK(launch_init<SAT_PREC_F64>(dim3(init_bx, ub, 2), s2, a));
      for (int m = 0; m < 2; m++) {
        if (!spec_mode[m]) continue;
        PhaseDev f = q;
        f.list = b->d_spec + (size_t)(1 + m) * ne1;
        f.list_count = spec_count + 1 + m;
        f.counter = spec_count + 3 + m;
        SAT_CHECK(launch_fwd_sweep<SAT_PREC_F64>(m == 0 ? kMetricIso : kMetricGeneral, Cr, Wr, sgrid, s2, f, 0, 0, 0));
      }
      if (nrec_c) SAT_CHECK(launch_recv(dim3(rec_blocks), s2, q, nrec_c));
      if (b->calc_adj) {
        SAT_CHECK(launch_adj_setup<SAT_PREC_F64>(dim3(node_bx, ub), s2, a));
        SAT_CHECK(launch_delta<SAT_PREC_F64>(dim3((unsigned)cn), s2, q));  // one wave per event
        a.counter = spec_count + 5;
        SAT_CHECK(launch_adj_sweep<SAT_PREC_F64>(Cr, Wr, sgrid, s2, a, 0, 0));
      }
      return SAT_OK;
    };

    // FP64 re-solves of the FP32 pairs whose stop decision was within the band,
    // on stream2 once the forward sweeps (which list them) are done (after the
    // speculated pairs' solves queued there at the start of the run).  Batched
    // runs: the FP64 phase kernels over the pool lists, the per-pair kernel over
    // the overflow lists; per-solve runs: the per-pair kernel.
    auto launch_resolves = [&](bool phase) -> int {
      SAT_CHECK(hipEventRecord(mk[2], b->stream));
      SAT_CHECK(hipStreamWaitEvent(b->stream2, mk[2], 0));
      if (!spec) SAT_CHECK(mark(5, b->stream2));
      if (f32 && cn) {
        hipStream_t s2 = b->stream2;
        if (phase) {
          const int rc = pool_pipeline(b->d_fb, fb_count, resf_heads, resa_heads, 1);
          if (rc) return rc;
        }
        for (size_t gi = g_lo; gi < g_hi; gi++) {
          BatchDev a = bd;
          a.grids = b->d_grids64;
          a.ws = b->d_ws64;
          a.ws_slot = b->sg64.slot;
          a.ws_seg = b->sg64.seg;
          a.ws_plain = b->sg64.plain;
          a.counter = phase ? ov_heads + gi : resf_heads + gi;
          a.f32_band = 0.0;
          a.adj_lds_bytes = b->group_lds64[gi];
          const int* list = (phase ? b->d_ov : b->d_fb) + b->group_first[gi];
          const int* cnt = phase ? ov_count + gi : fb_count + gi;
          const int nblk = std::min(b->group_count[gi], b->nslots64);
          SAT_CHECK(launch_pair<SAT_PREC_F64>(1, b->group_C[gi], wseg(gi), dim3(nblk), s2, a, 0, 0, (int)gi, list,
                                              cnt));
        }
      }
      SAT_CHECK(hipEventRecord(mk[6], b->stream2));
      return SAT_OK;
    };

    if (mode == SAT_RUN_PER_SOLVE) {
      // one launch per pair of the fused kernel (the plain per-solve path)
      bd.ws = b->d_ws;
      bd.ws_slot = b->sg.slot;
      bd.ws_seg = b->sg.seg;
      bd.ws_plain = b->sg.plain;
      for (size_t gi = g_lo; gi < g_hi; gi++)
        for (int i = 0; i < b->group_count[gi]; i++) {
          const int first = b->group_first[gi] + i;
          BatchDev a = bd;
          a.counter = solo_heads + first;
          a.fb_list = bd.fb_list + b->group_first[gi];
          a.adj_lds_bytes = b->group_lds[gi];
          if (!f32)
            SAT_CHECK(launch_pair<SAT_PREC_F64>(0, b->group_C[gi], wseg(gi), dim3(1), b->stream, a, first, 1, (int)gi,
                                                nullptr,
                                                nullptr));
          else
            SAT_CHECK(launch_pair<SAT_PREC_F32>(1, b->group_C[gi], wseg(gi), dim3(1), b->stream, a, first, 1, (int)gi,
                                                nullptr,
                                                nullptr));
        }
      SAT_CHECK(mark(1, b->stream));
      const int rc = launch_resolves(false);
      if (rc) return rc;
      SAT_CHECK(mark(3, b->stream));
      SAT_CHECK(mark(4, b->stream));
    } else {
      if (cn) {
        const dim3 col_grid((unsigned)((b->max_nx + 63) / 64), (unsigned)cn);
        SAT_CHECK(f32 ? launch_trig<SAT_PREC_F32>(col_grid, b->stream, pc) : launch_trig<SAT_PREC_F64>(col_grid, b->stream, pc));
        if (spec) {
          // the speculated pairs' FP64 solves start on stream2 once their column
          // tables exist, beside the FP32 init and forward sweeps
          SAT_CHECK(hipEventRecord(b->spec_ev, b->stream));
          SAT_CHECK(hipStreamWaitEvent(b->stream2, b->spec_ev, 0));
          SAT_CHECK(mark(5, b->stream2));
          const int rc = spec_pipeline();
          if (rc) return rc;
        }
        PhaseDev pf = pc;  // the forward geometry (two pairs per wave where a grid runs them)
        pf.grids = b->d_grids_fwd;
        SAT_CHECK(f32 ? launch_init<SAT_PREC_F32>(dim3(init_bx, cn, 3), b->stream, pf)
                      : launch_init<SAT_PREC_F64>(dim3(init_bx, cn, 2), b->stream, pf));
      }
      SAT_CHECK(mark(1, b->stream));
      {
        const int rc = fan_out([&](int gi, hipStream_t st) -> int {
          PhaseDev a = pc;
          a.grids = b->d_grids_fwd;
          a.counter = fwd_heads + gi;
          a.fb_list = pc.fb_list + b->group_first[gi];
          a.ov_list = pc.ov_list + b->group_first[gi];
          a.slots = b->d_slots + (size_t)b->group_slot0[gi] * b->slot_bytes;
          if (b->group_p2[gi] > 0) {
            const int nw = (b->group_nblk[gi] + 1) / 2;
            SAT_CHECK(launch_fwd_sweep_pair2(b->group_p2[gi], dim3(nw), st, a, b->group_first[gi], b->group_count[gi],
                                             (int)gi, (int)b->group_p2h[gi]));
            return SAT_OK;
          }
          const int C = b->group_wC[gi], W = b->group_W[gi], iso = b->group_iso[gi];
          SAT_CHECK(f32 ? launch_fwd_sweep<SAT_PREC_F32>(iso, C, W, dim3(b->group_nblk[gi]), st, a,
                                                         b->group_first[gi], b->group_count[gi], gi)
                        : launch_fwd_sweep<SAT_PREC_F64>(iso, C, W, dim3(b->group_nblk[gi]), st, a,
                                                         b->group_first[gi], b->group_count[gi], gi));
          return SAT_OK;
        });
        if (rc) return rc;
      }
      {
        const int rc = launch_resolves(true);
        if (rc) return rc;
      }
      if (nrec_c) SAT_CHECK(launch_recv(dim3(rec_blocks), b->stream, pc, nrec_c));
      if (b->calc_adj && cn) {
        SAT_CHECK(f32 ? launch_adj_setup<SAT_PREC_F32>(node_grid, b->stream, pc)
                      : launch_adj_setup<SAT_PREC_F64>(node_grid, b->stream, pc));
        const dim3 ev_grid((unsigned)cn);  // one wave per event
        SAT_CHECK(f32 ? launch_delta<SAT_PREC_F32>(ev_grid, b->stream, pc) : launch_delta<SAT_PREC_F64>(ev_grid, b->stream, pc));
      }
      SAT_CHECK(mark(3, b->stream));
      if (b->calc_adj && cn) {
        const int rc = fan_out([&](int gi, hipStream_t st) -> int {
          PhaseDev a = pc;
          a.counter = adj_heads + gi;
          a.slots = b->d_slots + (size_t)b->group_slot0[gi] * b->slot_bytes;
          const dim3 grid(b->group_nblk[gi]);
          const int C = adj_cols(gi), W = adj_waves(gi);
          SAT_CHECK(f32 ? launch_adj_sweep<SAT_PREC_F32>(C, W, grid, st, a, b->group_first[gi], b->group_count[gi])
                        : launch_adj_sweep<SAT_PREC_F64>(C, W, grid, st, a, b->group_first[gi], b->group_count[gi]));
          return SAT_OK;
        });
        if (rc) return rc;
      }
      SAT_CHECK(mark(4, b->stream));
    }
    // the chunk's re-solves are done before its Ta is reduced and before the
    // next chunk reuses the workspace
    SAT_CHECK(hipStreamWaitEvent(b->stream, mk[6], 0));
    if (b->calc_adj && b->K_total) {
      ReduceDev rd = b->rd;
      rd.ev_lo = b->chunk_first[c];
      rd.ev_hi = b->chunk_first[c] + cn;
      rd.accumulate = c > 0;
      const int threads = 256;
      const int blocks = (int)std::min<size_t>((b->K_total + threads - 1) / threads, 65535);
      if (!f32) hipLaunchKernelGGL(reduce_kernel<double>, dim3(blocks), dim3(threads), 0, b->stream, rd);
      else hipLaunchKernelGGL(reduce_kernel<float>, dim3(blocks), dim3(threads), 0, b->stream, rd);
      SAT_CHECK(hipGetLastError());
      // The anisotropic terms of accumulate_kernels (src/preproc.cpp:37-65) of the
      // grids with a registered metric, from the chunk's T and masked Ta.
      for (size_t gi = 0; gi < b->grids.size(); gi++) {
        if (!b->d_aniso[gi]) continue;
        const GridDev& g = b->grids[gi].g;
        int pl0 = 0;
        for (size_t k = 0; k < gi; k++) pl0 += b->grids[k].g.nperiod;
        const size_t tot = (size_t)g.nperiod * g.plane;
        AnisoDev ad{g, g.nperiod, pl0, b->rd.ev_off, b->rd.ev_ids, b->d_events, b->d_T, b->d_ta, b->d_aniso[gi],
                    b->d_aniso[gi] + 3 * tot, b->d_aniso[gi] + 4 * tot, rd.ev_lo, rd.ev_hi, c > 0};
        const unsigned ab = (unsigned)std::min<size_t>((tot + 255) / 256, 65535);
        if (f32) hipLaunchKernelGGL(aniso_kernel<float>, dim3(ab), dim3(256), 0, b->stream, ad);
        else hipLaunchKernelGGL(aniso_kernel<double>, dim3(ab), dim3(256), 0, b->stream, ad);
        SAT_CHECK(hipGetLastError());
      }
      // The density kernel (is_kden, src/preproc.cpp:127-136): a second adjoint
      // solve per pair from the receiver sources -w, on the same face fluxes,
      // not masked, accumulated as kden_s += field * svel^-3.  Every pair of
      // the chunk, in the batch's precision (also the FP64 re-solved ones, on
      // their FP64 T); Ta's planes are free again once K and the anisotropic
      // terms have read them.
      if (b->kden && cn) {
        PhaseDev pk = pc;
        pk.want_skip = -1;
        pk.kden_pass = 1;
        const int ng_c = (int)(g_hi - g_lo);
        hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(256), 0, b->stream, adj_heads + g_lo, ng_c);
        SAT_CHECK(hipGetLastError());
        if (nrec_c) SAT_CHECK(launch_recv(dim3(rec_blocks), b->stream, pk, nrec_c));
        SAT_CHECK(f32 ? launch_adj_setup<SAT_PREC_F32>(node_grid, b->stream, pk)
                      : launch_adj_setup<SAT_PREC_F64>(node_grid, b->stream, pk));
        SAT_CHECK(f32 ? launch_delta<SAT_PREC_F32>(dim3((unsigned)cn), b->stream, pk)
                      : launch_delta<SAT_PREC_F64>(dim3((unsigned)cn), b->stream, pk));
        const int rc = fan_out([&](int gi, hipStream_t st) -> int {
          PhaseDev a = pk;
          a.counter = adj_heads + gi;
          a.slots = b->d_slots + (size_t)b->group_slot0[gi] * b->slot_bytes;
          const int C = adj_cols(gi), W = adj_waves(gi);
          SAT_CHECK(f32 ? launch_adj_sweep<SAT_PREC_F32>(C, W, dim3(b->group_nblk[gi]), st, a, b->group_first[gi],
                                                         b->group_count[gi])
                        : launch_adj_sweep<SAT_PREC_F64>(C, W, dim3(b->group_nblk[gi]), st, a, b->group_first[gi],
                                                         b->group_count[gi]));
          return SAT_OK;
        });
        if (rc) return rc;
        ReduceDev rk = rd;
        rk.K = b->d_Kden;
        if (!f32) hipLaunchKernelGGL(reduce_kernel<double>, dim3(blocks), dim3(threads), 0, b->stream, rk);
        else hipLaunchKernelGGL(reduce_kernel<float>, dim3(blocks), dim3(threads), 0, b->stream, rk);
        SAT_CHECK(hipGetLastError());
      }
    }
  }
  hipLaunchKernelGGL(chi_kernel, dim3(1), dim3(kChiThreads), 0, b->stream, b->d_chi_ev, b->calc_adj ? ne : 0,
                     b->d_chi);
  SAT_CHECK(hipGetLastError());
  return SAT_OK;
}

// After a run: every event's fields are on the device, and the optional terms
// that were enabled are formed.
void mark_ran(sat_batch* b) {
  b->ran = true;
  for (size_t gi = 0; gi < b->d_aniso.size(); gi++) b->aniso_ran[gi] = b->d_aniso[gi] != nullptr && b->calc_adj;
  b->kden_ran = b->kden && b->calc_adj;
}

void drop_graph(sat_batch* b) {
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
}

}  // namespace

int sat_batch_run(sat_batch* b, int mode) {
  if (!b || (mode != SAT_RUN_BATCHED && mode != SAT_RUN_PER_SOLVE && mode != SAT_RUN_GRAPH)) return SAT_EINVAL;
  if (b->groups_dirty) {
    drop_graph(b);
    const int rc = build_groups(b);
    if (rc) return rc;
  }
  if (mode != SAT_RUN_GRAPH) {
    b->last_graph = false;
    SAT_CHECK(hipEventRecord(b->mk[0], b->stream));
    const int rc = enqueue_run(b, mode, false);
    if (rc) return rc;
    SAT_CHECK(hipEventRecord(b->mk[7], b->stream));
    mark_ran(b);
    return SAT_OK;
  }
  if (!b->graph_exec) {
    // Capture the batched run once: every kernel and stream dependency of the
    // pipeline, chunks included; replays reuse it (the device inputs change
    // in place: update_media / update_obs need no new capture).
    SAT_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_run(b, SAT_RUN_BATCHED, true);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    SAT_CHECK(ec);
    b->graph = g;
    SAT_CHECK(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
  }
  b->last_graph = true;
  SAT_CHECK(hipEventRecord(b->mk[0], b->stream));
  SAT_CHECK(hipGraphLaunch(b->graph_exec, b->stream));
  SAT_CHECK(hipEventRecord(b->mk[7], b->stream));
  mark_ran(b);
  return SAT_OK;
}

int sat_batch_sync(sat_batch* b) {
  if (!b) return SAT_EINVAL;
  SAT_CHECK(hipStreamSynchronize(b->stream));
  return SAT_OK;
}

int sat_batch_combine_kernels(sat_batch* b, int grid, int nz, int mode, const double* sen_vs, const double* sen_vp,
                              const double* sen_rho, const double* vs3d, const double* vp3d, double* ker_vs,
                              double* ker_vp, double* ker_rho, double* kernel_ms) {
  if (!b || !b->calc_adj || grid < 0 || grid >= (int)b->grids.size()) return SAT_EINVAL;
  size_t off = 0;  // the grid's per-period kernels in the batch's kernel planes (grids in order)
  for (int gi = 0; gi < grid; gi++) off += (size_t)b->grids[gi].g.nperiod * b->grids[gi].g.plane;
  const GridDev& g = b->grids[grid].g;
  SAT_CHECK(hipStreamSynchronize(b->stream));  // the last run's kernels are complete
  return sat::combine_impl(g.nx, g.ny, nz, g.nperiod, mode, nullptr, b->d_K + off, sen_vs, sen_vp, sen_rho, vs3d,
                           vp3d, ker_vs, ker_vp, ker_rho, kernel_ms);
}

int sat_batch_set_aniso_metric(sat_batch* b, int grid, const double* a, const double* bm, const double* c) {
  if (!b || !b->calc_adj || grid < 0 || grid >= (int)b->grids.size()) return SAT_EINVAL;
  const GridDev& g = b->grids[grid].g;
  const size_t tot = (size_t)g.nperiod * g.plane;
  if (!a && !bm && !c) {  // unregister
    if (b->d_aniso[grid]) {
      SAT_CHECK(hipStreamSynchronize(b->stream));
      SAT_CHECK(hipFree(b->d_aniso[grid]));
      b->d_aniso[grid] = nullptr;
      b->aniso_ran[grid] = 0;
      drop_graph(b);
    }
    return SAT_OK;
  }
  if (!a || !bm || !c) return SAT_EINVAL;
  SAT_CHECK(hipStreamSynchronize(b->stream));  // no run is using the planes
  if (!b->d_aniso[grid]) {
    SAT_CHECK(dalloc(&b->d_aniso[grid], 5 * tot));
    drop_graph(b);  // a captured run does not form these terms
  }
  SAT_CHECK(hipMemcpy(b->d_aniso[grid], a, sizeof(double) * tot, hipMemcpyDefault));
  SAT_CHECK(hipMemcpy(b->d_aniso[grid] + tot, bm, sizeof(double) * tot, hipMemcpyDefault));
  SAT_CHECK(hipMemcpy(b->d_aniso[grid] + 2 * tot, c, sizeof(double) * tot, hipMemcpyDefault));
  b->aniso_ran[grid] = 0;
  return SAT_OK;
}

int sat_batch_set_kden(sat_batch* b, int on) {
  if (!b || (on && !b->calc_adj)) return SAT_EINVAL;
  if ((on != 0) == b->kden) return SAT_OK;
  SAT_CHECK(hipStreamSynchronize(b->stream));
  if (on && !b->d_Kden) SAT_CHECK(dalloc(&b->d_Kden, std::max<size_t>(b->K_total, 1)));
  b->kden = on != 0;
  b->kden_ran = false;
  drop_graph(b);
  return SAT_OK;
}

int sat_batch_kden_results(sat_batch* b, double* kden_s) {
  if (!b || !kden_s || !b->kden_ran) return SAT_EINVAL;
  SAT_CHECK(hipStreamSynchronize(b->stream));
  SAT_CHECK(hipMemcpy(kden_s, b->d_Kden, sizeof(double) * b->K_total, hipMemcpyDefault));
  return SAT_OK;
}

int sat_batch_aniso_kernels(sat_batch* b, int grid, const double* a, const double* bm, const double* c,
                            double* adj_xi, double* adj_eta) {
  if (!b || !b->calc_adj || grid < 0 || grid >= (int)b->grids.size() || !adj_xi || !adj_eta) return SAT_EINVAL;
  if (!a && !bm && !c) {  // the terms the last run formed (sat_batch_set_aniso_metric)
    if (!b->aniso_ran[grid]) return SAT_EINVAL;
    const GridDev& g = b->grids[grid].g;
    const size_t tot = (size_t)g.nperiod * g.plane;
    SAT_CHECK(hipStreamSynchronize(b->stream));
    SAT_CHECK(hipMemcpy(adj_xi, b->d_aniso[grid] + 3 * tot, sizeof(double) * tot, hipMemcpyDefault));
    SAT_CHECK(hipMemcpy(adj_eta, b->d_aniso[grid] + 4 * tot, sizeof(double) * tot, hipMemcpyDefault));
    return SAT_OK;
  }
  if (!a || !bm || !c) return SAT_EINVAL;
  // every event's T and Ta must be on the device: after a run, and not split into chunks
  if (!b->ran || b->chunk_first.size() != 1) return SAT_EINVAL;
  const GridDev& g = b->grids[grid].g;
  const int nper = g.nperiod;
  int pl0 = 0;
  for (int gi = 0; gi < grid; gi++) pl0 += b->grids[gi].g.nperiod;
  const size_t total = (size_t)nper * g.plane;
  int rc = SAT_OK;
  double* buf = nullptr;
  SAT_CHECK(hipStreamSynchronize(b->stream));  // the last run's kernels are complete
  SAT_CHECK(hipMalloc(&buf, sizeof(double) * 5 * total));
  {
    AnisoDev ad{g, nper, pl0, b->rd.ev_off, b->rd.ev_ids, b->d_events, b->d_T, b->d_ta, buf, buf + 3 * total,
                buf + 4 * total, 0, b->nevent, 0};
    hipError_t e = hipMemcpy(buf, a, sizeof(double) * total, hipMemcpyDefault);
    if (e == hipSuccess) e = hipMemcpy(buf + total, bm, sizeof(double) * total, hipMemcpyDefault);
    if (e == hipSuccess) e = hipMemcpy(buf + 2 * total, c, sizeof(double) * total, hipMemcpyDefault);
    if (e == hipSuccess) {
      const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65535);
      if (b->precision == SAT_PREC_F32)
        hipLaunchKernelGGL(aniso_kernel<float>, dim3(blocks), dim3(256), 0, nullptr, ad);
      else
        hipLaunchKernelGGL(aniso_kernel<double>, dim3(blocks), dim3(256), 0, nullptr, ad);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(adj_xi, ad.xi, sizeof(double) * total, hipMemcpyDefault);
    if (e == hipSuccess) e = hipMemcpy(adj_eta, ad.eta, sizeof(double) * total, hipMemcpyDefault);
    if (e != hipSuccess) rc = (e == hipErrorOutOfMemory) ? SAT_ENOMEM : SAT_EDEVICE;
  }
  (void)hipFree(buf);
  return rc;
}

int sat_batch_last_ms(sat_batch* b, double* total_ms, double* solve_ms) {
  if (!b) return SAT_EINVAL;
  float t = 0.f, s = 0.f;
  SAT_CHECK(hipEventElapsedTime(&t, b->mk[0], b->mk[7]));
  if (b->last_graph) s = std::nanf("");  // a graph replay is timed as a whole
  else SAT_CHECK(hipEventElapsedTime(&s, b->mk[0], b->mk[4]));
  if (total_ms) *total_ms = t;
  if (solve_ms) *solve_ms = s;
  return SAT_OK;
}

int sat_batch_phase_ms(sat_batch* b, double* ms, int n) {
  if (!b || !ms || n < 1) return SAT_EINVAL;
  static const int from[7] = {0, 1, 2, 3, 5, 4, 0};
  static const int to[7] = {1, 2, 3, 4, 6, 7, 7};
  for (int i = 0; i < n && i < 7; i++) {
    if (b->last_graph && i < 6) {  // a graph replay is timed as a whole
      ms[i] = std::nan("");
      continue;
    }
    float t = 0.f;
    SAT_CHECK(hipEventElapsedTime(&t, b->mk[from[i]], b->mk[to[i]]));
    ms[i] = t;
  }
  return SAT_OK;
}

int sat_batch_results(sat_batch* b, double* syn, double* adj_s, double* chi, int* iters_fwd, int* iters_adj,
                      int* flags) {
  if (!b) return SAT_EINVAL;
  SAT_CHECK(hipStreamSynchronize(b->stream));
  const size_t nr = (size_t)std::max(b->nrec, 1);
  const size_t ne = (size_t)std::max(b->nevent, 1);
  // destinations in host or device memory (hipMemcpyDefault tells them apart)
  if (syn && b->nrec) SAT_CHECK(hipMemcpy(syn, b->d_rec + 4 * nr, sizeof(double) * b->nrec, hipMemcpyDefault));
  if (adj_s && b->K_total) {
    if (!b->calc_adj) SAT_CHECK(hipMemset(b->d_K, 0, sizeof(double) * b->K_total));  // no kernels: zeros
    SAT_CHECK(hipMemcpy(adj_s, b->d_K, sizeof(double) * b->K_total, hipMemcpyDefault));
  }
  if (chi) SAT_CHECK(hipMemcpy(chi, b->d_chi, sizeof(double), hipMemcpyDefault));
  if (iters_fwd && b->nevent) SAT_CHECK(hipMemcpy(iters_fwd, b->d_iters, sizeof(int) * b->nevent, hipMemcpyDefault));
  if (iters_adj && b->nevent) SAT_CHECK(hipMemcpy(iters_adj, b->d_iters + ne, sizeof(int) * b->nevent, hipMemcpyDefault));
  if (flags && b->nevent) SAT_CHECK(hipMemcpy(flags, b->d_iters + 2 * ne, sizeof(int) * b->nevent, hipMemcpyDefault));
  return SAT_OK;
}

int sat_batch_set_f32_band(sat_batch* b, double band) {
  if (!b || !(band >= 0.0 && band < 1.0)) return SAT_EINVAL;
  if (band != b->f32_band) drop_graph(b);  // the band is a value captured in the graph
  b->f32_band = band;
  return SAT_OK;
}

int sat_batch_set_option(sat_batch* b, int option, int value) {
  if (!b) return SAT_EINVAL;
  switch (option) {
    case SAT_OPT_PAIR2:
      if (value < -1 || value > 1) return SAT_EINVAL;
      if (value != b->opt_pair2) b->groups_dirty = true;
      b->opt_pair2 = value;
      return SAT_OK;
    case SAT_OPT_ADJ_COLS:
      if (value < 0 || value > 1) return SAT_EINVAL;
      if ((value != 0) != b->opt_adj_cols) b->groups_dirty = true;
      b->opt_adj_cols = value != 0;
      return SAT_OK;
    case SAT_OPT_ORDER:
      if (value < 0 || value > 1) return SAT_EINVAL;
      if ((value != 0) != b->order_by_rounds) b->groups_dirty = true;
      b->order_by_rounds = value != 0;
      return SAT_OK;
    case SAT_OPT_SPECULATE:
      if (value < 0 || value > 1) return SAT_EINVAL;
      if ((value != 0) != b->opt_spec) drop_graph(b);  // a different set of launches
      b->opt_spec = value != 0;
      return SAT_OK;
    default:
      return SAT_EINVAL;
  }
}

int sat_batch_get_option(sat_batch* b, int option, int* value) {
  if (!b || !value) return SAT_EINVAL;
  switch (option) {
    case SAT_OPT_PAIR2: *value = b->opt_pair2; return SAT_OK;
    case SAT_OPT_ADJ_COLS: *value = b->opt_adj_cols ? 1 : 0; return SAT_OK;
    case SAT_OPT_ORDER: *value = b->order_by_rounds ? 1 : 0; return SAT_OK;
    case SAT_OPT_SPECULATE: *value = b->opt_spec ? 1 : 0; return SAT_OK;
    default: return SAT_EINVAL;
  }
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
      const size_t ng = b->group_C.size();
      SAT_CHECK(hipMemcpy(cnt.data(), b->d_counter, sizeof(int) * ng, hipMemcpyDeviceToHost));
      for (int c : cnt) *n_resolved += c;
      SAT_CHECK(hipMemcpy(cnt.data(), b->d_counter + 5 * ng, sizeof(int) * ng, hipMemcpyDeviceToHost));
      for (int c : cnt) *n_resolved += c;
      // the speculated pairs (solved in FP64 from the start of the run)
      const size_t spec_off = 8 * (size_t)std::max(ne, 1) + 2 + b->chunk_first.size();
      int ns = 0;
      SAT_CHECK(hipMemcpy(&ns, b->d_counter + spec_off, sizeof(int), hipMemcpyDeviceToHost));
      *n_resolved += ns;
    }
  }
  return SAT_OK;
}

int sat_batch_waves(sat_batch* b, int* waves) {
  if (!b || !waves) return SAT_EINVAL;
  for (size_t gi = 0; gi < b->grids.size(); gi++) waves[gi] = b->grids[gi].W;
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
    std::snprintf(buf, buflen, "%s | %s | CUs=%d | LDS/block=%zu | build=gfx950", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount, (size_t)prop.sharedMemPerBlock);
  return SAT_OK;
}

}  // extern "C"
