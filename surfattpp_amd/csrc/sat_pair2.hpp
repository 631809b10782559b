// sat_pair2.hpp -- the batched FP32 forward sweeps with two pairs per wave.
//
// A grid of at most 32 lanes' worth of columns (La = ceil(nx / C) <= 32, C <=
// kPair2MaxC) leaves half of a 64-lane wave idle under the one-pair-per-wave
// sweeps, and its lane-skew ramp (La - 1 steps of partly idle lanes per
// sweep) is paid once per pair.  Here lanes [0, 32) sweep one pair and lanes
// [32, 64) another, each half exactly as sat_engine.hpp's single-wave sweep
// does with its own lane index lp = lane mod 32: the same neighbours, the same
// values in the same order, so every pair's iterates are those of the
// one-pair sweeps.  What differs per half is the data: the pair's own T0
// planes (per-lane pointers in the policy context), its half of each swept
// plane (an element offset), and its state -- a half whose pair has converged
// stores its result and takes the next pair from the work queue at the next
// round boundary while the other half keeps sweeping; all halves sweep the
// same directions in the same order, so a new pair simply starts at its first
// round.  A half with no pair left keeps issuing its loads (in bounds) and
// drops its stores.  The two halves never read each other's values: the
// cross-lane reads at lp = 0 and lp = 31 feed only the grid's first and last
// columns, which have no neighbour there.
#pragma once
#include "sat_phases.hpp"

namespace sat {

// Reductions over the 32 lanes of each half (every lane gets its half's value).
__device__ inline double half_max(double v) {
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o));
  return v;
}
__device__ inline double half_sum(double v) {
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// One sweep of both halves.  tin / tst: this lane's half of the input planes
// (per-lane pointers); tout: the plane's base (wave-uniform) with hoff the
// half's element offset into it and out_bytes the size of both halves; act:
// the half has a pair.
template <class Pol, int C, int XD, int YD, bool NORM, int OOUT>
__device__ __forceinline__ void sweep_pair2(const typename Pol::Ctx& cx, const typename Pol::V* tin,
                                            typename Pol::V* tout, const typename Pol::V* tst, int hoff,
                                            int out_bytes, bool act, double& lmax, double& lsum) {
  using V = typename Pol::V;
  using A = typename Pol::A;
  using Lay = WaveLayout;
  using Head = typename Pol::template Head<C>;
  using Body = typename Pol::template Body<C>;
  constexpr int O = (XD * YD > 0) ? 0 : 1;
  const SweepGeom g = uniform(Lay::geom(*cx.g));
  const int lane = threadIdx.x;
  const int lp = lane & (kHalfWave - 1);
  const int nx = g.nx;
  const SweepAddr<XD, YD> sa(g, lp);
  const int lag = sa.lag, nsteps = sa.nsteps;
  const __amdgpu_buffer_rsrc_t rs = plane_rsrc(tout, out_bytes);
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
    const V nw_u = shfl_any((XD > 0) ? prev_u[C - 1] : prev_u[0], lane - XD);
    const V od_u = shfl_any((XD > 0) ? nxt_h.u[0] : nxt_h.u[C - 1], lane + XD);
    const A nw_a = shfl_any((XD > 0) ? prev_a[C - 1] : prev_a[0], lane - XD);
    const A od_a = shfl_any((XD > 0) ? nxt_h.aux[0] : nxt_h.aux[C - 1], lane + XD);
    const bool live = act && sa.in_range(r);
    const int j = sa.row_clamped(r);
    const bool hS = j > 0, hN = j < g.ny - 1;
    V newv[C];
#pragma unroll
    for (int kk = 0; kk < C; kk++) {
      const int k = (XD > 0) ? kk : (C - 1 - kk);
      const int ix = lp * C + k;
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
      const V upd = Pol::template update<C>(cx, cur_h, cur_b, k, j, uW, aW, ix > 0, uE, aE, ix < nx - 1, uS, aS,
                                            hS, uN, aN, hN);
      newv[k] = (!valid || Pol::fixed(cx, ix, j)) ? cur_h.u[k] : upd;
      if (NORM) {
        const double d = valid ? Pol::template norm<C>(cur_h, cur_b, k, newv[k]) : 0.0;
        lsum += d;
        lmax = dmax(lmax, d);
      }
    }
    {
      const int off = live ? (Lay::template base<C, OOUT>(g, j, lp) + hoff) * (int)sizeof(V) : kDropOffset;
#pragma unroll
      for (int k = 0; k < C; k++) bst(rs, off + k * (int)sizeof(V), newv[k]);
    }
#pragma unroll
    for (int k = 0; k < C; k++) { prev_u[k] = newv[k]; prev_a[k] = cur_h.aux[k]; }
    cur_h = nxt_h;
    cur_b = nxt_b;
    nxt_h = nn_h;
  }
}

// The FP32 isotropic-metric forward sweeps, two pairs per wave (one
// work-group = one wave; slot: 3 swept planes of two halves each, half h at
// element h * hstride).  Per pair: the initial sigma, the rounds of four
// sweeps until L1 and Linf < tol (src/eikonal_solver.cpp:146-314), then the
// same epilogue as fwd_sweep_kernel (conv, iteration count, flags, the FP64
// re-solve decision, T).
template <int C>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2)))
void fwd_sweep_kernel_pair2(PhaseDev pd, int first, int count, int group, int hstride) {
  using PT = PhaseT<SAT_PREC_F32>;
  using Pol = typename PT::FwdIso;
  using U = typename PT::U;
  const int lane = threadIdx.x;
  const int half = lane / kHalfWave;
  const int lp = lane & (kHalfWave - 1);
  const int hoff = half * hstride;
  char* sb = pd.slots + (size_t)blockIdx.x * pd.slot_bytes;
  U* Q[3] = {reinterpret_cast<U*>(sb), reinterpret_cast<U*>(sb + pd.slot_plane),
             reinterpret_cast<U*>(sb + 2 * pd.slot_plane)};
  const int out_bytes = (int)pd.slot_plane;
  const int n_work = count;
  int e = -1;         // this half's pair
  int e_last = -1;    // the last pair this half had (its planes keep the idle half's loads in bounds)
  bool act = false;   // the half has a pair in progress
  bool more = true;   // the work queue may still hold pairs for this half
  int it = 0;
  double linf = -1.0, linf_prev = -1.0;
  int rot = 0;        // B0 = Q[rot], B1 = Q[rot + 1], B2 = Q[rot + 2] (mod 3)
  for (;;) {
    // a half without a pair takes the next one (passing over the pairs solved in
    // FP64 on the second stream) and writes its initial sigma into B0
    while (!act && more) {
      int idx = 0;
      if (lp == 0) idx = atomicAdd(pd.counter, 1);
      idx = __shfl(idx, half * kHalfWave);
      if (idx < n_work) {
        e = pd.order[first + idx];
        if (pd.skip[e] == kSkipSpec) continue;
        e_last = e;
        act = true;
        it = 0;
        linf = linf_prev = -1.0;
        const EventX ev = pd.ev[e];
        const GridDev& g = pd.grids[ev.grid];
        const int n = (int)g.wplane;
        U* B0 = Q[rot] + hoff;
        for (int q = lp; q < n; q += kHalfWave) {  // :119-129 initial sigma, orientation 0
          int ix, iy;
          const bool real = wave_decode0(g, q, ix, iy);
          const bool fixed = real && ix >= ev.fb[0] && ix <= ev.fb[1] && iy >= ev.fb[2] && iy <= ev.fb[3];
          B0[q] = !real ? U(0) : (fixed ? U(0) : U(9));
        }
      } else {
        more = false;
      }
    }
    const unsigned long long busy = __ballot(act);
    if (busy == 0) break;
    // an idle half sweeps the other half's pair (or its own last pair) and stores nothing
    const int e_other = __shfl(e_last, (half ^ 1) * kHalfWave);
    const int ec = act ? e : (e_last >= 0 ? e_last : e_other);
    __syncthreads();  // the initial planes are in memory before the sweeps read them
    const EventX ev = pd.ev[ec];
    const GridDev& g = pd.grids[ev.grid];
    const auto cx = phase_fwd_ctx<SAT_PREC_F32, kMetricIso, Pol>(pd, g, ev);
    U* B0 = Q[rot];
    U* B1 = Q[(rot + 1) % 3];
    U* B2 = Q[(rot + 2) % 3];
    double lmax = 0.0, lsum = 0.0;
    sweep_pair2<Pol, C, 1, 1, false, 1>(cx, B0 + hoff, B1, nullptr, hoff, out_bytes, act, lmax, lsum);
    __syncthreads();
    sweep_pair2<Pol, C, 1, -1, false, 1>(cx, B1 + hoff, B2, nullptr, hoff, out_bytes, act, lmax, lsum);
    __syncthreads();
    sweep_pair2<Pol, C, -1, 1, false, 0>(cx, B2 + hoff, B1, nullptr, hoff, out_bytes, act, lmax, lsum);
    __syncthreads();
    sweep_pair2<Pol, C, -1, -1, true, 0>(cx, B1 + hoff, B2, B0 + hoff, hoff, out_bytes, act, lmax, lsum);
    __syncthreads();
    const double Linf = half_max(lmax);
    const double L1 = half_sum(lsum) / (double)(g.nx * g.ny);
    if (act) {
      it++;
      linf_prev = linf;
      linf = Linf;
      const bool converged = L1 < kFwdTol && Linf < kFwdTol;
      if (converged || it >= pd.fwd_maxiter) {
        // the pair's epilogue (fwd_sweep_kernel): B2 holds its result, orientation 0
        if (lp == 0) {
          pd.conv[2 * (size_t)e] = linf;
          pd.conv[2 * (size_t)e + 1] = linf_prev;
          pd.iters_f[e] = it;
          if (!converged) atomicOr(pd.flags + e, FLAG_NOT_CONVERGED);
        }
        const double b = pd.f32_band;
        const bool near_last = converged && linf >= (1.0 - b) * kFwdTol;
        const bool near_prev = linf_prev >= 0.0 && linf_prev < (1.0 + b) * kFwdTol;
        const bool amb = b > 0.0 && (near_last || near_prev);
        if (lp == 0) {
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
        if (!amb) {
          // :317 T = T0 + sigma*T0, column-major
          double* T = pd.T + ev.pl_off;
          const U* u = B2 + hoff;
          for (int iy = 0; iy < g.ny; iy++)
            for (int ix = lp; ix < g.nx; ix += kHalfWave) {
              const size_t o = wave_index(g, 0, ix, iy);
              const size_t p = ix + (size_t)iy * g.nx;
              T[p] = pd.T0d[ev.pl_off + p] + (double)u[o] * pd.T0d[ev.pl_off + p];
            }
        }
        act = false;
      }
    }
    // the next round starts from B2
    rot = (rot + 2) % 3;
    __syncthreads();  // the results are read before a new pair's initial plane overwrites them
  }
}

}  // namespace sat
