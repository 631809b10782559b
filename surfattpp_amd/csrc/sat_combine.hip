// sat_combine.hip -- per-period 2-D sensitivity kernels combined with the depth
// sensitivities into 3-D model kernels: preproc::combine_kernels
// (src/preproc.cpp:196-342), its isotropic, anisotropic (Gc, Gs) and density
// branches.
//
// One thread per output element (ix, iy, k); it walks the periods in order
// (:239), so every output is the reference's FP64 sum in the reference's order.
// A batched dot product per element (nper terms), read once: HBM-bound, no MFMA.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "../../include/surfatt_eik.h"

namespace {

// include/utils.h:198-200 and :210-212, same operation order
__device__ inline double drho_dalpha(double vp) {
  return 1.6612 - 2 * 0.4721 * vp + 3 * 0.0671 * vp * vp - 4 * 0.0043 * vp * vp * vp +
         5 * 0.000106 * vp * vp * vp * vp;
}
__device__ inline double dalpha_dbeta(double vs) {
  return 2.0947 - 2 * 0.8206 * vs + 3 * 0.2683 * vs * vs - 4 * 0.0251 * vs * vs * vs;
}

struct CombineDev {
  int nx, ny, nz, nper, mode;
  const double* adj;    // [nper][ny][nx]
  const double* s_vs;   // [nx][ny][nz][nper]
  const double* s_vp;
  const double* s_rho;
  const double* vs3d;   // [nx][ny][nz]
  const double* vp3d;
  double* k_vs;         // [nx][ny][nz]
  double* k_vp;
  double* k_rho;
  // is_anisotropy (r1 != nullptr): Gc, Gs kernels and the scale of the vs/vp/rho terms
  const double* r1;     // [nx][ny][nper]
  const double* r2;
  const double* axi;    // [nper][ny][nx]
  const double* aeta;
  const double* svel;   // [nper][ny][nx]
  const double* s_gc;   // [nx][ny][nz][nper]
  const double* s_gs;
  double* k_gc;
  double* k_gs;
  // is_kden (aden != nullptr)
  const double* aden;   // [nper][ny][nx]
  double* k_den;
};

// A work-group of 256 threads takes R consecutive output elements (a tile of
// R * nper sensitivity entries, contiguous in every tensor).  Pass 1: the
// threads walk the tile with coalesced loads and form each period's term of
// each kernel (att * (...), the Gc/Gs terms, the density term) into LDS.
// Pass 2: one thread per element sums its terms in period order -- the
// reference's operations in the reference's order (:246-338), so the same
// bits.  R = 64 for up to three kernels, 32 with the anisotropic and density
// kernels (six).
constexpr int kThreads = 256;
constexpr int kMaxPeriods = 96;  // LDS: 6 * 32 * 96 or 3 * 64 * 96 doubles at most (144 KiB)
template <int R>
__global__ __launch_bounds__(kThreads) void combine_kernel(CombineDev d) {
  extern __shared__ double sh[];  // terms [nk][R * nper], then dab, dra [R] each
  const size_t n3 = (size_t)d.nx * d.ny * d.nz;
  const size_t plane = (size_t)d.nx * d.ny;
  const int np = d.nper;
  const bool aniso = d.r1 != nullptr, kden = d.aden != nullptr;
  // kernel slots in LDS: vs, vp, rho (modes 1, 2), gc, gs, den
  const int i_vp = 1, i_rho = 2, i_gc = (d.mode == 0) ? 1 : 3, i_gs = i_gc + 1;
  const int i_den = i_gc + (aniso ? 2 : 0);
  const int nk = i_den + (kden ? 1 : 0);
  const size_t ts = (size_t)R * np;  // terms per kernel slot
  double* fac = sh + (size_t)nk * ts;  // dab [R] | dra [R]
  for (size_t r0 = (size_t)blockIdx.x * R; r0 < n3; r0 += (size_t)gridDim.x * R) {
    const int rows = (int)min((size_t)R, n3 - r0);
    __syncthreads();  // the previous tile's readers are done
    if (d.mode == 0 && (int)threadIdx.x < rows) {
      fac[threadIdx.x] = dalpha_dbeta(d.vs3d[r0 + threadIdx.x]);
      fac[R + threadIdx.x] = drho_dalpha(d.vp3d[r0 + threadIdx.x]);
    }
    __syncthreads();
    const size_t e0 = r0 * np;
    for (int i = threadIdx.x; i < rows * np; i += kThreads) {
      const int row = i / np, ip = i - row * np;
      const size_t node = (r0 + row) / d.nz;  // ix * ny + iy
      const int ix = (int)(node / d.ny), iy = (int)(node - (size_t)ix * d.ny);
      const size_t o2 = ip * plane + ix + (size_t)iy * d.nx;
      const double att = d.adj[o2];
      const double aden = kden ? d.aden[o2] : 0.0;
      const double r1 = aniso ? d.r1[node * np + ip] : 0.0;
      const double r2 = aniso ? d.r2[node * np + ip] : 0.0;
      if (d.mode == 0) {
        const double dab = fac[row], dra = fac[R + row];
        sh[i] = att * (d.s_vs[e0 + i] + d.s_vp[e0 + i] * dab + d.s_rho[e0 + i] * dra * dab);
        if (kden) sh[i_den * ts + i] = aden * (d.s_vs[e0 + i] + d.s_vp[e0 + i] * dab + d.s_rho[e0 + i] * dra * dab);
      } else {
        const double scale = 1.0 / sqrt(1.0 + r1 * r1 + r2 * r2);  // :253-256 (1 without anisotropy)
        sh[i] = att * scale * d.s_vs[e0 + i];
        sh[i_vp * ts + i] = att * scale * d.s_vp[e0 + i];
        if (kden) sh[i_den * ts + i] = aden * scale * d.s_vs[e0 + i];
        if (d.mode == 1) sh[i_rho * ts + i] = att * scale * d.s_rho[e0 + i];
      }
      if (aniso) {  // :313-336
        const double D = 1.0 + r1 * r1 + r2 * r2;
        const double sqrtD = sqrt(D);
        const double D15 = D * sqrtD;
        const double D2 = D * D;
        const double sv = d.svel[o2];
        const double axi = d.axi[o2], aeta = d.aeta[o2];
        sh[i_gc * ts + i] = (-att * r1 / D15 + axi * (1.0 - r1 * r1 + r2 * r2) / (sv * D2) +
                             aeta * (-2.0 * r1 * r2) / (sv * D2)) * d.s_gc[e0 + i];
        sh[i_gs * ts + i] = (-att * r2 / D15 + axi * (-2.0 * r1 * r2) / (sv * D2) +
                             aeta * (1.0 + r1 * r1 - r2 * r2) / (sv * D2)) * d.s_gs[e0 + i];
      }
    }
    __syncthreads();
    if ((int)threadIdx.x < rows) {
      const int row = threadIdx.x;
      const size_t o3 = r0 + row;
      double kvs = 0.0, kvp = 0.0, krho = 0.0, kgc = 0.0, kgs = 0.0, kdn = 0.0;
      for (int ip = 0; ip < np; ip++) {
        const int t = row * np + ip;
        kvs -= sh[t];
        if (d.mode) kvp -= sh[i_vp * ts + t];
        if (d.mode == 1) krho -= sh[i_rho * ts + t];
        if (kden) kdn -= sh[i_den * ts + t];
        if (aniso) {
          kgc += sh[i_gc * ts + t];
          kgs += sh[i_gs * ts + t];
        }
      }
      d.k_vs[o3] = kvs;
      if (d.mode) {
        d.k_vp[o3] = kvp;
        d.k_rho[o3] = (d.mode == 2) ? kvs * 0.33 : krho;  // RHO_SCALING (include/config.h:87)
      }
      if (aniso) {
        d.k_gc[o3] = kgc;
        d.k_gs[o3] = kgs;
      }
      if (kden) d.k_den[o3] = kdn;
    }
  }
}

#define C_CHECK(x)                                                                 \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "[surfatt_eik] %s: %s\n", #x, hipGetErrorString(e_));  \
      rc = (e_ == hipErrorOutOfMemory) ? SAT_ENOMEM : SAT_EDEVICE;                 \
      goto done;                                                                   \
    }                                                                              \
  } while (0)

}  // namespace

namespace sat {

// The combination with the per-period kernels either from a caller's pointer
// (host or device memory: copied, the runtime tells them apart) or already on
// the device in the batch's workspace (adj_dev, [nper][ny][nx], read in place).
int combine_impl(const sat_combine_desc& in, const double* adj_dev, double* kernel_ms) {
  const int nx = in.nx, ny = in.ny, nz = in.nz, nper = in.nper, mode = in.mode;
  if (nx < 1 || ny < 1 || nz < 1 || nper < 1 || mode < 0 || mode > 2) return SAT_EINVAL;
  if (nper > kMaxPeriods) return SAT_EINVAL;  // the staged rows must fit in LDS
  if ((!in.adj_s && !adj_dev) || !in.sen_vs || !in.ker_vs) return SAT_EINVAL;
  if (mode == 0 && (!in.sen_vp || !in.sen_rho || !in.vs3d || !in.vp3d)) return SAT_EINVAL;
  if (mode > 0 && (!in.sen_vp || !in.ker_vp || !in.ker_rho || (mode == 1 && !in.sen_rho))) return SAT_EINVAL;
  const bool aniso = in.r1 != nullptr, kden = in.adj_den != nullptr;
  if (aniso && (!in.r2 || !in.adj_xi || !in.adj_eta || !in.svel || !in.sen_gc || !in.sen_gs || !in.ker_gc ||
                !in.ker_gs))
    return SAT_EINVAL;
  if (kden && !in.ker_den) return SAT_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SAT_EDEVICE;
  const size_t n2 = (size_t)nx * ny, n3 = n2 * nz, n4 = n3 * nper;
  int rc = SAT_OK;
  double* buf = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CombineDev d{};
  const size_t na2 = (adj_dev ? 0 : nper * n2) + (aniso ? 3 * nper * n2 + 2 * n2 * nper : 0) + (kden ? nper * n2 : 0);
  const size_t total = na2 + (3 + (aniso ? 2 : 0)) * n4 + (5 + (aniso ? 2 : 0) + (kden ? 1 : 0)) * n3;
  C_CHECK(hipMalloc(&buf, sizeof(double) * total));
  {
    double* p = buf;
    auto take = [&](size_t n) { double* q = p; p += n; return q; };
    auto up = [&](const double* src, size_t n) -> double* {  // a device copy (nullptr for no source)
      if (!src) return nullptr;
      double* q = take(n);
      return hipMemcpy(q, src, sizeof(double) * n, hipMemcpyDefault) == hipSuccess ? q : nullptr;
    };
    d.nx = nx; d.ny = ny; d.nz = nz; d.nper = nper; d.mode = mode;
    d.adj = adj_dev ? adj_dev : up(in.adj_s, nper * n2);
    d.s_vs = up(in.sen_vs, n4);
    d.s_vp = up(in.sen_vp, n4);
    d.s_rho = (mode != 2) ? up(in.sen_rho, n4) : nullptr;
    d.vs3d = (mode == 0) ? up(in.vs3d, n3) : nullptr;
    d.vp3d = (mode == 0) ? up(in.vp3d, n3) : nullptr;
    d.k_vs = take(n3);
    d.k_vp = take(n3);
    d.k_rho = take(n3);
    if (aniso) {
      d.r1 = up(in.r1, n2 * nper);
      d.r2 = up(in.r2, n2 * nper);
      d.axi = up(in.adj_xi, nper * n2);
      d.aeta = up(in.adj_eta, nper * n2);
      d.svel = up(in.svel, nper * n2);
      d.s_gc = up(in.sen_gc, n4);
      d.s_gs = up(in.sen_gs, n4);
      d.k_gc = take(n3);
      d.k_gs = take(n3);
    }
    if (kden) {
      d.aden = up(in.adj_den, nper * n2);
      d.k_den = take(n3);
    }
    if (!d.adj || !d.s_vs || (in.sen_vp && !d.s_vp) || (mode == 1 && !d.s_rho) ||
        (mode == 0 && (!d.vs3d || !d.vp3d)) ||
        (aniso && (!d.r1 || !d.r2 || !d.axi || !d.aeta || !d.svel || !d.s_gc || !d.s_gs)) || (kden && !d.aden)) {
      rc = SAT_EDEVICE;
      goto done;
    }
  }
  C_CHECK(hipEventCreate(&e0));
  C_CHECK(hipEventCreate(&e1));
  {
    const int nk = ((mode == 0) ? 1 : 3) + (aniso ? 2 : 0) + (kden ? 1 : 0);
    const int R = nk <= 3 ? 64 : 32;
    const unsigned blocks = (unsigned)std::min<size_t>((n3 + R - 1) / R, 1u << 20);
    const size_t lds = sizeof(double) * ((size_t)nk * R * (size_t)nper + 2 * R);
    C_CHECK(hipEventRecord(e0, nullptr));
    if (R == 64) hipLaunchKernelGGL(combine_kernel<64>, dim3(blocks), dim3(kThreads), lds, nullptr, d);
    else hipLaunchKernelGGL(combine_kernel<32>, dim3(blocks), dim3(kThreads), lds, nullptr, d);
    C_CHECK(hipGetLastError());
    C_CHECK(hipEventRecord(e1, nullptr));
    C_CHECK(hipEventSynchronize(e1));
    if (kernel_ms) {
      float t = 0.f;
      C_CHECK(hipEventElapsedTime(&t, e0, e1));
      *kernel_ms = t;
    }
  }
  C_CHECK(hipMemcpy(in.ker_vs, d.k_vs, sizeof(double) * n3, hipMemcpyDefault));
  if (mode > 0) {
    C_CHECK(hipMemcpy(in.ker_vp, d.k_vp, sizeof(double) * n3, hipMemcpyDefault));
    C_CHECK(hipMemcpy(in.ker_rho, d.k_rho, sizeof(double) * n3, hipMemcpyDefault));
  }
  if (aniso) {
    C_CHECK(hipMemcpy(in.ker_gc, d.k_gc, sizeof(double) * n3, hipMemcpyDefault));
    C_CHECK(hipMemcpy(in.ker_gs, d.k_gs, sizeof(double) * n3, hipMemcpyDefault));
  }
  if (kden) C_CHECK(hipMemcpy(in.ker_den, d.k_den, sizeof(double) * n3, hipMemcpyDefault));
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (buf) (void)hipFree(buf);
  return rc;
}

// the isotropic entry's arguments as a descriptor
int combine_impl(int nx, int ny, int nz, int nper, int mode, const double* adj_s, const double* adj_dev,
                 const double* sen_vs, const double* sen_vp, const double* sen_rho, const double* vs3d,
                 const double* vp3d, double* ker_vs, double* ker_vp, double* ker_rho, double* kernel_ms) {
  sat_combine_desc c{};
  c.nx = nx; c.ny = ny; c.nz = nz; c.nper = nper; c.mode = mode;
  c.adj_s = adj_s; c.sen_vs = sen_vs; c.sen_vp = sen_vp; c.sen_rho = sen_rho; c.vs3d = vs3d; c.vp3d = vp3d;
  c.ker_vs = ker_vs; c.ker_vp = ker_vp; c.ker_rho = ker_rho;
  return combine_impl(c, adj_dev, kernel_ms);
}

}  // namespace sat

extern "C" int sat_combine_kernels(int nx, int ny, int nz, int nper, int mode, const double* adj_s,
                                   const double* sen_vs, const double* sen_vp, const double* sen_rho,
                                   const double* vs3d, const double* vp3d, double* ker_vs, double* ker_vp,
                                   double* ker_rho, double* kernel_ms) {
  if (!adj_s) return SAT_EINVAL;
  return sat::combine_impl(nx, ny, nz, nper, mode, adj_s, nullptr, sen_vs, sen_vp, sen_rho, vs3d, vp3d, ker_vs,
                           ker_vp, ker_rho, kernel_ms);
}

extern "C" int sat_combine_kernels_ex(const sat_combine_desc* d, double* kernel_ms) {
  if (!d || !d->adj_s) return SAT_EINVAL;
  return sat::combine_impl(*d, nullptr, kernel_ms);
}
