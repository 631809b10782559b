// sat_combine.hip -- per-period 2-D sensitivity kernels combined with the depth
// sensitivities into 3-D model kernels: the isotropic part of
// preproc::combine_kernels (src/preproc.cpp:196-342).
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
};

// A work-group of 256 threads takes 64 consecutive output elements (a tile of
// 64 * nper sensitivity entries, contiguous in all three tensors).  Pass 1: the
// threads walk the tile with coalesced loads and form each period's term
// att * (...) of each element into LDS.  Pass 2: one thread per element
// subtracts its terms in period order -- the reference's operations in the
// reference's order (:258-301), so the same bits.
constexpr int kRows = 64;
constexpr int kThreads = 256;
constexpr int kMaxPeriods = 96;  // LDS: 3 * 64 * 96 doubles = 144 KiB at most
__global__ __launch_bounds__(kThreads) void combine_kernel(CombineDev d) {
  extern __shared__ double sh[];  // terms [nk][kRows * nper], then dab, dra [kRows] each
  const size_t n3 = (size_t)d.nx * d.ny * d.nz;
  const size_t plane = (size_t)d.nx * d.ny;
  const int np = d.nper;
  const int nk = (d.mode == 0) ? 1 : 3;  // kernels formed
  double* fac = sh + (size_t)nk * kRows * np;  // dab [kRows] | dra [kRows]
  const double scale = 1.0 / sqrt(1.0 + 0.0 * 0.0 + 0.0 * 0.0);  // no anisotropy: 1
  for (size_t r0 = (size_t)blockIdx.x * kRows; r0 < n3; r0 += (size_t)gridDim.x * kRows) {
    const int rows = (int)min((size_t)kRows, n3 - r0);
    __syncthreads();  // the previous tile's readers are done
    if (d.mode == 0 && (int)threadIdx.x < rows) {
      fac[threadIdx.x] = dalpha_dbeta(d.vs3d[r0 + threadIdx.x]);
      fac[kRows + threadIdx.x] = drho_dalpha(d.vp3d[r0 + threadIdx.x]);
    }
    __syncthreads();
    const size_t e0 = r0 * np;
    for (int i = threadIdx.x; i < rows * np; i += kThreads) {
      const int row = i / np, ip = i - row * np;
      const size_t node = (r0 + row) / d.nz;  // ix * ny + iy
      const int ix = (int)(node / d.ny), iy = (int)(node - (size_t)ix * d.ny);
      const double att = d.adj[ip * plane + ix + (size_t)iy * d.nx];
      if (d.mode == 0) {
        const double dab = fac[row], dra = fac[kRows + row];
        sh[i] = att * (d.s_vs[e0 + i] + d.s_vp[e0 + i] * dab + d.s_rho[e0 + i] * dra * dab);
      } else {
        sh[i] = att * scale * d.s_vs[e0 + i];
        sh[kRows * np + i] = att * scale * d.s_vp[e0 + i];
        if (d.mode == 1) sh[2 * kRows * np + i] = att * scale * d.s_rho[e0 + i];
      }
    }
    __syncthreads();
    if ((int)threadIdx.x < rows) {
      const int row = threadIdx.x;
      const size_t o3 = r0 + row;
      double kvs = 0.0, kvp = 0.0, krho = 0.0;
      for (int ip = 0; ip < np; ip++) {
        kvs -= sh[row * np + ip];
        if (d.mode) kvp -= sh[kRows * np + row * np + ip];
        if (d.mode == 1) krho -= sh[2 * kRows * np + row * np + ip];
      }
      d.k_vs[o3] = kvs;
      if (d.mode) {
        d.k_vp[o3] = kvp;
        d.k_rho[o3] = (d.mode == 2) ? kvs * 0.33 : krho;  // RHO_SCALING (include/config.h:87)
      }
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

// The combination with the per-period kernels either on the host (adj_s,
// uploaded) or already on the device (adj_dev, [nper][ny][nx], read in place).
int combine_impl(int nx, int ny, int nz, int nper, int mode, const double* adj_s, const double* adj_dev,
                 const double* sen_vs, const double* sen_vp, const double* sen_rho, const double* vs3d,
                 const double* vp3d, double* ker_vs, double* ker_vp, double* ker_rho, double* kernel_ms) {
  if (nx < 1 || ny < 1 || nz < 1 || nper < 1 || mode < 0 || mode > 2) return SAT_EINVAL;
  if (nper > kMaxPeriods) return SAT_EINVAL;  // the staged rows must fit in LDS
  if ((!adj_s && !adj_dev) || !sen_vs || !ker_vs) return SAT_EINVAL;
  if (mode == 0 && (!sen_vp || !sen_rho || !vs3d || !vp3d)) return SAT_EINVAL;
  if (mode > 0 && (!sen_vp || !ker_vp || !ker_rho || (mode == 1 && !sen_rho))) return SAT_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SAT_EDEVICE;
  const size_t n2 = (size_t)nx * ny, n3 = n2 * nz, n4 = n3 * nper;
  int rc = SAT_OK;
  double* buf = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CombineDev d{};
  // one allocation: adj | s_vs | s_vp | s_rho | vs3d | vp3d | k_vs | k_vp | k_rho
  const size_t total = (adj_dev ? 0 : nper * n2) + 3 * n4 + 5 * n3;
  C_CHECK(hipMalloc(&buf, sizeof(double) * total));
  {
    double* p = buf;
    auto take = [&](size_t n) { double* q = p; p += n; return q; };
    double* adj = take(adj_dev ? 0 : nper * n2);
    double* s_vs = take(n4);
    double* s_vp = take(n4);
    double* s_rho = take(n4);
    double* v_s = take(n3);
    double* v_p = take(n3);
    d = CombineDev{nx, ny, nz, nper, mode, adj_dev ? adj_dev : adj, s_vs, s_vp, s_rho, v_s, v_p, take(n3),
                   take(n3), take(n3)};
    if (!adj_dev) C_CHECK(hipMemcpy(adj, adj_s, sizeof(double) * nper * n2, hipMemcpyHostToDevice));
    C_CHECK(hipMemcpy(s_vs, sen_vs, sizeof(double) * n4, hipMemcpyHostToDevice));
    if (sen_vp) C_CHECK(hipMemcpy(s_vp, sen_vp, sizeof(double) * n4, hipMemcpyHostToDevice));
    if (sen_rho && mode != 2) C_CHECK(hipMemcpy(s_rho, sen_rho, sizeof(double) * n4, hipMemcpyHostToDevice));
    if (mode == 0) {
      C_CHECK(hipMemcpy(v_s, vs3d, sizeof(double) * n3, hipMemcpyHostToDevice));
      C_CHECK(hipMemcpy(v_p, vp3d, sizeof(double) * n3, hipMemcpyHostToDevice));
    }
  }
  C_CHECK(hipEventCreate(&e0));
  C_CHECK(hipEventCreate(&e1));
  {
    const unsigned blocks = (unsigned)std::min<size_t>((n3 + kRows - 1) / kRows, 1u << 20);
    const size_t lds = sizeof(double) * ((mode == 0 ? 1 : 3) * kRows * (size_t)nper + 2 * kRows);
    C_CHECK(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(combine_kernel, dim3(blocks), dim3(kThreads), lds, nullptr, d);
    C_CHECK(hipGetLastError());
    C_CHECK(hipEventRecord(e1, nullptr));
    C_CHECK(hipEventSynchronize(e1));
    if (kernel_ms) {
      float t = 0.f;
      C_CHECK(hipEventElapsedTime(&t, e0, e1));
      *kernel_ms = t;
    }
  }
  C_CHECK(hipMemcpy(ker_vs, d.k_vs, sizeof(double) * n3, hipMemcpyDeviceToHost));
  if (mode > 0) {
    C_CHECK(hipMemcpy(ker_vp, d.k_vp, sizeof(double) * n3, hipMemcpyDeviceToHost));
    C_CHECK(hipMemcpy(ker_rho, d.k_rho, sizeof(double) * n3, hipMemcpyDeviceToHost));
  }
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (buf) (void)hipFree(buf);
  return rc;
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
