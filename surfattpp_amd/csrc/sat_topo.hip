// sat_topo.hip -- the anisotropic metric of a topographic surface, per period,
// on the GPU: SurfGrid::build_media_matrix_with_topo (src/surf_grid.cpp:109-145)
// with gaussian_smooth_geo_2 (include/utils.h:505-573) and gradient_2_geo
// (include/utils.h:595-649); the planes the eikonal path takes as spha, sphb,
// sphc (m11 = a, m22 = b, m12 = -c, src/surf_grid.cpp:184-190).
//
// smooth_kernel: one work-group per (output row j, period).  The reference's
// weight of window node (m, n) for output (i, j) depends on n and m - i only
// (haversine a = sin^2(dlat/2) + cos(lat_j) sin^2(dlon/2) cos(lat_n),
// w = exp(-(2 atan(sqrt(a/(1-a))) in degrees)^2 / (2 sigma^2))), so the
// work-group evaluates each distinct weight once into LDS -- a chunk of window
// rows at a time -- and then every thread (one per output column i) sums
// w * z and w over its clipped window from LDS weights (a broadcast) and
// coalesced z rows.  The reference evaluates the same weight at every output
// node; the value is the same up to the rounding of lon(m) - lon(i).  FP64
// throughout; compute-bound (one FMA pair per window node), no MFMA: the
// weights vary per row j, so there is no shared dense operand.
//
// metric_kernel: one thread per node of each period: central / one-sided
// geographic differences of the smoothed surface, D = 1 + fx^2 + fy^2,
// a = (1 + fy^2)/D, b = (1 + fx^2)/D, c = fx fy / D, in the reference's
// operation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/surfatt_eik.h"
#include "sat_crmath.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kDeg2Rad = kPi / 180.0;
constexpr double kRad2Deg = 180.0 / kPi;
constexpr double kREarth = 6371.0;
constexpr int kThreads = 256;
constexpr int kLdsDoubles = 8192;  // 64 KiB: lon table + a chunk of weight rows

struct TopoDev {
  int nx, ny, nper;
  const double* lon;    // [nx]
  const double* lat;    // [ny]
  const double* z;      // [ny][nx]
  const double* sigma;  // [nper] degrees
  double* zs;           // [nper][ny][nx]
  double* m11;          // [nper][ny][nx]
  double* m12;
  double* m22;
};

// Window half-width in longitude, clipped to the grid (offsets beyond nx - 1
// never reach a node).
__host__ __device__ inline int lon_half_width(double sigma, double dx, int nx) {
  const double hx = 3.0 * sigma / dx;
  return hx < (double)(nx - 1) ? (int)hx : nx - 1;
}

__global__ __launch_bounds__(kThreads) void smooth_kernel(TopoDev d) {
  extern __shared__ double sh[];
  const int j = blockIdx.x, p = blockIdx.y, nx = d.nx, ny = d.ny;
  const double* z = d.z;
  double* out = d.zs + ((size_t)p * ny + j) * nx;
  const double sigma = d.sigma[p];
  if (!(sigma > 0.0)) {  // no smoothing: the surface itself
    for (int i = threadIdx.x; i < nx; i += blockDim.x) out[i] = z[(size_t)j * nx + i];
    return;
  }
  const double sigma3 = 3.0 * sigma;
  const double dx = d.lon[1] - d.lon[0];
  const double dy = d.lat[1] - d.lat[0];
  const int hx = lon_half_width(sigma, dx, nx);
  const int K = 2 * hx + 1;
  const double lat0 = d.lat[j];
  const double cos_lat0 = sat::cr::cos_cr(lat0 * kDeg2Rad);
  const double hyd = sigma3 / cos_lat0 / dy;
  const int hy = hyd < (double)ny ? (int)hyd : ny;
  const int n1 = max(0, j - hy), n2 = min(ny - 1, j + hy);
  const double inv_2sig2 = 1.0 / (2.0 * (sigma * sigma));
  constexpr double two_rad2deg = 2.0 * kRad2Deg;

  // lon_contrib(k) = cos(lat_j) * sin^2(dlon/2) for offset k - hx (even in the offset)
  const int rows_per_chunk = (kLdsDoubles - K) / (K + 2);  // >= 1 (checked on the host)
  double* lc = sh;                         // [K]
  double* s2lat = sh + K;                  // [rows_per_chunk] sin^2(dlat/2) of each window row
  double* clat = s2lat + rows_per_chunk;   // [rows_per_chunk] cos(lat) of each window row
  double* W = clat + rows_per_chunk;       // [rows_per_chunk][K]
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    const int o = k < hx ? hx - k : k - hx;
    const double s = sat::cr::sin_cr((d.lon[o] - d.lon[0]) * (kDeg2Rad * 0.5));
    lc[k] = cos_lat0 * (s * s);
  }

  // this thread's output columns
  constexpr int kMaxCols = 8;  // nx <= kThreads * kMaxCols (checked on the host)
  double num[kMaxCols], den[kMaxCols];
#pragma unroll
  for (int c = 0; c < kMaxCols; c++) num[c] = den[c] = 0.0;

  for (int r0 = n1; r0 <= n2; r0 += rows_per_chunk) {
    const int rows = min(rows_per_chunk, n2 - r0 + 1);
    __syncthreads();  // the previous chunk's readers are done
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
      const double latn = d.lat[r0 + r];
      const double s = sat::cr::sin_cr((latn - lat0) * (kDeg2Rad * 0.5));
      s2lat[r] = s * s;
      clat[r] = sat::cr::cos_cr(latn * kDeg2Rad);
    }
    __syncthreads();  // lc and the row table written
    for (int e = threadIdx.x; e < rows * K; e += blockDim.x) {
      const int r = e / K, k = e - r * K;
      double a = s2lat[r] + lc[k] * clat[r];
      a = fmin(a, 1.0);
      const double t = two_rad2deg * atan(sqrt(a / (1.0 - a)));
      W[e] = exp(-(t * t) * inv_2sig2);
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kMaxCols; c++) {
      const int i = threadIdx.x + c * kThreads;
      if (i < nx) {
        const int klo = max(0, hx - i), khi = min(K - 1, nx - 1 - i + hx);
        double sn = num[c], sd = den[c];
        for (int r = 0; r < rows; r++) {
          const double* __restrict__ zr = z + (size_t)(r0 + r) * nx + (i - hx);  // window column i - hx + k
          const double* wr = W + r * K;
          // unrolled so that the loads of several window nodes are in flight
          // at once; the sums keep the node order
#pragma unroll 8
          for (int k = klo; k <= khi; k++) {
            const double w = wr[k];
            sn += w * zr[k];
            sd += w;
          }
        }
        num[c] = sn;
        den[c] = sd;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < kMaxCols; c++) {
    const int i = threadIdx.x + c * kThreads;
    if (i < nx) out[i] = num[c] / den[c];
  }
}

__global__ void metric_kernel(TopoDev d) {
  const int nx = d.nx, ny = d.ny;
  const size_t plane = (size_t)nx * ny;
  const size_t total = plane * d.nper;
  const double km = kREarth * kDeg2Rad;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
    const size_t p = q / plane;
    const int j = (int)((q - p * plane) / nx), i = (int)(q - p * plane - (size_t)j * nx);
    const double* f = d.zs + p * plane;
    const double cos_lat = sat::cr::cos_cr(d.lat[j] * kDeg2Rad);
    // dx(i, j) = (lon(i) - lon(i-1)) * R * DEG2RAD * cos(lat j); dy(j) likewise without cos
    auto dxk = [&](int ii) { return ((d.lon[ii] - d.lon[ii - 1]) * km) * cos_lat; };
    auto dyk = [&](int jj) { return (d.lat[jj] - d.lat[jj - 1]) * km; };
    const double* row = f + (size_t)j * nx;
    double fx, fy;
    if (i == 0)
      fx = (row[1] - row[0]) / dxk(1);
    else if (i == nx - 1)
      fx = (row[nx - 1] - row[nx - 2]) / dxk(nx - 1);
    else
      fx = (row[i + 1] - row[i - 1]) / (2.0 * dxk(i));
    if (j == 0)
      fy = (f[nx + i] - f[i]) / dyk(1);
    else if (j == ny - 1)
      fy = (f[(size_t)(ny - 1) * nx + i] - f[(size_t)(ny - 2) * nx + i]) / dyk(ny - 1);
    else
      fy = (f[(size_t)(j + 1) * nx + i] - f[(size_t)(j - 1) * nx + i]) / (2.0 * dyk(j));
    const double fx2 = fx * fx, fy2 = fy * fy;
    const double denom = 1.0 + fx2 + fy2;
    const double a = (1.0 + fy2) / denom;
    const double b = (1.0 + fx2) / denom;
    const double c = fx * fy / denom;
    d.m11[q] = a;
    d.m22[q] = b;
    d.m12[q] = -c;
  }
}

#define T_CHECK(x)                                                                 \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "[surfatt_eik] %s: %s\n", #x, hipGetErrorString(e_));  \
      rc = (e_ == hipErrorOutOfMemory) ? SAT_ENOMEM : SAT_EDEVICE;                 \
      goto done;                                                                   \
    }                                                                              \
  } while (0)

}  // namespace

extern "C" int sat_topo_metric(int nx, int ny, int nper, const double* lon_deg, const double* lat_deg,
                               const double* z, const double* sigma_deg, double* m11, double* m12, double* m22,
                               double* zs_out, double* kernel_ms) {
  if (nx < 3 || ny < 3 || nper < 1 || nper > 65535) return SAT_EINVAL;
  if (!lon_deg || !lat_deg || !z || !sigma_deg || !m11 || !m12 || !m22) return SAT_EINVAL;
  if (nx > kThreads * 8) return SAT_EINVAL;  // columns per thread of smooth_kernel
  const double dx = lon_deg[1] - lon_deg[0], dy = lat_deg[1] - lat_deg[0];
  if (!(dx > 0.0) || !(dy > 0.0)) return SAT_EINVAL;
  for (int p = 0; p < nper; p++) {
    if (std::isnan(sigma_deg[p]) || std::isinf(sigma_deg[p])) return SAT_EINVAL;
    if (sigma_deg[p] > 0.0) {
      const int K = 2 * lon_half_width(sigma_deg[p], dx, nx) + 1;
      if (K + K + 2 > kLdsDoubles) return SAT_EINVAL;  // the lon table and one weight row must fit in LDS
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SAT_EDEVICE;
  int rc = SAT_OK;
  const size_t plane = (size_t)nx * ny, n3 = plane * nper;
  double* buf = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TopoDev d{};
  // one allocation: lon | lat | z | sigma | zs | m11 | m12 | m22
  T_CHECK(hipMalloc(&buf, sizeof(double) * (nx + ny + plane + nper + 4 * n3)));
  {
    double* q = buf;
    auto take = [&](size_t n) { double* r = q; q += n; return r; };
    double* lon = take(nx);
    double* lat = take(ny);
    double* zz = take(plane);
    double* sg = take(nper);
    d = TopoDev{nx, ny, nper, lon, lat, zz, sg, take(n3), take(n3), take(n3), take(n3)};
    T_CHECK(hipMemcpy(lon, lon_deg, sizeof(double) * nx, hipMemcpyHostToDevice));
    T_CHECK(hipMemcpy(lat, lat_deg, sizeof(double) * ny, hipMemcpyHostToDevice));
    T_CHECK(hipMemcpy(zz, z, sizeof(double) * plane, hipMemcpyHostToDevice));
    T_CHECK(hipMemcpy(sg, sigma_deg, sizeof(double) * nper, hipMemcpyHostToDevice));
  }
  T_CHECK(hipEventCreate(&e0));
  T_CHECK(hipEventCreate(&e1));
  T_CHECK(hipEventRecord(e0, nullptr));
  hipLaunchKernelGGL(smooth_kernel, dim3(ny, nper), dim3(kThreads), sizeof(double) * kLdsDoubles, nullptr, d);
  T_CHECK(hipGetLastError());
  {
    const unsigned blocks = (unsigned)std::min<size_t>((n3 + 255) / 256, 4096);
    hipLaunchKernelGGL(metric_kernel, dim3(blocks), dim3(256), 0, nullptr, d);
    T_CHECK(hipGetLastError());
  }
  T_CHECK(hipEventRecord(e1, nullptr));
  T_CHECK(hipEventSynchronize(e1));
  if (kernel_ms) {
    float t = 0.f;
    T_CHECK(hipEventElapsedTime(&t, e0, e1));
    *kernel_ms = t;
  }
  T_CHECK(hipMemcpy(m11, d.m11, sizeof(double) * n3, hipMemcpyDeviceToHost));
  T_CHECK(hipMemcpy(m12, d.m12, sizeof(double) * n3, hipMemcpyDeviceToHost));
  T_CHECK(hipMemcpy(m22, d.m22, sizeof(double) * n3, hipMemcpyDeviceToHost));
  if (zs_out) T_CHECK(hipMemcpy(zs_out, d.zs, sizeof(double) * n3, hipMemcpyDeviceToHost));
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (buf) (void)hipFree(buf);
  return rc;
}
