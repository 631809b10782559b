// Drives the input-producer shim (include/surfatt_surfker.hpp) on the GPU and
// prints, as %.17g: for the reference's unit-test model
// (tests/surfdisp_test/test_surfdisp.cpp:64-78) the float layers and the
// velocities of three requests; the [ix][iy][iper] maps of a small 3-D model by
// fwdsurf; the metric of a small surface by topo_media.
#include <cmath>
#include <cstdio>
#include <vector>

#include "surfatt_surfker.hpp"

static void put(const char* tag, const std::vector<double>& v) {
  std::printf("%s", tag);
  for (double x : v) std::printf(" %.17g", x);
  std::printf("\n");
}

int main() {
  const std::vector<double> dep{0.0, 10.0, 20.0, 30.0, 40.0}, vs{2.5, 3.5, 4.4, 4.5, 4.6};
  const std::vector<double> per{5.0, 10.0, 20.0, 30.0, 40.0};
  const int cases[3][3] = {{2, 0, 0}, {2, 1, 0}, {1, 0, 0}};  // iwave, igr, iflsph
  for (const auto& c : cases) {
    const auto req = surfker::build_disp_req(dep, vs, per, c[2], c[0], 1, c[1]);
    std::vector<double> f[4];
    for (size_t i = 0; i < req.thickness_km.size(); i++) {
      f[0].push_back((float)req.thickness_km[i]);
      f[1].push_back((float)req.vp_km_s[i]);
      f[2].push_back((float)req.vs_km_s[i]);
      f[3].push_back((float)req.rho_g_cm3[i]);
    }
    put("layers_thk", f[0]);
    put("layers_vp", f[1]);
    put("layers_vs", f[2]);
    put("layers_rho", f[3]);
    put("c", surfker::surfdisp(req));
  }
  // 3 x 2 columns, 6 nodes over 0-50 km
  const int nx = 3, ny = 2, nz = 6;
  std::vector<double> z(nz), vs3((size_t)nx * ny * nz);
  for (int k = 0; k < nz; k++) z[k] = 10.0 * k;
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int k = 0; k < nz; k++) vs3[((size_t)ix * ny + iy) * nz + k] = 2.9 + 0.3 * k + 0.05 * ix - 0.04 * iy;
  put("fwdsurf", surfatt::fwdsurf(nx, ny, z, vs3, {}, {}, per, 2, 0));
  // a Gaussian bump on a 9 x 7 grid at 0.02 degrees, two periods
  std::vector<double> xg(9), yg(7), zz(63);
  for (int i = 0; i < 9; i++) xg[i] = 120.0 + 0.02 * i;
  for (int j = 0; j < 7; j++) yg[j] = 30.0 + 0.02 * j;
  for (int j = 0; j < 7; j++)
    for (int i = 0; i < 9; i++) zz[i + 9 * j] = 1.5 * std::exp(-((i - 4) * (i - 4) + (j - 3) * (j - 3)) / 6.0);
  std::vector<double> m11, m12, m22;
  surfatt::topo_media(xg, yg, zz, {3.0, 3.4}, {4.0, 8.0}, 2.5, m11, m12, m22);
  put("m11", m11);
  put("m12", m12);
  put("m22", m22);
  return 0;
}
