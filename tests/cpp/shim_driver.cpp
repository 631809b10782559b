// Drives the C++ shim (include/surfatt_eik.hpp) the way the reference's call
// sites use namespace eikonal: forward, adjoint, mask on one grid, then one
// batched run.  Reads the case from argv[1] (binary: nx, ny, x0, y0, nr, then
// xx[nx], yy[ny], spha, sphb, sphc, fun [nx*ny each], xr, yr, adj [nr each])
// and writes T, Ta (masked) [nx*ny each] to argv[2].
#include <cstdio>
#include <vector>

#include "surfatt_eik.hpp"

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[2];
  double src[2];
  int nr = 0;
  if (std::fread(hdr, sizeof(int), 2, f) != 2 || std::fread(src, sizeof(double), 2, f) != 2 ||
      std::fread(&nr, sizeof(int), 1, f) != 1)
    return 2;
  const int nx = hdr[0], ny = hdr[1];
  std::vector<double> xx(nx), yy(ny), xr(nr), yr(nr), adj(nr);
  surfatt::Field spha(nx, ny), sphb(nx, ny), sphc(nx, ny), fun(nx, ny);
  bool ok = std::fread(xx.data(), sizeof(double), nx, f) == (size_t)nx &&
            std::fread(yy.data(), sizeof(double), ny, f) == (size_t)ny;
  for (surfatt::Field* m : {&spha, &sphb, &sphc, &fun}) ok = ok && std::fread(m->data(), sizeof(double), m->size(), f) == m->size();
  for (std::vector<double>* v : {&xr, &yr, &adj}) ok = ok && std::fread(v->data(), sizeof(double), nr, f) == (size_t)nr;
  std::fclose(f);
  if (!ok) return 2;
  try {
    const surfatt::Field T = eikonal::FSM_UW_PS_lonlat_2d(xx, yy, spha, sphb, sphc, fun, src[0], src[1]);
    surfatt::Field Ta = eikonal::FSM_O1_JSE_lonlat_2d(xx, yy, spha, sphb, sphc, T, xr, yr, adj);
    eikonal::mask_uniform_grid(xx, yy, Ta, src[0], src[1]);
    bool threw = false;
    try {
      eikonal::FSM_UW_PS_lonlat_2d(std::vector<double>{0.0, 1.0}, yy, spha, sphb, sphc, fun, src[0], src[1]);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    if (!threw) return 3;
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(T.data(), sizeof(double), T.size(), o);
    std::fwrite(Ta.data(), sizeof(double), Ta.size(), o);
    std::fclose(o);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
