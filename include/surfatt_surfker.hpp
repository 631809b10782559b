// surfatt_surfker.hpp -- C++20 shim for the producers of the eikonal path's
// inputs, over the C ABI (surfatt_eik.h), with the reference's names:
//
//   surfker::DispersionRequest, build_disp_req, surfdisp
//                         include/surfdisp.h:9-25,52-62; src/surfdisp.cpp:60-139,428-462
//   surfatt::fwdsurf       SurfGrid::fwdsurf, src/surf_grid.cpp:194-245 (every
//                          column of a 3-D model in one GPU call)
//   surfatt::topo_media    SurfGrid::build_media_matrix_with_topo,
//                          src/surf_grid.cpp:80-145 (per-period metric planes)
//
// Vectors replace Eigen vectors.  Errors: the reference's std::runtime_error
// messages for invalid requests; a column whose fundamental mode has no root
// throws std::runtime_error("Dispersion calculation failed") where the
// reference's Fortran stops the program.  Link with -lsurfatt_eik.
#pragma once
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "surfatt_eik.hpp"

namespace surfker {

constexpr int IFLSPH = 1;  // include/config.h:79
constexpr int IMODE = 1;   // include/config.h:80

// include/surfdisp.h:9-21
struct DispersionRequest {
  std::vector<double> thickness_km, depths_km, vp_km_s, vs_km_s, rho_g_cm3, periods_s;
  int iflsph = IFLSPH;
  int iwave = 2;
  int mode = IMODE;
  int igr = 0;
};

// src/surfdisp.cpp:60-111: nz depth nodes -> nz + 1 layers, the last node's
// layer repeated (thickness of the layer above it), then a zero-thickness
// half-space with the last node's properties.
inline DispersionRequest build_disp_req(const std::vector<double>& dep, const std::vector<double>& vs,
                                        const std::vector<double>& vp, const std::vector<double>& rho,
                                        const std::vector<double>& periods_s, int iflsph, int iwave, int mode,
                                        int igr) {
  if (dep.size() != vs.size() || dep.size() != vp.size() || dep.size() != rho.size())
    throw std::runtime_error("build_disp_req: dep, vs, vp, and rho must have the same length");
  const int nz = (int)dep.size(), mmax = nz + 1;
  DispersionRequest r;
  r.iflsph = iflsph;
  r.iwave = iwave;
  r.mode = mode;
  r.igr = igr;
  r.depths_km.assign(mmax, 0.0);
  r.thickness_km.assign(mmax, 0.0);
  r.vs_km_s.assign(mmax, 0.0);
  r.vp_km_s.assign(mmax, 0.0);
  r.rho_g_cm3.assign(mmax, 0.0);
  r.periods_s = periods_s;
  for (int k = 0; k < nz; k++) r.depths_km[k] = dep[k];
  if (nz > 0) r.depths_km[mmax - 1] = dep[nz - 1];
  for (int k = 0; k < nz; k++) {
    r.vs_km_s[k] = vs[k];
    r.vp_km_s[k] = vp[k];
    r.rho_g_cm3[k] = rho[k];
    if (k == nz - 1)
      r.thickness_km[k] = (nz > 1) ? r.thickness_km[k - 1] : 0.0;
    else
      r.thickness_km[k] = dep[k + 1] - dep[k];
  }
  if (nz > 0) {
    r.thickness_km[mmax - 1] = 0.0;
    r.vp_km_s[mmax - 1] = r.vp_km_s[nz - 1];
    r.vs_km_s[mmax - 1] = r.vs_km_s[nz - 1];
    r.rho_g_cm3[mmax - 1] = r.rho_g_cm3[nz - 1];
  }
  return r;
}

// src/surfdisp.cpp:114-127: Brocher vp and density from vs in the reference's
// element-wise array form (include/utils.h:189-195, :227-233: square, cube,
// pow(4), pow(5)).
inline DispersionRequest build_disp_req(const std::vector<double>& dep, const std::vector<double>& vs,
                                        const std::vector<double>& periods_s, int iflsph, int iwave, int mode,
                                        int igr) {
  if (dep.size() != vs.size()) throw std::runtime_error("build_disp_req: depths_km and vs_km_s must have same length");
  std::vector<double> vp(vs.size()), rho(vs.size());
  for (size_t i = 0; i < vs.size(); i++) {
    const double s = vs[i];
    vp[i] = 0.9409 + 2.0947 * s - 0.8206 * (s * s) + 0.2683 * (s * s * s) - 0.0251 * std::pow(s, 4);
    const double p = vp[i];
    rho[i] = 1.6612 * p - 0.4721 * (p * p) + 0.0671 * (p * p * p) - 0.0043 * std::pow(p, 4) +
             0.000106 * std::pow(p, 5);
  }
  return build_disp_req(dep, vs, vp, rho, periods_s, iflsph, iwave, mode, igr);
}

// src/surfdisp.cpp:35-57 and :428-462 on the GPU (sat_surfdisp96, one column).
inline std::vector<double> surfdisp(const DispersionRequest& req) {
  const size_t nlayer = req.thickness_km.size();
  if (nlayer == 0) throw std::runtime_error("surfker::surfdisp: model is empty");
  if (req.vp_km_s.size() != nlayer || req.vs_km_s.size() != nlayer || req.rho_g_cm3.size() != nlayer)
    throw std::runtime_error("surfker::surfdisp: model vectors must have same length");
  if (req.periods_s.empty()) throw std::runtime_error("surfker::surfdisp: periods_s is empty");
  if (req.mode < 1) throw std::runtime_error("surfker::surfdisp: mode must be >= 1");
  if (req.iwave != 1 && req.iwave != 2)
    throw std::runtime_error("surfker::surfdisp: iwave must be 1 (Love) or 2 (Rayleigh)");
  std::vector<float> thk(nlayer), vp(nlayer), vs(nlayer), rho(nlayer);
  for (size_t i = 0; i < nlayer; i++) {
    thk[i] = (float)req.thickness_km[i];
    vp[i] = (float)req.vp_km_s[i];
    vs[i] = (float)req.vs_km_s[i];
    rho[i] = (float)req.rho_g_cm3[i];
  }
  std::vector<double> cg(req.periods_s.size(), 0.0);
  int flag = 0;
  surfatt::check("sat_surfdisp96",
                 sat_surfdisp96(1, (int)nlayer, thk.data(), vp.data(), vs.data(), rho.data(), req.iflsph, req.iwave,
                                req.mode, req.igr, (int)req.periods_s.size(), req.periods_s.data(), cg.data(), &flag,
                                nullptr));
  if (flag) throw std::runtime_error("Dispersion calculation failed");
  return cg;
}

}  // namespace surfker

namespace surfatt {

// SurfGrid::fwdsurf (src/surf_grid.cpp:194-245) for every column at once:
// vs3d (and vp3d, rho3d, or both empty for Brocher's relations) [nx][ny][nz]
// at the depth nodes zgrids; returns the maps in the reference's
// [ix][iy][iper] layout (surf_idx, include/surf_grid.h:63-65).
inline std::vector<double> fwdsurf(int nx, int ny, const std::vector<double>& zgrids, const std::vector<double>& vs3d,
                                   const std::vector<double>& vp3d, const std::vector<double>& rho3d,
                                   const std::vector<double>& periods, int iwave, int itype,
                                   int iflsph = surfker::IFLSPH, int mode = surfker::IMODE) {
  const int nz = (int)zgrids.size(), np = (int)periods.size();
  const size_t n3 = (size_t)nx * ny * nz;
  if (vs3d.size() != n3 || vp3d.size() != rho3d.size() || (!vp3d.empty() && vp3d.size() != n3))
    throw std::invalid_argument("fwdsurf: model arrays must be nx*ny*nz");
  std::vector<double> pm((size_t)np * nx * ny);
  std::vector<int> flags((size_t)nx * ny);
  check("sat_fwdsurf", sat_fwdsurf(nx, ny, nz, zgrids.data(), vs3d.data(), vp3d.empty() ? nullptr : vp3d.data(),
                                   rho3d.empty() ? nullptr : rho3d.data(), np, periods.data(), iflsph, iwave, mode,
                                   itype, pm.data(), flags.data(), nullptr));
  for (int f : flags)
    if (f) throw std::runtime_error("Dispersion calculation failed");
  std::vector<double> svel(pm.size());
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int ip = 0; ip < np; ip++)
        svel[((size_t)ix * ny + iy) * np + ip] = pm[((size_t)ip * ny + iy) * nx + ix];
  return svel;
}

// SurfGrid::build_media_matrix_with_topo (src/surf_grid.cpp:80-145, :184-190)
// after Topography::grid: the gridded surface z (column-major (nx, ny), km),
// the per-period mean velocities and periods -> m11, m12, m22 in the
// reference's [ix][iy][iper] layout, smoothing width
// v(T) * T * wavelen_factor / (2 pi) km (:115-118).
inline void topo_media(const std::vector<double>& xgrids, const std::vector<double>& ygrids,
                       const std::vector<double>& z, const std::vector<double>& avg_svel,
                       const std::vector<double>& periods, double wavelen_factor, std::vector<double>& m11,
                       std::vector<double>& m12, std::vector<double>& m22) {
  const int nx = (int)xgrids.size(), ny = (int)ygrids.size(), np = (int)periods.size();
  if (z.size() != (size_t)nx * ny || avg_svel.size() != periods.size())
    throw std::invalid_argument("topo_media: z must be nx*ny and avg_svel one value per period");
  const double pi = 3.14159265358979323846264338327950288, deg = 6371.0 * (pi / 180.0);
  std::vector<double> sig(np);
  for (int p = 0; p < np; p++) sig[p] = (avg_svel[p] * periods[p] * wavelen_factor / (2.0 * pi)) / deg;
  const size_t n3 = (size_t)np * nx * ny;
  std::vector<double> a(n3), b(n3), c(n3);
  check("sat_topo_metric", sat_topo_metric(nx, ny, np, xgrids.data(), ygrids.data(), z.data(), sig.data(), a.data(),
                                           c.data(), b.data(), nullptr, nullptr));
  m11.resize(n3);
  m12.resize(n3);
  m22.resize(n3);
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int ip = 0; ip < np; ip++) {
        const size_t src = ((size_t)ip * ny + iy) * nx + ix, dst = ((size_t)ix * ny + iy) * np + ip;
        m11[dst] = a[src];
        m12[dst] = c[src];
        m22[dst] = b[src];
      }
}

}  // namespace surfatt
