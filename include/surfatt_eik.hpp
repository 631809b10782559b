// surfatt_eik.hpp -- C++20 shim over the C ABI (surfatt_eik.h) with the
// reference's names, argument order and error behaviour, for call sites that
// use namespace eikonal (include/eikonal_solver.h) and preproc
// (include/preproc.h) of SurfATTPP.
//
//   eikonal::FSM_UW_PS_lonlat_2d   include/eikonal_solver.h:32-40
//   eikonal::FSM_O1_JSE_lonlat_2d  include/eikonal_solver.h:74-83
//   eikonal::mask_uniform_grid     include/eikonal_solver.h:85-89
//   surfatt::Batch                 the batched form of preproc::forward_for_event
//                                  (include/preproc.h:14, src/preproc.cpp:71-148)
//
// Fields are surfatt::Field: an owning column-major nx-by-ny matrix of double
// (element (ix, iy) at ix + iy*nx), the storage of Eigen::MatrixX<real_t>, with
// rows() == nx and cols() == ny like the reference's matrices.  A grid smaller
// than 3x3 throws std::invalid_argument (src/eikonal_solver.cpp:35-36); other
// failures throw std::runtime_error naming the status.  Link with
// -lsurfatt_eik.
#pragma once
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "surfatt_eik.h"

namespace surfatt {

class Field {
 public:
  Field() = default;
  Field(int nx, int ny, double v = 0.0) : nx_(nx), ny_(ny), d_((size_t)nx * ny, v) {}
  int rows() const { return nx_; }
  int cols() const { return ny_; }
  double& operator()(int ix, int iy) { return d_[(size_t)ix + (size_t)iy * nx_]; }
  double operator()(int ix, int iy) const { return d_[(size_t)ix + (size_t)iy * nx_]; }
  double* data() { return d_.data(); }
  const double* data() const { return d_.data(); }
  size_t size() const { return d_.size(); }

 private:
  int nx_ = 0, ny_ = 0;
  std::vector<double> d_;
};

inline void check(const char* fn, int rc) {
  if (rc == SAT_OK) return;
  const char* name = rc == SAT_EINVAL ? "SAT_EINVAL" : rc == SAT_ENOMEM ? "SAT_ENOMEM"
                   : rc == SAT_EDEVICE ? "SAT_EDEVICE" : "unknown";
  throw std::runtime_error(std::string(fn) + " failed: " + name);
}

inline void check_shape(const Field& f, int nx, int ny) {
  if (f.rows() != nx || f.cols() != ny) throw std::invalid_argument("field shape does not match the grid");
}

// Precision of the per-solve calls below (SAT_PREC_F64 reproduces the reference).
inline int& precision() {
  static thread_local int p = SAT_PREC_F64;
  return p;
}

// Iterations and flags of the last per-solve call on this thread.
inline sat_info& last_info() {
  static thread_local sat_info info{};
  return info;
}

// Owning wrapper of a sat_batch: every (source, period) pair of an inversion
// iteration, resident on the GPU.  run() queues the whole pipeline; results()
// synchronises and copies out.
class Batch {
 public:
  explicit Batch(const sat_batch_desc& d) { check("sat_batch_create", sat_batch_create(&d, &b_)); }
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;
  ~Batch() { sat_batch_destroy(b_); }
  void run(int mode = SAT_RUN_BATCHED) { check("sat_batch_run", sat_batch_run(b_, mode)); }
  void sync() { check("sat_batch_sync", sat_batch_sync(b_)); }
  void update_media(int grid, const double* svel, const double* m11, const double* m12, const double* m22) {
    check("sat_batch_update_media", sat_batch_update_media(b_, grid, svel, m11, m12, m22));
  }
  void update_obs(const double* tt_obs, const double* weight) {
    check("sat_batch_update_obs", sat_batch_update_obs(b_, tt_obs, weight));
  }
  // syn[nrec_total], adj_s (every grid's [nperiod][ny][nx] in order); returns chi.
  double results(double* syn, double* adj_s, int* iters_fwd = nullptr, int* iters_adj = nullptr,
                 int* flags = nullptr) {
    double chi = 0.0;
    check("sat_batch_results", sat_batch_results(b_, syn, adj_s, &chi, iters_fwd, iters_adj, flags));
    return chi;
  }
  sat_batch* handle() { return b_; }

 private:
  sat_batch* b_ = nullptr;
};

}  // namespace surfatt

namespace eikonal {

using surfatt::Field;

// Travel time from a point source (degrees), src/eikonal_solver.cpp:22-318.
inline Field FSM_UW_PS_lonlat_2d(const std::vector<double>& xx_deg, const std::vector<double>& yy_deg,
                                 const Field& spha, const Field& sphb, const Field& sphc, const Field& fun,
                                 double x0_deg, double y0_deg) {
  const int nx = (int)xx_deg.size(), ny = (int)yy_deg.size();
  if (nx < 3 || ny < 3) throw std::invalid_argument("Grid must be at least 3x3");
  for (const Field* f : {&spha, &sphb, &sphc, &fun}) surfatt::check_shape(*f, nx, ny);
  Field T(nx, ny);
  surfatt::check("sat_eik_forward",
                 sat_eik_forward(surfatt::precision(), nx, ny, xx_deg.data(), yy_deg.data(), spha.data(),
                                 sphb.data(), sphc.data(), fun.data(), x0_deg, y0_deg, T.data(),
                                 &surfatt::last_info()));
  return T;
}

// Adjoint field for receiver adjoint sources, src/eikonal_solver.cpp:323-512.
inline Field FSM_O1_JSE_lonlat_2d(const std::vector<double>& xx_deg, const std::vector<double>& yy_deg,
                                  const Field& spha, const Field& sphb, const Field& sphc, const Field& T,
                                  const std::vector<double>& xrec_deg, const std::vector<double>& yrec_deg,
                                  const std::vector<double>& sourceAdj) {
  const int nx = (int)xx_deg.size(), ny = (int)yy_deg.size();
  if (nx < 3 || ny < 3) throw std::invalid_argument("Grid must be at least 3x3");
  for (const Field* f : {&spha, &sphb, &sphc, &T}) surfatt::check_shape(*f, nx, ny);
  const int nr = (int)xrec_deg.size();
  if ((int)yrec_deg.size() != nr || (int)sourceAdj.size() != nr)
    throw std::invalid_argument("receiver arrays differ in length");
  Field Ta(nx, ny);
  surfatt::check("sat_eik_backward",
                 sat_eik_backward(surfatt::precision(), nx, ny, xx_deg.data(), yy_deg.data(), spha.data(),
                                  sphb.data(), sphc.data(), T.data(), nr, xrec_deg.data(), yrec_deg.data(),
                                  sourceAdj.data(), Ta.data(), &surfatt::last_info()));
  return Ta;
}

// Zero Ta within 1.5 cells of (x0, y0), src/eikonal_solver.cpp:514-536.
inline void mask_uniform_grid(const std::vector<double>& xx, const std::vector<double>& yy, Field& Ta, double x0,
                              double y0) {
  surfatt::check_shape(Ta, (int)xx.size(), (int)yy.size());
  surfatt::check("sat_eik_mask", sat_eik_mask((int)xx.size(), (int)yy.size(), xx.data(), yy.data(), Ta.data(), x0, y0));
}

}  // namespace eikonal
