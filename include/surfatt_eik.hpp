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
//   preproc::forward_for_event     the same call surface over light views of the
//                                  reference's SrcRec (include/src_rec.h:15-22) and
//                                  SurfGrid (include/surf_grid.h:63-65) members
//
// Fields are surfatt::Field: an owning column-major nx-by-ny matrix of double
// (element (ix, iy) at ix + iy*nx), the storage of Eigen::MatrixX<real_t>, with
// rows() == nx and cols() == ny like the reference's matrices.  A grid smaller
// than 3x3 throws std::invalid_argument (src/eikonal_solver.cpp:35-36); other
// failures throw std::runtime_error naming the status.  Link with
// -lsurfatt_eik.
#pragma once
#include <cstddef>
#include <map>
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

namespace preproc {

// The members of the reference's event_info (include/src_rec.h:15-22) that
// forward_for_event reads and writes.
struct event_info {
  double evla = 0.0, evlo = 0.0, period = 0.0;
  int iper = 0;
  std::vector<int> rec_indices;  // into SrcRecView::stlo/stla/tt/weight
  std::vector<double> syn_data;  // written: synthetic travel time per receiver
};

// The members of SrcRec that forward_for_event uses: per-record station
// coordinates, observed times and weights, and the events in the order of the
// reference's events_local map (keys "{evt}_{iper}", src/src_rec.cpp:149-171).
struct SrcRecView {
  std::vector<double> stla, stlo, tt, weight;
  std::map<std::string, event_info> events_local;
};

// The members of SurfGrid that forward_for_event uses: grid axes (degrees) and
// the per-period planes in the reference's [ix][iy][iper] layout
// (surf_idx = ((ix*ny) + iy)*nperiod + iper, include/surf_grid.h:63-65), and the
// per-period kernel accumulators adj_s_local (nx-by-ny each, added to).
struct SurfGridView {
  std::vector<double> xgrids, ygrids;
  int nperiod = 0;
  const double *svel = nullptr, *m11 = nullptr, *m12 = nullptr, *m22 = nullptr;
  std::vector<surfatt::Field> adj_s_local;
};

// src/preproc.cpp:71-148 for every event of sr in one batched GPU run: the
// forward field, syn_data, and with is_calc_adj the adjoint, mask, kernel
// accumulation adj_s_local[iper] += Ta .* svel^-3 (in event order) and chi,
// which is returned (0 without is_calc_adj, as in the reference).  Precision:
// surfatt::precision().  The period planes are transposed to the library's
// period-major layout once per call.
inline double forward_for_event(SrcRecView& sr, SurfGridView& sg, bool is_calc_adj) {
  const int nx = (int)sg.xgrids.size(), ny = (int)sg.ygrids.size(), np = sg.nperiod;
  if (nx < 3 || ny < 3) throw std::invalid_argument("Grid must be at least 3x3");
  if (np < 1 || !sg.svel || !sg.m11 || !sg.m12 || !sg.m22) throw std::invalid_argument("no period planes");
  const size_t N = (size_t)nx * ny;
  std::vector<double> planes[4];
  const double* src[4] = {sg.svel, sg.m11, sg.m12, sg.m22};
  for (int k = 0; k < 4; k++) {
    planes[k].resize((size_t)np * N);
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++)
        for (int ip = 0; ip < np; ip++)
          planes[k][(size_t)ip * N + ix + (size_t)iy * nx] = src[k][((size_t)ix * ny + iy) * np + ip];
  }
  sat_grid_desc g{nx, ny, np, sg.xgrids.data(), sg.ygrids.data(),
                  planes[0].data(), planes[1].data(), planes[2].data(), planes[3].data()};
  std::vector<sat_event_desc> ev;
  std::vector<double> rlon, rlat, tt, w;
  for (auto& kv : sr.events_local) {
    const event_info& e = kv.second;
    if (e.iper < 0 || e.iper >= np) throw std::invalid_argument("event period out of range");
    ev.push_back(sat_event_desc{0, e.iper, e.evlo, e.evla, (int)rlon.size(), (int)e.rec_indices.size()});
    for (int r : e.rec_indices) {
      if (r < 0 || r >= (int)sr.stlo.size()) throw std::invalid_argument("receiver index out of range");
      rlon.push_back(sr.stlo[r]);
      rlat.push_back(sr.stla[r]);
      tt.push_back(sr.tt[r]);
      w.push_back(sr.weight[r]);
    }
  }
  sat_batch_desc d{};
  d.ngrid = 1;
  d.grids = &g;
  d.nevent = (int)ev.size();
  d.events = ev.data();
  d.nrec_total = (int)rlon.size();
  d.rec_lon_deg = rlon.data();
  d.rec_lat_deg = rlat.data();
  d.tt_obs = tt.data();
  d.weight = w.data();
  d.precision = surfatt::precision();
  d.calc_adj = is_calc_adj ? 1 : 0;
  surfatt::Batch batch(d);
  batch.run();
  std::vector<double> syn(rlon.size() + 1), K((size_t)np * N);
  const double chi = batch.results(syn.data(), K.data());
  size_t off = 0;
  for (auto& kv : sr.events_local) {
    event_info& e = kv.second;
    e.syn_data.assign(syn.begin() + off, syn.begin() + off + e.rec_indices.size());
    off += e.rec_indices.size();
  }
  if (is_calc_adj) {
    if ((int)sg.adj_s_local.size() != np) sg.adj_s_local.assign(np, surfatt::Field(nx, ny));
    for (int ip = 0; ip < np; ip++) {
      surfatt::check_shape(sg.adj_s_local[ip], nx, ny);
      double* a = sg.adj_s_local[ip].data();
      for (size_t n = 0; n < N; n++) a[n] += K[(size_t)ip * N + n];
    }
  }
  return chi;
}

}  // namespace preproc
