// surfatt_eik.hpp -- C++20 shim over the C ABI (surfatt_eik.h) with the
// reference's names, argument order and error behaviour, for call sites that
// use namespace eikonal (include/eikonal_solver.h) and preproc
// (include/preproc.h) of SurfATTPP.
//
//   eikonal::FSM_UW_PS_lonlat_2d   include/eikonal_solver.h:32-40
//   eikonal::FSM_O1_JSE_lonlat_2d  include/eikonal_solver.h:74-83
//   eikonal::mask_uniform_grid     include/eikonal_solver.h:85-89
//   surfatt::forward_for_event     preproc::forward_for_event (include/preproc.h:14,
//                                  src/preproc.cpp:71-148) over the caller's own
//                                  SrcRec / SurfGrid types, batched on the GPU
//   surfatt::Batch                 owning wrapper of a sat_batch
//
// The eikonal functions are templates over the caller's vector and matrix
// types, so Eigen::VectorX<double> / Eigen::MatrixX<double> and the
// expressions the reference passes (a RowMajor matrix, -m12_ij) bind without
// this header including Eigen:
//   - a vector is anything with size() and data(), or size() and operator[];
//   - a matrix is anything with rows(), cols() and either data() (read as
//     column-major unless the type says IsRowMajor, and only when
//     innerStride() == 1 and outerStride() == rows() where those exist) or
//     element access m(i, j); rows() == nx, cols() == ny as in the reference.
// Column-major double storage is passed to the library without a copy; any
// other matrix is copied once.  Results are surfatt::Field (an owning
// column-major nx-by-ny matrix), which converts implicitly to any matrix type
// constructible as M(rows, cols) with element access, so
//   Eigen::MatrixX<real_t> tfield = eikonal::FSM_UW_PS_lonlat_2d(...);
// compiles as written in the reference.  A grid smaller than 3x3 throws
// std::invalid_argument (src/eikonal_solver.cpp:35-36); other failures throw
// std::runtime_error naming the status.  Link with -lsurfatt_eik.
#pragma once
#include <algorithm>
#include <concepts>
#include <cstddef>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "surfatt_eik.h"

namespace surfatt {

inline void check(const char* fn, int rc) {
  if (rc == SAT_OK) return;
  const char* name = rc == SAT_EINVAL ? "SAT_EINVAL" : rc == SAT_ENOMEM ? "SAT_ENOMEM"
                   : rc == SAT_EDEVICE ? "SAT_EDEVICE" : "unknown";
  throw std::runtime_error(std::string(fn) + " failed: " + name);
}

namespace detail {

template <class M>
constexpr bool row_major() {
  if constexpr (requires { M::IsRowMajor; }) return static_cast<bool>(M::IsRowMajor);
  else return false;
}

template <class M>
concept Matrix = requires(const M& m) {
  m.rows();
  m.cols();
};

template <class M>
concept ColMajorData = Matrix<M> && !row_major<M>() && requires(const M& m) {
  { m.data() } -> std::convertible_to<const double*>;
};

template <class M>
concept MutableColMajorData = ColMajorData<M> && requires(M& m) {
  { m.data() } -> std::convertible_to<double*>;
};

template <class M>
concept Elements = Matrix<M> && requires(const M& m) {
  { m(0, 0) } -> std::convertible_to<double>;
};

template <class M>
bool dense_columns(const M& m) {
  if constexpr (requires { m.innerStride(); m.outerStride(); })
    return m.innerStride() == 1 && (long)m.outerStride() == (long)m.rows();
  else
    return true;
}

template <class V>
concept Vector = requires(const V& v) { v.size(); } &&
                 (requires(const V& v) { { v.data() } -> std::convertible_to<const double*>; } ||
                  requires(const V& v) { { v[0] } -> std::convertible_to<double>; });

// A read-only column-major nx-by-ny image of the caller's matrix: its own
// storage when that already is column-major doubles, else a copy.
class MatIn {
 public:
  template <class M>
  MatIn(const M& m, int nx, int ny) {
    static_assert(Matrix<M>, "a matrix needs rows() and cols()");
    if ((long)m.rows() != nx || (long)m.cols() != ny)
      throw std::invalid_argument("field shape does not match the grid");
    if constexpr (ColMajorData<M>) {
      if (dense_columns(m)) {
        p_ = m.data();
        return;
      }
    }
    if constexpr (Elements<M>) {
      own_.resize((size_t)nx * ny);
      for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) own_[(size_t)ix + (size_t)iy * nx] = static_cast<double>(m(ix, iy));
      p_ = own_.data();
    } else {
      throw std::invalid_argument("matrix storage is neither dense column-major nor element-addressable");
    }
  }
  const double* data() const { return p_; }

 private:
  const double* p_ = nullptr;
  std::vector<double> own_;
};

// A read-only contiguous image of the caller's vector.
class VecIn {
 public:
  template <class V>
  explicit VecIn(const V& v) {
    static_assert(Vector<V>, "a vector needs size() and data() or operator[]");
    n_ = (int)v.size();
    if constexpr (requires { { v.data() } -> std::convertible_to<const double*>; }) {
      bool dense = true;
      if constexpr (requires { v.innerStride(); }) dense = v.innerStride() == 1;
      if (dense) {
        p_ = v.data();
        return;
      }
    }
    if constexpr (requires { { v[0] } -> std::convertible_to<double>; }) {
      own_.resize(n_);
      for (int i = 0; i < n_; i++) own_[i] = static_cast<double>(v[i]);
      p_ = own_.data();
    } else {
      throw std::invalid_argument("vector storage is not contiguous");
    }
  }
  int size() const { return n_; }
  const double* data() const { return p_; }

 private:
  int n_ = 0;
  const double* p_ = nullptr;
  std::vector<double> own_;
};

// Element (ix, iy) of a caller's accumulator matrix, writable.
template <class M>
void add_plane(M& acc, const double* plane, int nx, int ny) {
  if ((long)acc.rows() != nx || (long)acc.cols() != ny) {
    if constexpr (requires { acc.resize(nx, ny); }) {
      if (acc.rows() == 0 || acc.cols() == 0) {
        acc.resize(nx, ny);
        for (int iy = 0; iy < ny; iy++)
          for (int ix = 0; ix < nx; ix++) acc(ix, iy) = 0.0;
      } else {
        throw std::invalid_argument("kernel accumulator shape does not match the grid");
      }
    } else {
      throw std::invalid_argument("kernel accumulator shape does not match the grid");
    }
  }
  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix < nx; ix++) acc(ix, iy) += plane[(size_t)ix + (size_t)iy * nx];
}

}  // namespace detail

// Owning column-major nx-by-ny matrix of double (element (ix, iy) at
// ix + iy*nx), the storage of Eigen::MatrixX<real_t>; rows() == nx.
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
  void resize(int nx, int ny) {
    nx_ = nx;
    ny_ = ny;
    d_.assign((size_t)nx * ny, 0.0);
  }

  // Into the caller's matrix type (Eigen::MatrixX<double>, a RowMajor matrix, ...).
  template <class M>
    requires(!std::is_same_v<M, Field> && detail::Matrix<M> &&
             requires(M m) {
               M(1, 1);
               m(0, 0) = 0.0;
             })
  operator M() const {
    M m(nx_, ny_);
    if constexpr (detail::MutableColMajorData<M>) {
      if (detail::dense_columns(m)) {
        std::copy(d_.begin(), d_.end(), m.data());
        return m;
      }
    }
    for (int iy = 0; iy < ny_; iy++)
      for (int ix = 0; ix < nx_; ix++) m(ix, iy) = (*this)(ix, iy);
    return m;
  }

 private:
  int nx_ = 0, ny_ = 0;
  std::vector<double> d_;
};

inline void check_shape(const Field& f, int nx, int ny) {
  if (f.rows() != nx || f.cols() != ny) throw std::invalid_argument("field shape does not match the grid");
}

// Precision of the per-solve calls below and the default of forward_for_event
// (SAT_PREC_F64 reproduces the reference).
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
  void set_kden(bool on) { check("sat_batch_set_kden", sat_batch_set_kden(b_, on ? 1 : 0)); }
  void kden_results(double* kden_s) { check("sat_batch_kden_results", sat_batch_kden_results(b_, kden_s)); }
  // a, b, c: [nperiod][ny][nx]; all null unregisters
  void set_aniso_metric(int grid, const double* a, const double* b, const double* c) {
    check("sat_batch_set_aniso_metric", sat_batch_set_aniso_metric(b_, grid, a, b, c));
  }
  void aniso_results(int grid, double* adj_xi, double* adj_eta) {
    check("sat_batch_aniso_kernels", sat_batch_aniso_kernels(b_, grid, nullptr, nullptr, nullptr, adj_xi, adj_eta));
  }
  sat_batch* handle() { return b_; }

 private:
  sat_batch* b_ = nullptr;
};

}  // namespace surfatt

namespace eikonal {

using surfatt::Field;

// Travel time from a point source (degrees), src/eikonal_solver.cpp:22-318.
template <class VX, class VY, class MA, class MB, class MC, class MF>
Field FSM_UW_PS_lonlat_2d(const VX& xx_deg, const VY& yy_deg, const MA& spha, const MB& sphb, const MC& sphc,
                          const MF& fun, double x0_deg, double y0_deg) {
  const surfatt::detail::VecIn xx(xx_deg), yy(yy_deg);
  const int nx = xx.size(), ny = yy.size();
  if (nx < 3 || ny < 3) throw std::invalid_argument("Grid must be at least 3x3");
  const surfatt::detail::MatIn a(spha, nx, ny), b(sphb, nx, ny), c(sphc, nx, ny), f(fun, nx, ny);
  Field T(nx, ny);
  surfatt::check("sat_eik_forward",
                 sat_eik_forward(surfatt::precision(), nx, ny, xx.data(), yy.data(), a.data(), b.data(), c.data(),
                                 f.data(), x0_deg, y0_deg, T.data(), &surfatt::last_info()));
  return T;
}

// Adjoint field for receiver adjoint sources, src/eikonal_solver.cpp:323-512.
template <class VX, class VY, class MA, class MB, class MC, class MT, class VR, class VS, class VQ>
Field FSM_O1_JSE_lonlat_2d(const VX& xx_deg, const VY& yy_deg, const MA& spha, const MB& sphb, const MC& sphc,
                           const MT& T, const VR& xrec_deg, const VS& yrec_deg, const VQ& sourceAdj) {
  const surfatt::detail::VecIn xx(xx_deg), yy(yy_deg), xr(xrec_deg), yr(yrec_deg), sa(sourceAdj);
  const int nx = xx.size(), ny = yy.size();
  if (nx < 3 || ny < 3) throw std::invalid_argument("Grid must be at least 3x3");
  const surfatt::detail::MatIn a(spha, nx, ny), b(sphb, nx, ny), c(sphc, nx, ny), t(T, nx, ny);
  const int nr = xr.size();
  if (yr.size() != nr || sa.size() != nr) throw std::invalid_argument("receiver arrays differ in length");
  Field Ta(nx, ny);
  surfatt::check("sat_eik_backward",
                 sat_eik_backward(surfatt::precision(), nx, ny, xx.data(), yy.data(), a.data(), b.data(), c.data(),
                                  t.data(), nr, xr.data(), yr.data(), sa.data(), Ta.data(), &surfatt::last_info()));
  return Ta;
}

// Zero Ta within 1.5 cells of (x0, y0), src/eikonal_solver.cpp:514-536, in the
// caller's matrix (in place when it is dense column-major doubles).
template <class VX, class VY, class M>
void mask_uniform_grid(const VX& xx_deg, const VY& yy_deg, M& Ta, double x0, double y0) {
  const surfatt::detail::VecIn xx(xx_deg), yy(yy_deg);
  const int nx = xx.size(), ny = yy.size();
  if ((long)Ta.rows() != nx || (long)Ta.cols() != ny) throw std::invalid_argument("field shape does not match the grid");
  if constexpr (surfatt::detail::MutableColMajorData<M>) {
    if (surfatt::detail::dense_columns(Ta)) {
      surfatt::check("sat_eik_mask", sat_eik_mask(nx, ny, xx.data(), yy.data(), Ta.data(), x0, y0));
      return;
    }
  }
  Field tmp(nx, ny);
  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix < nx; ix++) tmp(ix, iy) = Ta(ix, iy);
  surfatt::check("sat_eik_mask", sat_eik_mask(nx, ny, xx.data(), yy.data(), tmp.data(), x0, y0));
  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix < nx; ix++) Ta(ix, iy) = tmp(ix, iy);
}

}  // namespace eikonal

namespace surfatt {

// The two InputParams switches forward_for_event reads (src/preproc.cpp:37,
// :127), and the batch's precision and residency.
struct forward_options {
  bool is_kden = false;        // IP.postproc().is_kden: kden_s_local too
  bool is_anisotropy = false;  // IP.inversion().is_anisotropy: adj_xi_local, adj_eta_local too
  int precision = -1;          // SAT_PREC_*; -1: surfatt::precision()
  bool keep_resident = true;   // keep the device batch between calls on the same (sr, sg)
};

namespace detail {

template <class SG>
int nperiod_of(const SG& sg) {
  if constexpr (requires { sg.nperiod(); }) return (int)sg.nperiod();
  else return (int)sg.nperiod;
}

// [ix][iy][iper] (SurfGrid::surf_idx, include/surf_grid.h:63-65) -> [iper][iy][ix]
template <class P>
void period_major(const P& src, int nx, int ny, int np, double sign, std::vector<double>& out) {
  const size_t N = (size_t)nx * ny;
  out.resize((size_t)np * N);
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int ip = 0; ip < np; ip++)
        out[(size_t)ip * N + ix + (size_t)iy * nx] = sign * static_cast<double>(src[((size_t)ix * ny + iy) * np + ip]);
}

// What a resident batch was built for: a call with the same layout reuses it
// (new planes and observations are uploaded in place).
struct Layout {
  int nx = 0, ny = 0, np = 0, precision = 0, calc_adj = 0;
  std::vector<double> axes;
  std::vector<int> ev_iper, ev_nrec;
  std::vector<double> ev_src, rec_pos;
  bool operator==(const Layout&) const = default;
};

struct Resident {
  Layout layout;
  std::unique_ptr<Batch> batch;
  bool kden = false, aniso = false;
};

inline std::map<std::pair<const void*, const void*>, Resident>& residents() {
  static thread_local std::map<std::pair<const void*, const void*>, Resident> r;
  return r;
}

}  // namespace detail

// Free every device batch kept by forward_for_event on this thread (before the
// device or the process goes away).
inline void release_resident_batches() { detail::residents().clear(); }

// preproc::forward_for_event (src/preproc.cpp:71-148) for every event of
// sr.events_local in one batched GPU run, over the caller's own types:
//   sr.events_local   a map of event_info (evla, evlo, iper, rec_indices,
//                     syn_data: written, resized to the event's receivers)
//   sr.stlo, sr.stla, sr.tt, sr.weight   per-record arrays (pointers or vectors)
//   sg.svel, sg.m11, sg.m12, sg.m22       [ix][iy][iper] planes (surf_idx)
//   sg.nperiod() (or an int member nperiod)
//   sg.adj_s_local[iper]                  nx-by-ny accumulators, added to
//   with is_kden:       sg.kden_s_local[iper] (added to, :32-34)
//   with is_anisotropy: sg.a, sg.b, sg.c ([ix][iy][iper]) and
//                       sg.adj_xi_local, sg.adj_eta_local (added to, :37-65)
// xgrids, ygrids: ModelGrid::MG().xgrids / ygrids.  Returns this call's chi
// (0 without is_calc_adj, as in the reference); the reference's MPI barrier and
// chi sum (:145-146) stay with the caller.  The accumulators receive the
// batch's per-period sums, so they equal the reference's bit for bit when they
// start at zero (reset_kernel_accumulators).  With keep_resident the device
// batch of this (sr, sg) is kept and reused while the grid, the events and
// their receivers stay the same: later calls upload only the planes and the
// observations.
template <class SR, class SG, class VX, class VY>
double forward_for_event(SR& sr, SG& sg, const VX& xgrids, const VY& ygrids, bool is_calc_adj,
                         const forward_options& o = {}) {
  const detail::VecIn xx(xgrids), yy(ygrids);
  const int nx = xx.size(), ny = yy.size(), np = detail::nperiod_of(sg);
  if (nx < 3 || ny < 3) throw std::invalid_argument("Grid must be at least 3x3");
  if (np < 1) throw std::invalid_argument("no period planes");
  if ((o.is_kden || o.is_anisotropy) && !is_calc_adj)
    throw std::invalid_argument("the density and anisotropic kernels need is_calc_adj");
  constexpr bool has_kden = requires { sg.kden_s_local[0]; };
  constexpr bool has_aniso = requires { sg.a[0]; sg.b[0]; sg.c[0]; sg.adj_xi_local[0]; sg.adj_eta_local[0]; };
  if (o.is_kden && !has_kden) throw std::invalid_argument("is_kden needs sg.kden_s_local");
  if (o.is_anisotropy && !has_aniso)
    throw std::invalid_argument("is_anisotropy needs sg.a, sg.b, sg.c, sg.adj_xi_local, sg.adj_eta_local");
  const size_t N = (size_t)nx * ny;
  std::vector<double> planes[4];
  detail::period_major(sg.svel, nx, ny, np, 1.0, planes[0]);
  detail::period_major(sg.m11, nx, ny, np, 1.0, planes[1]);
  detail::period_major(sg.m12, nx, ny, np, 1.0, planes[2]);
  detail::period_major(sg.m22, nx, ny, np, 1.0, planes[3]);

  detail::Layout lay;
  lay.nx = nx;
  lay.ny = ny;
  lay.np = np;
  lay.precision = o.precision >= 0 ? o.precision : precision();
  lay.calc_adj = is_calc_adj ? 1 : 0;
  lay.axes.assign(xx.data(), xx.data() + nx);
  lay.axes.insert(lay.axes.end(), yy.data(), yy.data() + ny);
  std::vector<sat_event_desc> ev;
  std::vector<double> rlon, rlat, tt, w;
  for (auto& kv : sr.events_local) {
    const auto& e = kv.second;
    if (e.iper < 0 || e.iper >= np) throw std::invalid_argument("event period out of range");
    ev.push_back(sat_event_desc{0, e.iper, (double)e.evlo, (double)e.evla, (int)rlon.size(), (int)e.rec_indices.size()});
    lay.ev_iper.push_back(e.iper);
    lay.ev_nrec.push_back((int)e.rec_indices.size());
    lay.ev_src.push_back((double)e.evlo);
    lay.ev_src.push_back((double)e.evla);
    for (int r : e.rec_indices) {
      if (r < 0) throw std::invalid_argument("receiver index out of range");
      if constexpr (requires { sr.stlo.size(); })
        if ((size_t)r >= (size_t)sr.stlo.size()) throw std::invalid_argument("receiver index out of range");
      rlon.push_back((double)sr.stlo[r]);
      rlat.push_back((double)sr.stla[r]);
      tt.push_back((double)sr.tt[r]);
      w.push_back((double)sr.weight[r]);
    }
  }
  lay.rec_pos = rlon;
  lay.rec_pos.insert(lay.rec_pos.end(), rlat.begin(), rlat.end());

  const auto key = std::make_pair((const void*)&sr, (const void*)&sg);
  detail::Resident fresh;
  detail::Resident* res = &fresh;
  if (o.keep_resident) {
    res = &detail::residents()[key];
    if (res->batch && !(res->layout == lay)) res->batch.reset();
  }
  if (!res->batch) {
    sat_grid_desc g{nx, ny, np, xx.data(), yy.data(), planes[0].data(), planes[1].data(), planes[2].data(),
                    planes[3].data(), 0};
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
    d.precision = lay.precision;
    d.calc_adj = lay.calc_adj;
    res->batch = std::make_unique<Batch>(d);
    res->layout = std::move(lay);
    res->kden = res->aniso = false;
  } else {
    res->batch->update_media(0, planes[0].data(), planes[1].data(), planes[2].data(), planes[3].data());
    res->batch->update_obs(tt.data(), w.data());
  }
  Batch& batch = *res->batch;
  if (res->kden != o.is_kden) {
    batch.set_kden(o.is_kden);
    res->kden = o.is_kden;
  }
  if (o.is_anisotropy) {
    if constexpr (has_aniso) {
      std::vector<double> a, b, c;
      detail::period_major(sg.a, nx, ny, np, 1.0, a);
      detail::period_major(sg.b, nx, ny, np, 1.0, b);
      detail::period_major(sg.c, nx, ny, np, 1.0, c);
      batch.set_aniso_metric(0, a.data(), b.data(), c.data());
      res->aniso = true;
    }
  } else if (res->aniso) {
    batch.set_aniso_metric(0, nullptr, nullptr, nullptr);
    res->aniso = false;
  }
  batch.run();
  std::vector<double> syn(rlon.size() + 1), K(is_calc_adj ? (size_t)np * N : 1);
  const double chi = batch.results(syn.data(), is_calc_adj ? K.data() : nullptr);
  size_t off = 0;
  for (auto& kv : sr.events_local) {
    auto& e = kv.second;
    const size_t n = e.rec_indices.size();
    e.syn_data.resize(n);
    for (size_t r = 0; r < n; r++) e.syn_data[r] = syn[off + r];
    off += n;
  }
  if (is_calc_adj) {
    if constexpr (requires { sg.adj_s_local.resize((size_t)np); })
      if (sg.adj_s_local.size() < (size_t)np) sg.adj_s_local.resize(np);
    for (int ip = 0; ip < np; ip++) detail::add_plane(sg.adj_s_local[ip], K.data() + (size_t)ip * N, nx, ny);
    if constexpr (has_kden) {
      if (o.is_kden) {
        batch.kden_results(K.data());
        if constexpr (requires { sg.kden_s_local.resize((size_t)np); })
          if (sg.kden_s_local.size() < (size_t)np) sg.kden_s_local.resize(np);
        for (int ip = 0; ip < np; ip++) detail::add_plane(sg.kden_s_local[ip], K.data() + (size_t)ip * N, nx, ny);
      }
    }
    if constexpr (has_aniso) {
      if (o.is_anisotropy) {
        std::vector<double> xi((size_t)np * N), eta((size_t)np * N);
        batch.aniso_results(0, xi.data(), eta.data());
        if constexpr (requires { sg.adj_xi_local.resize((size_t)np); sg.adj_eta_local.resize((size_t)np); }) {
          if (sg.adj_xi_local.size() < (size_t)np) sg.adj_xi_local.resize(np);
          if (sg.adj_eta_local.size() < (size_t)np) sg.adj_eta_local.resize(np);
        }
        for (int ip = 0; ip < np; ip++) {
          detail::add_plane(sg.adj_xi_local[ip], xi.data() + (size_t)ip * N, nx, ny);
          detail::add_plane(sg.adj_eta_local[ip], eta.data() + (size_t)ip * N, nx, ny);
        }
      }
    }
  }
  return chi;
}

}  // namespace surfatt

namespace preproc {

// Light stand-ins of the reference's event_info (include/src_rec.h:15-22),
// SrcRec and SurfGrid members that forward_for_event reads and writes, for
// callers without those classes; surfatt::forward_for_event takes the
// reference's own classes as they are.
struct event_info {
  double evla = 0.0, evlo = 0.0, period = 0.0;
  int iper = 0;
  std::vector<int> rec_indices;  // into SrcRecView::stlo/stla/tt/weight
  std::vector<double> syn_data;  // written: synthetic travel time per receiver
};

struct SrcRecView {
  std::vector<double> stla, stlo, tt, weight;
  std::map<std::string, event_info> events_local;  // keys "{evt}_{iper}" (src/src_rec.cpp:149-171)
};

struct SurfGridView {
  std::vector<double> xgrids, ygrids;
  int nperiod = 0;
  const double *svel = nullptr, *m11 = nullptr, *m12 = nullptr, *m22 = nullptr;
  std::vector<surfatt::Field> adj_s_local;
};

// src/preproc.cpp:71-148 on the views: surfatt::forward_for_event with the
// view's own axes.
inline double forward_for_event(SrcRecView& sr, SurfGridView& sg, bool is_calc_adj) {
  if (!sg.svel || !sg.m11 || !sg.m12 || !sg.m22) throw std::invalid_argument("no period planes");
  return surfatt::forward_for_event(sr, sg, sg.xgrids, sg.ygrids, is_calc_adj);
}

}  // namespace preproc
