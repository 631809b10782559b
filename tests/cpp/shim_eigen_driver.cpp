// Compiles the shim (include/surfatt_eik.hpp) against Eigen-shaped types
// (eigen_standin.hpp) and the reference's own data structures, and runs the
// two ways a maintainer would bind it:
//
//  A. the reference's per-event loop of preproc::forward_for_event
//     (src/preproc.cpp:80-143) with its call expressions as written there --
//     extract_period_ij into RowMajor matrices, svel_ij.cwiseInverse() * _1_CR,
//     -m12_ij as an expression, the eikonal:: calls on Eigen::VectorX /
//     Eigen::MatrixX, the adjoint sources as .array() expressions, the kden
//     branch -- with accumulate_kernels' adj_s / kden_s sums on the host;
//  B. surfatt::forward_for_event(sr, sg, mg.xgrids, mg.ygrids, is_calc_adj,
//     options) on the same SrcRec / SurfGrid objects: one batched GPU run with
//     kden_s, adj_xi and adj_eta, then a second call with a new velocity model
//     (the resident batch reused).
//
// SrcRec, SurfGrid, ModelGrid and event_info below carry the members of the
// reference's classes that these paths touch, with the reference's types
// (include/src_rec.h:15-22, :94-107; include/surf_grid.h:35-48;
// include/model_grid.h:41-42).  The caller-side helpers extract_period_ij
// (src/preproc.cpp:7-17) and interp2d (include/utils.h:358-402) are
// restated for the driver.
//
// Input argv[1] (binary): nx, ny, nper, ne, nrec (int32); xx[nx], yy[ny];
// svel, m11, m12, m22 [nx*ny*nper each, [ix][iy][iper]]; per event iper,
// nrec_e (int32), evlo, evla (double); per receiver stlo, stla, tt, weight.
// Output argv[2]: for A then B then B's second call: chi, syn[nrec] (map
// order), adj_s, kden_s [nper][ny][nx]; B and its second call also adj_xi,
// adj_eta [nper][ny][nx].
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "eigen_standin.hpp"
#include "surfatt_eik.hpp"

using real_t = double;
constexpr real_t _0_CR = 0.0, _1_CR = 1.0, _3_CR = 3.0;
static int ngrid_i = 0, ngrid_j = 0;

struct event_info {
  real_t evla;
  real_t evlo;
  real_t period;
  int iper;
  std::vector<int> rec_indices;
  Eigen::VectorX<real_t> syn_data;
};

struct SrcRec {
  real_t* stla = nullptr;
  real_t* stlo = nullptr;
  real_t* tt = nullptr;
  real_t* weight = nullptr;
  std::map<std::string, event_info> events_local;
};

struct SurfGrid {
  real_t *svel = nullptr, *a = nullptr, *b = nullptr, *c = nullptr;
  real_t *m11 = nullptr, *m12 = nullptr, *m22 = nullptr;
  std::vector<Eigen::MatrixX<real_t>> adj_s_local, adj_xi_local, adj_eta_local, kden_s_local;
  int nperiod_ = 0;
  int nperiod() const { return nperiod_; }
};

struct ModelGrid {
  Eigen::VectorX<real_t> xgrids, ygrids;
};

using MatRM = Eigen::Matrix<real_t, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor>;

// period iper of a [ix][iy][iper] buffer as an ngrid_i x ngrid_j RowMajor matrix
static MatRM extract_period_ij(const real_t* buf, int np, int iper) {
  MatRM m(ngrid_i, ngrid_j);
  for (int ix = 0; ix < ngrid_i; ix++)
    for (int iy = 0; iy < ngrid_j; iy++) m(ix, iy) = buf[((size_t)ix * ngrid_j + iy) * np + iper];
  return m;
}

// bilinear interpolation of z (nx x ny) at (xq, yq): the cell whose lower
// corner is the last grid value <= the query (clamped to the last cell), NaN
// outside the grid
static Eigen::VectorX<real_t> interp2d(const Eigen::VectorX<real_t>& xg, const Eigen::VectorX<real_t>& yg,
                                       const Eigen::MatrixX<real_t>& z, const Eigen::VectorX<real_t>& xq,
                                       const Eigen::VectorX<real_t>& yq) {
  const int nx = (int)xg.size(), ny = (int)yg.size();
  Eigen::VectorX<real_t> out(xq.size());
  for (Eigen::Index k = 0; k < xq.size(); k++) {
    const double qx = xq[k], qy = yq[k];
    if (qx < xg[0] || qx > xg[nx - 1] || qy < yg[0] || qy > yg[ny - 1]) {
      out[k] = std::nan("");
      continue;
    }
    int i = 0, j = 0;
    while (i + 1 < nx && xg[i + 1] <= qx) i++;
    while (j + 1 < ny && yg[j + 1] <= qy) j++;
    i = std::min(i, nx - 2);
    j = std::min(j, ny - 2);
    const double tx = (qx - xg[i]) / (xg[i + 1] - xg[i]);
    const double ty = (qy - yg[j]) / (yg[j + 1] - yg[j]);
    out[k] = (1 - tx) * (1 - ty) * z(i, j) + tx * (1 - ty) * z(i + 1, j) + (1 - tx) * ty * z(i, j + 1) +
             tx * ty * z(i + 1, j + 1);
  }
  return out;
}

// A: the reference's event loop with its call expressions
static real_t forward_for_event_per_solve(SrcRec& sr, SurfGrid& sg, const ModelGrid& mg, bool is_calc_adj,
                                          bool is_kden) {
  real_t chi = _0_CR;
  for (auto& evt : sr.events_local) {
    real_t evla = evt.second.evla;
    real_t evlo = evt.second.evlo;
    int iper = evt.second.iper;
    auto svel_ij = extract_period_ij(sg.svel, sg.nperiod(), iper);
    auto m11_ij = extract_period_ij(sg.m11, sg.nperiod(), iper);
    auto m12_ij = extract_period_ij(sg.m12, sg.nperiod(), iper);
    auto m22_ij = extract_period_ij(sg.m22, sg.nperiod(), iper);
    Eigen::MatrixX<real_t> svel_ij_inv = svel_ij.cwiseInverse();
    Eigen::MatrixX<real_t> svel_ij_inv_scaled = svel_ij_inv * _1_CR;
    Eigen::MatrixX<real_t> tfield = eikonal::FSM_UW_PS_lonlat_2d(
        mg.xgrids, mg.ygrids,
        m11_ij, m22_ij, -m12_ij,
        svel_ij_inv_scaled, evlo, evla
    );
    const int n_rec = (int)evt.second.rec_indices.size();
    Eigen::VectorX<real_t> stlo_rec(n_rec), stla_rec(n_rec), tt_rec(n_rec), weight_rec(n_rec);
    for (int r = 0; r < n_rec; r++) {
      const int i = evt.second.rec_indices[r];
      stlo_rec[r] = sr.stlo[i];
      stla_rec[r] = sr.stla[i];
      tt_rec[r] = sr.tt[i];
      weight_rec[r] = sr.weight[i];
    }
    Eigen::VectorX<real_t> ttsyn = interp2d(mg.xgrids, mg.ygrids, tfield, stlo_rec, stla_rec);
    evt.second.syn_data = ttsyn;
    if (is_calc_adj) {
      Eigen::VectorX<real_t> adjoint_source = weight_rec.array() * (ttsyn - tt_rec).array();
      Eigen::MatrixX<real_t> adj_field = eikonal::FSM_O1_JSE_lonlat_2d(
          mg.xgrids, mg.ygrids,
          m11_ij, m22_ij, -m12_ij,
          tfield, stlo_rec, stla_rec, adjoint_source
      );
      eikonal::mask_uniform_grid(mg.xgrids, mg.ygrids, adj_field, evlo, evla);
      Eigen::MatrixX<real_t> kden_field;
      if (is_kden) {
        Eigen::VectorX<real_t> adjoint_source_kden =
            -weight_rec.array() * Eigen::VectorX<real_t>::Ones(n_rec).array();
        kden_field = eikonal::FSM_O1_JSE_lonlat_2d(
            mg.xgrids, mg.ygrids,
            m11_ij, m22_ij, -m12_ij,
            tfield, stlo_rec, stla_rec, adjoint_source_kden
        );
        eikonal::mask_uniform_grid(mg.xgrids, mg.ygrids, adj_field, evlo, evla);
      }
      // accumulate_kernels' sums (src/preproc.cpp:28-34)
      for (int iy = 0; iy < ngrid_j; iy++)
        for (int ix = 0; ix < ngrid_i; ix++) {
          const double inv3 = std::pow(svel_ij(ix, iy), -_3_CR);
          sg.adj_s_local[iper](ix, iy) += adj_field(ix, iy) * inv3;
          if (is_kden) sg.kden_s_local[iper](ix, iy) += kden_field(ix, iy) * inv3;
        }
      double s = 0.0;
      for (int r = 0; r < n_rec; r++) s += weight_rec[r] * ((ttsyn[r] - tt_rec[r]) * (ttsyn[r] - tt_rec[r]));
      chi += 0.5 * s;
    }
  }
  return chi;
}

// reset_kernel_accumulators: nperiod zero nx-by-ny matrices
static void reset(std::vector<Eigen::MatrixX<real_t>>& v, int np) {
  v.assign(np, Eigen::MatrixX<real_t>(ngrid_i, ngrid_j));
  for (auto& m : v)
    for (int iy = 0; iy < ngrid_j; iy++)
      for (int ix = 0; ix < ngrid_i; ix++) m(ix, iy) = 0.0;
}

static void put(FILE* o, const double* p, size_t n) { std::fwrite(p, sizeof(double), n, o); }

static void put_planes(FILE* o, const std::vector<Eigen::MatrixX<real_t>>& v) {
  for (const auto& m : v)
    for (int iy = 0; iy < ngrid_j; iy++)
      for (int ix = 0; ix < ngrid_i; ix++) {
        const double x = m(ix, iy);
        put(o, &x, 1);
      }
}

static void put_result(FILE* o, real_t chi, SrcRec& sr, SurfGrid& sg, bool aniso) {
  put(o, &chi, 1);
  for (auto& kv : sr.events_local) put(o, kv.second.syn_data.data(), (size_t)kv.second.syn_data.size());
  put_planes(o, sg.adj_s_local);
  put_planes(o, sg.kden_s_local);
  if (aniso) {
    put_planes(o, sg.adj_xi_local);
    put_planes(o, sg.adj_eta_local);
  }
}

template <class T>
static bool rd(FILE* f, T* p, size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int h[5];
  if (!rd(f, h, 5)) return 2;
  const int nx = h[0], ny = h[1], np = h[2], ne = h[3], nrec = h[4];
  ngrid_i = nx;
  ngrid_j = ny;
  const size_t n3 = (size_t)nx * ny * np;
  ModelGrid mg;
  mg.xgrids.resize(nx);
  mg.ygrids.resize(ny);
  std::vector<double> pl[4];
  bool ok = rd(f, mg.xgrids.data(), nx) && rd(f, mg.ygrids.data(), ny);
  for (auto& p : pl) {
    p.resize(n3);
    ok = ok && rd(f, p.data(), n3);
  }
  std::vector<double> mc(n3);  // SurfGrid c = -m12 (src/surf_grid.cpp:186-189)
  for (size_t i = 0; i < n3; i++) mc[i] = -pl[2][i];
  SurfGrid sg;
  sg.nperiod_ = np;
  sg.svel = pl[0].data();
  sg.m11 = sg.a = pl[1].data();
  sg.m12 = pl[2].data();
  sg.m22 = sg.b = pl[3].data();
  sg.c = mc.data();
  SrcRec sr;
  int rec = 0;
  for (int e = 0; e < ne && ok; e++) {
    int ip_n[2];
    double lo_la[2];
    ok = rd(f, ip_n, 2) && rd(f, lo_la, 2);
    char key[32];
    std::snprintf(key, sizeof key, "%08d", e);
    event_info& ev = sr.events_local[key];
    ev.iper = ip_n[0];
    ev.evlo = lo_la[0];
    ev.evla = lo_la[1];
    ev.period = 0.0;
    for (int r = 0; r < ip_n[1]; r++) ev.rec_indices.push_back(rec++);
  }
  std::vector<double> rv[4];
  for (auto& v : rv) {
    v.resize(nrec);
    ok = ok && rd(f, v.data(), nrec);
  }
  std::fclose(f);
  if (!ok || rec != nrec) return 2;
  sr.stlo = rv[0].data();
  sr.stla = rv[1].data();
  sr.tt = rv[2].data();
  sr.weight = rv[3].data();
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  try {
    reset(sg.adj_s_local, np);
    reset(sg.kden_s_local, np);
    const real_t chi_a = forward_for_event_per_solve(sr, sg, mg, true, true);
    put_result(o, chi_a, sr, sg, false);

    surfatt::forward_options opt;
    opt.is_kden = true;
    opt.is_anisotropy = true;
    for (int call = 0; call < 2; call++) {
      if (call == 1)  // the next model of an inversion
        for (size_t i = 0; i < n3; i++) pl[0][i] *= 1.01;
      for (auto* v : {&sg.adj_s_local, &sg.kden_s_local, &sg.adj_xi_local, &sg.adj_eta_local}) reset(*v, np);
      for (auto& kv : sr.events_local) kv.second.syn_data.resize(0);
      const real_t chi_b = surfatt::forward_for_event(sr, sg, mg.xgrids, mg.ygrids, true, opt);
      put_result(o, chi_b, sr, sg, true);
    }
    surfatt::release_resident_batches();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    std::fclose(o);
    return 1;
  }
  std::fclose(o);
  return 0;
}
