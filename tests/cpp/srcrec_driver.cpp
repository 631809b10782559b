// Drives surfatt::SrcRec (include/surfatt_srcrec.hpp).
//   srcrec_driver table.csv out_prefix dlon dlat margin [run]
// writes out_prefix.txt (periods, events, stations and model axes, numbers
// as %.17g) and out_prefix.csv (SrcRec::write).  With "run" it also builds
// the station-box grid, fills every period with the table's mean velocity
// (identity metric), runs preproc::forward_for_event in FP64 on the GPU and
// writes chi, the residual statistics and out_prefix_fwd.csv
// (write with is_fwd after gather_syn_tt).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "surfatt_srcrec.hpp"

int main(int argc, char** argv) {
  if (argc < 6) return 2;
  const std::string table = argv[1], out = argv[2];
  const double dlon = std::atof(argv[3]), dlat = std::atof(argv[4]);
  const int margin = std::atoi(argv[5]);
  const bool run = argc > 6 && std::string(argv[6]) == "run";
  surfatt::SrcRec sr;
  sr.load(table);
  FILE* f = std::fopen((out + ".txt").c_str(), "w");
  if (!f) return 2;
  std::fprintf(f, "n_obs %d\nnsrc_total %d\nnperiod %d\n", sr.n_obs(), sr.nsrc_total(), sr.periods_info.nperiod);
  for (int i = 0; i < sr.periods_info.nperiod; i++)
    std::fprintf(f, "period %.17g %.17g %d\n", sr.periods_info.periods[i], sr.periods_info.meanvel[i],
                 sr.periods_info.n_obs[i]);
  for (auto& kv : sr.events_local) {
    const auto& e = kv.second;
    std::fprintf(f, "event %s %d %.17g %.17g %.17g", kv.first.c_str(), e.iper, e.evla, e.evlo, e.period);
    for (int r : e.rec_indices) std::fprintf(f, " %d", r);
    std::fprintf(f, "\n");
  }
  for (int i = 0; i < sr.n_obs(); i++)
    std::fprintf(f, "row %.17g %.17g %.17g\n", sr.weight[i], sr.dist[i], sr.vel[i]);
  const auto st = surfatt::SrcRec::build_stas(&sr, nullptr);
  for (size_t i = 0; i < st.stnm.size(); i++)
    std::fprintf(f, "station %s %.17g %.17g\n", st.stnm[i].c_str(), st.stla[i], st.stlo[i]);
  const auto axes = surfatt::build_model_axes(st, dlon, dlat, margin);
  for (double x : axes.first) std::fprintf(f, "x %.17g\n", x);
  for (double y : axes.second) std::fprintf(f, "y %.17g\n", y);
  sr.write(out + ".csv");
  if (run) {
    const int nx = (int)axes.first.size(), ny = (int)axes.second.size(), np = sr.periods_info.nperiod;
    std::vector<double> svel((size_t)nx * ny * np), m11(svel.size(), 1.0), m12(svel.size(), 0.0),
        m22(svel.size(), 1.0);
    for (size_t n = 0; n < svel.size(); n++) svel[n] = sr.periods_info.meanvel[n % np];  // [ix][iy][iper]
    preproc::SurfGridView sg;
    sg.xgrids = axes.first;
    sg.ygrids = axes.second;
    sg.nperiod = np;
    sg.svel = svel.data();
    sg.m11 = m11.data();
    sg.m12 = m12.data();
    sg.m22 = m22.data();
    surfatt::precision() = SAT_PREC_F64;
    const double chi = preproc::forward_for_event(sr, sg, true);
    sr.gather_syn_tt();
    const auto rs = sr.compute_residual_stats();
    std::fprintf(f, "chi %.17g\nresidual %.17g %.17g\n", chi, rs.mean, rs.stddev);
    for (int i = 0; i < sr.n_obs(); i++) std::fprintf(f, "tt_fwd %.17g\n", sr.tt_fwd[i]);
    for (int ip = 0; ip < np; ip++)
      for (size_t n = 0; n < (size_t)nx * ny; n++) std::fprintf(f, "adj_s %.17g\n", sg.adj_s_local[ip].data()[n]);
    sr.write(out + "_fwd.csv", true);
  }
  std::fclose(f);
  return 0;
}
