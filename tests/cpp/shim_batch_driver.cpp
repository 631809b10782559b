// Drives preproc::forward_for_event of the C++ shim (include/surfatt_eik.hpp)
// with the reference's data layout: period planes [ix][iy][iper], events in a
// map keyed like events_local.  Reads argv[1] (binary: nx, ny, nper, ne, nrec;
// xx[nx], yy[ny]; svel, m11, m12, m22 [nx*ny*nper each, [ix][iy][iper]];
// per event iper, nrec_e (int32) and evlo, evla (double); per receiver stlo,
// stla, tt, weight (double)) and writes chi, syn[nrec], adj_s [nper][ny][nx]
// to argv[2].
#include <cstdio>
#include <vector>

#include "surfatt_eik.hpp"

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
  const size_t n3 = (size_t)nx * ny * np;
  preproc::SurfGridView sg;
  sg.xgrids.resize(nx);
  sg.ygrids.resize(ny);
  sg.nperiod = np;
  std::vector<double> pl[4];
  bool ok = rd(f, sg.xgrids.data(), nx) && rd(f, sg.ygrids.data(), ny);
  for (auto& p : pl) {
    p.resize(n3);
    ok = ok && rd(f, p.data(), n3);
  }
  sg.svel = pl[0].data();
  sg.m11 = pl[1].data();
  sg.m12 = pl[2].data();
  sg.m22 = pl[3].data();
  preproc::SrcRecView sr;
  int rec = 0;
  for (int e = 0; e < ne && ok; e++) {
    int ip_n[2];
    double lo_la[2];
    ok = rd(f, ip_n, 2) && rd(f, lo_la, 2);
    char key[32];
    std::snprintf(key, sizeof key, "%08d", e);
    preproc::event_info& ev = sr.events_local[key];
    ev.iper = ip_n[0];
    ev.evlo = lo_la[0];
    ev.evla = lo_la[1];
    for (int r = 0; r < ip_n[1]; r++) ev.rec_indices.push_back(rec++);
  }
  for (auto* v : {&sr.stlo, &sr.stla, &sr.tt, &sr.weight}) {
    v->resize(nrec);
    ok = ok && rd(f, v->data(), nrec);
  }
  std::fclose(f);
  if (!ok || rec != nrec) return 2;
  try {
    const double chi = preproc::forward_for_event(sr, sg, true);
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(&chi, sizeof(double), 1, o);
    for (auto& kv : sr.events_local) std::fwrite(kv.second.syn_data.data(), sizeof(double), kv.second.syn_data.size(), o);
    for (int ip = 0; ip < np; ip++) std::fwrite(sg.adj_s_local[ip].data(), sizeof(double), (size_t)nx * ny, o);
    std::fclose(o);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
