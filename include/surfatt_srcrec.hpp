// Source-receiver observation table for the C++20 shim: the reference's
// SrcRec (include/src_rec.h:44-136, src/src_rec.cpp:20-474) on one process,
// without MPI shared-memory windows.  Header-only, standard library only.
//
// surfatt::SrcRec derives from preproc::SrcRecView, so a loaded table goes
// straight into preproc::forward_for_event (include/surfatt_eik.hpp), which
// visits events_local in the reference's std::map key order and writes
// syn_data back.  rank/size deal the events like select_rank_for_src
// (include/parallel.h:39-41): event i of src_name_list belongs to rank
// i % size.
//
// Differences from the reference, all at error paths: a file that cannot be
// opened throws std::runtime_error (the reference logs and calls MPI_Abort);
// the other errors are the reference's own exceptions (std::out_of_range for a
// missing column or a short row, std::invalid_argument / std::out_of_range
// from std::stod for a number that does not convert).
#pragma once

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "surfatt_eik.hpp"

namespace surfatt {

// include/src_rec.h:30-35
struct PeriodInfo {
  std::vector<double> periods;
  std::vector<double> meanvel;
  std::vector<int> n_obs;
  int nperiod = 0;
};

// include/src_rec.h:24-28
struct Stations {
  std::vector<std::string> stnm;
  std::vector<double> stla, stlo;
};

// include/src_rec.h:37-40
struct ResidualStats {
  double mean = 0.0, stddev = 0.0;
};

namespace csv {

// Rows of a CSV document split the way rapidcsv does with its default
// parameters (include/rapidcsv.h:1557-1760): ',' separates, '"' toggles
// quoting only at the start of a cell, CR is dropped, every LF ends a row
// (empty lines included), cells are not trimmed, a cell wrapped in quotes is
// unquoted with doubled quotes collapsed, a UTF-8 byte-order mark is skipped.
inline std::vector<std::vector<std::string>> parse(const std::string& data) {
  std::vector<std::vector<std::string>> rows;
  std::vector<std::string> row;
  std::string cell;
  bool quoted = false;
  auto unquote = [](const std::string& c) {
    if (c.size() >= 2 && c.front() == '"' && c.back() == '"') {
      std::string s = c.substr(1, c.size() - 2), out;
      for (size_t i = 0; i < s.size(); i++) {
        out += s[i];
        if (s[i] == '"' && i + 1 < s.size() && s[i + 1] == '"') i++;
      }
      return out;
    }
    return c;
  };
  size_t i = (data.size() >= 3 && data.compare(0, 3, "\xEF\xBB\xBF") == 0) ? 3 : 0;
  for (; i < data.size(); i++) {
    const char ch = data[i];
    if (ch == '"') {
      if (cell.empty() || cell[0] == '"') quoted = !quoted;
      cell += ch;
    } else if (ch == ',') {
      if (!quoted) {
        row.push_back(unquote(cell));
        cell.clear();
      } else {
        cell += ch;
      }
    } else if (ch == '\r') {
    } else if (ch == '\n') {
      row.push_back(unquote(cell));
      rows.push_back(std::move(row));
      row.clear();
      cell.clear();
      quoted = false;
    } else {
      cell += ch;
    }
  }
  if (!row.empty() || !cell.empty()) {
    row.push_back(unquote(cell));
    rows.push_back(std::move(row));
  }
  return rows;
}

// A cell as rapidcsv writes it: quoted (quotes doubled) when it holds the
// separator, a space or a line feed (include/rapidcsv.h:1815-1844).
inline std::string cell(const std::string& s) {
  if (s.find(',') == std::string::npos && s.find(' ') == std::string::npos && s.find('\n') == std::string::npos)
    return s;
  std::string out = "\"";
  for (char c : s) {
    out += c;
    if (c == '"') out += '"';
  }
  return out + "\"";
}

}  // namespace csv

// Haversine distance in km (include/utils.h:286-299).
inline double gps2dist(double lat0, double lon0, double lat1, double lon1) {
  const double deg2rad = 3.14159265358979323846264338327950288 / 180.0;
  const double dlat = (lat0 - lat1) * deg2rad;
  const double dlon = (lon1 - lon0) * deg2rad;
  const double a = std::sin(dlat * 0.5) * std::sin(dlat * 0.5) +
                   std::sin(dlon * 0.5) * std::sin(dlon * 0.5) * std::cos(lat0 * deg2rad) * std::cos(lat1 * deg2rad);
  return 6371.0 * (2.0 * std::atan2(std::sqrt(a), std::sqrt(1.0 - a)));
}

class SrcRec : public preproc::SrcRecView {
 public:
  // stla, stlo, tt, weight and events_local come from SrcRecView.
  std::vector<double> evla, evlo, dist, period_all, vel;
  std::vector<double> tt_fwd, vel_fwd;
  std::vector<std::string> evtname, staname;
  std::vector<std::string> src_name_list_local, src_name_list;
  PeriodInfo periods_info;

  explicit SrcRec(int rank = 0, int size = 1) : rank_(rank), size_(size) {
    if (size < 1 || rank < 0 || rank >= size) throw std::invalid_argument("rank must lie in [0, size)");
  }

  int n_obs() const { return n_obs_; }
  int nsrc_total() const { return nsrc_total_; }

  // src/src_rec.cpp:20-131: columns by header name; weight (1), dist
  // (haversine) and vel (dist/tt, 0 where tt <= 0) are optional, and a column
  // that does not convert in full counts as absent.
  void load(const std::string& filepath) {
    std::ifstream f(filepath, std::ios::binary);
    if (!f) throw std::runtime_error("SrcRec::load: failed to open file " + filepath);
    std::ostringstream ss;
    ss << f.rdbuf();
    load_text(ss.str());
  }

  void load_text(const std::string& text) {
    const auto rows = csv::parse(text);
    std::map<std::string, size_t> col;
    if (!rows.empty())
      for (size_t i = 0; i < rows[0].size(); i++) col[rows[0][i]] = i;  // a repeated name keeps its last column
    n_obs_ = rows.empty() ? 0 : (int)rows.size() - 1;
    auto text_col = [&](const std::string& name) {
      auto it = col.find(name);
      if (it == col.end()) throw std::out_of_range("column not found: " + name);
      std::vector<std::string> out;
      for (size_t r = 1; r < rows.size(); r++) {
        if (it->second >= rows[r].size())
          throw std::out_of_range("requested column index " + std::to_string(it->second) + " >= " +
                                  std::to_string(rows[r].size()) + " (number of columns on row index " +
                                  std::to_string(r - 1) + ")");
        out.push_back(rows[r][it->second]);
      }
      return out;
    };
    auto num_col = [&](const std::string& name) {
      std::vector<double> out;
      for (const auto& s : text_col(name)) out.push_back(std::stod(s));
      return out;
    };
    auto optional = [&](const std::string& name, std::vector<double>& dst) {
      try {
        dst = num_col(name);
        return true;
      } catch (const std::exception&) {
        return false;
      }
    };
    stla = num_col("stla");
    stlo = num_col("stlo");
    evla = num_col("evla");
    evlo = num_col("evlo");
    period_all = num_col("period");
    tt = num_col("tt");
    evtname = text_col("evtname");
    staname = text_col("staname");
    if (!optional("weight", weight)) weight.assign(n_obs_, 1.0);
    if (!optional("dist", dist)) {
      dist.resize(n_obs_);
      for (int i = 0; i < n_obs_; i++) dist[i] = gps2dist(stla[i], stlo[i], evla[i], evlo[i]);
    }
    if (!optional("vel", vel)) {
      vel.resize(n_obs_);
      for (int i = 0; i < n_obs_; i++) vel[i] = (tt[i] > 0) ? (dist[i] / tt[i]) : 0.0;
    }
    get_periods();
    get_events();
  }

  // src/src_rec.cpp:325-368 on one process: tt_fwd[row] = syn of the
  // event's receiver at that row (this rank's events), vel_fwd = dist / tt_fwd.
  void gather_syn_tt() {
    tt_fwd.assign(n_obs_, 0.0);
    for (const auto& key : src_name_list) {
      auto it = events_local.find(key);
      if (it == events_local.end()) continue;
      const auto& e = it->second;
      for (size_t i = 0; i < e.rec_indices.size() && i < e.syn_data.size(); i++) tt_fwd[e.rec_indices[i]] = e.syn_data[i];
    }
    vel_fwd.resize(n_obs_);
    for (int i = 0; i < n_obs_; i++) vel_fwd[i] = dist[i] / tt_fwd[i];
  }

  // src/src_rec.cpp:371-398 on one process.
  ResidualStats compute_residual_stats() const {
    double n = 0.0, sum = 0.0, sq = 0.0;
    for (const auto& kv : events_local) {
      const auto& e = kv.second;
      for (size_t i = 0; i < e.rec_indices.size(); i++) {
        const double r = e.syn_data[i] - tt[e.rec_indices[i]];
        n += 1.0;
        sum += r;
        sq += r * r;
      }
    }
    if (n == 0.0) return {};
    const double mean = sum / n;
    const double var = sq / n - mean * mean;
    return {mean, std::sqrt(std::max(var, 0.0))};
  }

  // src/src_rec.cpp:403-459: the table in the reference's column order with
  // fixed decimals; tt and vel are the synthetic ones when is_fwd.
  void write(const std::string& filepath, bool is_fwd = false) const {
    if (n_obs_ == 0) throw std::runtime_error("SrcRec::write: no observations to write");
    if (is_fwd && (tt_fwd.empty() || vel_fwd.empty()))
      throw std::runtime_error("Forward-modeled travel times (tt_fwd) are not assigned");
    auto fmt = [](double v, int prec) {
      std::ostringstream o;
      o << std::fixed << std::setprecision(prec) << v;
      return o.str();
    };
    std::ofstream f(filepath, std::ios::binary | std::ios::trunc);
    if (!f) throw std::runtime_error("SrcRec::write: failed to open file " + filepath);
    f << "tt,staname,stla,stlo,evtname,evla,evlo,period,weight,dist,vel\n";
    const auto& T = is_fwd ? tt_fwd : tt;
    const auto& V = is_fwd ? vel_fwd : vel;
    for (int i = 0; i < n_obs_; i++) {
      f << fmt(T[i], 6) << ',' << csv::cell(staname[i]) << ',' << fmt(stla[i], 4) << ',' << fmt(stlo[i], 4) << ','
        << csv::cell(evtname[i]) << ',' << fmt(evla[i], 4) << ',' << fmt(evlo[i], 4) << ','
        << fmt(period_all[i], 2) << ',' << fmt(weight[i], 4) << ',' << fmt(dist[i], 3) << ',' << fmt(V[i], 6)
        << '\n';
    }
  }

  // SrcRec::build_stas (src/src_rec.cpp:274-323): station name -> coordinates
  // of its first row; the names present in both tables when both hold rows,
  // else those of the one that does; in name order.
  static Stations build_stas(const SrcRec* ph, const SrcRec* gr) {
    auto make_map = [](const SrcRec& sr) {
      std::map<std::string, std::pair<double, double>> m;
      for (int i = 0; i < sr.n_obs(); i++) m.try_emplace(sr.staname[i], sr.stla[i], sr.stlo[i]);
      return m;
    };
    const bool have_ph = ph && ph->n_obs() > 0, have_gr = gr && gr->n_obs() > 0;
    std::map<std::string, std::pair<double, double>> merged;
    if (have_ph && have_gr) {
      auto mp = make_map(*ph), mg = make_map(*gr);
      for (auto& [name, c] : mp)
        if (mg.count(name)) merged[name] = c;
    } else if (have_ph) {
      merged = make_map(*ph);
    } else if (have_gr) {
      merged = make_map(*gr);
    }
    Stations st;
    for (auto& [name, c] : merged) {
      st.stnm.push_back(name);
      st.stla.push_back(c.first);
      st.stlo.push_back(c.second);
    }
    return st;
  }

 private:
  int rank_, size_;
  int n_obs_ = 0, nsrc_total_ = 0;

  // src/src_rec.cpp:235-266: distinct periods in ascending order; per period
  // the number and mean velocity of the rows within 1e-6 of it.
  void get_periods() {
    std::vector<double> p(period_all);
    std::sort(p.begin(), p.end());
    p.erase(std::unique(p.begin(), p.end()), p.end());
    periods_info = PeriodInfo{};
    periods_info.nperiod = (int)p.size();
    periods_info.periods = p;
    periods_info.meanvel.assign(p.size(), 0.0);
    for (size_t i = 0; i < p.size(); i++) {
      int count = 0;
      double sum = 0.0;
      for (int j = 0; j < n_obs_; j++)
        if (std::abs(period_all[j] - p[i]) < 1e-6) {
          count++;
          sum += vel[j];
        }
      periods_info.n_obs.push_back(count);
      periods_info.meanvel[i] = sum / count;
    }
  }

  // src/src_rec.cpp:141-231: one event per distinct event name (byte order)
  // and period index with rows, keyed "{evtname}_{iper}", receivers in row
  // order, evla/evlo/period from its last row.
  void get_events() {
    std::vector<std::string> keys(evtname);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::map<std::string, preproc::event_info> events;
    src_name_list.clear();
    for (const auto& src : keys) {
      for (int iper = 0; iper < periods_info.nperiod; iper++) {
        preproc::event_info info;
        for (int i = 0; i < n_obs_; i++) {
          if (evtname[i] == src && std::abs(periods_info.periods[iper] - period_all[i]) < 1e-6) {
            info.rec_indices.push_back(i);
            info.evla = evla[i];
            info.evlo = evlo[i];
            info.period = period_all[i];
            info.iper = iper;
          }
        }
        if (!info.rec_indices.empty()) {
          info.syn_data.assign(info.rec_indices.size(), 0.0);
          const std::string key = src + "_" + std::to_string(iper);
          events[key] = info;
          src_name_list.push_back(key);
        }
      }
    }
    nsrc_total_ = (int)src_name_list.size();
    events_local.clear();
    src_name_list_local.clear();
    for (int i = 0; i < nsrc_total_; i++) {
      if (i % size_ != rank_) continue;
      src_name_list_local.push_back(src_name_list[i]);
      events_local[src_name_list[i]] = events[src_name_list[i]];
    }
  }
};

// ModelGrid's lon/lat axes for grid_method 0 (src/model_grid.cpp:127-165):
// the station box widened by num_grid_margin cells, llround((end-beg)/d)+1
// nodes, spaced like Eigen 3.4's LinSpaced (low + i*step with the last node
// exactly high; counted back from high when |high| < |low|).
inline std::vector<double> linspaced(int n, double low, double high) {
  std::vector<double> v(n);
  if (n == 1) {
    v[0] = high;
    return v;
  }
  const double step = (high - low) / double(n - 1);
  const bool flip = std::abs(high) < std::abs(low);
  for (int i = 0; i < n; i++) {
    if (flip)
      v[i] = (i == 0) ? low : high - double(n - 1 - i) * step;
    else
      v[i] = (i == n - 1) ? high : low + double(i) * step;
  }
  return v;
}

inline std::pair<std::vector<double>, std::vector<double>> build_model_axes(const Stations& st, double dlon,
                                                                             double dlat, int num_grid_margin) {
  if (st.stlo.empty()) throw std::invalid_argument("no stations");
  const double xbeg = *std::min_element(st.stlo.begin(), st.stlo.end()) - num_grid_margin * dlon;
  const double xend = *std::max_element(st.stlo.begin(), st.stlo.end()) + num_grid_margin * dlon;
  const double ybeg = *std::min_element(st.stla.begin(), st.stla.end()) - num_grid_margin * dlat;
  const double yend = *std::max_element(st.stla.begin(), st.stla.end()) + num_grid_margin * dlat;
  const int nx = (int)(std::llround((xend - xbeg) / dlon) + 1);
  const int ny = (int)(std::llround((yend - ybeg) / dlat) + 1);
  return {linspaced(nx, xbeg, xend), linspaced(ny, ybeg, yend)};
}

}  // namespace surfatt
