// layout_check.cpp -- host-side check of every address the sweep loops issue.
//
// Built with hipcc for the host only: it calls the same __host__ __device__
// addressing code the kernels inline (SweepAddr::load_base, store_offset,
// WaveLayout / SegLayout, wave_geom, wave_index, wave_decode), over many grid
// shapes, waves per pair and the four sweep directions, and checks:
//   - every load (all lanes, all steps, the ring's rows past the end included)
//     lies inside its plane: [0, wplane) wavefront-major, [0, P*ny) segmented;
//   - wavefront-major: every load of the step a row belongs to lies in that
//     step's contiguous run of La*C elements, and a lane outside the active
//     lanes loads exactly what the last active lane loads;
//   - every live store lies below the raw buffer's num_records (wplane
//     elements), every dropped store at or past it; the live stores of a sweep
//     hit each (row, lane, column) slot exactly once;
//   - segmented: a lane with no live row stores into the plane's trash row,
//     [P*ny, P*ny + kTrashElems);
//   - wave_decode inverts wave_index over every node, in both orientations.
// Prints "ok <checks>" and exits 0, or prints the first failure and exits 1.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../surfattpp_amd/csrc/sat_phases.hpp"

using namespace sat;

static long g_checks = 0;
__attribute__((format(printf, 3, 4))) static void require(bool cond, int line, const char* fmt, ...) {
  ++g_checks;
  if (cond) return;
  std::printf("FAIL layout_check.cpp:%d: ", line);
  va_list ap;
  va_start(ap, fmt);
  std::vprintf(fmt, ap);
  va_end(ap);
  std::printf("\n");
  std::exit(1);
}
#define REQUIRE(cond, ...) require((cond), __LINE__, __VA_ARGS__)

// The deepest register ring of the sweeps (rows in flight past the step's own).
constexpr int kRingMax = 8;

template <class Lay, int C, int XD, int YD>
static void check_loads(const SweepGeom& g, int nlanes, long plane) {
  const long run = (long)g.La * C;
  for (int lane = 0; lane < nlanes; lane++) {
    const SweepAddr<XD, YD> sa(g, lane);
    const SweepAddr<XD, YD> last(g, g.La - 1);
    // the first row a lane loads is -lag (step 0); the head/body pipeline loads
    // up to nsteps - lag + 1, the ring up to nsteps - lag + kRingMax (its
    // unrolled tail)
    for (int r = -sa.lag; r <= sa.nsteps - sa.lag + kRingMax; r++) {
      const int base = sa.template load_base<Lay, C>(g, r);
      REQUIRE(base >= 0 && base + C <= plane, "nx %d ny %d C %d lane %d r %d base %d plane %ld", g.nx, g.ny, C,
              lane, r, base, plane);
      const int j = sa.row_clamped(r);
      REQUIRE(j >= 0 && j < g.ny, "row %d", j);
      if constexpr (Lay::kStepRun) {
        const int ta = imin(r + sa.lag, sa.nsteps - 1);
        // the diagonal of step ta: q = ta when the rows ascend, nsteps - 1 - ta when they descend
        const int q = (YD > 0) ? ta : sa.nsteps - 1 - ta;
        REQUIRE(base >= q * run && base + C <= (q + 1) * run,
                "nx %d ny %d C %d lane %d r %d: base %d outside the run of step %d", g.nx, g.ny, C, lane, r, base,
                ta);
        if (lane >= g.La) {
          // the last active lane's row of the same step
          const int rl = ta - last.lag;
          const int lb = last.template load_base<Lay, C>(g, rl);
          REQUIRE(base == lb, "idle lane %d step %d", lane, ta);
        }
      }
    }
  }
}

template <class Lay, int C, int XD, int YD, int OOUT, class V>
static void check_stores(const SweepGeom& g, int nlanes, long plane, long trash = kTrashElems) {
  std::vector<unsigned char> hit((size_t)plane + trash, 0);
  long live_n = 0;
  for (int lane = 0; lane < nlanes; lane++) {
    const SweepAddr<XD, YD> sa(g, lane);
    for (int t = 0; t < sa.nsteps + kRingMax + 1; t++) {  // the ring runs up to R - 1 steps past the end
      const int r = t - sa.lag;
      const bool live = sa.in_range(r);
      const int j = sa.row_clamped(r);
      const int off = store_offset<Lay, C, OOUT, V>(g, live, j, lane);
      for (int k = 0; k < C; k++) {
        if constexpr (Lay::kStepRun) {
          const long nrec = (long)g.trash0 * (long)sizeof(V);  // the raw buffer's num_records, bytes
          const long b = (long)off + (long)k * (long)sizeof(V);
          REQUIRE(nrec < (long)kDropOffset, "plane too large for the drop offset: %ld", nrec);
          if (live) {
            REQUIRE(b >= 0 && b + (long)sizeof(V) <= nrec, "live store lane %d r %d k %d at %ld of %ld", lane, r, k,
                    b, nrec);
            REQUIRE(b % (long)sizeof(V) == 0, "misaligned");
            const long e = b / (long)sizeof(V);
            REQUIRE(hit[e] == 0, "slot %ld stored twice (lane %d r %d k %d)", e, lane, r, k);
            hit[e] = 1;
            live_n++;
          } else {
            REQUIRE(b >= nrec, "dropped store lane %d r %d k %d at %ld below %ld", lane, r, k, b, nrec);
            REQUIRE(b + (long)sizeof(V) <= 0x7fffffffL, "drop offset overflows");
          }
        } else {
          const long e = (long)off + k;
          if (live) {
            REQUIRE(e >= 0 && e < plane, "live store lane %d r %d k %d at %ld of %ld", lane, r, k, e, plane);
            REQUIRE(hit[e] == 0, "slot %ld stored twice", e);
            hit[e] = 1;
            live_n++;
          } else {
            REQUIRE(e >= plane && e < plane + trash, "trash store lane %d k %d at %ld", lane, k, e);
          }
        }
      }
    }
  }
  REQUIRE(live_n == (long)g.ny * g.La * C, "live stores %ld of %ld slots", live_n, (long)g.ny * g.La * C);
}

// --plant: check against planes one element short, which must fail (the
// checks are not vacuous).
static int g_short = 0;

template <int C>
static void check_shape_C(int nx, int ny, int W) {
  WaveGeom wg = wave_geom(nx, ny, W);
  wg.wplane -= g_short;
  REQUIRE(wg.C == C, "C %d != %d", wg.C, C);
  REQUIRE((long)wg.La * C >= nx && (long)(wg.La - 1) * C < nx && wg.La <= kWave * W, "La %d", wg.La);
  REQUIRE(wg.wplane + g_short == (size_t)(ny + wg.La - 1) * wg.La * C, "wplane");
  const int nlanes = kWave * W;
  // wavefront-major, both value widths (FP64 swept values, FP32 planes)
  SweepGeom gw{nx, ny, wg.La, wg.La * C, (int)wg.wplane};
  const long wp = (long)wg.wplane;
  check_loads<WaveLayout, C, 1, 1>(gw, nlanes, wp);
  check_loads<WaveLayout, C, 1, -1>(gw, nlanes, wp);
  check_loads<WaveLayout, C, -1, 1>(gw, nlanes, wp);
  check_loads<WaveLayout, C, -1, -1>(gw, nlanes, wp);
  check_stores<WaveLayout, C, 1, 1, 1, double>(gw, nlanes, wp);
  check_stores<WaveLayout, C, 1, -1, 1, double>(gw, nlanes, wp);
  check_stores<WaveLayout, C, -1, 1, 0, double>(gw, nlanes, wp);
  check_stores<WaveLayout, C, -1, -1, 0, double>(gw, nlanes, wp);
  check_stores<WaveLayout, C, 1, 1, 1, float>(gw, nlanes, wp);
  check_stores<WaveLayout, C, -1, -1, 0, float>(gw, nlanes, wp);
  // wave_index / wave_decode round trip, and the padding count
  GridDev g{};
  g.nx = nx; g.ny = ny; g.C = C; g.La = wg.La; g.wplane = wg.wplane;
  for (int o = 0; o < 2; o++) {
    long real = 0;
    for (long q = 0; q < wp; q++) {
      int ix, iy;
      if (wave_decode(g, o, (int)q, ix, iy)) {
        real++;
        REQUIRE((long)wave_index(C, wg.La, o, ix, iy) == q, "o %d q %ld -> (%d, %d)", o, q, ix, iy);
      }
    }
    REQUIRE(real == (long)nx * ny, "decoded %ld nodes of %ld", real, (long)nx * ny);
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++) {
        const size_t q = wave_index(C, wg.La, o, ix, iy);
        int dx, dy;
        REQUIRE(q < wg.wplane && wave_decode(g, o, (int)q, dx, dy) && dx == ix && dy == iy, "(%d, %d)", ix, iy);
        if (o == 0) {
          REQUIRE(wave_decode0(g, (int)q, dx, dy) && dx == ix && dy == iy, "decode0 (%d, %d)", ix, iy);
        }
      }
  }
  // segmented rows (one wave: the per-pair and per-solve kernels)
  if (W == 1) {
    const int P = wg.La * C;
    SweepGeom gs{nx, ny, wg.La, P, P * ny};
    const long sp = (long)P * ny;
    check_loads<SegLayout, C, 1, 1>(gs, nlanes, sp);
    check_loads<SegLayout, C, -1, -1>(gs, nlanes, sp);
    check_loads<SegLayout, C, 1, -1>(gs, nlanes, sp);
    check_loads<SegLayout, C, -1, 1>(gs, nlanes, sp);
    check_stores<SegLayout, C, 1, 1, 1, double>(gs, nlanes, sp);
    check_stores<SegLayout, C, -1, 1, 0, double>(gs, nlanes, sp);
    check_stores<SegLayout, C, 1, -1, 1, float>(gs, nlanes, sp);
  }
}

// Grids wider than 512 columns: W = wide_waves(nx) waves of kMaxC columns per
// lane, wavefront-major (batched sweeps) and segmented (per-pair and per-solve
// kernels, trash row of seg_trash_elems(W) elements).
static int check_wide(int nx, int ny) {
  const int W = wide_waves(nx);
  if (W <= 1) return 0;
  constexpr int C = kMaxC;
  const int La = (nx + C - 1) / C;
  REQUIRE(La <= kWave * W && (long)(La - 1) * C < nx, "wide La %d W %d", La, W);
  const int nlanes = kWave * W;
  const long wp = (long)(ny + La - 1) * La * C - g_short;
  SweepGeom gw{nx, ny, La, La * C, (int)wp};
  check_loads<WaveLayout, C, 1, 1>(gw, nlanes, wp);
  check_loads<WaveLayout, C, -1, -1>(gw, nlanes, wp);
  check_stores<WaveLayout, C, 1, -1, 1, double>(gw, nlanes, wp);
  check_stores<WaveLayout, C, -1, 1, 0, float>(gw, nlanes, wp);
  const int P = La * C;
  const long sp = (long)P * ny;
  SweepGeom gs{nx, ny, La, P, (int)sp};
  check_loads<SegLayout, C, 1, 1>(gs, nlanes, sp);
  check_loads<SegLayout, C, -1, 1>(gs, nlanes, sp);
  check_stores<SegLayout, C, 1, 1, 1, double>(gs, nlanes, sp, seg_trash_elems(W) - g_short);
  check_stores<SegLayout, C, -1, -1, 0, double>(gs, nlanes, sp, seg_trash_elems(W) - g_short);
  check_stores<SegLayout, C, 1, -1, 1, float>(gs, nlanes, sp, seg_trash_elems(W) - g_short);
  return 1;
}

static int check_shape(int nx, int ny, int W) {
  const int C = wave_geom(nx, ny, W).C;
  switch (C) {
    case 1: check_shape_C<1>(nx, ny, W); break;
    case 2: check_shape_C<2>(nx, ny, W); break;
    case 3: check_shape_C<3>(nx, ny, W); break;
    case 4: check_shape_C<4>(nx, ny, W); break;
    case 5: check_shape_C<5>(nx, ny, W); break;
    case 6: check_shape_C<6>(nx, ny, W); break;
    case 7: check_shape_C<7>(nx, ny, W); break;
    case 8: check_shape_C<8>(nx, ny, W); break;
    default: return 0;  // wider than the kernels' kMaxC columns per lane: refused at set-up
  }
  return 1;
}

int main(int argc, char** argv) {
  const bool all = argc > 1 && std::string(argv[1]) == "--all";
  if (argc > 1 && std::string(argv[1]) == "--plant") g_short = 1;
  // widths 3..512 (every one with --all) at a few heights, plus the workloads' shapes
  std::vector<int> nxs, nys = {3, 4, 7, 65, 201};
  for (int nx = 3; nx <= 512; nx += (all ? 1 : 4)) nxs.push_back(nx);
  for (int nx : {61, 64, 65, 101, 121, 127, 128, 129, 161, 192, 193, 201, 241, 256, 257, 321, 384, 385, 401, 448,
                 449, 511, 512})
    nxs.push_back(nx);
  long shapes = 0;
  for (int nx : nxs)
    for (int ny : nys)
      for (int W : {1, 2, 4, 8})
        if (W == 1 || nx > kWave * (W / 2)) shapes += check_shape(nx, ny, W);
  for (int s : {61, 71, 101, 161, 241, 321, 401, 481})
    for (int W : {1, 2, 4, 8}) shapes += check_shape(s, s + 20, W);
  for (int nx : {513, 520, 600, 640, 1000, 1024, 1025, 1500, 2048, 2049, 3000, 4096})
    for (int ny : {3, 7, 64}) shapes += check_wide(nx, ny);
  REQUIRE(wide_waves(4097) == 0 && wide_waves(512) == 1 && wide_waves(513) == 2, "wide_waves");
  std::printf("ok %ld shapes, %ld checks\n", shapes, g_checks);
  return 0;
}
