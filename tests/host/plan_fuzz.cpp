// plan_fuzz.cpp -- the launch planner (csrc/sarbp_plan.cpp) on random grids,
// tracks and parameters, under AddressSanitizer + UndefinedBehaviorSanitizer
// (tests/test_host_sanitizers.py builds and runs it; no GPU).  For every case
// that plans, the launch descriptor must satisfy what the kernels assume:
//   * the tile shape is compiled for the arithmetic, and REF falls back to
//     DF only when no REF shape fits the radius limit;
//   * the LDS layout of (window, pulses per stage) fits the budget, P >= 1;
//   * the tiles cover the image;
//   * the window holds every range bin a pixel of any tile needs, with the
//     interpolator's taps -- checked by brute force over all pixels and
//     pulses of small cases, in FP64;
//   * the DF quantum makes alpha_h di + beta_h dj exact for the tile;
//   * the stage depth fitted to occupancy (a stand-in occupancy source:
//     registers allow 4 workgroups per CU, 6 for 1 x 2 tiles; LDS in 512-B
//     blocks of 160 KiB) is at most the tuned one, never costs a workgroup
//     per CU, stops at 4 pulses for a plain launch and goes no deeper for a
//     batched one, with the same window;
// and the batch grouping must put equal keys together and stay within its
// limit.  Exit status 0 when every check passes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "sarbp.h"
#include "sarbp_host.h"

using namespace sarbp;

namespace {

int g_failures = 0, g_planned = 0, g_df = 0, g_refused = 0, g_fit_plain = 0, g_fit_batch = 0;

#define CHECK(cond, ...)                                      \
  do {                                                        \
    if (!(cond)) {                                            \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);                      \
      std::fprintf(stderr, "\n");                             \
      ++g_failures;                                           \
    }                                                         \
  } while (0)

double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Stand-in for the device library's occupancy query.
int stand_in_occupancy(int, int, int, int px, int py, size_t shm) {
  const int regs = px == 1 && py == 2 ? 6 : 4;
  const size_t blocks = (shm + 16 + 511) / 512;
  return std::max(1, std::min<int>(regs, (int)((160 * 1024 / 512) / blocks)));
}

// Every window-relative bin of every (tile, pulse, pixel) lies where the
// interpolator reads inside the window: [lo, hi] in window samples.
void check_window(const BpArgs& a, const std::vector<double>& pos, int px, int py) {
  const int TX = kTileX * px, TY = kWaves * py;
  double lo_tap = 0, hi_tap = 0;  // samples the interpolator reads around the bin
  if (a.kernel_interp == kNearest) { lo_tap = 0; hi_tap = 0; }
  if (a.kernel_interp == kLinear) { lo_tap = 0; hi_tap = 1; }
  if (a.kernel_interp == kSinc8 || a.kernel_interp == kSinc8W) { lo_tap = 3; hi_tap = 4; }
  for (int ty = 0; ty < a.tiles_y; ++ty)
    for (int tx = 0; tx < a.tiles_x; ++tx) {
      const int ic = tx * TX + TX / 2, jc = ty * TY + TY / 2;
      double c[3];
      for (int k = 0; k < 3; ++k) c[k] = a.origin[k] + ic * a.ux[k] + jc * a.uy[k];
      for (int m = 0; m < a.n_pulses; m += std::max(1, a.n_pulses / 7)) {
        const double* p = &pos[3 * (size_t)m];
        const double rc = std::sqrt((c[0] - p[0]) * (c[0] - p[0]) + (c[1] - p[1]) * (c[1] - p[1]) +
                                    (c[2] - p[2]) * (c[2] - p[2]));
        const double start = std::floor((rc - a.r0) * a.inv_dr) - a.half;
        for (int j = 0; j < TY; j += std::max(1, TY / 4))
          for (int i = 0; i < TX; i += std::max(1, TX / 8)) {
            const int x = tx * TX + i, y = ty * TY + j;
            if (x >= a.nx || y >= a.ny) continue;
            double q[3];
            for (int k = 0; k < 3; ++k) q[k] = a.origin[k] + x * a.ux[k] + y * a.uy[k] - p[k];
            const double u = (std::sqrt(dot(q, q)) - a.r0) * a.inv_dr - start;
            const bool ok = u - lo_tap >= -1e-6 && u + hi_tap <= a.win - 1 + 1e-6;
            if (!ok) {
              CHECK(ok, "tile (%d,%d) pulse %d pixel (%d,%d): bin %.3f outside window %d (interp %d)", tx, ty, m,
                    x, y, u, a.win, a.kernel_interp);
              return;
            }
          }
      }
    }
}

void one_case(std::mt19937_64& rng, bool brute) {
  std::uniform_real_distribution<double> U(0.0, 1.0);
  sarbp_grid g{};
  const double sp = std::exp(std::log(0.05) + U(rng) * (std::log(8.0) - std::log(0.05)));
  const double ang = U(rng) * 6.283185307179586, tilt = (U(rng) - 0.5) * 0.4;
  g.ux[0] = sp * std::cos(ang);
  g.ux[1] = sp * std::sin(ang);
  g.ux[2] = sp * tilt * 0.1;
  g.uy[0] = -sp * std::sin(ang) * (0.8 + 0.4 * U(rng));
  g.uy[1] = sp * std::cos(ang) * (0.8 + 0.4 * U(rng));
  g.uy[2] = sp * tilt;
  g.nx = brute ? 1 + (int)(U(rng) * 300) : (int)(U(rng) * 5000);
  g.ny = brute ? 1 + (int)(U(rng) * 120) : (int)(U(rng) * 5000);
  const double range = std::exp(std::log(300.0) + U(rng) * (std::log(1.0e6) - std::log(300.0)));
  g.origin[0] = -0.5 * g.nx * sp;
  g.origin[1] = range * (0.5 + 0.5 * U(rng));
  g.origin[2] = 0.0;
  sarbp_params p{};
  p.c = 299792458.0;
  p.fc = U(rng) < 0.5 ? 9.6e9 : 5.3e9;
  p.dr = 0.05 + U(rng) * 2.0;
  p.n_pulses = 1 + (int)(U(rng) * (brute ? 64 : 6000));
  p.n_samples = 16 + (int)(U(rng) * 16384);
  p.r0 = range * 0.4;
  p.kaiser_beta = U(rng) * 12.0;
  p.interp = (int)(U(rng) * 3);
  p.precision = (int)(U(rng) * 4);
  p.dtype = (int)(U(rng) * 2);
  const bool with_pos = U(rng) < 0.7;
  std::vector<double> pos(3 * (size_t)p.n_pulses);
  const double alt = range * (0.2 + 0.6 * U(rng));
  for (int m = 0; m < p.n_pulses; ++m) {
    pos[3 * m] = (m - 0.5 * p.n_pulses) * 0.15 + std::sin(m * 0.01) * U(rng);
    pos[3 * m + 1] = U(rng) * 2.0;
    pos[3 * m + 2] = alt + U(rng);
  }
  const double* h_pos = with_pos ? pos.data() : nullptr;
  if (validate(&g, &p) != SARBP_OK) return;
  const int cls = size_class(&g, &p, h_pos);
  CHECK(cls >= 0 && cls < SARBP_SIZE_CLASSES, "size class %d", cls);
  // the built-in default of the call's kind: with positions, or the
  // position-free table (windows over the disc of each tile's radius)
  sarbp_tuning t = default_tuning(p.interp, cls, p.precision, h_pos != nullptr);
  if (U(rng) < 0.5) {  // a random compiled shape and stage depth
    static const int shapes[][2] = {{1, 2}, {1, 4}, {1, 8}, {2, 2}, {2, 4}, {2, 8}, {4, 2}, {4, 4}};
    const int* sh = shapes[(int)(U(rng) * 8)];
    t.cols_per_lane = sh[0];
    t.rows_per_lane = sh[1];
    t.pulses_per_stage = 1 + (int)(U(rng) * 64);
    if (check_tuning(p.precision, &t) != SARBP_OK) t = default_tuning(p.interp, cls, p.precision, h_pos != nullptr);
  }
  const Fitted f = fit_tuning(t, p.precision, &g, &p, h_pos);
  CHECK(tile_shape_supported(f.prec, f.t.cols_per_lane, f.t.rows_per_lane), "shape %dx%d not compiled for %d",
        f.t.cols_per_lane, f.t.rows_per_lane, f.prec);
  CHECK(f.prec == p.precision || (p.precision == SARBP_PREC_REF && f.prec == kDF), "arithmetic %d for %d", f.prec,
        p.precision);
  BpArgs a{};
  const sarbp_status st = plan_args(h_pos, &g, &p, f.t, f.prec, &a, nullptr);
  if (st != SARBP_OK) {
    CHECK(st == SARBP_ERR_UNSUPPORTED, "plan_args status %d", st);
    ++g_refused;
    return;
  }
  ++g_planned;
  if (f.prec == kDF) ++g_df;
  CHECK(a.pulses_per_stage >= 1, "P %d", a.pulses_per_stage);
  CHECK(lds_layout(a.kernel_interp, a.win, a.pulses_per_stage).total <= kLdsBudget, "LDS over budget");
  CHECK(a.win >= 8 && a.win % 4 == 0, "window %d", a.win);
  const int TX = kTileX * f.t.cols_per_lane, TY = kWaves * f.t.rows_per_lane;
  CHECK((long long)a.tiles_x * TX >= g.nx && (long long)a.tiles_y * TY >= g.ny, "tiles do not cover the image");
  CHECK(a.kernel_prec == f.prec, "descriptor arithmetic");
  // DF quantum: the span bound the kernel's exactness relies on
  const double span = std::sqrt(dot(g.ux, g.ux)) * (TX / 2) + std::sqrt(dot(g.uy, g.uy)) * (TY / 2) + 1.0;
  CHECK(span * a.s_quant <= 8388608.0, "quantum 2^e = %g too fine for span %g", a.s_quant, span);
  if (brute && with_pos) check_window(a, pos, f.t.cols_per_lane, f.t.rows_per_lane);

  // stage depth: P0 is the tuned depth fitted to the LDS budget alone
  int P0 = std::max(1, f.t.pulses_per_stage);
  while (P0 > 1 && lds_layout(a.kernel_interp, a.win, P0).total > kLdsBudget) --P0;
  auto occ = [&](int q) {
    return stand_in_occupancy(0, 0, 0, f.t.cols_per_lane, f.t.rows_per_lane,
                              lds_layout(a.kernel_interp, a.win, q).total);
  };
  BpArgs b{};
  CHECK(plan_args(h_pos, &g, &p, f.t, f.prec, &b, nullptr, true) == SARBP_OK, "batched plan refused");
  const int Pp = a.pulses_per_stage, Pb = b.pulses_per_stage;
  CHECK(b.win == a.win && b.kernel_interp == a.kernel_interp, "batched plan's window differs");
  CHECK(Pp <= P0 && Pb <= Pp && Pb >= 1, "depths: tuned %d, plain %d, batched %d", P0, Pp, Pb);
  CHECK(Pp >= std::min(P0, 4), "plain depth %d below 4 (tuned %d)", Pp, P0);
  CHECK(occ(Pp) >= occ(P0) && occ(Pb) >= occ(Pp), "occupancy lost: %d %d %d", occ(P0), occ(Pp), occ(Pb));
  // each pulse taken off gained a workgroup per CU, and the stop is where the
  // next one would not (or at the floor)
  g_fit_plain += Pp < P0;
  g_fit_batch += Pb < Pp;
  for (int q = Pp; q < P0; ++q) CHECK(occ(q) > occ(q + 1), "plain step %d -> %d gained nothing", q + 1, q);
  CHECK(Pp <= 4 || occ(Pp - 1) <= occ(Pp), "plain fitting stopped early at %d", Pp);
  for (int q = Pb; q < Pp; ++q) CHECK(occ(q) > occ(q + 1), "batched step %d -> %d gained nothing", q + 1, q);
  CHECK(Pb == 1 || occ(Pb - 1) <= occ(Pb), "batched fitting stopped early at %d", Pb);
}

void group_cases(std::mt19937_64& rng) {
  std::uniform_int_distribution<int> D(0, 3);
  for (int rep = 0; rep < 2000; ++rep) {
    const int n = 1 + (int)(rng() % 300);
    std::vector<GroupKey> k((size_t)n);
    for (GroupKey& x : k) x = GroupKey{1 + D(rng) % 2, 2 << (D(rng) % 3), 2 + D(rng) % 2, D(rng) % 2, 1 + D(rng)};
    std::vector<int> out((size_t)n);
    const int max_groups = 1 + (int)(rng() % 10);
    const int ng = assign_groups(k.data(), n, max_groups, out.data());
    if (ng < 0) {  // even without occupancy keys too many kernels
      int kinds = 0;
      for (int i = 0; i < n; ++i) {
        bool seen = false;
        for (int j = 0; j < i && !seen; ++j)
          seen = k[j].px == k[i].px && k[j].py == k[i].py && k[j].kinterp == k[i].kinterp && k[j].kprec == k[i].kprec;
        kinds += !seen;
      }
      CHECK(kinds > max_groups, "refused %d kernels with %d groups allowed", kinds, max_groups);
      continue;
    }
    CHECK(ng <= max_groups, "%d groups > %d", ng, max_groups);
    for (int i = 0; i < n; ++i) {
      CHECK(out[i] >= 0 && out[i] < ng, "group index %d", out[i]);
      for (int j = 0; j < i; ++j) {
        const bool same_kernel = k[j].px == k[i].px && k[j].py == k[i].py && k[j].kinterp == k[i].kinterp &&
                                 k[j].kprec == k[i].kprec;
        if (out[i] == out[j]) CHECK(same_kernel, "jobs %d and %d share a group across kernels", i, j);
        if (same_kernel && k[j].occ == k[i].occ) CHECK(out[i] == out[j], "equal keys %d, %d split", i, j);
      }
    }
  }
}

}  // namespace

int main() {
  set_occupancy_source(&stand_in_occupancy);
  std::mt19937_64 rng(20261203);
  for (int i = 0; i < 3000; ++i) one_case(rng, false);
  for (int i = 0; i < 400; ++i) one_case(rng, true);
  group_cases(rng);
  std::printf("planned %d (DF fallback or request %d), refused %d; stage depth fitted %d plain, %d more batched\n",
              g_planned, g_df, g_refused, g_fit_plain, g_fit_batch);
  if (g_failures) {
    std::printf("%d checks failed\n", g_failures);
    return 1;
  }
  std::printf("all plan checks passed\n");
  return 0;
}
