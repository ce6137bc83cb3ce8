// sarbp_plan.cpp -- launch planning on the host, shared by the C ABI
// (sarbp_api.hip), the batched launches and the server, and free of HIP
// runtime calls so that the server core and its tests build with a plain
// host compiler: argument validation, the sinc8 tap polynomials, tile-shape
// fitting, the LDS range window, size classes, the launch descriptor and the
// batch grouping.  Nothing here exits or throws.
#include "sarbp.h"
#include "sarbp_host.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace sarbp {

// Detail text of the calling thread's last error: a fixed buffer, so that
// reporting an error allocates nothing.
thread_local char g_detail[512];

const char* last_detail() { return g_detail; }

sarbp_status fail(sarbp_status st, const char* fmt, ...) {
  char buf[sizeof g_detail];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  std::memcpy(g_detail, buf, sizeof buf);
  return st;
}

constexpr double kPi = 3.14159265358979323846;

std::atomic<OccupancyFn> g_occupancy_source{nullptr};

void set_occupancy_source(OccupancyFn fn) { g_occupancy_source.store(fn); }

int kernel_occupancy(int kinterp, int kprec, int dtype, int px, int py, size_t shm) {
  const OccupancyFn fn = g_occupancy_source.load();
  return fn ? fn(kinterp, kprec, dtype, px, py, shm) : 0;
}

double bessel_i0(double x) {
  double q = 0.25 * x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; ++k) {
    term *= q / ((double)k * (double)k);
    sum += term;
    if (term < 1e-18 * sum) break;
  }
  return sum;
}

double kaiser_sinc8(double x, double beta, double inv_i0b) {
  const double ax = std::fabs(x);
  if (ax >= 4.0) return 0.0;
  const double t = x / 4.0;
  const double w = bessel_i0(beta * std::sqrt(1.0 - t * t)) * inv_i0b;
  const double s = (x == 0.0) ? 1.0 : std::sin(kPi * x) / (kPi * x);
  return s * w;
}

// Tap polynomials of the sinc8 interpolator, in the centred variable e of
// each interpolation interval.  Whole bins (`halves` = 1): e = f - 1/2 over
// f in [0, 1), C[k * 8 + t] multiplies e^k in the weight K(f + 3 - t) of tap
// t (sample i0 - 3 + t), degree 7.  Half bins (`halves` = 2): half h =
// floor(2 f), e = 2 f - h - 1/2, C[(h * 6 + k) * 8 + t], degree 5.  Each
// polynomial interpolates its weight at the N Chebyshev nodes of
// [-1/2, 1/2] (near-minimax; max error ~2e-6 for beta = 6), an N x N
// Vandermonde solve in FP64 with partial pivoting, into `out` (FP64).
//
// K(x) = K(-x) maps (half h, tap t, e) to (half 1 - h, tap 7 - t, -e), so
// C[1 - h][k][7 - t] = (-1)^k C[h][k][t]: half 0 (whole bins: taps 0..3) is
// fitted and the rest is set by that identity, exactly, which the kernels'
// coefficient precompute relies on.
template <int N>
void sinc8_poly_fit(double beta, int halves, double* out) {
  const double inv_i0b = 1.0 / bessel_i0(beta);
  double nodes[N];
  for (int n = 0; n < N; ++n) nodes[n] = 0.5 * std::cos(kPi * (n + 0.5) / N);
  const int fitted = halves == 2 ? kSincTaps : kSincTaps / 2;  // taps of half 0 fitted
  for (int t = 0; t < fitted; ++t) {
    double A[N][N + 1];
    for (int n = 0; n < N; ++n) {
      double v = 1.0;
      for (int k = 0; k < N; ++k, v *= nodes[n]) A[n][k] = v;
      A[n][N] = kaiser_sinc8((0.5 + nodes[n]) / halves + 3.0 - t, beta, inv_i0b);
    }
    for (int c = 0; c < N; ++c) {
      int piv = c;
      for (int r = c + 1; r < N; ++r)
        if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
      for (int k = 0; k <= N; ++k) std::swap(A[c][k], A[piv][k]);
      for (int r = 0; r < N; ++r) {
        if (r == c) continue;
        const double m = A[r][c] / A[c][c];
        for (int k = c; k <= N; ++k) A[r][k] -= m * A[c][k];
      }
    }
    for (int k = 0; k < N; ++k) {
      const double c = A[k][N] / A[k][k];
      out[k * kSincTaps + t] = c;
      out[((halves - 1) * N + k) * kSincTaps + (kSincTaps - 1 - t)] = (k & 1) ? -c : c;
    }
  }
}

void sinc8_poly(double beta, float* out) {
  double c[kSincHalves * kSincCoef * kSincTaps];
  sinc8_poly_fit<kSincCoef>(beta, kSincHalves, c);
  for (int i = 0; i < kSincHalves * kSincCoef * kSincTaps; ++i) out[i] = (float)c[i];
}

void sinc8_poly_whole(double beta, float* out) {
  double c[kSincCoefW * kSincTaps];
  sinc8_poly_fit<kSincCoefW>(beta, 1, c);
  for (int i = 0; i < kSincCoefW * kSincTaps; ++i) out[i] = (float)c[i];
}

// The kernels' form of the polynomials: with sigma_t = s_t + s_(7-t) and
// delta_t = s_t - s_(7-t) (t < 4) the mirror sums and differences of the 8
// staged samples, the interpolant's coefficients at one window position are
//   half bins:  D[0][k] = P_k . sigma + Q_k . delta,
//               D[1][k] = (-1)^k (P_k . sigma - Q_k . delta),
//               P_k = (C[0][k][t] + C[0][k][7-t]) / 2,
//               Q_k = (C[0][k][t] - C[0][k][7-t]) / 2;
//   whole bins: D[k] = C[k][t<4] . sigma (k even), . delta (k odd),
// 4 FMA per term instead of 8.  Rows of 4: half bins [k][P, Q], whole bins
// [k].
void sinc8_kernel_rows(int kinterp, double beta, float* out) {
  if (kinterp == kSinc8) {
    double c[kSincHalves * kSincCoef * kSincTaps];
    sinc8_poly_fit<kSincCoef>(beta, kSincHalves, c);
    for (int k = 0; k < kSincCoef; ++k)
      for (int t = 0; t < 4; ++t) {
        const double a = c[k * kSincTaps + t], b = c[k * kSincTaps + kSincTaps - 1 - t];
        out[(2 * k) * 4 + t] = (float)(0.5 * (a + b));
        out[(2 * k + 1) * 4 + t] = (float)(0.5 * (a - b));
      }
  } else if (kinterp == kSinc8W) {
    double c[kSincCoefW * kSincTaps];
    sinc8_poly_fit<kSincCoefW>(beta, 1, c);
    for (int k = 0; k < kSincCoefW; ++k)
      for (int t = 0; t < 4; ++t) out[k * 4 + t] = (float)c[k * kSincTaps + t];
  }
}

bool finite_all(const double* v, int n) {
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

double norm3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

sarbp_status check_tuning(int precision, const sarbp_tuning* t) {
  if (!tile_shape_supported(precision, t->cols_per_lane, t->rows_per_lane))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "unsupported tile shape: %d cols x %d rows per lane",
                t->cols_per_lane, t->rows_per_lane);
  if (t->pulses_per_stage < 1 || t->pulses_per_stage > 256)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "pulses_per_stage must be in [1, 256]");
  return SARBP_OK;
}

sarbp_status validate(const sarbp_grid* g, const sarbp_params* p) {
  if (!g || !p) return fail(SARBP_ERR_INVALID_ARGUMENT, "grid or params is NULL");
  if (g->nx < 0 || g->ny < 0) return fail(SARBP_ERR_INVALID_ARGUMENT, "negative image size %d x %d", g->nx, g->ny);
  if (!finite_all(g->origin, 3) || !finite_all(g->ux, 3) || !finite_all(g->uy, 3))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "grid vectors must be finite");
  if (p->n_pulses < 0) return fail(SARBP_ERR_INVALID_ARGUMENT, "n_pulses < 0");
  if (p->n_samples < 1) return fail(SARBP_ERR_INVALID_ARGUMENT, "n_samples < 1");
  if (p->interp < 0 || p->interp > 2) return fail(SARBP_ERR_INVALID_ARGUMENT, "interp %d", p->interp);
  if (p->precision < 0 || p->precision > 3) return fail(SARBP_ERR_INVALID_ARGUMENT, "precision %d", p->precision);
  if (p->dtype < 0 || p->dtype > 1) return fail(SARBP_ERR_INVALID_ARGUMENT, "dtype %d", p->dtype);
  if (p->accumulate != 0 && p->accumulate != 1) return fail(SARBP_ERR_INVALID_ARGUMENT, "accumulate %d", p->accumulate);
  if (!(std::isfinite(p->c) && p->c > 0)) return fail(SARBP_ERR_INVALID_ARGUMENT, "c must be > 0");
  if (!(std::isfinite(p->fc) && p->fc >= 0)) return fail(SARBP_ERR_INVALID_ARGUMENT, "fc must be >= 0");
  if (!(std::isfinite(p->dr) && p->dr > 0)) return fail(SARBP_ERR_INVALID_ARGUMENT, "dr must be > 0");
  if (!std::isfinite(p->r0)) return fail(SARBP_ERR_INVALID_ARGUMENT, "r0 must be finite");
  if (p->interp == SARBP_INTERP_SINC8 && !(std::isfinite(p->kaiser_beta) && p->kaiser_beta >= 0 && p->kaiser_beta <= 50))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "kaiser_beta must be in [0, 50]");
  if (norm3(g->ux) == 0.0 || norm3(g->uy) == 0.0) return fail(SARBP_ERR_INVALID_ARGUMENT, "zero grid step");
  return SARBP_OK;
}

sarbp_tuning default_tuning(int interp, int cls, int precision, bool positions) {
  // Built-in defaults: per size class (range windows of a 64 x 32 tile of
  // < 80, < 120, < 176, < 272 and >= 272 samples), the (tile shape x pulses
  // per stage) point of MI355X sweeps that is closest to the best on every
  // swept geometry -- 2048 pulses over 2048^2 pixels and 8192 pulses over
  // 1536^2, 0.25 / 0.35 / 0.5 / 0.7 / 1 m pixels at X band, 5 km: within 0-9%
  // of each geometry's best.  Wider windows fit fewer pulses in a stage at
  // the occupancy that wins.  Two-column tiles win on some grids but lose up
  // to 37% under long apertures, where their width adds range spread; the
  // autotuner picks them where they pay.  Without positions a tile's window
  // covers the disc of its radius, so the class of a grid is higher and the
  // squarer 64 x 32 tile (most pixels per radius) wins in most classes
  // (within 0-8% of the best of its swept geometries; at 0.25 m under 8192
  // pulses, 18%).  The planner lowers P where the LDS budget needs it.
  // Linear and nearest stage 32 pulses of their 8-byte windows.
  sarbp_tuning t;
  t.cols_per_lane = 1;
  t.rows_per_lane = 4;
  t.pulses_per_stage = 32;
  if (interp == SARBP_INTERP_SINC8) {
    static const int kShape[SARBP_SIZE_CLASSES][3] = {{1, 4, 8}, {1, 4, 6}, {1, 4, 4}, {1, 4, 3}, {1, 4, 2}};
    static const int kShapeNoPos[SARBP_SIZE_CLASSES][3] = {{1, 8, 4}, {1, 8, 3}, {1, 4, 3}, {1, 8, 3}, {1, 8, 2}};
    const int c = std::min(std::max(cls, 0), SARBP_SIZE_CLASSES - 1);
    const int* sh = positions ? kShape[c] : kShapeNoPos[c];
    t.cols_per_lane = sh[0];
    t.rows_per_lane = sh[1];
    t.pulses_per_stage = sh[2];
  }
  // the F64 and F32 modes compile one column per lane
  if (!tile_shape_supported(precision, t.cols_per_lane, t.rows_per_lane)) t.cols_per_lane = 1;
  return t;
}

// Largest |pixel - tile reference| the FP32 differential supports.  Two FP32
// roundings grow with it: pp = |d|^2 - s^2 carries ~ulp(|d|^2) into
// q = pp / (2R), a phase error ~ k ulp(|d|^2) / (2R) (at 48 m and R = 5 km,
// 5e-6 rad at X band), and s = e.d carries ~ulp(|d|) into the range bin
// (at 48 m and dr = 0.19 m, 1.5e-5 bins).  The limit is 48 m, or, when the
// antenna positions bound the range from below, what keeps both errors at
// those levels: 48 m x sqrt(R_min / 5 km), at most 256 range bins.
constexpr double kMaxTileRadius = 48.0;

double radius_limit(double rmin, double dr) {
  if (!(rmin > 5000.0)) return kMaxTileRadius;
  return std::max(kMaxTileRadius, std::min(kMaxTileRadius * std::sqrt(rmin / 5000.0), 256.0 * dr));
}

double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Smallest distance from an antenna position to the imaged area: the grid's
// parallelogram extended by one tile of px x py pixels per lane on every side
// (tiles overhang ragged edges).  0 without positions or for a degenerate grid.
double min_range(const sarbp_grid* g, int px, int py, const double* h_pos, int n) {
  if (!h_pos || n <= 0) return 0.0;
  const double tx = kTileX * px, ty = kWaves * py;
  const double a11 = dot3(g->ux, g->ux), a12 = dot3(g->ux, g->uy), a22 = dot3(g->uy, g->uy);
  const double det = a11 * a22 - a12 * a12;
  if (!(det > 0.0)) return 0.0;
  const double i0 = -tx, i1 = std::max(g->nx - 1, 0) + tx, j0 = -ty, j1 = std::max(g->ny - 1, 0) + ty;
  // in-plane distance from plane coordinates (i, j) to the segment (pi, pj) + t (qi, qj), t in [0, 1]
  auto seg = [&](double i, double j, double pi, double pj, double qi, double qj) {
    const double di = i - pi, dj = j - pj;
    const double qq = qi * qi * a11 + 2 * qi * qj * a12 + qj * qj * a22;
    double t = qq > 0 ? (di * qi * a11 + (di * qj + dj * qi) * a12 + dj * qj * a22) / qq : 0.0;
    t = std::min(1.0, std::max(0.0, t));
    const double ei = di - t * qi, ej = dj - t * qj;
    return std::sqrt(std::max(0.0, ei * ei * a11 + 2 * ei * ej * a12 + ej * ej * a22));
  };
  double best = INFINITY;
  for (int m = 0; m < n; ++m) {
    const double* a = h_pos + 3 * (size_t)m;
    const double w[3] = {a[0] - g->origin[0], a[1] - g->origin[1], a[2] - g->origin[2]};
    const double b1 = dot3(w, g->ux), b2 = dot3(w, g->uy);
    const double i = (b1 * a22 - b2 * a12) / det, j = (a11 * b2 - a12 * b1) / det;
    const double off2 = std::max(0.0, dot3(w, w) - (i * b1 + j * b2));  // squared distance to the plane
    double in = 0.0;
    if (i < i0 || i > i1 || j < j0 || j > j1)
      in = std::min(std::min(seg(i, j, i0, j0, i1 - i0, 0), seg(i, j, i0, j1, i1 - i0, 0)),
                    std::min(seg(i, j, i0, j0, 0, j1 - j0), seg(i, j, i1, j0, 0, j1 - j0)));
    best = std::min(best, std::sqrt(off2 + in * in));
  }
  return std::isfinite(best) ? best : 0.0;
}

// Largest |pixel - tile reference pixel| over a tile of px x py pixels per lane.
double tile_radius(const sarbp_grid* g, int px, int py) {
  const int tx = kTileX * px, ty = kWaves * py;
  double r = 0.0;
  for (int ci = 0; ci < 4; ++ci) {
    const int di = (ci & 1) ? tx / 2 - 1 : -tx / 2, dj = (ci & 2) ? ty / 2 - 1 : -ty / 2;
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = di * g->ux[k] + dj * g->uy[k];
    r = std::max(r, norm3(d));
  }
  return r;
}

int plan_window(const sarbp_grid* g, const sarbp_params* p, int px, int py, const double* h_pos,
                int* kinterp);

// The tuning with its tile shape reduced, if needed, to the largest compiled
// shape no wider and no taller than the requested one whose tiles stay within
// the radius limit (SARBP_PREC_REF only).  A grid too coarse for every shape
// runs the double-float differential instead, with one column per lane and
// the most rows per lane (at most the request's) whose range window fits LDS.
Fitted fit_tuning(sarbp_tuning t, int precision, const sarbp_grid* g, const sarbp_params* p,
                  const double* h_pos) {
  if (precision != SARBP_PREC_REF) return Fitted{t, precision};
  auto fits = [&](int px, int py) {
    return tile_radius(g, px, py) <= radius_limit(min_range(g, px, py, h_pos, p->n_pulses), p->dr);
  };
  if (fits(t.cols_per_lane, t.rows_per_lane)) return Fitted{t, kRef};
  int best_px = 0, best_py = 0;
  for (int px = t.cols_per_lane; px >= 1; px /= 2)
    for (int py = t.rows_per_lane; py >= 2; py /= 2)
      if (tile_shape_supported(precision, px, py) && px * py > best_px * best_py && fits(px, py)) {
        best_px = px;
        best_py = py;
      }
  if (best_px) {
    t.cols_per_lane = best_px;
    t.rows_per_lane = best_py;
    return Fitted{t, kRef};
  }
  t.cols_per_lane = 1;
  int py = std::max(2, std::min(8, t.rows_per_lane));
  while (py > 2) {
    int kinterp = 0;
    const int win = plan_window(g, p, 1, py, h_pos, &kinterp);
    if (win > 0 && lds_layout(kinterp, win, 1).total <= kLdsBudget) break;
    py /= 2;
  }
  t.rows_per_lane = py;
  return Fitted{t, kDF};
}

// Range extent (metres, half-width) of one tile seen from any pulse, and the
// minimum slant range over the scene.  Without positions: triangle inequality
// |R - Rc| <= |d|.  With positions: |R - Rc| <= |e.ux| hx + |e.uy| hy + q with
// |e.u| <= min(|u|, max|(p - a).u| / dist(a, image plane)), maxima over the
// image corners (linear in p), and q <= dmax^2 / Rmin (twice the leading
// term pp / (2 R), as margin).
struct Extent {
  double half_m;   // bound on |R - Rc| over a tile
  double dmax;     // max |pixel - reference| over a tile
  double rmin;     // lower bound on slant range (0 if unknown)
};

Extent tile_extent(const sarbp_grid* g, int px, int py, const double* h_pos, int n_pulses) {
  const int tx = kTileX * px, ty = kWaves * py;
  // Offsets of tile pixels from the reference pixel: di in [-tx/2, tx/2-1], dj in [-ty/2, ty/2-1].
  const int di_lo = -tx / 2, di_hi = tx / 2 - 1, dj_lo = -ty / 2, dj_hi = ty / 2 - 1;
  double dmax = 0.0;
  for (int ci = 0; ci < 4; ++ci) {
    const int di = (ci & 1) ? di_hi : di_lo, dj = (ci & 2) ? dj_hi : dj_lo;
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = di * g->ux[k] + dj * g->uy[k];
    dmax = std::max(dmax, norm3(d));
  }
  Extent e{dmax, dmax, 0.0};
  if (!h_pos || n_pulses <= 0) return e;
  // image plane normal
  double n[3] = {g->ux[1] * g->uy[2] - g->ux[2] * g->uy[1], g->ux[2] * g->uy[0] - g->ux[0] * g->uy[2],
                 g->ux[0] * g->uy[1] - g->ux[1] * g->uy[0]};
  const double nn = norm3(n);
  if (nn == 0.0) return e;
  for (double& v : n) v /= nn;
  double corners[4][3];
  for (int ci = 0; ci < 4; ++ci) {
    const double ix = (ci & 1) ? std::max(g->nx - 1, 0) + tx : -tx;
    const double iy = (ci & 2) ? std::max(g->ny - 1, 0) + ty : -ty;
    for (int k = 0; k < 3; ++k) corners[ci][k] = g->origin[k] + ix * g->ux[k] + iy * g->uy[k];
  }
  const double lux = norm3(g->ux), luy = norm3(g->uy);
  double amax = 0.0, bmax = 0.0, rmin = INFINITY;
  int side = 0;
  for (int m = 0; m < n_pulses; ++m) {
    const double* a = h_pos + 3 * (size_t)m;
    const double sep = (a[0] - g->origin[0]) * n[0] + (a[1] - g->origin[1]) * n[1] + (a[2] - g->origin[2]) * n[2];
    const int sd = sep > 0 ? 1 : -1;
    if (side == 0) side = sd;
    if (sd != side || std::fabs(sep) == 0.0) return e;  // antennas on both sides: no bound
    double nx_ = 0.0, ny_ = 0.0;
    for (int ci = 0; ci < 4; ++ci) {
      double px = 0.0, py = 0.0;
      for (int k = 0; k < 3; ++k) {
        px += (corners[ci][k] - a[k]) * g->ux[k];
        py += (corners[ci][k] - a[k]) * g->uy[k];
      }
      nx_ = std::max(nx_, std::fabs(px));
      ny_ = std::max(ny_, std::fabs(py));
    }
    amax = std::max(amax, std::min(lux, nx_ / std::fabs(sep)));
    bmax = std::max(bmax, std::min(luy, ny_ / std::fabs(sep)));
    rmin = std::min(rmin, std::fabs(sep));
  }
  const double hx = tx / 2, hy = ty / 2;
  const double q = dmax * dmax / rmin;
  e.half_m = std::min(dmax, amax * hx + bmax * hy + q);
  e.rmin = rmin;
  return e;
}

// The size class of a job: the range window (samples) a 64 x 32-pixel tile
// needs, planned as plan_args plans it (with the host positions when given).
// The window sets a stage's LDS footprint per pulse and the sinc8 precompute
// per pixel, so it decides the pulses per stage and the tile shape that win;
// it grows with the pixel spacing over dr, with the aperture and with
// proximity.
int size_class(const sarbp_grid* g, const sarbp_params* p, const double* h_pos) {
  const Extent ext = tile_extent(g, 1, 8, h_pos, p->n_pulses);
  const double win = 2.0 * (std::ceil(ext.half_m / p->dr) + 6.0) + 2.0;
  const double bounds[SARBP_SIZE_CLASSES - 1] = {80.0, 120.0, 176.0, 272.0};
  int c = 0;
  while (c < SARBP_SIZE_CLASSES - 1 && win >= bounds[c]) ++c;
  return c;
}

// LDS range window (samples) of a tile of px x py pixels per lane and the
// interpolation kernel (kernel_interp) it runs; 0 if the window exceeds 8192
// samples.  Sinc8 runs half bins unless the tile is coarse against the range
// bins: per pixel-pulse the half-bin kernel issues 4 FMA fewer in the pixel
// loop and ~0.6 * 60 / (tile pixels) more issue slots per window position
// (about 60% of the window's positions are touched per pulse; 60 slots is
// where measured grids cross over: 64 x 16 tiles at 0.5 m, 100-120-sample
// windows, run 1.5-1.7% faster in half bins, at 0.7 m, 152 samples, 9%
// slower).
int plan_window(const sarbp_grid* g, const sarbp_params* p, int px, int py, const double* h_pos,
                int* kinterp) {
  const Extent ext = tile_extent(g, px, py, h_pos, p->n_pulses);
  // half window in samples: extent + sinc taps (3 left / 4 right) + 2 margin
  const double half_bins = std::ceil(ext.half_m / p->dr) + 6.0;
  *kinterp = p->interp;
  if (!(half_bins <= 4096.0)) return 0;
  const int win = ((2 * (int)half_bins + 2) + 3) & ~3;  // multiple of 4 samples
  if (p->interp == SARBP_INTERP_SINC8) {
    const double tile_px = (double)kTileX * px * kWaves * py;
    *kinterp = 0.6 * 60.0 * (win - 7) < 4.0 * tile_px ? kSinc8 : kSinc8W;
  }
  return win;
}

// Inverse quantum 2^e of the DF direction cosines: every alpha_h di + beta_h dj
// of a tile (|di| <= tx / 2, |dj| <= ty / 2, |alpha_h| <= |ux|, |beta_h| <= |uy|)
// is then an integer multiple of 2^-e below 2^(23 - e), exact in FP32.
double df_quant(const sarbp_grid* g, int px, int py) {
  const double span = norm3(g->ux) * (kTileX * px / 2) + norm3(g->uy) * (kWaves * py / 2) + 1.0;
  return std::ldexp(1.0, 23 - (int)std::ceil(std::log2(span)));
}

// Fills a launch descriptor.  Caller holds ctx->mu and the device guard.
sarbp_status plan_args(const double* h_pos, const sarbp_grid* g,
                       const sarbp_params* p, const sarbp_tuning& tune, int kprec, BpArgs* a,
                       sarbp_plan_info* info, bool batched) {
  const int px = tune.cols_per_lane, py = tune.rows_per_lane;
  if (!tile_shape_supported(kprec, px, py))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "unsupported tile shape: %d cols x %d rows per lane", px, py);
  const Extent ext = tile_extent(g, px, py, h_pos, p->n_pulses);
  const double rlim = radius_limit(min_range(g, px, py, h_pos, p->n_pulses), p->dr);
  if (kprec == kRef && ext.dmax > rlim)
    return fail(SARBP_ERR_UNSUPPORTED,
                "tile of %d x %d pixels spans %.1f m from its centre (limit %.0f m): use fewer pixels per lane",
                kTileX * px, kWaves * py, ext.dmax, rlim);
  int kinterp = p->interp;
  const int win = plan_window(g, p, px, py, h_pos, &kinterp);
  if (win == 0)
    return fail(SARBP_ERR_UNSUPPORTED, "range window of a tile too long (%.0f m): pixel spacing is too coarse for dr",
                2 * ext.half_m);
  int P = std::max(1, tune.pulses_per_stage);
  while (P > 1 && lds_layout(kinterp, win, P).total > kLdsBudget) --P;
  if (lds_layout(kinterp, win, P).total > kLdsBudget)
    return fail(SARBP_ERR_UNSUPPORTED, "range window of %d samples does not fit LDS", win);
  // A tuning's stage depth is set for its class's typical window; a longer
  // window can push the stage's LDS over a workgroups-per-CU step (measured on
  // an 84-sample window, 1 x 4 px per lane: 6 pulses per stage at 3 workgroups
  // per CU 541 ms, 5 pulses at 4 per CU 495 ms).  Take the shallower stage
  // while it gains a workgroup per CU.  A plain launch stops at 4 pulses:
  // below that its barrier per stage costs more than the occupancy gains (a
  // 152-sample window, 2048^2 x 2048 pulses: 3 pulses at 3 per CU 11.4 ms, 2
  // at 4 per CU 12.3 ms; a 248-sample window: 2 pulses 16.7 ms, 1 pulse
  // 18.0 ms).  A batched launch's persistent workgroups gain more from
  // occupancy and fit down to 1 pulse (64 mixed jobs: 191.7 ms that way,
  // 217.0 ms stopping at 4, same box).  The depth changes no result bit:
  // each pixel adds its pulses in pulse order whatever the stages.
  const int min_stage = batched ? 1 : 4;
  auto per_cu = [&](int q) {
    return kernel_occupancy(kinterp, kprec, p->dtype, px, py, lds_layout(kinterp, win, q).total);
  };
  while (P > min_stage && per_cu(P - 1) > per_cu(P)) --P;

  std::memset(a, 0, sizeof *a);
  for (int k = 0; k < 3; ++k) {
    a->origin[k] = g->origin[k];
    a->ux[k] = g->ux[k];
    a->uy[k] = g->uy[k];
  }
  a->r0 = p->r0;
  a->dr = p->dr;
  a->inv_dr = 1.0 / p->dr;
  a->k_rev = 2.0 * p->fc / p->c;
  a->inv_dr_f = (float)a->inv_dr;
  a->k_rev_f = (float)a->k_rev;
  a->k_rev_lo = (float)(a->k_rev - (double)a->k_rev_f);
  a->s_quant = df_quant(g, px, py);
  a->n_pulses = p->n_pulses;
  a->n_samples = p->n_samples;
  a->nx = g->nx;
  a->ny = g->ny;
  const int tx = kTileX * px, ty = kWaves * py;
  a->tiles_x = (g->nx + tx - 1) / tx;
  a->tiles_y = (g->ny + ty - 1) / ty;
  a->win = win;
  a->half = win / 2;
  a->pulses_per_stage = P;
  a->pulse_begin = 0;
  a->pulse_end = p->n_pulses;
  a->accumulate = p->accumulate;
  a->dmax = ext.dmax;
  a->order_tol = 1e-5;
  a->kernel_interp = kinterp;
  a->kernel_prec = kprec;
  sinc8_kernel_rows(kinterp, p->kaiser_beta, a->sinc_poly);
  if (info) {
    info->tiles_x = a->tiles_x;
    info->tiles_y = a->tiles_y;
    info->tile_nx = tx;
    info->tile_ny = ty;
    info->window = win;
    info->pulses_per_stage = P;
    info->rows_per_lane = py;
    info->cols_per_lane = px;
    info->lds_bytes = (int32_t)lds_layout(kinterp, win, P).total;
    info->sinc8_halves = kinterp == kSinc8 ? 2 : (kinterp == kSinc8W ? 1 : 0);
    info->precision = kprec;
    const double rm = ext.rmin > ext.dmax ? ext.rmin - ext.dmax : 0.0;
    const double d2 = ext.dmax * ext.dmax;
    info->range_order = rm > 0 && 2 * kPi * a->k_rev * 0.625 * d2 * d2 / (rm * rm * rm) <= a->order_tol ? 1 : 0;
    info->max_phase_err_rad = rm > 0 ? 2 * kPi * a->k_rev * 0.625 * d2 * d2 / (rm * rm * rm) : INFINITY;
  }
  return SARBP_OK;
}

GroupKey group_key(const BpArgs& a, const sarbp_tuning& t, int dtype) {
  const size_t shm = lds_layout(a.kernel_interp, a.win, a.pulses_per_stage).total;
  int occ = kernel_occupancy(a.kernel_interp, a.kernel_prec, dtype, t.cols_per_lane, t.rows_per_lane, shm);
  if (occ <= 0) occ = (int)std::max<size_t>(1, kCuLds / (shm + kBatchStaticLds));
  return GroupKey{t.cols_per_lane, t.rows_per_lane, a.kernel_interp, a.kernel_prec, occ};
}

int assign_groups(const GroupKey* k, int n, int max_groups, int* out) {
  for (int pass = 0; pass < 2; ++pass) {
    const bool by_occ = pass == 0;
    int ng = 0;
    bool over = false;
    for (int i = 0; i < n && !over; ++i) {
      int g = -1;
      for (int j = 0; j < i && g < 0; ++j)
        if (k[j].px == k[i].px && k[j].py == k[i].py && k[j].kinterp == k[i].kinterp &&
            k[j].kprec == k[i].kprec && (!by_occ || k[j].occ == k[i].occ))
          g = out[j];
      if (g < 0) {
        if (ng == max_groups)
          over = true;
        else
          g = ng++;
      }
      out[i] = g;
    }
    if (!over) return ng;
  }
  return -1;
}
}  // namespace sarbp
