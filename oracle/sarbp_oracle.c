/*
 * sarbp_oracle.c -- FP64 CPU back-projection oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Restates the BP loops of code/SAR_BP/matlab/BPA_SAR.m:133-153 and :178-197
 * (reference repository) in plain C, with the deviations listed in
 * sarbp_oracle.h.  Nothing in the product library links this file.
 */
#include "sarbp_oracle.h"

#include <math.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

double orc_bessel_i0(double x) {
  /* I0(x) = sum_k ((x/2)^k / k!)^2 ; terms decay once k > x/2. */
  double q = 0.25 * x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; ++k) {
    term *= q / ((double)k * (double)k);
    sum += term;
    if (term < 1e-18 * sum) break;
  }
  return sum;
}

static double orc_sinc(double x) {
  if (x == 0.0) return 1.0;
  double px = ORC_PI * x;
  return sin(px) / px;
}

double orc_kaiser_sinc8(double x, double beta) {
  double ax = fabs(x);
  if (ax >= 4.0) return 0.0;
  double t = x / 4.0;
  double w = orc_bessel_i0(beta * sqrt(1.0 - t * t)) / orc_bessel_i0(beta);
  return orc_sinc(x) * w;
}

static inline void orc_sample(const float* p32, const double* p64, int ns,
                              long k, double* re, double* im) {
  if (k < 0 || k >= ns) { *re = 0.0; *im = 0.0; return; }
  if (p32) { *re = (double)p32[2 * k]; *im = (double)p32[2 * k + 1]; }
  else     { *re = p64[2 * k];         *im = p64[2 * k + 1]; }
}

/* MATLAB round(): halves away from zero (BPA_SAR.m:141). */
static inline double orc_round_half_away(double u) {
  return (u >= 0.0) ? floor(u + 0.5) : -floor(-u + 0.5);
}

void orc_interp(const float* prof32, const double* prof64, int ns, double u,
                const orc_params* p, double out[2]) {
  double re = 0.0, im = 0.0, sr, si;
  if (p->interp == ORC_INTERP_NEAREST) {
    long k = (long)orc_round_half_away(u);
    orc_sample(prof32, prof64, ns, k, &re, &im);
  } else if (p->interp == ORC_INTERP_LINEAR) {
    double fl = floor(u);
    long k = (long)fl;
    double w = u - fl;
    double ar, ai, br, bi;
    orc_sample(prof32, prof64, ns, k, &ar, &ai);
    orc_sample(prof32, prof64, ns, k + 1, &br, &bi);
    re = (1.0 - w) * ar + w * br;
    im = (1.0 - w) * ai + w * bi;
  } else {
    double fl = floor(u);
    long k0 = (long)fl;
    for (long k = k0 - 3; k <= k0 + 4; ++k) {
      double wk = orc_kaiser_sinc8(u - (double)k, p->kaiser_beta);
      orc_sample(prof32, prof64, ns, k, &sr, &si);
      re += wk * sr;
      im += wk * si;
    }
  }
  out[0] = re;
  out[1] = im;
}

void orc_backproject(const float* prof32, const double* prof64, int n_pulses,
                     int n_samples, const double* pos, const double* origin,
                     const double* ux, const double* uy, int nx, int row0,
                     int row1, const orc_params* p, double* img, int accumulate,
                     int n_threads) {
  const double k_rev = 2.0 * p->fc / p->c; /* revolutions of phase per metre of range */
  const double inv_dr = 1.0 / p->dr;
  const long npix = (long)(row1 - row0) * nx;
#ifdef _OPENMP
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(n_threads)
#endif
  for (long q = 0; q < npix; ++q) {
    const int iy = row0 + (int)(q / nx), ix = (int)(q % nx);
    const double px = origin[0] + ix * ux[0] + iy * uy[0];
    const double py = origin[1] + ix * ux[1] + iy * uy[1];
    const double pz = origin[2] + ix * ux[2] + iy * uy[2];
    double acc_r = 0.0, acc_i = 0.0;
    for (int m = 0; m < n_pulses; ++m) {
      const double dx = pos[3 * m] - px, dy = pos[3 * m + 1] - py,
                   dz = pos[3 * m + 2] - pz;
      const double R = sqrt(dx * dx + dy * dy + dz * dz);
      const double u = (R - p->r0) * inv_dr;
      double v[2];
      const float* row32 = prof32 ? prof32 + 2 * (size_t)m * n_samples : NULL;
      const double* row64 = prof64 ? prof64 + 2 * (size_t)m * n_samples : NULL;
      orc_interp(row32, row64, n_samples, u, p, v);
      /* exp(+j 4 pi fc R / c) = exp(+j 2 pi phi), phi = 2 fc R / c revolutions;
         the integer part of phi is removed exactly before the trig call. */
      const double phi = k_rev * R;
      const double fr = phi - floor(phi);
      const double cs = cos(2.0 * ORC_PI * fr), sn = sin(2.0 * ORC_PI * fr);
      acc_r += v[0] * cs - v[1] * sn;
      acc_i += v[0] * sn + v[1] * cs;
    }
    if (accumulate) {
      img[2 * q] += acc_r;
      img[2 * q + 1] += acc_i;
    } else {
      img[2 * q] = acc_r;
      img[2 * q + 1] = acc_i;
    }
  }
}

/* ---- FP32 CPU baseline (not part of the oracle proper) -------------------
 * A conventional single-precision CPU BP: FP32 geometry and phase, and the
 * Kaiser-sinc kernel read from a table (1024 entries per sample, linear
 * interpolation between entries) instead of evaluating the Bessel series. */

#define ORC_TAB_Q 1024

void orc_backproject_f32(const float* prof32, int n_pulses, int n_samples,
                         const double* pos, const double* origin,
                         const double* ux, const double* uy, int nx, int row0,
                         int row1, const orc_params* p, float* img,
                         int n_threads) {
  const float k_rev = (float)(2.0 * p->fc / p->c);
  const float inv_dr = (float)(1.0 / p->dr);
  const float r0 = (float)p->r0;
  static float tab[8 * ORC_TAB_Q + 2];
  for (int i = 0; i <= 8 * ORC_TAB_Q + 1; ++i)
    tab[i] = (float)orc_kaiser_sinc8((double)i / ORC_TAB_Q - 4.0, p->kaiser_beta);
  const long npix = (long)(row1 - row0) * nx;
#ifdef _OPENMP
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(n_threads)
#endif
  for (long q = 0; q < npix; ++q) {
    const int iy = row0 + (int)(q / nx), ix = (int)(q % nx);
    const float px = (float)(origin[0] + ix * ux[0] + iy * uy[0]);
    const float py = (float)(origin[1] + ix * ux[1] + iy * uy[1]);
    const float pz = (float)(origin[2] + ix * ux[2] + iy * uy[2]);
    float acc_r = 0.0f, acc_i = 0.0f;
    for (int m = 0; m < n_pulses; ++m) {
      const float dx = (float)pos[3 * m] - px, dy = (float)pos[3 * m + 1] - py,
                  dz = (float)pos[3 * m + 2] - pz;
      const float R = sqrtf(dx * dx + dy * dy + dz * dz);
      const float u = (R - r0) * inv_dr;
      const float* row = prof32 + 2 * (size_t)m * n_samples;
      float vr = 0.0f, vi = 0.0f;
      const float fl = floorf(u);
      const long k0 = (long)fl;
      if (p->interp == ORC_INTERP_NEAREST) {
        long k = (long)((u >= 0.0f) ? floorf(u + 0.5f) : -floorf(-u + 0.5f));
        if (k >= 0 && k < n_samples) { vr = row[2 * k]; vi = row[2 * k + 1]; }
      } else if (p->interp == ORC_INTERP_LINEAR) {
        const float w = u - fl;
        if (k0 >= 0 && k0 < n_samples) { vr += (1.0f - w) * row[2 * k0]; vi += (1.0f - w) * row[2 * k0 + 1]; }
        if (k0 + 1 >= 0 && k0 + 1 < n_samples) { vr += w * row[2 * k0 + 2]; vi += w * row[2 * k0 + 3]; }
      } else {
        const float f = u - fl;
        for (int t = 0; t < 8; ++t) {
          const long k = k0 - 3 + t;
          if (k < 0 || k >= n_samples) continue;
          /* x = u - k = f + 3 - t, table index (x + 4) * Q */
          const float xi = (f + (float)(7 - t)) * (float)ORC_TAB_Q;
          const int i0 = (int)xi;
          const float a = xi - (float)i0;
          const float wk = tab[i0] + a * (tab[i0 + 1] - tab[i0]);
          vr += wk * row[2 * k];
          vi += wk * row[2 * k + 1];
        }
      }
      const float phi = k_rev * R;
      const float fr = phi - floorf(phi);
      const float cs = cosf(2.0f * (float)ORC_PI * fr), sn = sinf(2.0f * (float)ORC_PI * fr);
      acc_r += vr * cs - vi * sn;
      acc_i += vr * sn + vi * cs;
    }
    img[2 * q] = acc_r;
    img[2 * q + 1] = acc_i;
  }
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
