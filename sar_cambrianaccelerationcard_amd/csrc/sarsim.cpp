// sarsim.cpp -- FP64 SAR scene simulator (tracks, raw LFM echoes, FFT matched
// filter, zero-pad upsampling).  Semantics and deviations: sarsim.h.
#include "sarsim.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;

int ilog2_exact(int32_t n) {
  if (n <= 0 || (n & (n - 1))) return -1;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// Twiddles w_k = exp(-2 pi j k / n), k < n/2, each evaluated directly (no
// recurrence) so every factor carries ~1 ulp of error.
void fft_inplace(std::complex<double>* a, int32_t n, bool inverse,
                 const std::vector<std::complex<double>>& tw) {
  const int lg = ilog2_exact(n);
  for (int32_t i = 1, j = 0; i < n; ++i) {
    int32_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int s = 1; s <= lg; ++s) {
    const int32_t len = 1 << s, half = len >> 1, step = n / len;
    for (int32_t i = 0; i < n; i += len) {
      for (int32_t k = 0; k < half; ++k) {
        std::complex<double> w = tw[(size_t)k * step];
        if (inverse) w = std::conj(w);
        const std::complex<double> u = a[i + k], v = a[i + k + half] * w;
        a[i + k] = u + v;
        a[i + k + half] = u - v;
      }
    }
  }
}

std::vector<std::complex<double>> twiddles(int32_t n) {
  std::vector<std::complex<double>> tw((size_t)n / 2 + 1);
  for (int32_t k = 0; k <= n / 2; ++k) {
    const double a = -2.0 * kPi * (double)k / (double)n;
    tw[k] = std::complex<double>(std::cos(a), std::sin(a));
  }
  return tw;
}

int chirp_samples(const sarsim_radar* r) {
  return (int)std::ceil(r->pulse_width * r->fs);
}

}  // namespace

extern "C" {

int32_t sarsim_fft_size(const sarsim_radar* r) {
  int32_t need = r->n_raw + chirp_samples(r);
  int32_t n = 1;
  while (n < need) n <<= 1;
  return n;
}

int sarsim_fft(double* data, int32_t n, int inverse) {
  if (ilog2_exact(n) < 0) return 1;
  auto tw = twiddles(n);
  fft_inplace(reinterpret_cast<std::complex<double>*>(data), n, inverse != 0, tw);
  return 0;
}

void sarsim_track_straight(double* pos, int32_t n_pulses, double v, double prf,
                           double x_center, double y, double z) {
  for (int32_t m = 0; m < n_pulses; ++m) {
    pos[3 * m + 0] = x_center + v * ((double)m - 0.5 * (n_pulses - 1)) / prf;
    pos[3 * m + 1] = y;
    pos[3 * m + 2] = z;
  }
}

void sarsim_track_perturb(double* pos, int32_t n_pulses, double sigma,
                          int32_t lowpass_pulses, uint64_t seed) {
  if (n_pulses <= 0) return;
  if (lowpass_pulses < 1) lowpass_pulses = 1;
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int axis = 1; axis <= 2; ++axis) {
    const int32_t pad = lowpass_pulses;
    std::vector<double> white((size_t)n_pulses + 2 * pad);
    for (double& w : white) w = nd(rng);
    std::vector<double> sm((size_t)n_pulses);
    double mean = 0.0;
    for (int32_t m = 0; m < n_pulses; ++m) {
      double s = 0.0;
      const int32_t c = m + pad, lo = c - lowpass_pulses / 2;
      for (int32_t t = 0; t < lowpass_pulses; ++t) s += white[(size_t)lo + t];
      sm[m] = s / lowpass_pulses;
      mean += sm[m];
    }
    mean /= n_pulses;
    double var = 0.0;
    for (double& s : sm) { s -= mean; var += s * s; }
    var /= n_pulses;
    const double scale = var > 0.0 ? sigma / std::sqrt(var) : 0.0;
    for (int32_t m = 0; m < n_pulses; ++m) pos[3 * m + axis] += scale * sm[m];
  }
}

void sarsim_echo_raw(const double* pos, int32_t n_pulses, const double* targets,
                     int32_t n_targets, const sarsim_radar* r, double* out,
                     int n_threads) {
  const double kr = r->bandwidth / r->pulse_width;
  const double k_rev = 2.0 * r->fc / r->c;  // phase revolutions per metre of range
#ifdef _OPENMP
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(n_threads)
#endif
  for (int32_t m = 0; m < n_pulses; ++m) {
    double* row = out + 2 * (size_t)m * r->n_raw;
    std::memset(row, 0, sizeof(double) * 2 * (size_t)r->n_raw);
    for (int32_t k = 0; k < n_targets; ++k) {
      const double* t = targets + 5 * (size_t)k;
      if (r->beam_half_width > 0.0 &&
          std::fabs(pos[3 * m] - t[0]) >= r->beam_half_width)
        continue;
      const double dx = pos[3 * m] - t[0], dy = pos[3 * m + 1] - t[1],
                   dz = pos[3 * m + 2] - t[2];
      const double R = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double tau = 2.0 * R / r->c;
      // carrier term exp(-j 4 pi R / lambda), integer revolutions removed exactly
      const double cr = k_rev * R, crf = cr - std::floor(cr);
      const std::complex<double> amp =
          std::complex<double>(t[3], t[4]) *
          std::complex<double>(std::cos(2.0 * kPi * crf), -std::sin(2.0 * kPi * crf));
      const double nlo = std::ceil((tau - r->t0) * r->fs);
      const double nhi = std::floor((tau + r->pulse_width - r->t0) * r->fs);
      const int32_t n0 = (int32_t)std::max(0.0, nlo);
      const int32_t n1 = (int32_t)std::min((double)r->n_raw - 1, nhi);
      for (int32_t n = n0; n <= n1; ++n) {
        const double tk = r->t0 + (double)n / r->fs - tau;
        if (!(tk > 0.0 && tk < r->pulse_width)) continue;
        const double tc = tk - 0.5 * r->pulse_width;
        const double ph = 0.5 * kr * tc * tc;  // revolutions of pi Kr (tk - Tp/2)^2
        const double phf = ph - std::floor(ph);
        const std::complex<double> e(std::cos(2.0 * kPi * phf), std::sin(2.0 * kPi * phf));
        const std::complex<double> v = amp * e;
        row[2 * n] += v.real();
        row[2 * n + 1] += v.imag();
      }
    }
  }
}

void sarsim_add_noise(double* data, int64_t n_complex, double sigma2,
                      uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, std::sqrt(0.5 * sigma2));
  for (int64_t i = 0; i < 2 * n_complex; ++i) data[i] += nd(rng);
}

int sarsim_compress(const double* raw, int32_t n_pulses, const sarsim_radar* r,
                    double* out64, float* out32, int n_threads) {
  const int32_t nf = sarsim_fft_size(r);
  const int32_t L = r->upsample < 1 ? 1 : r->upsample;
  const int32_t nu = nf * L;
  if (r->crop_offset < 0 || r->crop_offset + r->n_out > nu) return 1;
  const double kr = r->bandwidth / r->pulse_width;
  const auto tw = twiddles(nf);
  const auto twu = twiddles(nu);
  // Reference chirp starting at the window start (BPA_SAR.m:93-94) and its FFT.
  std::vector<std::complex<double>> H((size_t)nf, 0.0);
  for (int32_t n = 0; n < nf && n < r->n_raw; ++n) {
    const double t = (double)n / r->fs;
    if (!(t > 0.0 && t < r->pulse_width)) continue;
    const double tc = t - 0.5 * r->pulse_width;
    const double ph = 0.5 * kr * tc * tc;
    const double phf = ph - std::floor(ph);
    H[n] = std::complex<double>(std::cos(2.0 * kPi * phf), std::sin(2.0 * kPi * phf));
  }
  fft_inplace(H.data(), nf, false, tw);
  const int32_t half = (nf + 1) / 2;  // bins [0, half) are kept at the front
  const double scale = 1.0 / (double)nf;
#ifdef _OPENMP
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#endif
  {
    std::vector<std::complex<double>> a((size_t)nf), b((size_t)nu);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
    for (int32_t m = 0; m < n_pulses; ++m) {
      const double* row = raw + 2 * (size_t)m * r->n_raw;
      for (int32_t n = 0; n < nf; ++n)
        a[n] = n < r->n_raw ? std::complex<double>(row[2 * n], row[2 * n + 1]) : 0.0;
      fft_inplace(a.data(), nf, false, tw);
      std::fill(b.begin(), b.end(), std::complex<double>(0.0, 0.0));
      for (int32_t k = 0; k < half; ++k) b[k] = a[k] * std::conj(H[k]);
      for (int32_t k = half; k < nf; ++k) b[(size_t)k + (nu - nf)] = a[k] * std::conj(H[k]);
      fft_inplace(b.data(), nu, true, twu);
      for (int32_t i = 0; i < r->n_out; ++i) {
        const std::complex<double> v = b[(size_t)r->crop_offset + i] * scale;
        const size_t o = 2 * ((size_t)m * r->n_out + i);
        if (out64) { out64[o] = v.real(); out64[o + 1] = v.imag(); }
        if (out32) { out32[o] = (float)v.real(); out32[o + 1] = (float)v.imag(); }
      }
    }
  }
  return 0;
}

void sarsim_compressed_points(const double* pos, int32_t n_pulses,
                              const double* scat, int32_t n_scat,
                              const sarsim_radar* r, int32_t half_width,
                              double* out, int n_threads) {
  const double r0 = sarsim_out_r0(r), dr = sarsim_out_dr(r);
  const double k_rev = 2.0 * r->fc / r->c;
  const double peak = r->pulse_width * r->fs;
  const double bw = 2.0 * r->bandwidth / r->c;  // sinc argument per metre
  const double hw = (double)half_width;
#ifdef _OPENMP
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(n_threads)
#endif
  for (int32_t m = 0; m < n_pulses; ++m) {
    double* row = out + 2 * (size_t)m * r->n_out;
    const double* a = pos + 3 * (size_t)m;
    for (int32_t k = 0; k < n_scat; ++k) {
      const double* s = scat + 5 * (size_t)k;
      const double dx = a[0] - s[0], dy = a[1] - s[1], dz = a[2] - s[2];
      const double R = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double u = (R - r0) / dr;
      const int32_t lo = (int32_t)std::max(0.0, std::ceil(u - hw));
      const int32_t hi = (int32_t)std::min((double)r->n_out - 1, std::floor(u + hw));
      if (lo > hi) continue;
      const double cr = k_rev * R, crf = cr - std::floor(cr);
      const std::complex<double> amp =
          std::complex<double>(s[3], s[4]) * peak *
          std::complex<double>(std::cos(2.0 * kPi * crf), -std::sin(2.0 * kPi * crf));
      for (int32_t i = lo; i <= hi; ++i) {
        const double x = (r0 + i * dr - R);
        const double arg = kPi * bw * x;
        const double sn = (arg == 0.0) ? 1.0 : std::sin(arg) / arg;
        const double taper = 0.5 + 0.5 * std::cos(kPi * (i - u) / (hw + 1.0));
        const std::complex<double> v = amp * (sn * taper);
        row[2 * i] += v.real();
        row[2 * i + 1] += v.imag();
      }
    }
  }
}

double sarsim_out_r0(const sarsim_radar* r) {
  const int32_t L = r->upsample < 1 ? 1 : r->upsample;
  return 0.5 * r->c * (r->t0 + (double)r->crop_offset / (L * r->fs));
}

double sarsim_out_dr(const sarsim_radar* r) {
  const int32_t L = r->upsample < 1 ? 1 : r->upsample;
  return 0.5 * r->c / (L * r->fs);
}

}  // extern "C"
