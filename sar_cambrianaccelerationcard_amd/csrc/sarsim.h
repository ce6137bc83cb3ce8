/*
 * sarsim.h -- FP64 stripmap SAR scene simulator: antenna tracks, raw LFM echoes
 * of point scatterers, FFT matched-filter range compression and amplitude-
 * preserving FFT zero-pad range upsampling.  Produces the range-compressed
 * pulses that back-projection consumes.
 *
 * Follows code/SAR_BP/matlab/BPA_SAR.m (reference repository):
 *   :81-88   echo model  s(t) = sum_k a_k exp(j(pi Kr (t-tau_k)^2 - 4 pi R_k/lambda))
 *            gated to 0 < t - tau_k < Tp, and (optionally) by the azimuth beam
 *            |x_pulse - x_target| < La/2 (also code/SAR_BP/CUDA/BPA_CUDA_kernels.cu:27)
 *   :93-100  matched filter IFFT(FFT(s) . conj(FFT(h))), h = reference chirp
 *            starting at the window start t0
 *   :106-111 upsampling by zero insertion after bin floor((N+1)/2)
 * Deviations: the chirp is centred at baseband, exp(j pi Kr (t - tau - Tp/2)^2),
 * so that a bandwidth B up to fs fits the complex sampling band (the script's
 * chirp sweeps [0, B], which only fits because it samples at Fr = 2 B,
 * BPA_SAR.m:30,87); the matched filter is a linear (zero-padded) correlation, and the
 * upsampled IFFT is scaled by 1/N instead of 1/(L N) so that upsampled samples
 * at multiples of L equal the compressed samples (the script divides the
 * amplitude by L, BPA_SAR.m:111).
 */
#ifndef SARSIM_H
#define SARSIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double c;            /* propagation speed, m/s                     */
  double fc;           /* carrier, Hz                                */
  double bandwidth;    /* chirp bandwidth B, Hz                      */
  double pulse_width;  /* chirp duration Tp, s (Kr = B / Tp)         */
  double fs;           /* raw complex sampling rate, Hz              */
  double t0;           /* fast time of raw sample 0, s               */
  int32_t n_raw;       /* raw samples per pulse                      */
  int32_t upsample;    /* L >= 1                                     */
  int32_t crop_offset; /* first upsampled sample kept                */
  int32_t n_out;       /* upsampled samples kept per pulse           */
  double beam_half_width; /* azimuth gate |x_pulse - x_target| < this; <= 0: off */
} sarsim_radar;

/* Size of the FFT used for range compression: the smallest power of two
 * >= n_raw + chirp samples. */
int32_t sarsim_fft_size(const sarsim_radar* r);

/* In-place radix-2 complex FP64 FFT (interleaved), n a power of two.
 * inverse != 0 computes the unscaled inverse transform.  Returns 0 on success. */
int sarsim_fft(double* data, int32_t n, int inverse);

/* Straight track along +x: pulse m at (x_center + v*(m - (n-1)/2)/prf, y, z). */
void sarsim_track_straight(double* pos, int32_t n_pulses, double v, double prf,
                           double x_center, double y, double z);

/* Adds seeded Gaussian deviations to the y and z coordinates: white N(0,1)
 * smoothed by a centred moving average of `lowpass_pulses` pulses, then scaled
 * to standard deviation sigma over the track. */
void sarsim_track_perturb(double* pos, int32_t n_pulses, double sigma,
                          int32_t lowpass_pulses, uint64_t seed);

/* Raw echoes: out[n_pulses][n_raw] complex128 (interleaved), overwritten.
 * targets[k] = {x, y, z, amp_re, amp_im}. */
void sarsim_echo_raw(const double* pos, int32_t n_pulses, const double* targets,
                     int32_t n_targets, const sarsim_radar* r, double* out,
                     int n_threads);

/* Adds complex AWGN of total variance sigma2 (sigma2/2 per component). */
void sarsim_add_noise(double* data, int64_t n_complex, double sigma2,
                      uint64_t seed);

/* Range compression + upsampling + crop of raw echoes [n_pulses][n_raw].
 * out64 ([n_pulses][n_out] complex128) and/or out32 (complex64) may be NULL. */
int sarsim_compress(const double* raw, int32_t n_pulses, const sarsim_radar* r,
                    double* out64, float* out32, int n_threads);

/* Adds, to out[n_pulses][n_out] (complex128, range grid of sarsim_out_r0 /
 * sarsim_out_dr), the compressed response of point scatterers
 * scat[k] = {x, y, z, amp_re, amp_im}:
 *   amp * Tp fs * sinc(2 B (r - R) / c) * exp(-j 4 pi fc R / c)
 * over the output samples within +-half_width of the scatterer's range, with a
 * Hann taper over that span.  This is the band-limited compressed pulse of the
 * centred chirp near its peak (unit chirp: Tp fs at the peak), without its
 * range sidelobes beyond half_width samples. */
void sarsim_compressed_points(const double* pos, int32_t n_pulses,
                              const double* scat, int32_t n_scat,
                              const sarsim_radar* r, int32_t half_width,
                              double* out, int n_threads);

/* Slant range of output sample 0 and output sample spacing, m. */
double sarsim_out_r0(const sarsim_radar* r);
double sarsim_out_dr(const sarsim_radar* r);

#ifdef __cplusplus
}
#endif
#endif
