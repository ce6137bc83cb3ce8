/*
 * sarbp_oracle.h -- FP64 CPU restatement of time-domain back-projection (BP).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the GPU imager is compared
 * against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it; the product library never links or calls it.
 *
 * Semantics follow code/SAR_BP/matlab/BPA_SAR.m (reference repository):
 *   - BPA_SAR.m:139,182   R_m = |antenna_m - pixel|
 *   - BPA_SAR.m:141,184   range bin = (2R/c - t0) * Fr   == (R - r0) / dr
 *   - BPA_SAR.m:144-147   gather of the range-compressed sample (MATLAB: nearest
 *                         bin by round(); here also linear and 8-tap Kaiser sinc)
 *   - BPA_SAR.m:148,192   rotation by exp(+j 4 pi fc R / c)
 *   - BPA_SAR.m:150,194   coherent sum over all pulses
 * with these deliberate deviations from the script (documented quirks):
 *   - bins are 0-based and used 0-based (the script indexes SAR1(m, nr) with a
 *     0-based nr into a 1-based array, BPA_SAR.m:141,145);
 *   - samples outside [0, n_samples) read as zero instead of erroring;
 *   - antenna positions are arbitrary per-pulse FP64 3-vectors and the image
 *     grid is origin + ix*ux + iy*uy (the script: straight track, flat grid).
 *
 * Pinned by tests/test_oracle.py: mpmath brute force at 50 digits on tiny scenes
 * (tests/golden/), the point-target closed form N*A*e^{j0}, integer-bin exactness
 * of every interpolator, linearity, pulse additivity and global phase covariance.
 */
#ifndef SARBP_ORACLE_H
#define SARBP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_INTERP_NEAREST = 0, ORC_INTERP_LINEAR = 1, ORC_INTERP_SINC8 = 2 };

typedef struct {
  double c;            /* propagation speed, m/s                          */
  double fc;           /* carrier frequency, Hz                           */
  double r0;           /* slant range of range sample 0, m                */
  double dr;           /* range sample spacing, m                         */
  double kaiser_beta;  /* Kaiser window shape for ORC_INTERP_SINC8        */
  int interp;          /* ORC_INTERP_*                                    */
} orc_params;

/* Modified Bessel function of the first kind, order 0 (power series). */
double orc_bessel_i0(double x);

/* 8-tap Kaiser-windowed sinc kernel K(x), zero for |x| >= 4. */
double orc_kaiser_sinc8(double x, double beta);

/* Interpolate one complex profile (interleaved re,im; FP32 or FP64 samples,
 * exactly one pointer non-NULL) at fractional bin u.  out = {re, im}. */
void orc_interp(const float* prof32, const double* prof64, int n_samples,
                double u, const orc_params* p, double out[2]);

/*
 * Back-project rows [row0, row1) of an nx x ny image.
 *   prof32 / prof64 : [n_pulses][n_samples] complex, interleaved (exactly one non-NULL)
 *   pos             : [n_pulses][3] antenna positions, m
 *   origin, ux, uy  : pixel(ix,iy) = origin + ix*ux + iy*uy, m
 *   img             : [(row1-row0)][nx] complex128 interleaved; overwritten
 *                     unless accumulate != 0
 *   n_threads       : OpenMP threads (<= 0: runtime default)
 */
void orc_backproject(const float* prof32, const double* prof64,
                     int n_pulses, int n_samples, const double* pos,
                     const double* origin, const double* ux, const double* uy,
                     int nx, int row0, int row1, const orc_params* p,
                     double* img, int accumulate, int n_threads);

/*
 * Single-precision CPU BP (same interpolators, FP32 range and phase), the
 * "FP32 CPU" baseline row.  img is complex64 interleaved.
 */
void orc_backproject_f32(const float* prof32, int n_pulses, int n_samples,
                         const double* pos, const double* origin,
                         const double* ux, const double* uy, int nx, int row0,
                         int row1, const orc_params* p, float* img,
                         int n_threads);

/* Threads OpenMP would use by default (for reporting). */
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
