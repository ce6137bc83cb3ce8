// sarbp_wk.hip -- omega-K (wavenumber-domain) image formation: zero padding
// and azimuth down-conversion, 2-D FFT, reference function, Stolt
// interpolation along range frequency with a P-tap truncated sinc, 2-D inverse
// FFT (code/SAR_WK/wk_v5.2.3/wk.cc:43-343, code/SAR_WK/matlab/mian_wka_sinc.m).
//
// The 2-D transforms are separable, so the image is formed line by line in
// the order azimuth FFT -> [range DFT, reference function, Stolt, range
// inverse DFT] -> azimuth inverse FFT, and the bracket is ONE kernel per
// azimuth-frequency line (wk_rows): the line is read from HBM once, its
// length-nr range DFT and inverse DFT run as Bluestein chirp-z transforms on
// 2^LG-point FFTs in LDS (nr need not be a power of two: the reference's
// scene pads to 3414 = 2 * 3 * 569), and it is written once.  The line
// kernel reads its line as a column of an [nr][na] workspace (lines of
// neighbouring workgroups share the cache lines in the XCD's L2).  For
// power-of-two na the azimuth transforms are LDS kernels: the forward one
// writes workspace rows, the line kernel writes image rows and the inverse
// transforms image columns in place; for other na they are hipFFT row
// transforms of the workspace, the line kernel writes workspace columns and
// one tiled transpose writes the [na][nr] image.
// Kernels written here:
//   wk_az_fwd    (power-of-two na) per echo range column: padding, azimuth
//                down-conversion (FP64 phase) and the azimuth FFT in LDS,
//                written as a row of the [nr][na] workspace.
//   wk_az_inv_col / wk_az_inv_col2  (power-of-two na) the azimuth inverse
//                FFT in LDS on one or two image columns, in place.
//   wk_az_fwd_bs / wk_az_inv_bs  (other na up to 4096) the same with the
//                azimuth DFT as a chirp-z transform in LDS.
//   wk_prep_t    (other na) zeros, or the echo sample times its line's
//                down-conversion factor, written transposed ([nr][na]) for
//                hipFFT row transforms, in 64 x 64 tiles through LDS.
//   wk_transpose (other na) 64 x 64 tiles through LDS.
//   wk_prep_rows, wk_ref_table, wk_scale  the task-level steps
//                (sarbp_wk_preprocess, sarbp_wk_reference_function, the
//                scale of sarbp_wk_fft2's inverse).
//   wk_stolt_grid  the GridSample Stolt step (sarbp_stolt_gridsample).
//   wk_rows      per line: stage x[n] conj(b[n]) (b[n] = exp(j pi n^2 / nr)),
//                FFT, times the chirp filter spectrum, inverse FFT, times
//                conj(b[k]): the range spectrum X[k]; times the reference
//                function; Stolt taps into registers; then the same chirp-z
//                steps on conj(Y[n] b[n]) give the inverse DFT.
//   wk_stolt     the Stolt step alone on a line staged in LDS (the
//                sarbp_stolt_sinc entry point).
// Per sample one FP64 sqrt s = sqrt(max(0, (f0+fr)^2 - c^2 fa^2/(4 vr^2)))
// gives both the reference phase (revolutions 2 r0/c s + fr^2/(2 kr), plus a
// linear fr term for the window-centred range reference; reduced in FP64) and
// the Stolt centre n - (s - f0 - fr) nr/fs.
#include "sarbp.h"
#include "sarbp_fft.h"
#include "sarbp_host.h"

#include <hipfft/hipfft.h>

#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

namespace sarbp {

namespace {

constexpr int kWkThreads = 256;
constexpr int kWkMaxRg = 8192;  // a line of complex64 in 64 KiB of LDS
constexpr int kWkMaxTaps = 16;
constexpr int kRowsMaxRg = 8192;  // fused line kernel: 2 nr - 1 <= 16384 points (128 KiB of LDS)
constexpr double kPiD = 3.14159265358979323846;

using fft::cmul;
using fft::conjf2;

struct StoltArgs {
  const float2* in;
  float2* out;            // may equal in
  const double* fr_axis;  // [nr] device, or NULL: fftshift formula
  const double* fa_axis;  // [na] device, or NULL: fftshift formula
  int na, nr, taps, window;
  double f0;
  double kk;          // c^2 / (4 vr^2)
  double nr_over_fs;  // samples per Hz of range frequency
  double fr_gap;      // fs / nr (formula axis)
  double fa_gap;      // prf / na (formula axis)
  double f_nc;        // formula azimuth axis offset
  int rfm;            // 1: multiply by the reference function while staging
  double rfm_a;       // 2 r0 / c: s term of the reference phase, revolutions
  double rfm_b;       // 1 / (2 kr): fr^2 term, revolutions
  double rfm_c;       // fr term of the reference phase, revolutions per Hz
  float scale;        // applied to every output sample
};

// MATLAB fftshift(-n/2 : n/2-1)[i]: element (i + ceil(n/2)) mod n of the ramp.
__device__ __forceinline__ double shifted_bin(int i, int n) {
  int k = i + (n + 1) / 2;
  if (k >= n) k -= n;
  return (double)k - 0.5 * (double)n;
}

__device__ __forceinline__ float sinc_at(float x, float sin_pi_x) {
  // sin(pi x) / (pi x) with sin(pi x) given; 1 at x = 0.
  return x == 0.0f ? 1.0f : sin_pi_x * __builtin_amdgcn_rcpf(3.14159265358979f * x);
}

// LDS slot of line sample i: identity (LG = 0) or the FFT buffer's mapping.
template <int LG>
__device__ __forceinline__ int lslot(int i) {
  if constexpr (LG == 0)
    return i;
  else
    return fft::slot<LG>(i);
}

// Azimuth frequency of line m and c^2 fa^2 / (4 vr^2).
__device__ __forceinline__ double line_kfa2(const StoltArgs& a, int m) {
  const double fa = a.fa_axis ? a.fa_axis[m] : a.f_nc + shifted_bin(m, a.na) * a.fa_gap;
  return a.kk * fa * fa;
}

// Range frequency of sample j, Hz.
__device__ __forceinline__ double line_fr(const StoltArgs& a, int j) {
  return a.fr_axis ? a.fr_axis[j] : shifted_bin(j, a.nr) * a.fr_gap;
}

// v exp(j 2 pi rev): the phase reduced to [-1/2, 1/2] revolutions in FP64;
// the hardware sine and cosine take revolutions (absolute error ~1.2e-7 there).
__device__ __forceinline__ float2 rotate_rev(float2 v, double rev) {
  const float ph = (float)(rev - rint(rev));
  const float sn = __builtin_amdgcn_sinf(ph), cs = __builtin_amdgcn_cosf(ph);
  return make_float2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
}

// Range-frequency sample j of a line: applies the reference function to v
// (if a.rfm) and returns the Stolt centre as (c0, f), c0 + f = centre.
__device__ __forceinline__ float2 stolt_sample(const StoltArgs& a, int j, double kfa2, float2 v, int* c0,
                                               float* f) {
  const double fr = line_fr(a, j);
  const double fv = a.f0 + fr;
  const double s = sqrt(fmax(fma(fv, fv, -kfa2), 0.0));
  if (a.rfm) v = rotate_rev(v, fma(a.rfm_a, s, (a.rfm_b * fr + a.rfm_c) * fr));
  const double cen = (double)j - (s - fv) * a.nr_over_fs;
  const double fl = floor(cen);
  // Within 2^-25 below an integer the fraction would round to 1 in FP32 (a
  // zero divisor in the weights); the largest float below 1 keeps the taps of
  // floor(cen).
  *c0 = (int)fl;
  *f = fminf((float)(cen - fl), 0x1.fffffep-1f);
  return v;
}

// sin(pi f) for f in [0, 1]: sin(pi g), g = min(f, 1 - f) (1 - f is exact for
// f >= 1/2), as g P(g^2) with a degree-4 P fitted for least maximum relative
// error on g in [0, 1/2] (5.3e-9 before rounding, about 2 ulp in FP32).
__device__ __forceinline__ float sinpi01(float f) {
  const float g = fminf(f, 1.0f - f);
  const float u = g * g;
  float p = 0.07756070702744297f;
  p = fmaf(p, u, -0.5982422909074577f);
  p = fmaf(p, u, 2.550069753062404f);
  p = fmaf(p, u, -5.167709686218649f);
  p = fmaf(p, u, 3.14159263690788f);
  return g * p;
}

// Sum over the taps of one output at centre c0 + f (f in [0, 1)) for any tap
// count, with per-tap bounds handling: the floor window clamps, the centred
// window skips taps outside the line.
template <int LG>
__device__ __forceinline__ float2 stolt_taps_general(const float2* line, int nr, int c0, float f, float sf,
                                                     int half, int window) {
  float2 acc = make_float2(0.0f, 0.0f);
  if (window == SARBP_STOLT_FLOOR) {
    for (int p = -half; p < half; ++p) {
      int idx = c0 + p;
      float w;
      if ((unsigned)idx < (unsigned)nr) {
        const float x = f - (float)p;
        w = sinc_at(x, (p & 1) ? -sf : sf);
      } else {
        idx = idx < 0 ? 0 : nr - 1;
        const float x = (float)(c0 - idx) + f;
        w = sinc_at(x, sinpif(x));
      }
      const float2 v = line[lslot<LG>(idx)];
      acc.x = fmaf(w, v.x, acc.x);
      acc.y = fmaf(w, v.y, acc.y);
    }
  } else {
    // Integers q with |f - q| <= half: q = -half .. half when f == 0, else
    // -half+1 .. half (the q = -half tap then lies beyond half).
    for (int q = -half; q <= half; ++q) {
      const int idx = c0 + q;
      const float x = f - (float)q;
      if (fabsf(x) > (float)half || (unsigned)idx >= (unsigned)nr) continue;
      const float w = sinc_at(x, (q & 1) ? -sf : sf);
      const float2 v = line[lslot<LG>(idx)];
      acc.x = fmaf(w, v.x, acc.x);
      acc.y = fmaf(w, v.y, acc.y);
    }
  }
  return acc;
}

// The same sum for P taps at offsets q0 .. q0+P-1 that all lie inside the
// line, 0 < f < 1 (no tap at distance 0): weights (-1)^q (sin(pi f)/pi) / (f - q),
// reciprocals taken in pairs (1/a = b/(ab), 1/b = a/(ab)) with sin(pi f)/pi
// and the sign folded into the pair's factor.
template <int LG, int P>
__device__ __forceinline__ float2 stolt_taps_inner(const float2* line, int c0, float f, float sf, int q0) {
  const float amp = sf * 0.318309886183790672f;
  float d[P];
#pragma unroll
  for (int t = 0; t < P; ++t) d[t] = f - (float)(q0 + t);
  float2 acc = make_float2(0.0f, 0.0f);
#pragma unroll
  for (int t = 0; t < P; t += 2) {
    float r = amp * __builtin_amdgcn_rcpf(d[t] * d[t + 1]);
    if ((q0 + t) & 1) r = -r;
    const float w0 = d[t + 1] * r, w1 = -d[t] * r;
    const float2 v0 = line[lslot<LG>(c0 + q0 + t)], v1 = line[lslot<LG>(c0 + q0 + t + 1)];
    acc.x = fmaf(w0, v0.x, acc.x);
    acc.y = fmaf(w0, v0.y, acc.y);
    acc.x = fmaf(w1, v1.x, acc.x);
    acc.y = fmaf(w1, v1.y, acc.y);
  }
  return acc;
}

// One Stolt output at centre c0 + f from a line held in LDS.
template <int LG, int P>
__device__ __forceinline__ float2 stolt_output(const StoltArgs& a, const float2* line, int c0, float f) {
  const int half = P > 0 ? P / 2 : a.taps >> 1;
  // First tap offset and the range of c0 for which every tap is inside.
  const int q0 = a.window == SARBP_STOLT_FLOOR ? -half : -half + 1;
  const int c_lo = -q0, c_hi = a.nr - 1 - (q0 + 2 * half - 1);
  // sin(pi (f - p)) = (-1)^p sin(pi f) for integer p.
  const float sf = sinpi01(f);
  if (P > 0 && f > 0.0f && c0 >= c_lo && c0 <= c_hi)
    return stolt_taps_inner<LG, (P > 0 ? P : 2)>(line, c0, f, sf, q0);
  if (a.window == SARBP_STOLT_FLOOR) {
    // Every tap clamped to the same end sample e: 2 half equal terms
    // sinc(c - e) line[e], sin(pi (c - e)) = (-1)^(c0 - e) sin(pi f).
    int e = -1;
    if (c0 + half - 1 <= 0)
      e = 0;
    else if (c0 - half >= a.nr - 1)
      e = a.nr - 1;
    if (e >= 0) {
      const int d = c0 - e;
      const float w = (float)(2 * half) * sinc_at((float)d + f, (d & 1) ? -sf : sf);
      const float2 v = line[lslot<LG>(e)];
      return make_float2(w * v.x, w * v.y);
    }
  } else if (c0 + half < 0 || c0 - half > a.nr - 1) {
    return make_float2(0.0f, 0.0f);  // no tap inside the line
  }
  return stolt_taps_general<LG>(line, a.nr, c0, f, sf, half, a.window);
}

// K = samples per thread (nr <= K * kWkThreads); P = taps (0: any, runtime).
template <int K, int P>
__global__ __launch_bounds__(kWkThreads) void wk_stolt(StoltArgs a) {
  extern __shared__ float2 line[];
  const int m = blockIdx.x;
  const int nr = a.nr;
  const size_t base = (size_t)m * (size_t)nr;
  const double kfa2 = line_kfa2(a, m);
  int fl[K];
  float fr_off[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = threadIdx.x + k * kWkThreads;
    fl[k] = 0;
    fr_off[k] = 0.0f;
    if (j < nr) line[j] = stolt_sample(a, j, kfa2, a.in[base + j], &fl[k], &fr_off[k]);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = threadIdx.x + k * kWkThreads;
    if (j >= nr) continue;
    const float2 acc = stolt_output<0, P>(a, line, fl[k], fr_off[k]);
    a.out[base + j] = make_float2(acc.x * a.scale, acc.y * a.scale);
  }
}

// ---- GridSample Stolt ------------------------------------------------------------

// The reference's GridSample Stolt step (performStoltInterpolation,
// task/stolt.cc:38-318): line m rotated by nr/2 (column w holds input
// (w + nr - nr/2) mod nr), output n sampled from it by linear interpolation
// with zeros outside at ix = ((x + 1) / 2) (nr - 1),
// x = float(float(sqrt((f0 + fr[n])^2 - kk fa[m]^2) - f0) / (fr nr / 2)) --
// the bilinear, zero-padded, align-corners grid sample with the row coordinate
// on row m.  Every step up to the weights is evaluated in the reference's
// precision and order (FP64 mapping rounded to FP32, FP32 grid arithmetic,
// no contraction), so the sample positions are the reference's.
struct GridArgs {
  const float2* in;
  float2* out;  // may equal in
  const double* fr_axis;
  const double* fa_axis;
  int na, nr;
  double f0, kk;      // kk = c^2 / (4 vr^2)
  double half_band;   // fr nr / 2
};

template <int K>
__global__ __launch_bounds__(kWkThreads) void wk_stolt_grid(GridArgs a) {
#pragma clang fp contract(off)
  extern __shared__ float2 line[];
  const int m = blockIdx.x;
  const int nr = a.nr, mid = nr / 2;
  const size_t base = (size_t)m * (size_t)nr;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int w = threadIdx.x + k * kWkThreads;
    if (w < nr) {
      int src = w + nr - mid;
      if (src >= nr) src -= nr;
      line[w] = a.in[base + src];
    }
  }
  __syncthreads();
  const double fa = a.fa_axis[m];
  const double kfa2 = a.kk * (fa * fa);
  const float span = (float)(nr - 1);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int n = threadIdx.x + k * kWkThreads;
    if (n >= nr) continue;
    const double fv = a.f0 + a.fr_axis[n];
    const float fr_new = (float)(sqrt(fv * fv - kfa2) - a.f0);
    const float x = (float)((double)fr_new / a.half_band);
    const float ix = ((x + 1.0f) / 2.0f) * span;
    const float f0i = floorf(ix);
    const float t = ix - f0i;
    const int i0 = (int)f0i;
    float2 v0 = make_float2(0.0f, 0.0f), v1 = v0;
    if (i0 >= 0 && i0 < nr) v0 = line[i0];
    if (i0 + 1 >= 0 && i0 + 1 < nr) v1 = line[i0 + 1];
    const float u = 1.0f - t;
    a.out[base + n] = make_float2(u * v0.x + t * v1.x, u * v0.y + t * v1.y);
  }
}

hipError_t launch_stolt_grid(const GridArgs& a, hipStream_t s) {
  const int k = (a.nr + kWkThreads - 1) / kWkThreads;
  const size_t lds = (size_t)a.nr * sizeof(float2);
  const dim3 grid((unsigned)a.na), block(kWkThreads);
  if (k <= 4)
    hipLaunchKernelGGL(wk_stolt_grid<4>, grid, block, lds, s, a);
  else if (k <= 8)
    hipLaunchKernelGGL(wk_stolt_grid<8>, grid, block, lds, s, a);
  else if (k <= 16)
    hipLaunchKernelGGL(wk_stolt_grid<16>, grid, block, lds, s, a);
  else
    hipLaunchKernelGGL(wk_stolt_grid<32>, grid, block, lds, s, a);
  return hipGetLastError();
}

// ---- fused line kernel ---------------------------------------------------------

// Chirp-z tables of one (nr, LG): b[n] = exp(j pi n^2 / nr) for n < nr;
// bh[k] = FFT_M(bb)[k] / M with bb[n] = b[|n|] at n and M - n (|n| < nr),
// M = 2^LG >= 2 nr - 1; tw[t] = exp(-j 2 pi t / M).
// Sample j of input line m is st.in[m in_line + j in_sample] (rows of an
// [na][nr] array: in_line = nr, in_sample = 1; columns of an [nr][na] array:
// in_line = 1, in_sample = na), likewise for st.out; samples outside
// [nz_lo, nz_hi) are known zeros and not read.  Output sample n of a line
// goes to (n + out_shift) mod nr, times out_rot (which includes the
// 1/(na nr) scale): a circular shift with a constant phase is the
// range-frequency ramp exp(-j 2 pi fr h / fs) of the centred range reference
// on the fftshift axis.
struct RowArgs {
  StoltArgs st;
  const float2* tw;   // [M]
  const float2* bh;   // [M]
  const float2* chirp;  // [nr]
  int out_shift;      // 0 <= out_shift < nr
  int nz_lo, nz_hi;   // 0 <= nz_lo <= nz_hi <= nr
  int in_line, in_sample, out_line, out_sample;
  float2 out_rot;
};

// Line of workgroup b: the dispatcher deals workgroups to the 8 XCDs in turn
// (b -> XCD b % 8), so XCD x gets a contiguous run of lines and the lines in
// flight on one XCD share the 128-byte cache lines of the [nr][na] array in
// that XCD's L2.
__device__ __forceinline__ int xcd_line(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7, i = b >> 3;
  return x * q + min(x, r) + i;
}

constexpr int rows_threads(int lg) { return (1 << lg) / 16; }

// Chirp-z core of a length-nr DFT on a line held in registers (x[e] = point
// tid + e T, fft::fft_regs).  On entry x holds z[n] conj(b[n]) for n < nr and
// zeros above; on return it holds that sequence circularly convolved with b,
// so DFT(z)[k] = conj(b[k]) x[k] for k < nr.  The chirp-spectrum product
// between the two transforms needs no LDS pass.  The caller's reads of buf end
// before the call, and it synchronises before its next writes of buf.
template <int LG>
__device__ __forceinline__ void chirp_convolve_regs(float2* buf, const RowArgs& r, float2 (&x)[16]) {
  constexpr int T = rows_threads(LG);
  fft::fft_regs<LG, false>(buf, r.tw, x);
  // An opaque index keeps the 16 table loads after the transform (hoisted
  // above it they would hold 32 registers through it).
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
#pragma unroll
  for (int e = 0; e < 16; ++e) x[e] = cmul(x[e], r.bh[tid + e * T]);
  __syncthreads();
  fft::fft_regs<LG, true>(buf, r.tw, x);
}

// One workgroup per azimuth-frequency line m; nr <= K T, 2 nr - 1 <= 2^LG.
// Thread tid holds samples tid + e T of the line in registers from the read
// to the write; LDS carries the transforms' inner passes and the Stolt taps.
template <int LG, int K, int P>
__global__ __launch_bounds__(rows_threads(LG), 4) void wk_rows(RowArgs r) {
  constexpr int T = rows_threads(LG);
  static_assert(K <= 16, "K samples per thread of the 16 held");
  extern __shared__ float2 buf[];
  const StoltArgs& a = r.st;
  const int m = xcd_line(blockIdx.x, a.na);
  const int nr = a.nr;
  const float2* __restrict__ in = a.in + (size_t)m * r.in_line;
  float2* out = a.out + (size_t)m * r.out_line;
  const int tid = threadIdx.x;

  // Forward range DFT of z conj(b), zeros outside [nz_lo, nz_hi).
  float2 x[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int i = tid + e * T;
    x[e] = i >= r.nz_lo && i < r.nz_hi ? cmul(in[(size_t)i * r.in_sample], conjf2(r.chirp[i]))
                                       : make_float2(0.0f, 0.0f);
  }
  chirp_convolve_regs<LG>(buf, r, x);

  // Range spectrum X[k] = conj(b[k]) c[k], times the reference function;
  // Stolt centres kept per output.
  const double kfa2 = line_kfa2(a, m);
  int fl[K];
  float fr_off[K];
  __syncthreads();  // the transform's last reads of buf are done
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const int j = tid + t * T;
    fl[t] = 0;
    fr_off[t] = 0.0f;
    if (j < nr) {
      const float2 v = cmul(x[t], conjf2(r.chirp[j]));
      buf[fft::slot<LG>(j)] = stolt_sample(a, j, kfa2, v, &fl[t], &fr_off[t]);
    }
  }
  __syncthreads();
  float2 y[K];
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const int j = tid + t * T;
    y[t] = j < nr ? stolt_output<LG, P>(a, buf, fl[t], fr_off[t]) : make_float2(0.0f, 0.0f);
  }
  __syncthreads();  // the Stolt taps' reads of buf are done

  // Inverse range DFT: x[n] = conj(DFT(conj(Y))[n]) = b[n] conj(c[n]) with the
  // line staged as conj(Y[n] b[n]).
#pragma unroll
  for (int e = 0; e < 16; ++e) x[e] = make_float2(0.0f, 0.0f);
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const int j = tid + t * T;
    if (j < nr) x[t] = conjf2(cmul(y[t], r.chirp[j]));
  }
  chirp_convolve_regs<LG>(buf, r, x);
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const int j = tid + t * T;
    if (j < nr) {
      const float2 v = cmul(r.chirp[j], conjf2(x[t]));
      int o = j + r.out_shift;
      if (o >= nr) o -= nr;
      out[(size_t)o * r.out_sample] = cmul(v, r.out_rot);
    }
  }
}

// out[c][r] = in[r][c] for in [rows][cols]: 64 x 64 tiles, 256 threads.
__global__ __launch_bounds__(256) void wk_transpose(const float2* __restrict__ in, float2* __restrict__ out,
                                                     int rows, int cols) {
  __shared__ float2 tile[64][65];
  const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int r = r0 + ty + 4 * k, c = c0 + tx;
    if (r < rows && c < cols) tile[ty + 4 * k][tx] = in[(size_t)r * cols + c];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int c = c0 + ty + 4 * k, r = r0 + tx;
    if (r < rows && c < cols) out[(size_t)c * rows + r] = tile[tx][ty + 4 * k];
  }
}

struct PrepArgs {
  const float2* echo;  // [na_in][nr_in]
  float2* out;         // [nr][na], transposed
  int na, nr, na_in, nr_in, pad_az, pad_rg;
  double rev_per_line;  // -f_nc / prf: down-conversion phase per line, revolutions
};

// Padding and azimuth down-conversion, written transposed: out[j][m] (=
// [nr][na]) for 64 x 64 tiles of (range sample j, padded line m), through LDS,
// for the rows j in [pad_rg, pad_rg + nr_in) that hold echo (the zero rows
// are neither transformed nor read before the line kernel writes them).
__global__ __launch_bounds__(256) void wk_prep_t(PrepArgs a) {
  __shared__ float2 tile[64][65];
  __shared__ float2 rot[64];
  const int j0 = a.pad_rg + blockIdx.x * 64, m0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  if (threadIdx.x < 64) {
    // exp(-j 2 pi f_nc ta) of line m0 + t, ta = (m - na/2) / prf.
    const double rev = a.rev_per_line * ((double)(m0 + (int)threadIdx.x) - 0.5 * (double)a.na);
    float sn, cs;
    sincospif(2.0f * (float)(rev - rint(rev)), &sn, &cs);
    rot[threadIdx.x] = make_float2(cs, sn);
  }
  __syncthreads();
  // Read lines m0 + ty + 4k, samples j0 + tx (coalesced along the echo line).
#pragma unroll 4
  for (int k = 0; k < 16; ++k) {
    const int m = m0 + ty + 4 * k, j = j0 + tx;
    const int mi = m - a.pad_az, ji = j - a.pad_rg;
    float2 v = make_float2(0.0f, 0.0f);
    if (mi >= 0 && mi < a.na_in && ji >= 0 && ji < a.nr_in) {
      const float2 e = a.echo[(size_t)mi * a.nr_in + ji], r = rot[ty + 4 * k];
      v = make_float2(e.x * r.x - e.y * r.y, e.x * r.y + e.y * r.x);
    }
    tile[ty + 4 * k][tx] = v;
  }
  __syncthreads();
#pragma unroll 4
  for (int k = 0; k < 16; ++k) {
    const int j = j0 + ty + 4 * k, m = m0 + tx;
    if (j < a.pad_rg + a.nr_in && m < a.na) a.out[(size_t)j * a.na + m] = tile[tx][ty + 4 * k];
  }
}

// Padding and azimuth down-conversion into [na][nr] (preprocess_wk,
// preprocess.cc:17-166): zeros, or the echo sample times its line's factor
// exp(-j 2 pi f_nc ta), ta = (m - na/2) / prf, as wk_prep_t computes it.
__global__ __launch_bounds__(256) void wk_prep_rows(PrepArgs a) {
  const long long total = (long long)a.na * a.nr;
  const long long step = (long long)gridDim.x * blockDim.x;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += step) {
    const int m = (int)(idx / a.nr), j = (int)(idx - (long long)m * a.nr);
    const int mi = m - a.pad_az, ji = j - a.pad_rg;
    float2 v = make_float2(0.0f, 0.0f);
    if (mi >= 0 && mi < a.na_in && ji >= 0 && ji < a.nr_in) {
      const double rev = a.rev_per_line * ((double)m - 0.5 * (double)a.na);
      float sn, cs;
      sincospif(2.0f * (float)(rev - rint(rev)), &sn, &cs);
      const float2 e = a.echo[(size_t)mi * a.nr_in + ji];
      v = make_float2(e.x * cs - e.y * sn, e.x * sn + e.y * cs);
    }
    a.out[idx] = v;
  }
}

// exp(j theta) on [na][nr] (generateReferenceFunction_CPU, compression.cc:17-58):
// theta = 4 pi r0/c sqrt(max(0, (f0 + fr)^2 - c^2 fa^2 / (4 vr^2))) - pi fr^2 / kr,
// evaluated in revolutions in FP64 and reduced before the hardware sine and
// cosine.
struct RefArgs {
  float2* out;
  const double* fr_axis;
  const double* fa_axis;
  int na, nr;
  double f0, bb, rev_s, rev_q;  // bb = c^2/(4 vr^2); rev_s = 2 r0 / c; rev_q = -1 / (2 kr)
};

__global__ __launch_bounds__(256) void wk_ref_table(RefArgs a) {
  const long long total = (long long)a.na * a.nr;
  const long long step = (long long)gridDim.x * blockDim.x;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += step) {
    const int m = (int)(idx / a.nr), j = (int)(idx - (long long)m * a.nr);
    const double fa = a.fa_axis[m], fr = a.fr_axis[j];
    const double f = a.f0 + fr;
    const double s = sqrt(fmax(f * f - a.bb * fa * fa, 0.0));
    const double rev = a.rev_s * s + a.rev_q * fr * fr;
    const float ph = (float)(rev - rint(rev));
    a.out[idx] = make_float2(__builtin_amdgcn_cosf(ph), __builtin_amdgcn_sinf(ph));
  }
}

__global__ __launch_bounds__(256) void wk_scale(float2* d, long long n, float sc) {
  const long long step = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
    const float2 v = d[i];
    d[i] = make_float2(v.x * sc, v.y * sc);
  }
}

unsigned stride_grid(long long n) {
  const long long b = (n + 255) / 256;
  return (unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

// Padding, azimuth down-conversion and the azimuth FFT of one echo range
// column c, for a power-of-two padded azimuth length na = 2^LG: the column is
// read down the echo (the columns in flight on one XCD share the cache lines
// of its L2), transformed in LDS and written as row pad_rg + c of the [nr][na]
// workspace.  Replaces wk_prep_t and the hipFFT row transforms for these na.
struct AzArgs {
  const float2* echo;   // [na_in][nr_in]
  float2* out;          // [nr][na]
  const float2* tw;     // [na]: exp(-j 2 pi t / na)
  int na, na_in, nr_in, pad_az, pad_rg;
  double rev_per_line;  // -f_nc / prf, revolutions
};

template <int LG>
__global__ __launch_bounds__((1 << LG) / 16) void wk_az_fwd(AzArgs a) {
  constexpr int NF = 1 << LG, T = NF / 16;
  extern __shared__ float2 buf[];
  const int c = xcd_line(blockIdx.x, a.nr_in);
  const int tid = threadIdx.x;
  float2 x[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int m = tid + e * T;
    const int mi = m - a.pad_az;
    float2 v = make_float2(0.0f, 0.0f);
    if (mi >= 0 && mi < a.na_in) {
      // exp(-j 2 pi f_nc ta), ta = (m - na/2) / prf, as in wk_prep_t.
      const double rev = a.rev_per_line * ((double)m - 0.5 * (double)NF);
      float sn, cs;
      sincospif(2.0f * (float)(rev - rint(rev)), &sn, &cs);
      const float2 z = a.echo[(size_t)mi * a.nr_in + c];
      v = make_float2(z.x * cs - z.y * sn, z.x * sn + z.y * cs);
    }
    x[e] = v;
  }
  fft::fft_regs<LG, false>(buf, a.tw, x);
  float2* row = a.out + (size_t)(a.pad_rg + c) * NF;
#pragma unroll
  for (int e = 0; e < 16; ++e) row[tid + e * T] = x[e];
}

// Two adjacent echo columns (2p, 2p + 1) per workgroup for even nr_in and a
// 16-byte aligned echo: 16-byte reads per echo row, one down-conversion
// factor per row for both, and the two transforms taking turns in one LDS
// buffer; written as workspace rows pad_rg + 2p and pad_rg + 2p + 1.
template <int LG>
__global__ __launch_bounds__((1 << LG) / 16) void wk_az_fwd2(AzArgs a) {
  constexpr int NF = 1 << LG, T = NF / 16;
  extern __shared__ float2 buf[];
  const int c = 2 * xcd_line(blockIdx.x, a.nr_in >> 1);
  const int tid = threadIdx.x;
  float2 x[16], y[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int m = tid + e * T;
    const int mi = m - a.pad_az;
    x[e] = y[e] = make_float2(0.0f, 0.0f);
    if (mi >= 0 && mi < a.na_in) {
      const double rev = a.rev_per_line * ((double)m - 0.5 * (double)NF);
      float sn, cs;
      sincospif(2.0f * (float)(rev - rint(rev)), &sn, &cs);
      const float4 z = *reinterpret_cast<const float4*>(a.echo + (size_t)mi * a.nr_in + c);
      x[e] = make_float2(z.x * cs - z.y * sn, z.x * sn + z.y * cs);
      y[e] = make_float2(z.z * cs - z.w * sn, z.z * sn + z.w * cs);
    }
  }
  fft::fft_regs<LG, false>(buf, a.tw, x);
  __syncthreads();  // the first transform's last reads of buf are done
  fft::fft_regs<LG, false>(buf, a.tw, y);
  float2* row = a.out + (size_t)(a.pad_rg + c) * NF;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    row[tid + e * T] = x[e];
    row[NF + tid + e * T] = y[e];
  }
}

// Azimuth inverse FFT of image column r ([na][nr], na = 2^LG), in place: the
// column is read whole before it is written; columns in flight on one XCD
// share cache lines in its L2.  Unscaled (the line kernel applies
// 1/(na nr)).
template <int LG>
__global__ __launch_bounds__((1 << LG) / 16) void wk_az_inv_col(float2* img, const float2* __restrict__ tw, int nr) {
  constexpr int NF = 1 << LG, T = NF / 16;
  extern __shared__ float2 buf[];
  const int r = xcd_line(blockIdx.x, nr);
  const int tid = threadIdx.x;
  float2 x[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) x[e] = img[(size_t)(tid + e * T) * nr + r];
  fft::fft_regs<LG, true>(buf, tw, x);
#pragma unroll
  for (int e = 0; e < 16; ++e) img[(size_t)(tid + e * T) * nr + r] = x[e];
}

// Two adjacent columns (2p, 2p + 1) per workgroup for even nr: 16-byte
// accesses per image row; the two transforms take turns in one LDS buffer.
template <int LG>
__global__ __launch_bounds__((1 << LG) / 16) void wk_az_inv_col2(float2* img, const float2* __restrict__ tw, int nr) {
  constexpr int NF = 1 << LG, T = NF / 16;
  extern __shared__ float2 buf[];
  const int r = 2 * xcd_line(blockIdx.x, nr >> 1);
  const int tid = threadIdx.x;
  float2 x[16], y[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const float4 v = *reinterpret_cast<const float4*>(img + (size_t)(tid + e * T) * nr + r);
    x[e] = make_float2(v.x, v.y);
    y[e] = make_float2(v.z, v.w);
  }
  fft::fft_regs<LG, true>(buf, tw, x);
  __syncthreads();  // the first transform's last reads of buf are done
  fft::fft_regs<LG, true>(buf, tw, y);
#pragma unroll
  for (int e = 0; e < 16; ++e)
    *reinterpret_cast<float4*>(img + (size_t)(tid + e * T) * nr + r) = make_float4(x[e].x, x[e].y, y[e].x, y[e].y);
}

template <int LG>
hipError_t launch_az_inv_lg(float2* img, const float2* tw, int nr, hipStream_t s) {
  // 16-byte accesses need an even row length and a 16-byte aligned image.
  if ((nr & 1) == 0 && (reinterpret_cast<uintptr_t>(img) & 15) == 0) {
    hipLaunchKernelGGL(wk_az_inv_col2<LG>, dim3((unsigned)(nr / 2)), dim3((1 << LG) / 16),
                       (size_t)fft::lds_words<LG>() * sizeof(float2), s, img, tw, nr);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(wk_az_inv_col<LG>, dim3((unsigned)nr), dim3((1 << LG) / 16),
                     (size_t)fft::lds_words<LG>() * sizeof(float2), s, img, tw, nr);
  return hipGetLastError();
}

hipError_t launch_az_inv(float2* img, const float2* tw, int nr, int lg, hipStream_t s) {
  switch (lg) {
    case 10: return launch_az_inv_lg<10>(img, tw, nr, s);
    case 11: return launch_az_inv_lg<11>(img, tw, nr, s);
    case 12: return launch_az_inv_lg<12>(img, tw, nr, s);
    case 13: return launch_az_inv_lg<13>(img, tw, nr, s);
    default: return hipErrorInvalidValue;
  }
}

// The azimuth transforms for na that is not a power of two (na <= 4096): the
// length-na DFT of a column as a chirp-z transform on 2^LG >= 2 na - 1
// points in LDS (t: the chirp tables of length na, as for the range lines).
// Forward: pad, down-convert and transform echo column c, written as row
// pad_rg + c of the [nr][na] workspace.
template <int LG>
__global__ __launch_bounds__(rows_threads(LG)) void wk_az_fwd_bs(AzArgs a, RowArgs t) {
  constexpr int T = rows_threads(LG);
  extern __shared__ float2 buf[];
  const int c = xcd_line(blockIdx.x, a.nr_in);
  const int tid = threadIdx.x;
  const int na = a.na;
  float2 x[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int m = tid + e * T;
    float2 v = make_float2(0.0f, 0.0f);
    if (m < na) {
      const int mi = m - a.pad_az;
      if (mi >= 0 && mi < a.na_in) {
        const double rev = a.rev_per_line * ((double)m - 0.5 * (double)na);
        float sn, cs;
        sincospif(2.0f * (float)(rev - rint(rev)), &sn, &cs);
        const float2 z = a.echo[(size_t)mi * a.nr_in + c];
        v = make_float2(z.x * cs - z.y * sn, z.x * sn + z.y * cs);
      }
      v = cmul(v, conjf2(t.chirp[m]));
    }
    x[e] = v;
  }
  chirp_convolve_regs<LG>(buf, t, x);
  float2* row = a.out + (size_t)(a.pad_rg + c) * na;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int k = tid + e * T;
    if (k < na) row[k] = cmul(x[e], conjf2(t.chirp[k]));
  }
}

// Inverse (unscaled) of image column r in place: conj(DFT(conj(x))).
template <int LG>
__global__ __launch_bounds__(rows_threads(LG)) void wk_az_inv_bs(float2* img, RowArgs t, int na, int nr) {
  constexpr int T = rows_threads(LG);
  extern __shared__ float2 buf[];
  const int r = xcd_line(blockIdx.x, nr);
  const int tid = threadIdx.x;
  float2 x[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int m = tid + e * T;
    x[e] = m < na ? cmul(conjf2(img[(size_t)m * nr + r]), conjf2(t.chirp[m])) : make_float2(0.0f, 0.0f);
  }
  chirp_convolve_regs<LG>(buf, t, x);
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int k = tid + e * T;
    if (k < na) img[(size_t)k * nr + r] = conjf2(cmul(x[e], conjf2(t.chirp[k])));
  }
}

template <int LG>
hipError_t launch_az_bs_lg(const AzArgs& a, const RowArgs& t, hipStream_t s) {
  hipLaunchKernelGGL(wk_az_fwd_bs<LG>, dim3((unsigned)a.nr_in), dim3(rows_threads(LG)),
                     (size_t)fft::lds_words<LG>() * sizeof(float2), s, a, t);
  return hipGetLastError();
}

template <int LG>
hipError_t launch_az_inv_bs_lg(float2* img, const RowArgs& t, int na, int nr, hipStream_t s) {
  hipLaunchKernelGGL(wk_az_inv_bs<LG>, dim3((unsigned)nr), dim3(rows_threads(LG)),
                     (size_t)fft::lds_words<LG>() * sizeof(float2), s, img, t, na, nr);
  return hipGetLastError();
}

hipError_t launch_az_bs(const AzArgs& a, const RowArgs& t, int lg, hipStream_t s) {
  switch (lg) {
    case 11: return launch_az_bs_lg<11>(a, t, s);
    case 12: return launch_az_bs_lg<12>(a, t, s);
    case 13: return launch_az_bs_lg<13>(a, t, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_az_inv_bs(float2* img, const RowArgs& t, int na, int nr, int lg, hipStream_t s) {
  switch (lg) {
    case 11: return launch_az_inv_bs_lg<11>(img, t, na, nr, s);
    case 12: return launch_az_inv_bs_lg<12>(img, t, na, nr, s);
    case 13: return launch_az_inv_bs_lg<13>(img, t, na, nr, s);
    default: return hipErrorInvalidValue;
  }
}

// log2 na when na is a power of two the fused azimuth kernel takes, else 0.
int az_lg(int na) {
  for (int lg = 10; lg <= 13; ++lg)
    if (na == (1 << lg)) return lg;
  return 0;
}

template <int LG>
hipError_t launch_az_lg(const AzArgs& a, hipStream_t s) {
  // 16-byte reads need an even echo row length and a 16-byte aligned echo.
  if ((a.nr_in & 1) == 0 && (reinterpret_cast<uintptr_t>(a.echo) & 15) == 0)
    hipLaunchKernelGGL(wk_az_fwd2<LG>, dim3((unsigned)(a.nr_in / 2)), dim3((1 << LG) / 16),
                       (size_t)fft::lds_words<LG>() * sizeof(float2), s, a);
  else
    hipLaunchKernelGGL(wk_az_fwd<LG>, dim3((unsigned)a.nr_in), dim3((1 << LG) / 16),
                       (size_t)fft::lds_words<LG>() * sizeof(float2), s, a);
  return hipGetLastError();
}

hipError_t launch_az(const AzArgs& a, int lg, hipStream_t s) {
  switch (lg) {
    case 10: return launch_az_lg<10>(a, s);
    case 11: return launch_az_lg<11>(a, s);
    case 12: return launch_az_lg<12>(a, s);
    case 13: return launch_az_lg<13>(a, s);
    default: return hipErrorInvalidValue;
  }
}

// Smallest LG in [11, 14] with 2 nr - 1 <= 2^LG (nr <= kRowsMaxRg).
int rows_lg(int nr) {
  static_assert(kRowsMaxRg >= kWkMaxRg && 2 * kRowsMaxRg - 1 <= (1 << 14),
                "every line check_stolt admits fits the 16384-point line kernel");
  int lg = 11;
  while ((1 << lg) < 2 * nr - 1) ++lg;
  return lg;
}

template <int LG, int P>
hipError_t launch_rows_lp(const RowArgs& r, hipStream_t s) {
  constexpr size_t lds = (size_t)fft::lds_words<LG>() * sizeof(float2);
  if constexpr (lds > 64 * 1024) {
    // Above 64 KiB of dynamic LDS the kernel has to ask for it (the device
    // allows 160 KiB per workgroup).
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&wk_rows<LG, 8, P>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((wk_rows<LG, 8, P>), dim3((unsigned)r.st.na), dim3(rows_threads(LG)), lds, s, r);
  return hipGetLastError();
}

hipError_t launch_rows(const RowArgs& r, int lg, hipStream_t s) {
  const bool p6 = r.st.taps == 6;
  switch (lg) {
    case 11: return p6 ? launch_rows_lp<11, 6>(r, s) : launch_rows_lp<11, 0>(r, s);
    case 12: return p6 ? launch_rows_lp<12, 6>(r, s) : launch_rows_lp<12, 0>(r, s);
    case 13: return p6 ? launch_rows_lp<13, 6>(r, s) : launch_rows_lp<13, 0>(r, s);
    case 14: return p6 ? launch_rows_lp<14, 6>(r, s) : launch_rows_lp<14, 0>(r, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_transpose(const float2* in, float2* out, int rows, int cols, hipStream_t s) {
  hipLaunchKernelGGL(wk_transpose, dim3((unsigned)((cols + 63) / 64), (unsigned)((rows + 63) / 64)), dim3(256), 0,
                     s, in, out, rows, cols);
  return hipGetLastError();
}

template <int P>
void launch_stolt_p(const StoltArgs& a, hipStream_t s) {
  const int k = (a.nr + kWkThreads - 1) / kWkThreads;
  const size_t lds = (size_t)a.nr * sizeof(float2);
  const dim3 grid((unsigned)a.na), block(kWkThreads);
  if (k <= 4)
    hipLaunchKernelGGL((wk_stolt<4, P>), grid, block, lds, s, a);
  else if (k <= 8)
    hipLaunchKernelGGL((wk_stolt<8, P>), grid, block, lds, s, a);
  else if (k <= 16)
    hipLaunchKernelGGL((wk_stolt<16, P>), grid, block, lds, s, a);
  else
    hipLaunchKernelGGL((wk_stolt<32, P>), grid, block, lds, s, a);
}

// The reference's 6 taps get an unrolled kernel; other tap counts the
// generic one.
hipError_t launch_stolt(const StoltArgs& a, hipStream_t s) {
  if (a.taps == 6)
    launch_stolt_p<6>(a, s);
  else
    launch_stolt_p<0>(a, s);
  return hipGetLastError();
}

sarbp_status check_stolt(double f0, double c, double vr, int na, int nr, double fs, int taps, int window) {
  if (na < 0 || nr < 0) return fail(SARBP_ERR_INVALID_ARGUMENT, "negative size %d x %d", na, nr);
  if (nr > kWkMaxRg) return fail(SARBP_ERR_UNSUPPORTED, "nr = %d exceeds %d", nr, kWkMaxRg);
  if (taps < 2 || taps > kWkMaxTaps || (taps & 1))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "taps = %d (even, 2..%d)", taps, kWkMaxTaps);
  if (window != SARBP_STOLT_FLOOR && window != SARBP_STOLT_CENTERED)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "window %d", window);
  if (!(std::isfinite(f0) && std::isfinite(c) && c > 0.0 && std::isfinite(vr) && vr != 0.0 &&
        std::isfinite(fs) && fs > 0.0))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "f0, c, vr or fs not finite/positive");
  return SARBP_OK;
}

}  // namespace

struct WkCache {
  std::map<std::pair<int, int>, hipfftHandle> row_plans;  // 1-D (length, batch)
  std::map<int, float2*> chirp_tables;                    // nr -> [tw | bh | chirp]
  std::map<int, float2*> az_twiddles;                     // na -> exp(-j 2 pi t / na)
  std::map<std::pair<int, int>, hipfftHandle> plans_2d;   // (na, nr) of sarbp_wk_fft2
  float2* work = nullptr;  // [nr][na] workspace of sarbp_omega_k
  size_t work_cap = 0;     // complex64 words
  double* d_axes = nullptr;  // [nr | na] of sarbp_stolt_sinc
  double* h_axes = nullptr;  // pinned staging: the axes last uploaded
  size_t cap = 0;
  int axes_nr = -1, axes_na = -1;  // split of the uploaded axes
  hipEvent_t done = nullptr;  // after the last sarbp_stolt_sinc kernel
  // After the last sarbp_omega_k: the next call's stream waits for it, so
  // calls on different streams do not overlap on the shared workspace and
  // transform plans.
  hipEvent_t wk_done = nullptr;
  bool wk_pending = false;
};

void wk_cache_destroy(WkCache* wc) {
  if (!wc) return;
  if (wc->done) (void)hipEventSynchronize(wc->done);
  (void)hipDeviceSynchronize();  // the work buffer and tables may still be in use
  for (auto& kv : wc->row_plans) (void)hipfftDestroy(kv.second);
  for (auto& kv : wc->plans_2d) (void)hipfftDestroy(kv.second);
  for (auto& kv : wc->chirp_tables) (void)hipFree(kv.second);
  for (auto& kv : wc->az_twiddles) (void)hipFree(kv.second);
  (void)hipFree(wc->work);
  (void)hipFree(wc->d_axes);
  (void)hipHostFree(wc->h_axes);
  if (wc->done) (void)hipEventDestroy(wc->done);
  if (wc->wk_done) (void)hipEventDestroy(wc->wk_done);
  delete wc;
}

namespace {

sarbp_status wk_cache(sarbp_context* ctx, WkCache** out) {
  if (!ctx->wk) {
    ctx->wk = new (std::nothrow) WkCache;
    if (!ctx->wk) return fail(SARBP_ERR_OUT_OF_MEMORY, "omega-K cache");
  }
  *out = ctx->wk;
  return SARBP_OK;
}

// Batched 1-D transforms of `batch` contiguous rows of length n.
sarbp_status row_plan(WkCache* wc, int n, int batch, hipfftHandle* out) {
  const auto key = std::make_pair(n, batch);
  auto it = wc->row_plans.find(key);
  if (it != wc->row_plans.end()) {
    *out = it->second;
    return SARBP_OK;
  }
  hipfftHandle h = nullptr;
  int len[1] = {n};
  const hipfftResult r = hipfftPlanMany(&h, 1, len, len, 1, n, len, 1, n, HIPFFT_C2C, batch);
  if (r != HIPFFT_SUCCESS)
    return fail(r == HIPFFT_ALLOC_FAILED ? SARBP_ERR_OUT_OF_MEMORY : SARBP_ERR_DEVICE,
                "hipfftPlanMany(%d x %d) failed (%d)", n, batch, (int)r);
  wc->row_plans[key] = h;
  *out = h;
  return SARBP_OK;
}

// Chirp-z tables of length nr on 2^lg points (see RowArgs), built in FP64.
sarbp_status chirp_tables(WkCache* wc, int nr, int lg, const float2** out) {
  auto it = wc->chirp_tables.find(nr);
  if (it != wc->chirp_tables.end()) {
    *out = it->second;
    return SARBP_OK;
  }
  const int M = 1 << lg;
  using cd = std::complex<double>;
  std::vector<float2> h((size_t)2 * M + nr);
  std::vector<cd> b((size_t)nr), bb((size_t)M, cd(0.0, 0.0));
  for (int n = 0; n < nr; ++n) {
    // pi n^2 / nr with n^2 reduced modulo 2 nr (exact in 64-bit integers).
    const long long q = ((long long)n * n) % (2LL * nr);
    const double ang = kPiD * (double)q / (double)nr;
    b[n] = cd(std::cos(ang), std::sin(ang));
  }
  for (int n = 0; n < nr; ++n) {
    bb[n] = b[n];
    if (n > 0) bb[M - n] = b[n];
  }
  fft_host(bb);
  for (int t = 0; t < M; ++t) {
    const double ang = -2.0 * kPiD * (double)t / (double)M;
    h[t] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    const cd v = bb[t] / (double)M;
    h[M + t] = make_float2((float)v.real(), (float)v.imag());
  }
  for (int n = 0; n < nr; ++n) h[2 * (size_t)M + n] = make_float2((float)b[n].real(), (float)b[n].imag());
  float2* d = nullptr;
  hipError_t e = hipMalloc(&d, h.size() * sizeof(float2));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(chirp tables)");
  if ((e = hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess) {
    (void)hipFree(d);
    return hip_fail(e, "hipMemcpy(chirp tables)");
  }
  wc->chirp_tables[nr] = d;
  *out = d;
  return SARBP_OK;
}

// exp(-j 2 pi t / n), t < n, built in FP64.
sarbp_status az_twiddles(WkCache* wc, int n, const float2** out) {
  auto it = wc->az_twiddles.find(n);
  if (it != wc->az_twiddles.end()) {
    *out = it->second;
    return SARBP_OK;
  }
  std::vector<float2> h((size_t)n);
  for (int t = 0; t < n; ++t) {
    const double ang = -2.0 * kPiD * (double)t / (double)n;
    h[t] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  float2* d = nullptr;
  hipError_t e = hipMalloc(&d, h.size() * sizeof(float2));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(azimuth twiddles)");
  if ((e = hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess) {
    (void)hipFree(d);
    return hip_fail(e, "hipMemcpy(azimuth twiddles)");
  }
  wc->az_twiddles[n] = d;
  *out = d;
  return SARBP_OK;
}

sarbp_status work_buffer(WkCache* wc, size_t words) {
  if (words <= wc->work_cap) return SARBP_OK;
  hipError_t e = hipDeviceSynchronize();  // the old buffer may be in use
  if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
  (void)hipFree(wc->work);
  wc->work = nullptr;
  wc->work_cap = 0;
  if ((e = hipMalloc(&wc->work, words * sizeof(float2))) != hipSuccess) return hip_fail(e, "hipMalloc(omega-K work)");
  wc->work_cap = words;
  return SARBP_OK;
}

sarbp_status exec_fft(hipfftHandle h, float2* d, int dir, hipStream_t s) {
  hipfftResult r = hipfftSetStream(h, s);
  if (r == HIPFFT_SUCCESS)
    r = hipfftExecC2C(h, reinterpret_cast<hipfftComplex*>(d), reinterpret_cast<hipfftComplex*>(d), dir);
  if (r != HIPFFT_SUCCESS) return fail(SARBP_ERR_DEVICE, "hipfftExecC2C failed (%d)", (int)r);
  return SARBP_OK;
}

// Enqueues the omega-K steps of sarbp_omega_k (arguments validated) on s.
sarbp_status omega_k_enqueue(WkCache* wc, const sarbp_cf32* d_echo, int na_in, int nr_in, int na, int nr,
                             const sarbp_wk_params* wp, sarbp_cf32* d_image, hipStream_t s);

}  // namespace
}  // namespace sarbp

using namespace sarbp;

namespace sarbp {
namespace {

// Stages the Stolt steps' axes (host FP64, fr [nr] then fa [na]) on the
// device; a call whose axes equal the last upload only reuses them, any other
// first waits for the previous Stolt kernel (which may still read them).
sarbp_status stage_axes(WkCache* wc, const double* h_fr_axis, const double* h_fa_axis, int na, int nr,
                        hipStream_t s) {
  hipError_t e;
  const size_t n = (size_t)na + (size_t)nr;
  // Axes equal to the last upload are already on the device: the call then
  // only enqueues the kernel.  Otherwise the previous call's kernel, which
  // may still read the axis buffers, is waited for before they are rewritten.
  const bool same = wc->axes_nr == nr && wc->axes_na == na &&
                    std::memcmp(wc->h_axes, h_fr_axis, (size_t)nr * sizeof(double)) == 0 &&
                    std::memcmp(wc->h_axes + nr, h_fa_axis, (size_t)na * sizeof(double)) == 0;
  if (!same && wc->done && (e = hipEventSynchronize(wc->done)) != hipSuccess)
    return hip_fail(e, "hipEventSynchronize");
  if (!same && n > wc->cap) {
    (void)hipFree(wc->d_axes);
    (void)hipHostFree(wc->h_axes);
    wc->d_axes = nullptr;
    wc->h_axes = nullptr;
    wc->cap = 0;
    if ((e = hipMalloc(&wc->d_axes, n * sizeof(double))) != hipSuccess) return hip_fail(e, "hipMalloc(axes)");
    if ((e = hipHostMalloc(&wc->h_axes, n * sizeof(double), hipHostMallocDefault)) != hipSuccess)
      return hip_fail(e, "hipHostMalloc(axes)");
    wc->cap = n;
  }
  if (!wc->done && (e = hipEventCreateWithFlags(&wc->done, hipEventDisableTiming)) != hipSuccess)
    return hip_fail(e, "hipEventCreate");
  if (!same) {
    wc->axes_nr = wc->axes_na = -1;
    std::memcpy(wc->h_axes, h_fr_axis, (size_t)nr * sizeof(double));
    std::memcpy(wc->h_axes + nr, h_fa_axis, (size_t)na * sizeof(double));
    if ((e = hipMemcpyAsync(wc->d_axes, wc->h_axes, n * sizeof(double), hipMemcpyHostToDevice, s)) != hipSuccess)
      return hip_fail(e, "hipMemcpyAsync(axes)");
    wc->axes_nr = nr;
    wc->axes_na = na;
  }
  return SARBP_OK;
}

}  // namespace
}  // namespace sarbp

extern "C" {

sarbp_status sarbp_wk_padded_size(int32_t na_in, int32_t nr_in, int32_t* na, int32_t* nr) {
  if (na_in < 0 || nr_in < 0) return fail(SARBP_ERR_INVALID_ARGUMENT, "negative echo size");
  if (na_in > (1 << 28) || nr_in > (1 << 28)) return fail(SARBP_ERR_UNSUPPORTED, "echo too large");
  // round(): half away from zero, as std::round in preprocess.cc:21-24.
  const int pa = (int)std::floor(na_in / 6.0 + 0.5), pr = (int)std::floor(nr_in / 3.0 + 0.5);
  if (na) *na = na_in + 2 * pa;
  if (nr) *nr = nr_in + 2 * pr;
  return SARBP_OK;
}

sarbp_status sarbp_stolt_sinc(sarbp_context* ctx, const sarbp_cf32* d_in, sarbp_cf32* d_out,
                              const double* h_fr_axis, const double* h_fa_axis, double f0, double c,
                              double vr, int32_t na, int32_t nr, double fs, int32_t taps,
                              int32_t window, void* stream) {
  if (!ctx) return fail(SARBP_ERR_INVALID_ARGUMENT, "context is NULL");
  sarbp_status st = check_stolt(f0, c, vr, na, nr, fs, taps, window);
  if (st != SARBP_OK) return st;
  if (na == 0 || nr == 0) return SARBP_OK;
  if (!d_in || !d_out || !h_fr_axis || !h_fa_axis)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "data or axis pointer is NULL");
  std::lock_guard<std::mutex> lk(ctx->mu);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  WkCache* wc = nullptr;
  if ((st = wk_cache(ctx, &wc)) != SARBP_OK) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if ((st = stage_axes(wc, h_fr_axis, h_fa_axis, na, nr, s)) != SARBP_OK) return st;
  hipError_t e;
  StoltArgs a{};
  a.in = reinterpret_cast<const float2*>(d_in);
  a.out = reinterpret_cast<float2*>(d_out);
  a.fr_axis = wc->d_axes;
  a.fa_axis = wc->d_axes + nr;
  a.na = na;
  a.nr = nr;
  a.taps = taps;
  a.window = window;
  a.f0 = f0;
  a.kk = c * c / (4.0 * vr * vr);
  a.nr_over_fs = (double)nr / fs;
  a.rfm = 0;
  a.scale = 1.0f;
  if ((e = launch_stolt(a, s)) != hipSuccess) return hip_fail(e, "wk_stolt launch");
  if ((e = hipEventRecord(wc->done, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  return SARBP_OK;
}

sarbp_status sarbp_stolt_gridsample(sarbp_context* ctx, const sarbp_cf32* d_in, sarbp_cf32* d_out,
                                   const double* h_fr_axis, const double* h_fa_axis, double f0, double c,
                                   double vr, int32_t na, int32_t nr, double fr, void* stream) {
  if (!ctx) return fail(SARBP_ERR_INVALID_ARGUMENT, "context is NULL");
  sarbp_status st = check_stolt(f0, c, vr, na, nr, 1.0, 2, SARBP_STOLT_FLOOR);
  if (st != SARBP_OK) return st;
  if (!(std::isfinite(fr) && fr > 0.0)) return fail(SARBP_ERR_INVALID_ARGUMENT, "fr = %g: must be > 0", fr);
  if (na == 0 || nr == 0) return SARBP_OK;
  if (!d_in || !d_out || !h_fr_axis || !h_fa_axis)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "data or axis pointer is NULL");
  std::lock_guard<std::mutex> lk(ctx->mu);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  WkCache* wc = nullptr;
  if ((st = wk_cache(ctx, &wc)) != SARBP_OK) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if ((st = stage_axes(wc, h_fr_axis, h_fa_axis, na, nr, s)) != SARBP_OK) return st;
  GridArgs a{};
  a.in = reinterpret_cast<const float2*>(d_in);
  a.out = reinterpret_cast<float2*>(d_out);
  a.fr_axis = wc->d_axes;
  a.fa_axis = wc->d_axes + nr;
  a.na = na;
  a.nr = nr;
  a.f0 = f0;
  a.kk = c * c / (4.0 * vr * vr);
  a.half_band = fr * (double)nr * 0.5;
  hipError_t e;
  if ((e = launch_stolt_grid(a, s)) != hipSuccess) return hip_fail(e, "wk_stolt_grid launch");
  if ((e = hipEventRecord(wc->done, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  return SARBP_OK;
}

sarbp_status sarbp_wk_preprocess(sarbp_context* ctx, const sarbp_cf32* d_echo, int32_t na_in, int32_t nr_in,
                                double prf, double f_nc, sarbp_cf32* d_out, void* stream) {
  if (!ctx) return fail(SARBP_ERR_INVALID_ARGUMENT, "context is NULL");
  int32_t na = 0, nr = 0;
  sarbp_status st = sarbp_wk_padded_size(na_in, nr_in, &na, &nr);
  if (st != SARBP_OK) return st;
  if (!(std::isfinite(prf) && prf > 0.0 && std::isfinite(f_nc)))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "prf or f_nc invalid");
  if (na == 0 || nr == 0) return SARBP_OK;
  if (!d_out || (na_in > 0 && nr_in > 0 && !d_echo))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "echo or output pointer is NULL");
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  PrepArgs pa{};
  pa.echo = reinterpret_cast<const float2*>(d_echo);
  pa.out = reinterpret_cast<float2*>(d_out);
  pa.na = na;
  pa.nr = nr;
  pa.na_in = na_in;
  pa.nr_in = nr_in;
  pa.pad_az = (na - na_in) / 2;
  pa.pad_rg = (nr - nr_in) / 2;
  pa.rev_per_line = -f_nc / prf;
  hipLaunchKernelGGL(wk_prep_rows, dim3(stride_grid((long long)na * nr)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), pa);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SARBP_OK : hip_fail(e, "wk_prep_rows launch");
}

sarbp_status sarbp_wk_fft2(sarbp_context* ctx, sarbp_cf32* d_data, int32_t na, int32_t nr, int32_t inverse,
                           void* stream) {
  if (!ctx) return fail(SARBP_ERR_INVALID_ARGUMENT, "context is NULL");
  if (na < 0 || nr < 0) return fail(SARBP_ERR_INVALID_ARGUMENT, "negative size %d x %d", na, nr);
  if (na == 0 || nr == 0) return SARBP_OK;
  if (!d_data) return fail(SARBP_ERR_INVALID_ARGUMENT, "data pointer is NULL");
  std::lock_guard<std::mutex> lk(ctx->mu);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  WkCache* wc = nullptr;
  sarbp_status st = wk_cache(ctx, &wc);
  if (st != SARBP_OK) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e;
  // The plan's work area is shared by calls on this context: order them.
  if (!wc->wk_done && (e = hipEventCreateWithFlags(&wc->wk_done, hipEventDisableTiming)) != hipSuccess)
    return hip_fail(e, "hipEventCreate");
  if (wc->wk_pending && (e = hipStreamWaitEvent(s, wc->wk_done, 0)) != hipSuccess)
    return hip_fail(e, "hipStreamWaitEvent");
  const auto key = std::make_pair((int)na, (int)nr);
  hipfftHandle h = nullptr;
  auto it = wc->plans_2d.find(key);
  if (it != wc->plans_2d.end()) {
    h = it->second;
  } else {
    const hipfftResult r = hipfftPlan2d(&h, na, nr, HIPFFT_C2C);
    if (r != HIPFFT_SUCCESS)
      return fail(r == HIPFFT_ALLOC_FAILED ? SARBP_ERR_OUT_OF_MEMORY : SARBP_ERR_DEVICE,
                  "hipfftPlan2d(%d x %d) failed (%d)", na, nr, (int)r);
    wc->plans_2d[key] = h;
  }
  if ((st = exec_fft(h, reinterpret_cast<float2*>(d_data), inverse ? HIPFFT_BACKWARD : HIPFFT_FORWARD, s)) !=
      SARBP_OK)
    return st;
  if (inverse) {
    const long long n = (long long)na * nr;
    hipLaunchKernelGGL(wk_scale, dim3(stride_grid(n)), dim3(256), 0, s, reinterpret_cast<float2*>(d_data), n,
                       (float)(1.0 / ((double)na * (double)nr)));
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "wk_scale launch");
  }
  if ((e = hipEventRecord(wc->wk_done, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  wc->wk_pending = true;
  return SARBP_OK;
}

sarbp_status sarbp_wk_reference_function(sarbp_context* ctx, sarbp_cf32* d_theta, const double* h_fr_axis,
                                         const double* h_fa_axis, double f0, double c, double vr, double kr,
                                         double r0, int32_t na, int32_t nr, void* stream) {
  if (!ctx) return fail(SARBP_ERR_INVALID_ARGUMENT, "context is NULL");
  if (na < 0 || nr < 0) return fail(SARBP_ERR_INVALID_ARGUMENT, "negative size %d x %d", na, nr);
  if (!(std::isfinite(f0) && std::isfinite(c) && c > 0.0 && std::isfinite(vr) && vr != 0.0 && std::isfinite(kr) &&
        kr != 0.0 && std::isfinite(r0)))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "f0, c, vr, kr or r0 invalid");
  if (na == 0 || nr == 0) return SARBP_OK;
  if (!d_theta || !h_fr_axis || !h_fa_axis) return fail(SARBP_ERR_INVALID_ARGUMENT, "pointer is NULL");
  std::lock_guard<std::mutex> lk(ctx->mu);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  WkCache* wc = nullptr;
  sarbp_status st = wk_cache(ctx, &wc);
  if (st != SARBP_OK) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if ((st = stage_axes(wc, h_fr_axis, h_fa_axis, na, nr, s)) != SARBP_OK) return st;
  RefArgs a{};
  a.out = reinterpret_cast<float2*>(d_theta);
  a.fr_axis = wc->d_axes;
  a.fa_axis = wc->d_axes + nr;
  a.na = na;
  a.nr = nr;
  a.f0 = f0;
  a.bb = c * c / (4.0 * vr * vr);
  a.rev_s = 2.0 * r0 / c;
  a.rev_q = -0.5 / kr;
  hipLaunchKernelGGL(wk_ref_table, dim3(stride_grid((long long)na * nr)), dim3(256), 0, s, a);
  hipError_t e;
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "wk_ref_table launch");
  if ((e = hipEventRecord(wc->done, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  return SARBP_OK;
}

sarbp_status sarbp_omega_k(sarbp_context* ctx, const sarbp_cf32* d_echo, int32_t na_in, int32_t nr_in,
                           const sarbp_wk_params* wp, sarbp_cf32* d_image, void* stream) {
  if (!ctx || !wp) return fail(SARBP_ERR_INVALID_ARGUMENT, "context or params is NULL");
  int32_t na = 0, nr = 0;
  sarbp_status st = sarbp_wk_padded_size(na_in, nr_in, &na, &nr);
  if (st != SARBP_OK) return st;
  if ((st = check_stolt(wp->f0, wp->c, wp->vr, na, nr, wp->fs, wp->taps, wp->window)) != SARBP_OK) return st;
  if (!(std::isfinite(wp->prf) && wp->prf > 0.0 && std::isfinite(wp->kr) && wp->kr != 0.0 &&
        std::isfinite(wp->r0) && std::isfinite(wp->f_nc)))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "prf, kr, r0 or f_nc invalid");
  if ((wp->range_ref != SARBP_WK_RANGE_REF_R0 && wp->range_ref != SARBP_WK_RANGE_REF_CENTRE) || wp->reserved != 0)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "range_ref %d / reserved %d invalid", wp->range_ref, wp->reserved);
  if (na_in == 0 || nr_in == 0) return fail(SARBP_ERR_INVALID_ARGUMENT, "empty echo");
  if (!d_echo || !d_image) return fail(SARBP_ERR_INVALID_ARGUMENT, "echo or image pointer is NULL");
  std::lock_guard<std::mutex> lk(ctx->mu);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  WkCache* wc = nullptr;
  if ((st = wk_cache(ctx, &wc)) != SARBP_OK) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e;
  if (!wc->wk_done && (e = hipEventCreateWithFlags(&wc->wk_done, hipEventDisableTiming)) != hipSuccess)
    return hip_fail(e, "hipEventCreate");
  if (wc->wk_pending && (e = hipStreamWaitEvent(s, wc->wk_done, 0)) != hipSuccess)
    return hip_fail(e, "hipStreamWaitEvent");
  st = omega_k_enqueue(wc, d_echo, na_in, nr_in, na, nr, wp, d_image, s);
  if ((e = hipEventRecord(wc->wk_done, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  wc->wk_pending = true;
  return st;
}

}  // extern "C"

namespace sarbp {
namespace {

sarbp_status omega_k_enqueue(WkCache* wc, const sarbp_cf32* d_echo, int na_in, int nr_in, int na, int nr,
                             const sarbp_wk_params* wp, sarbp_cf32* d_image, hipStream_t s) {
  sarbp_status st;
  float2* img = reinterpret_cast<float2*>(d_image);
  const int pad_az = (na - na_in) / 2, pad_rg = (nr - nr_in) / 2;

  PrepArgs pa{};
  pa.echo = reinterpret_cast<const float2*>(d_echo);
  pa.na = na;
  pa.nr = nr;
  pa.na_in = na_in;
  pa.nr_in = nr_in;
  pa.pad_az = pad_az;
  pa.pad_rg = pad_rg;
  pa.rev_per_line = -wp->f_nc / wp->prf;

  StoltArgs a{};
  a.in = img;
  a.out = img;
  a.na = na;
  a.nr = nr;
  a.taps = wp->taps;
  a.window = wp->window;
  a.f0 = wp->f0;
  a.kk = wp->c * wp->c / (4.0 * wp->vr * wp->vr);
  a.nr_over_fs = (double)nr / wp->fs;
  a.fr_gap = wp->fs / (double)nr;
  a.fa_gap = wp->prf / (double)na;
  a.f_nc = wp->f_nc;
  a.rfm = 1;
  a.rfm_a = 2.0 * wp->r0 / wp->c;
  a.rfm_b = 0.5 / wp->kr;
  a.scale = (float)(1.0 / ((double)na * (double)nr));
  int out_shift = 0;
  double out_phase = 0.0;  // radians
  if (wp->range_ref == SARBP_WK_RANGE_REF_CENTRE) {
    // Padded sample 0 at fast time ts: the spectrum goes to absolute time
    // (exp(-j 2 pi fr ts), folded into the reference phase), the reference
    // range is that of padded sample h, and after the Stolt step the origin
    // returns to padded sample 0 by exp(-j 2 pi fr h / fs).  On the axis
    // fr[k] = (((k + ceil(nr/2)) mod nr) - nr/2) fs/nr that ramp is
    // W^(k h) exp(j pi h) W^(ceil(nr/2) h), W = exp(-j 2 pi / nr): a circular
    // shift of the range line by h times a constant.
    const double ts = 2.0 * wp->r0 / wp->c - (double)pad_rg / wp->fs;
    const int h = nr / 2;
    a.rfm_a = ts + (double)h / wp->fs;
    a.rfm_c = -ts;
    out_shift = h;
    const long long ch = (long long)((nr + 1) / 2) * h % nr;
    out_phase = kPiD * (double)(h % 2) - 2.0 * kPiD * (double)ch / (double)nr;
  }
  hipError_t e;

  const int lg = rows_lg(nr);
  if (lg == 14) {
    // 128 KiB of LDS per workgroup has to be granted up front.
    for (const void* k : {reinterpret_cast<const void*>(&wk_rows<14, 8, 6>),
                          reinterpret_cast<const void*>(&wk_rows<14, 8, 0>)}) {
      if ((e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024)) != hipSuccess)
        return fail(SARBP_ERR_UNSUPPORTED, "lines of %d samples need 128 KiB of LDS per workgroup: %s", nr,
                    hipGetErrorString(e));
    }
  }

  // Azimuth FFT (rows of the transposed workspace; only the nr_in rows
  // holding echo are non-zero), one fused kernel per line (a column of the
  // workspace, in place) for the range DFT / reference function / Stolt /
  // range inverse DFT, azimuth inverse FFT, transpose into the image.
  const float2* tab = nullptr;
  hipfftHandle az_fwd = nullptr, az_inv = nullptr;
  // Azimuth transforms: LDS FFTs for power-of-two na (1024..8192), LDS
  // chirp-z transforms for other na up to 4096, hipFFT beyond.
  const int alg = az_lg(na);
  const int blg = !alg && na <= 4096 ? rows_lg(na) : 0;
  const bool fast = alg || blg;
  const float2* az_tw = nullptr;
  const float2* az_tab = nullptr;
  if ((st = chirp_tables(wc, nr, lg, &tab)) != SARBP_OK) return st;
  if (alg) {
    if ((st = az_twiddles(wc, na, &az_tw)) != SARBP_OK) return st;
  } else if (blg) {
    if ((st = chirp_tables(wc, na, blg, &az_tab)) != SARBP_OK) return st;
  } else if ((st = row_plan(wc, na, nr_in, &az_fwd)) != SARBP_OK) {
    return st;
  }
  if (!fast && (st = row_plan(wc, na, nr, &az_inv)) != SARBP_OK) return st;
  RowArgs azt{};
  if (blg) {
    azt.tw = az_tab;
    azt.bh = az_tab + ((size_t)1 << blg);
    azt.chirp = az_tab + ((size_t)2 << blg);
  }
  if ((st = work_buffer(wc, (size_t)na * nr)) != SARBP_OK) return st;
  float2* w = wc->work;
  if (fast) {
    AzArgs az{};
    az.echo = pa.echo;
    az.out = w;
    az.tw = az_tw;
    az.na = na;
    az.na_in = na_in;
    az.nr_in = nr_in;
    az.pad_az = pad_az;
    az.pad_rg = pad_rg;
    az.rev_per_line = pa.rev_per_line;
    e = alg ? launch_az(az, alg, s) : launch_az_bs(az, azt, blg, s);
    if (e != hipSuccess) return hip_fail(e, "azimuth forward launch");
  } else {
    pa.out = w;
    hipLaunchKernelGGL(wk_prep_t, dim3((unsigned)((nr_in + 63) / 64), (unsigned)((na + 63) / 64)), dim3(256), 0,
                       s, pa);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "wk_prep_t launch");
    if ((st = exec_fft(az_fwd, w + (size_t)pad_rg * na, HIPFFT_FORWARD, s)) != SARBP_OK) return st;
  }
  RowArgs r{};
  r.st = a;
  r.st.in = w;
  r.st.out = w;
  r.nz_lo = pad_rg;
  r.nz_hi = pad_rg + nr_in;
  r.in_line = 1;
  r.in_sample = na;
  r.out_line = 1;
  r.out_sample = na;
  if (fast) {
    // lines written as rows of the image; the azimuth inverse runs on its columns
    r.st.out = img;
    r.out_line = nr;
    r.out_sample = 1;
  }
  r.tw = tab;
  r.bh = tab + ((size_t)1 << lg);
  r.chirp = tab + ((size_t)2 << lg);
  r.out_shift = out_shift;
  r.out_rot = make_float2((float)(std::cos(out_phase) * a.scale), (float)(std::sin(out_phase) * a.scale));
  if ((e = launch_rows(r, lg, s)) != hipSuccess) return hip_fail(e, "wk_rows launch");
  if (fast) {
    e = alg ? launch_az_inv(img, az_tw, nr, alg, s) : launch_az_inv_bs(img, azt, na, nr, blg, s);
    if (e != hipSuccess) return hip_fail(e, "azimuth inverse launch");
    return SARBP_OK;
  }
  if ((st = exec_fft(az_inv, w, HIPFFT_BACKWARD, s)) != SARBP_OK) return st;
  if ((e = launch_transpose(w, img, nr, na, s)) != hipSuccess) return hip_fail(e, "wk_transpose launch");
  return SARBP_OK;
}

}  // namespace
}  // namespace sarbp
