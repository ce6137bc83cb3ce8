// sarbp_wk.hip -- omega-K (wavenumber-domain) image formation: zero padding
// and azimuth down-conversion, 2-D FFT, reference function, Stolt
// interpolation along range frequency with a P-tap truncated sinc, 2-D inverse
// FFT (code/SAR_WK/wk_v5.2.3/wk.cc:43-343, code/SAR_WK/matlab/mian_wka_sinc.m).
//
// Kernels written here:
//   wk_prep   one workgroup per padded azimuth line: zeros or the echo sample
//             times the line's down-conversion factor (FP64 phase, one
//             sincos per line).  Reads the echo once, writes the padded array
//             once.
//   wk_stolt  one workgroup per azimuth-frequency line: the line is read from
//             HBM into LDS (optionally times the reference function, whose
//             phase ~1e8 rad is formed and reduced in FP64), then each output
//             sample gathers its P taps from LDS.  The Stolt square root is
//             the reference function's, so both come from one FP64 sqrt per
//             sample.  One read and one write of the line: in place is safe
//             (the workgroup owns its line and stages it before writing).
// The 2-D transforms are hipFFT plans (the padded range length of the
// reference's scene, 3414 = 2 * 3 * 569, needs a prime-size algorithm).
#include "sarbp.h"
#include "sarbp_host.h"

#include <hipfft/hipfft.h>

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <utility>

namespace sarbp {

namespace {

constexpr int kWkThreads = 256;
constexpr int kWkMaxRg = 8192;  // a line of complex64 in 64 KiB of LDS
constexpr int kWkMaxTaps = 16;
constexpr double kPiD = 3.14159265358979323846;

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

// Sum over the taps of one output at centre c0 + f (f in [0, 1)) for any tap
// count, with per-tap bounds handling: the floor window clamps, the centred
// window skips taps outside the line.
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
      const float2 v = line[idx];
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
      const float2 v = line[idx];
      acc.x = fmaf(w, v.x, acc.x);
      acc.y = fmaf(w, v.y, acc.y);
    }
  }
  return acc;
}

// The same sum for P taps at offsets q0 .. q0+P-1 that all lie inside the
// line, 0 < f < 1 (no tap at distance 0): weights (-1)^q sin(pi f) / (pi (f - q)),
// reciprocals taken in pairs (1/a = b/(ab), 1/b = a/(ab)).
template <int P>
__device__ __forceinline__ float2 stolt_taps_inner(const float2* line, int c0, float f, float sf, int q0) {
  float d[P];
#pragma unroll
  for (int t = 0; t < P; ++t) d[t] = 3.14159265358979f * (f - (float)(q0 + t));
  float2 acc = make_float2(0.0f, 0.0f);
#pragma unroll
  for (int t = 0; t < P; t += 2) {
    const float r = __builtin_amdgcn_rcpf(d[t] * d[t + 1]);
    const float s0 = ((q0 + t) & 1) ? -sf : sf;
    const float w0 = s0 * d[t + 1] * r, w1 = -s0 * d[t] * r;
    const float2 v0 = line[c0 + q0 + t], v1 = line[c0 + q0 + t + 1];
    acc.x = fmaf(w0, v0.x, acc.x);
    acc.y = fmaf(w0, v0.y, acc.y);
    acc.x = fmaf(w1, v1.x, acc.x);
    acc.y = fmaf(w1, v1.y, acc.y);
  }
  return acc;
}

// K = samples per thread (nr <= K * kWkThreads); P = taps (0: any, runtime).
template <int K, int P>
__global__ __launch_bounds__(kWkThreads) void wk_stolt(StoltArgs a) {
  extern __shared__ float2 line[];
  const int m = blockIdx.x;
  const int nr = a.nr;
  const size_t base = (size_t)m * (size_t)nr;
  const double fa = a.fa_axis ? a.fa_axis[m] : a.f_nc + shifted_bin(m, a.na) * a.fa_gap;
  const double kfa2 = a.kk * fa * fa;

  int fl[K];
  float fr_off[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = threadIdx.x + k * kWkThreads;
    fl[k] = 0;
    fr_off[k] = 0.0f;
    if (j < nr) {
      float2 v = a.in[base + j];
      const double fr = a.fr_axis ? a.fr_axis[j] : shifted_bin(j, nr) * a.fr_gap;
      const double f = a.f0 + fr;
      const double s = sqrt(fmax(fma(f, f, -kfa2), 0.0));
      if (a.rfm) {
        // Phase in revolutions, reduced to [-1/2, 1/2]: the hardware sine and
        // cosine take revolutions (absolute error ~1.2e-7 there).
        const double rev = fma(a.rfm_a, s, a.rfm_b * fr * fr);
        const float ph = (float)(rev - rint(rev));
        const float sn = __builtin_amdgcn_sinf(ph), cs = __builtin_amdgcn_cosf(ph);
        v = make_float2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
      }
      line[j] = v;
      const double cen = (double)j - (s - f) * a.nr_over_fs;
      const double c0 = floor(cen);
      fl[k] = (int)c0;
      fr_off[k] = (float)(cen - c0);
    }
  }
  __syncthreads();

  const int half = P > 0 ? P / 2 : a.taps >> 1;
  // First tap offset and the range of c0 for which every tap is inside.
  const int q0 = a.window == SARBP_STOLT_FLOOR ? -half : -half + 1;
  const int c_lo = -q0, c_hi = nr - 1 - (q0 + 2 * half - 1);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = threadIdx.x + k * kWkThreads;
    if (j >= nr) continue;
    const int c0 = fl[k];
    const float f = fr_off[k];
    // sin(pi (f - p)) = (-1)^p sin(pi f) for integer p.
    const float sf = sinpif(f);
    float2 acc;
    if (P > 0 && f > 0.0f && c0 >= c_lo && c0 <= c_hi)
      acc = stolt_taps_inner<(P > 0 ? P : 2)>(line, c0, f, sf, q0);
    else
      acc = stolt_taps_general(line, nr, c0, f, sf, half, a.window);
    a.out[base + j] = make_float2(acc.x * a.scale, acc.y * a.scale);
  }
}

struct PrepArgs {
  const float2* echo;  // [na_in][nr_in]
  float2* out;         // [na][nr]
  int na, nr, na_in, nr_in, pad_az, pad_rg;
  double rev_per_line;  // -f_nc / prf: down-conversion phase per line, revolutions
};

__global__ __launch_bounds__(kWkThreads) void wk_prep(PrepArgs a) {
  const int m = blockIdx.x;
  const int mi = m - a.pad_az;
  float2* dst = a.out + (size_t)m * (size_t)a.nr;
  if (mi < 0 || mi >= a.na_in) {
    for (int j = threadIdx.x; j < a.nr; j += kWkThreads) dst[j] = make_float2(0.0f, 0.0f);
    return;
  }
  // exp(-j 2 pi f_nc ta), ta = (m - na/2) / prf.
  const double rev = a.rev_per_line * ((double)m - 0.5 * (double)a.na);
  float sn, cs;
  sincospif(2.0f * (float)(rev - rint(rev)), &sn, &cs);
  const float2* src = a.echo + (size_t)mi * (size_t)a.nr_in;
  for (int j = threadIdx.x; j < a.nr; j += kWkThreads) {
    const int ji = j - a.pad_rg;
    float2 v = make_float2(0.0f, 0.0f);
    if (ji >= 0 && ji < a.nr_in) {
      const float2 e = src[ji];
      v = make_float2(e.x * cs - e.y * sn, e.x * sn + e.y * cs);
    }
    dst[j] = v;
  }
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
  std::map<std::pair<int, int>, hipfftHandle> plans;
  double* d_axes = nullptr;  // [nr | na] of sarbp_stolt_sinc
  double* h_axes = nullptr;  // pinned staging: the axes last uploaded
  size_t cap = 0;
  int axes_nr = -1, axes_na = -1;  // split of the uploaded axes
  hipEvent_t done = nullptr;  // after the last sarbp_stolt_sinc kernel
};

void wk_cache_destroy(WkCache* wc) {
  if (!wc) return;
  if (wc->done) (void)hipEventSynchronize(wc->done);
  for (auto& kv : wc->plans) (void)hipfftDestroy(kv.second);
  (void)hipFree(wc->d_axes);
  (void)hipHostFree(wc->h_axes);
  if (wc->done) (void)hipEventDestroy(wc->done);
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

sarbp_status fft_plan(WkCache* wc, int na, int nr, hipfftHandle* out) {
  const auto key = std::make_pair(na, nr);
  auto it = wc->plans.find(key);
  if (it != wc->plans.end()) {
    *out = it->second;
    return SARBP_OK;
  }
  hipfftHandle h = nullptr;
  const hipfftResult r = hipfftPlan2d(&h, na, nr, HIPFFT_C2C);
  if (r != HIPFFT_SUCCESS)
    return fail(r == HIPFFT_ALLOC_FAILED ? SARBP_ERR_OUT_OF_MEMORY : SARBP_ERR_DEVICE,
                "hipfftPlan2d(%d, %d) failed (%d)", na, nr, (int)r);
  wc->plans[key] = h;
  *out = h;
  return SARBP_OK;
}

sarbp_status exec_fft(hipfftHandle h, float2* d, int dir, hipStream_t s) {
  hipfftResult r = hipfftSetStream(h, s);
  if (r == HIPFFT_SUCCESS)
    r = hipfftExecC2C(h, reinterpret_cast<hipfftComplex*>(d), reinterpret_cast<hipfftComplex*>(d), dir);
  if (r != HIPFFT_SUCCESS) return fail(SARBP_ERR_DEVICE, "hipfftExecC2C failed (%d)", (int)r);
  return SARBP_OK;
}

}  // namespace
}  // namespace sarbp

using namespace sarbp;

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
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (!same) {
    wc->axes_nr = wc->axes_na = -1;
    std::memcpy(wc->h_axes, h_fr_axis, (size_t)nr * sizeof(double));
    std::memcpy(wc->h_axes + nr, h_fa_axis, (size_t)na * sizeof(double));
    if ((e = hipMemcpyAsync(wc->d_axes, wc->h_axes, n * sizeof(double), hipMemcpyHostToDevice, s)) != hipSuccess)
      return hip_fail(e, "hipMemcpyAsync(axes)");
    wc->axes_nr = nr;
    wc->axes_na = na;
  }
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
  if (na_in == 0 || nr_in == 0) return fail(SARBP_ERR_INVALID_ARGUMENT, "empty echo");
  if (!d_echo || !d_image) return fail(SARBP_ERR_INVALID_ARGUMENT, "echo or image pointer is NULL");
  std::lock_guard<std::mutex> lk(ctx->mu);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  WkCache* wc = nullptr;
  if ((st = wk_cache(ctx, &wc)) != SARBP_OK) return st;
  hipfftHandle plan = nullptr;
  if ((st = fft_plan(wc, na, nr, &plan)) != SARBP_OK) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  float2* img = reinterpret_cast<float2*>(d_image);

  PrepArgs pa{};
  pa.echo = reinterpret_cast<const float2*>(d_echo);
  pa.out = img;
  pa.na = na;
  pa.nr = nr;
  pa.na_in = na_in;
  pa.nr_in = nr_in;
  pa.pad_az = (na - na_in) / 2;
  pa.pad_rg = (nr - nr_in) / 2;
  pa.rev_per_line = -wp->f_nc / wp->prf;
  hipLaunchKernelGGL(wk_prep, dim3((unsigned)na), dim3(kWkThreads), 0, s, pa);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "wk_prep launch");
  if ((st = exec_fft(plan, img, HIPFFT_FORWARD, s)) != SARBP_OK) return st;

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
  if ((e = launch_stolt(a, s)) != hipSuccess) return hip_fail(e, "wk_stolt launch");
  return exec_fft(plan, img, HIPFFT_BACKWARD, s);
}

}  // extern "C"
