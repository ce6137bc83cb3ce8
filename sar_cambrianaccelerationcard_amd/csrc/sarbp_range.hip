// sarbp_range.hip -- range compression on the device: FFT matched filter and
// upsampling by spectral zero insertion (code/SAR_BP/matlab/BPA_SAR.m:93-100,
// :106-111), the step that turns raw LFM echoes into the profiles BP reads.
//
// One workgroup per pulse (256 threads, up to 1024 for the largest FFTs).
// Below 2048 points every transform pass runs in LDS; from 2048 points each
// thread keeps 8 points in registers (fft::fft_regs) and only the inner
// passes go through LDS.  The LDS buffer is nf complex64 words plus padding,
// at most 64 KiB (nf = 8192).  A pulse costs one HBM read of its raw echo
// and one write of its output window.
//
// The L-times upsampled inverse transform is split into L polyphase inverse
// FFTs of size nf instead of one of size L*nf: with k~ the signed bin of the
// zero-inserted spectrum Y (k~ = k for k < (nf+1)/2, k - nf otherwise),
//   x[L q + r] = sum_k Y[k] e^{j 2 pi k~ (L q + r) / (L nf)}
//              = IFFT_nf( Y[k] e^{j 2 pi k~ r / (L nf)} )[q],
// so phase r is an nf-point inverse FFT of the spectrum times a ramp.  The
// FFTs are the radix-8 Stockham transforms of sarbp_fft.h.
//
// Twiddles, the matched filter conj(FFT(h))/nf and the ramp are FP32 tables
// computed in FP64 on the host and kept per context.
#include "sarbp.h"
#include "sarbp_fft.h"
#include "sarbp_host.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

namespace sarbp {

namespace {

constexpr int kRcMaxLog2 = 13;  // nf <= 8192: 64 KiB of LDS

// Threads per workgroup for nf = 2^LG: 256, or nf / 8 above 2048 points so
// that a thread keeps at most 8 spectrum bins and 2 butterflies per pass in
// registers.
constexpr int rc_threads(int lg) { return lg >= 11 ? (1 << lg) / 8 : 256; }
constexpr double kPiD = 3.14159265358979323846;

using fft::cadd;
using fft::cmul;
using fft::csub;

__device__ __forceinline__ uint32_t bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

struct RcArgs {
  const float2* raw;
  const float2* tw;
  const float2* hc;
  const float2* ramp;
  void* out;
  int n_raw, L, crop, n_out, dtype;
};

__device__ __forceinline__ void store_out(void* out, size_t o, float2 v, int dtype) {
  if (dtype == SARBP_DTYPE_C64)
    static_cast<float2*>(out)[o] = v;
  else
    static_cast<uint32_t*>(out)[o] = bf16_rne(v.x) | (bf16_rne(v.y) << 16);
}

// Upsampled samples [c NF, (c + 1) NF) = q in [c NF / LT, (c + 1) NF / LT):
// the LT phases (keep[e][r] = phase r at q = tid + e T) are interleaved in LDS
// one chunk at a time, then the chunk's part of the output window is stored
// with whole cache lines per store instruction.
template <int LG, int LT, int T, int E>
__device__ __forceinline__ void rc_interleave_store(const RcArgs& a, float2* buf, const float2 (&keep)[E][LT],
                                                    size_t obase) {
  constexpr int NF = 1 << LG;
  constexpr int QC = NF / LT;  // q per chunk
  const int tid = threadIdx.x;
#pragma unroll
  for (int c = 0; c < LT; ++c) {
    const int u0 = c * NF;  // first upsampled index of the chunk
    if (u0 - a.crop >= a.n_out || u0 + NF - a.crop <= 0) continue;  // uniform
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int q = tid + e * T;
      if (q >= c * QC && q < (c + 1) * QC) {
#pragma unroll
        for (int r = 0; r < LT; ++r) buf[fft::slot<LG>((q - c * QC) * LT + r)] = keep[e][r];
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid + e * T;
      const int o = u0 + i - a.crop;
      if ((NF % T == 0 || i < NF) && o >= 0 && o < a.n_out)
        store_out(a.out, obase + (size_t)o, buf[fft::slot<LG>(i)], a.dtype);
    }
  }
}

// Block m compresses pulse m: raw[m][0, n_raw) -> out[m][0, n_out).
// LT > 0 (= L in {1, 2, 4}): every phase's samples are kept in registers and
// the L interleaved phases are written back through LDS in contiguous chunks,
// so each store instruction covers whole cache lines.  LT = 0: any L, each
// phase stored as it is computed (L-strided stores).
template <int LG, int LT>
__global__ __launch_bounds__(rc_threads(LG)) void rc_kernel(RcArgs a) {
  constexpr int NF = 1 << LG;
  constexpr int kRcThreads = rc_threads(LG);
  constexpr int E = (NF + kRcThreads - 1) / kRcThreads;
  extern __shared__ float2 buf[];
  const int tid = threadIdx.x;
  const size_t m = blockIdx.x;
  const float2* row = a.raw + m * (size_t)a.n_raw;
  // 8 points per thread (NF >= 2048): the transforms run on registers
  // (fft::fft_regs), only their inner passes through LDS.
  constexpr bool kRegs = kRcThreads * 8 == NF && LT > 0;
  float2 spec[E];
  if constexpr (kRegs) {
    float2 x[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int i = tid + e * kRcThreads;
      x[e] = i < a.n_raw ? row[i] : make_float2(0.f, 0.f);
    }
    fft::fft_regs<LG, false, 8>(buf, a.tw, x);
#pragma unroll
    for (int e = 0; e < 8; ++e) spec[e] = cmul(x[e], a.hc[tid + e * kRcThreads]);
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid + e * kRcThreads;
      if (NF % kRcThreads == 0 || i < NF) buf[fft::slot<LG>(i)] = i < a.n_raw ? row[i] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft::fft_lds<LG, rc_threads(LG), false>(buf, a.tw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid + e * kRcThreads;
      if (NF % kRcThreads == 0 || i < NF) spec[e] = cmul(buf[fft::slot<LG>(i)], a.hc[i]);
    }
  }
  const size_t obase = m * (size_t)a.n_out;
  if constexpr (kRegs) {
    float2 keep[8][LT];
#pragma unroll
    for (int r = 0; r < LT; ++r) {
      __syncthreads();  // the previous transform's last reads of buf are done
      float2 x[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) x[e] = spec[e];
      fft::fft_regs<LG, true, 8>(buf, a.tw, x);
#pragma unroll
      for (int e = 0; e < 8; ++e) keep[e][r] = x[e];
      if (r + 1 < LT) {
#pragma unroll
        for (int e = 0; e < 8; ++e) spec[e] = cmul(spec[e], a.ramp[tid + e * kRcThreads]);
      }
    }
    rc_interleave_store<LG, LT, kRcThreads, 8>(a, buf, keep, obase);
  } else if constexpr (LT > 0) {
    float2 keep[E][LT];
#pragma unroll
    for (int r = 0; r < LT; ++r) {
      __syncthreads();  // previous phase's reads of buf are done
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int i = tid + e * kRcThreads;
        if (NF % kRcThreads == 0 || i < NF) buf[fft::slot<LG>(i)] = spec[e];
      }
      __syncthreads();
      fft::fft_lds<LG, rc_threads(LG), true>(buf, a.tw);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int q = tid + e * kRcThreads;
        if (NF % kRcThreads == 0 || q < NF) keep[e][r] = buf[fft::slot<LG>(q)];
      }
      if (r + 1 < LT) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int i = tid + e * kRcThreads;
          if (NF % kRcThreads == 0 || i < NF) spec[e] = cmul(spec[e], a.ramp[i]);
        }
      }
    }
    rc_interleave_store<LG, LT, kRcThreads, E>(a, buf, keep, obase);
  } else {
    for (int r = 0; r < a.L; ++r) {
      // outputs o = L q + r - crop in [0, n_out) for q in [q_lo, q_hi]
      const int num_lo = a.crop - r;
      const int q_lo = num_lo <= 0 ? 0 : (num_lo + a.L - 1) / a.L;
      const int q_hi = min(NF - 1, (a.n_out - 1 + a.crop - r) / a.L);
      const bool live = a.n_out - 1 + a.crop - r >= 0 && q_lo <= q_hi;  // uniform over the block
      if (live) {
        __syncthreads();  // previous phase's reads of buf are done
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int i = tid + e * kRcThreads;
          if (NF % kRcThreads == 0 || i < NF) buf[fft::slot<LG>(i)] = spec[e];
        }
        __syncthreads();
        fft::fft_lds<LG, rc_threads(LG), true>(buf, a.tw);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int q = tid + e * kRcThreads;
          if (q >= q_lo && q <= q_hi)
            store_out(a.out, obase + (size_t)(a.L * q + r - a.crop), buf[fft::slot<LG>(q)], a.dtype);
        }
      }
      if (r + 1 < a.L) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int i = tid + e * kRcThreads;
          if (NF % kRcThreads == 0 || i < NF) spec[e] = cmul(spec[e], a.ramp[i]);
        }
      }
    }
  }
}

template <int LG, int LT>
hipError_t launch_rc(const RcArgs& a, int n_pulses, hipStream_t s) {
  hipLaunchKernelGGL((rc_kernel<LG, LT>), dim3((unsigned)n_pulses), dim3(rc_threads(LG)),
                     (size_t)fft::lds_words<LG>() * sizeof(float2), s, a);
  return hipGetLastError();
}

template <int LG>
hipError_t launch_rc_l(const RcArgs& a, int n_pulses, hipStream_t s) {
  // The gathered path needs NF / L whole q per chunk.
  if (a.L == 1) return launch_rc<LG, 1>(a, n_pulses, s);
  if (a.L == 2 && LG >= 1) return launch_rc<LG, 2>(a, n_pulses, s);
  if (a.L == 4 && LG >= 2) return launch_rc<LG, 4>(a, n_pulses, s);
  return launch_rc<LG, 0>(a, n_pulses, s);
}

hipError_t launch_rc_lg(int lg, const RcArgs& a, int n_pulses, hipStream_t s) {
  switch (lg) {
    case 1: return launch_rc_l<1>(a, n_pulses, s);
    case 2: return launch_rc_l<2>(a, n_pulses, s);
    case 3: return launch_rc_l<3>(a, n_pulses, s);
    case 4: return launch_rc_l<4>(a, n_pulses, s);
    case 5: return launch_rc_l<5>(a, n_pulses, s);
    case 6: return launch_rc_l<6>(a, n_pulses, s);
    case 7: return launch_rc_l<7>(a, n_pulses, s);
    case 8: return launch_rc_l<8>(a, n_pulses, s);
    case 9: return launch_rc_l<9>(a, n_pulses, s);
    case 10: return launch_rc_l<10>(a, n_pulses, s);
    case 11: return launch_rc_l<11>(a, n_pulses, s);
    case 12: return launch_rc_l<12>(a, n_pulses, s);
    case 13: return launch_rc_l<13>(a, n_pulses, s);
    default: return hipErrorInvalidValue;
  }
}

// ---- host tables ---------------------------------------------------------------

using cd = std::complex<double>;

int chirp_samples(const sarbp_range_params* rp) { return (int)std::ceil(rp->pulse_width * rp->fs); }

sarbp_status check_params(const sarbp_range_params* rp, int* lg_out) {
  if (!rp) return fail(SARBP_ERR_INVALID_ARGUMENT, "range params is NULL");
  if (!(rp->c > 0.0) || !(rp->fs > 0.0) || !(rp->pulse_width > 0.0) || !std::isfinite(rp->c) ||
      !std::isfinite(rp->fs) || !std::isfinite(rp->pulse_width) || !std::isfinite(rp->bandwidth) ||
      !std::isfinite(rp->t0))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "c, fs, pulse_width must be finite and > 0; bandwidth, t0 finite");
  if (rp->n_raw < 1) return fail(SARBP_ERR_INVALID_ARGUMENT, "n_raw %d < 1", rp->n_raw);
  if (rp->upsample < 1 || rp->upsample > 16)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "upsample %d outside [1, 16]", rp->upsample);
  if (rp->crop_offset < 0 || rp->n_out < 0)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "crop_offset %d, n_out %d", rp->crop_offset, rp->n_out);
  if (rp->dtype != SARBP_DTYPE_C64 && rp->dtype != SARBP_DTYPE_BF16X2)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "dtype %d", rp->dtype);
  const double need = (double)rp->n_raw + (double)chirp_samples(rp);
  if (need > (double)(1 << 30)) return fail(SARBP_ERR_UNSUPPORTED, "n_raw + chirp samples too large");
  int lg = 1;
  while ((double)(1 << lg) < need) ++lg;
  if ((int64_t)rp->crop_offset + rp->n_out > ((int64_t)rp->upsample << lg))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "crop_offset + n_out = %lld exceeds L * nf = %lld",
                (long long)rp->crop_offset + rp->n_out, (long long)rp->upsample << lg);
  if (lg > kRcMaxLog2)
    return fail(SARBP_ERR_UNSUPPORTED, "FFT size %d exceeds %d (n_raw + chirp samples too long)", 1 << lg,
                1 << kRcMaxLog2);
  *lg_out = lg;
  return SARBP_OK;
}

}  // namespace

// Iterative radix-2 FFT, FP64, forward (exp(-j ...)), n a power of two.
void fft_host(std::vector<cd>& a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const double ang = -2.0 * kPiD / (double)len;
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const cd w(std::cos(ang * (double)k), std::sin(ang * (double)k));
        const cd u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

struct RangeTables {
  float2* d = nullptr;  // [tw | hc | ramp], nf entries each
  int nf = 0;
};

struct RangeCache {
  std::map<std::tuple<int, int, int, double, double, double>, RangeTables> tables;
};

void range_cache_destroy(RangeCache* rc) {
  if (!rc) return;
  for (auto& kv : rc->tables) (void)hipFree(kv.second.d);
  delete rc;
}

namespace {

// Builds (or finds) the tables of a parameter set.  Caller holds ctx->mu and
// the device guard.
sarbp_status get_tables(sarbp_context* ctx, const sarbp_range_params* rp, int lg, const float2** out) {
  if (!ctx->range) {
    ctx->range = new (std::nothrow) RangeCache;
    if (!ctx->range) return fail(SARBP_ERR_OUT_OF_MEMORY, "range cache");
  }
  const int nf = 1 << lg, L = rp->upsample;
  const auto key = std::make_tuple(nf, L, rp->n_raw, rp->bandwidth, rp->pulse_width, rp->fs);
  auto it = ctx->range->tables.find(key);
  if (it != ctx->range->tables.end()) {
    *out = it->second.d;
    return SARBP_OK;
  }
  std::vector<float2> h((size_t)3 * nf);
  for (int k = 0; k < nf; ++k) {
    const double ang = -2.0 * kPiD * (double)k / (double)nf;
    h[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  // Reference chirp starting at the window start, centred at baseband
  // (exp(j pi Kr (t - Tp/2)^2) for 0 < t < Tp), and its spectrum.
  const double kr = rp->bandwidth / rp->pulse_width;
  std::vector<cd> H((size_t)nf, cd(0.0, 0.0));
  for (int n = 0; n < nf && n < rp->n_raw; ++n) {
    const double t = (double)n / rp->fs;
    if (!(t > 0.0 && t < rp->pulse_width)) continue;
    const double tc = t - 0.5 * rp->pulse_width;
    const double ph = 0.5 * kr * tc * tc;  // revolutions
    const double phf = ph - std::floor(ph);
    H[n] = cd(std::cos(2.0 * kPiD * phf), std::sin(2.0 * kPiD * phf));
  }
  fft_host(H);
  const double inv_nf = 1.0 / (double)nf;
  const int half = (nf + 1) / 2;
  for (int k = 0; k < nf; ++k) {
    const cd v = std::conj(H[k]) * inv_nf;
    h[nf + k] = make_float2((float)v.real(), (float)v.imag());
    const double ks = k < half ? (double)k : (double)k - nf;
    const double ang = 2.0 * kPiD * ks / ((double)L * nf);
    h[2 * (size_t)nf + k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  RangeTables t;
  t.nf = nf;
  hipError_t e = hipMalloc(&t.d, h.size() * sizeof(float2));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(range tables)");
  if ((e = hipMemcpy(t.d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess) {
    (void)hipFree(t.d);
    return hip_fail(e, "hipMemcpy(range tables)");
  }
  ctx->range->tables[key] = t;
  *out = t.d;
  return SARBP_OK;
}

}  // namespace
}  // namespace sarbp

using namespace sarbp;

extern "C" {

sarbp_status sarbp_range_info(const sarbp_range_params* rp, int32_t* nf, double* r0, double* dr) {
  int lg = 0;
  const sarbp_status st = check_params(rp, &lg);
  if (st != SARBP_OK) return st;
  const double L = (double)rp->upsample;
  if (nf) *nf = 1 << lg;
  if (r0) *r0 = 0.5 * rp->c * (rp->t0 + (double)rp->crop_offset / (L * rp->fs));
  if (dr) *dr = 0.5 * rp->c / (L * rp->fs);
  return SARBP_OK;
}

sarbp_status sarbp_range_compress(sarbp_context* ctx, const sarbp_cf32* d_raw, int32_t n_pulses,
                                  const sarbp_range_params* rp, void* d_profiles, void* stream) {
  if (!ctx) return fail(SARBP_ERR_INVALID_ARGUMENT, "context is NULL");
  int lg = 0;
  sarbp_status st = check_params(rp, &lg);
  if (st != SARBP_OK) return st;
  if (n_pulses < 0) return fail(SARBP_ERR_INVALID_ARGUMENT, "n_pulses %d < 0", n_pulses);
  if (n_pulses == 0 || rp->n_out == 0) return SARBP_OK;
  if (!d_raw || !d_profiles) return fail(SARBP_ERR_INVALID_ARGUMENT, "raw or profiles pointer is NULL");
  std::lock_guard<std::mutex> lk(ctx->mu);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  const float2* tab = nullptr;
  if ((st = get_tables(ctx, rp, lg, &tab)) != SARBP_OK) return st;
  const size_t nf = (size_t)1 << lg;
  RcArgs a;
  a.raw = reinterpret_cast<const float2*>(d_raw);
  a.tw = tab;
  a.hc = tab + nf;
  a.ramp = tab + 2 * nf;
  a.out = d_profiles;
  a.n_raw = rp->n_raw;
  a.L = rp->upsample;
  a.crop = rp->crop_offset;
  a.n_out = rp->n_out;
  a.dtype = rp->dtype;
  const hipError_t e = launch_rc_lg(lg, a, n_pulses, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "range compression launch");
  return SARBP_OK;
}

}  // extern "C"
