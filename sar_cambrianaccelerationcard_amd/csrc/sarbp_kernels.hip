// sarbp_kernels.hip -- gfx950 time-domain back-projection kernels.
//
// Arithmetic (per pixel p, pulse m; code/SAR_BP/matlab/BPA_SAR.m:139-150 of the
// reference repository):
//   R   = |antenna_m - p|
//   u   = (R - r0) / dr                       fractional range bin
//   I_p = sum_m interp(profile_m, u) * exp(+j 2 pi (2 fc R / c))
//
// Work decomposition.  A workgroup (4 x wave64) owns one image tile of
// 64 x (4*PPL) pixels: lane = x offset, each lane accumulates PPL rows.  Pulses
// are consumed in stages of P: for every pulse of a stage one lane computes, in
// FP64, the tile-reference geometry (below) and the start of the pulse's range
// window; the workgroup copies each pulse's window of `win` samples from HBM
// into LDS with coalesced loads (zero outside the profile), and every lane
// back-projects its pixels against the staged pulses.  The pulse index is
// wavefront-uniform, so the per-pulse record is an LDS broadcast and the
// sample gathers hit a short, contiguous window.
//
// Sinc8.  Each tap weight K(f + 3 - t) of the 8-tap Kaiser sinc is, over the
// fractional bin f in [0, 1), a degree-7 polynomial in e = f - 1/2 (host-fitted
// coefficients C[k][t], error ~2e-6).  The interpolant at window position i0 is
// then v(e) = sum_k D_k e^k with D_k = sum_t C[k][t] s[i0 - 3 + t].  After
// staging, the workgroup computes D_0..D_7 for every (pulse, position) of the
// stage into LDS once; a pixel then reads its 8 coefficients (64 B) and runs a
// complex Horner scheme (14 FMA) -- no weight table and no per-pixel tap loop.
//
// Precision (SARBP_PREC_REF).  With c the tile's reference pixel, e the unit
// vector from antenna to c, Rc = |c - antenna| and d = p - c:
//   s = e.d,  pp = |d|^2 - s^2 >= 0,  R = Rc + s + q,  q = pp / (R + Rc + s)
// Rc, e and the phase of the reference, frac(2 fc Rc / c), are FP64 per
// (tile, pulse).  s, pp and q are small (metres) and FP32 carries them to
// ~1e-7 m.  The linear phase term k.s = (k e.ux) di + (k e.uy) dj (k = 2fc/c,
// di, dj integer pixel offsets) is evaluated with the per-step phase
// increments reduced modulo one revolution in FP64, which is exact for
// integer di, dj; FP32 then only ever holds a few revolutions.  q is expanded
// to first or second order in s / Rc, or evaluated with one reciprocal, as
// the FP64 setup's truncation bound for the stage requires.
//
// sin/cos: the hardware v_sin_f32 / v_cos_f32 take their argument in
// revolutions, which is what the phase already is.
#include "sarbp_internal.h"

namespace sarbp {
namespace {

__device__ __forceinline__ float sin_rev(float x) { return __builtin_amdgcn_sinf(x); }
__device__ __forceinline__ float cos_rev(float x) { return __builtin_amdgcn_cosf(x); }

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8);
// give each XCD a contiguous run of tiles so neighbouring tiles, which read
// the same range windows, share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
  return x * q + min(x, r) + i;
}

template <int DT>
__device__ __forceinline__ float2 load_sample(const void* prof, size_t idx) {
  if constexpr (DT == 0) {
    return reinterpret_cast<const float2*>(prof)[idx];
  } else {
    const uint32_t v = reinterpret_cast<const uint32_t*>(prof)[idx];
    return make_float2(__uint_as_float(v << 16), __uint_as_float(v & 0xffff0000u));
  }
}

// Largest first-tap index that keeps every tap inside a window of `win`.
template <int INTERP>
__host__ __device__ constexpr int bin_hi(int win) {
  return INTERP == kSinc8 ? win - 5 : (INTERP == kLinear ? win - 2 : win - 1);
}

// Interpolates pulse t of the stage at window-relative bin `bin`.  The tap
// index is clamped as an integer (v_cvt_i32_f32 maps NaN to 0), so no input,
// however malformed, addresses LDS outside the pulse's window.
template <int INTERP>
__device__ __forceinline__ void interp(const float2* __restrict__ raw,
                                       const float4* __restrict__ coef, int t, int W,
                                       float bin, int lim, float& vr, float& vi) {
  if constexpr (INTERP == kNearest) {
    const float2 s = raw[t * W + min(max((int)(bin + 0.5f), 0), lim)];
    vr = s.x;
    vi = s.y;
  } else if constexpr (INTERP == kLinear) {
    const float2* w = raw + t * W + min(max((int)bin, 0), lim);
    const float f = __builtin_amdgcn_fractf(bin);
    const float2 s0 = w[0], s1 = w[1];
    vr = fmaf(f, s1.x - s0.x, s0.x);
    vi = fmaf(f, s1.y - s0.y, s0.y);
  } else {
    const int j = min(max((int)bin, 3), lim) - 3;
    const float e = __builtin_amdgcn_fractf(bin) - 0.5f;
    const float4* d = coef + ((size_t)t * (W - 7) + j) * 4;
    const float4 d67 = d[3], d45 = d[2], d23 = d[1], d01 = d[0];
    vr = fmaf(d67.z, e, d67.x);
    vi = fmaf(d67.w, e, d67.y);
    vr = fmaf(vr, e, d45.z);
    vi = fmaf(vi, e, d45.w);
    vr = fmaf(vr, e, d45.x);
    vi = fmaf(vi, e, d45.y);
    vr = fmaf(vr, e, d23.z);
    vi = fmaf(vi, e, d23.w);
    vr = fmaf(vr, e, d23.x);
    vi = fmaf(vi, e, d23.y);
    vr = fmaf(vr, e, d01.z);
    vi = fmaf(vi, e, d01.w);
    vr = fmaf(vr, e, d01.x);
    vi = fmaf(vi, e, d01.y);
  }
}

// MAC of one interpolated sample rotated by exp(+j 2 pi ph).
__device__ __forceinline__ void rotate_acc(float vr, float vi, float ph, float& ar, float& ai) {
  const float sn = sin_rev(ph), cs = cos_rev(ph);
  ar = fmaf(vr, cs, ar);
  ar = fmaf(-vi, sn, ar);
  ai = fmaf(vr, sn, ai);
  ai = fmaf(vi, cs, ai);
}

// Back-projects the `cnt` staged pulses (REF precision) with range-expansion
// order ORD for the differential q = R - Rc - s:
//   1: q = pp (g - 2g^2 s)                               (g = 1 / (2 Rc))
//   2: q = pp (g - 2g^2 s + 4g^3 s^2 - g^3 pp)
//   3: q = pp / (2 (Rc + s) + q1), q1 the order-1 value  (exact up to rounding:
//      R - Rc - s = pp / (R + Rc + s) identically)
// Per-pulse LDS record: {alpha, beta, pa, pb | c0, g, 2g^2, bin(ref)}.
template <int INTERP, int PPL, int ORD>
__device__ __forceinline__ void stage_ref(const float4* __restrict__ par,
                                          const float* __restrict__ rc2,
                                          const float2* __restrict__ raw,
                                          const float4* __restrict__ coef, int W, int cnt,
                                          float fdi, const float* fdj, const float* dd,
                                          float inv_dr, float k_rev, int lim, float* acc_r,
                                          float* acc_i) {
  for (int t = 0; t < cnt; ++t) {
    const float4 p0 = par[2 * t], p1 = par[2 * t + 1];
    const float s_x = p0.x * fdi;
    const float ph_x = fmaf(p0.z, fdi, p1.x);
    const float g3 = 0.5f * p1.y * p1.z;
    const float g3x4 = 4.0f * g3;
    const float r2 = (ORD == 3) ? rc2[t] : 0.0f;
#pragma unroll
    for (int n = 0; n < PPL; ++n) {
      const float s = fmaf(p0.y, fdj[n], s_x);
      const float pp = fmaf(-s, s, dd[n]);
      float q;
      if constexpr (ORD == 1) {
        q = pp * fmaf(-p1.z, s, p1.y);
      } else if constexpr (ORD == 2) {
        const float u = fmaf(g3x4, s, -p1.z);
        const float v = fmaf(u, s, p1.y);
        q = pp * fmaf(-g3, pp, v);
      } else {
        const float q1 = pp * fmaf(-p1.z, s, p1.y);
        q = pp * __builtin_amdgcn_rcpf(fmaf(2.0f, s, r2) + q1);
      }
      const float bin = fmaf(s + q, inv_dr, p1.w);
      float vr, vi;
      interp<INTERP>(raw, coef, t, W, bin, lim, vr, vi);
      float ph = fmaf(p0.w, fdj[n], ph_x);
      ph = fmaf(k_rev, q, ph);
      rotate_acc(vr, vi, ph, acc_r[n], acc_i[n]);
    }
  }
}

// Back-projects pulses [a.pulse_begin, a.pulse_end) into one tile.
template <int INTERP, int PREC, int DT, int PPL>
__device__ __forceinline__ void bp_tile(const BpArgs& a, int tile, char* lds) {
  constexpr int TY = kWaves * PPL;
  const int tid = threadIdx.x;
  const int lx = tid & 63, wv = tid >> 6;
  const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
  const int x0 = tx * kTileX, y0 = ty * TY;
  const int P = a.pulses_per_stage, W = a.win;
  const int NPOS = W - 7;

  const LdsLayout L = lds_layout(INTERP, W, P);
  float2* lds_raw = reinterpret_cast<float2*>(lds + L.raw);
  float4* lds_coef = reinterpret_cast<float4*>(lds + L.coef);
  float* lds_par = reinterpret_cast<float*>(lds + L.par);
  int* lds_start = reinterpret_cast<int*>(lds + L.start);
  float* lds_rc2 = reinterpret_cast<float*>(lds + L.rc2);
  int* lds_order = reinterpret_cast<int*>(lds + L.order);

  // Tile reference pixel (integer pixel near the tile centre).
  const int ic = x0 + kTileX / 2, jc = y0 + TY / 2;
  const double cx = a.origin[0] + ic * a.ux[0] + jc * a.uy[0];
  const double cy = a.origin[1] + ic * a.ux[1] + jc * a.uy[1];
  const double cz = a.origin[2] + ic * a.ux[2] + jc * a.uy[2];

  // Per-lane pixel constants.
  const int di = lx - kTileX / 2;
  const float fdi = (float)di;
  float fdj[PPL], dd[PPL];
  double px[PPL], py[PPL], pz[PPL];
  float acc_r[PPL], acc_i[PPL];
#pragma unroll
  for (int n = 0; n < PPL; ++n) {
    const int dj = wv * PPL + n - TY / 2;
    fdj[n] = (float)dj;
    const double dx = di * a.ux[0] + dj * a.uy[0];
    const double dy = di * a.ux[1] + dj * a.uy[1];
    const double dz = di * a.ux[2] + dj * a.uy[2];
    dd[n] = (float)(dx * dx + dy * dy + dz * dz);
    px[n] = cx + dx;
    py[n] = cy + dy;
    pz[n] = cz + dz;
    acc_r[n] = 0.0f;
    acc_i[n] = 0.0f;
  }
  const float inv_dr = a.inv_dr_f, k_rev = a.k_rev_f;
  const int lim = bin_hi<INTERP>(W);
  if (tid == 0) *lds_order = 1;
  __syncthreads();

  for (int m0 = a.pulse_begin; m0 < a.pulse_end; m0 += P) {
    const int cnt = min(P, a.pulse_end - m0);
    // ---- per-pulse setup (FP64), one lane per pulse ----
    for (int t = tid; t < cnt; t += kThreads) {
      const int m = m0 + t;
      const double ax = a.pos[3 * (size_t)m], ay = a.pos[3 * (size_t)m + 1],
                   az = a.pos[3 * (size_t)m + 2];
      const double ex0 = cx - ax, ey0 = cy - ay, ez0 = cz - az;
      const double Rc = sqrt(ex0 * ex0 + ey0 * ey0 + ez0 * ez0);
      const double binc = (Rc - a.r0) * a.inv_dr;
      const double start = floor(binc) - (double)a.half;
      lds_start[t] = (int)fmax(fmin(start, 2.0e9), -2.0e9);
      float* pr = lds_par + 8 * t;
      if constexpr (PREC == kRef) {
        const double inv = 1.0 / Rc;
        const double ex = ex0 * inv, ey = ey0 * inv, ez = ez0 * inv;
        const double al = ex * a.ux[0] + ey * a.ux[1] + ez * a.ux[2];
        const double be = ex * a.uy[0] + ey * a.uy[1] + ez * a.uy[2];
        double pa = a.k_rev * al;
        pa -= rint(pa);
        double pb = a.k_rev * be;
        pb -= rint(pb);
        double c0 = a.k_rev * Rc;
        c0 -= floor(c0);
        const double g = 0.5 * inv;
        pr[0] = (float)al;
        pr[1] = (float)be;
        pr[2] = (float)pa;
        pr[3] = (float)pb;
        pr[4] = (float)c0;
        pr[5] = (float)g;
        pr[6] = (float)(g * inv);  // 2 g^2
        pr[7] = (float)(binc - start);
        lds_rc2[t] = (float)(2.0 * Rc);
        // Truncation bounds in radians over the tile (|d| <= dmax, R >= Rc - dmax):
        // order 1 drops pp s^2 / (2 R^3) + pp^2 / (8 R^3) <= (5/8) dmax^4 / R^3;
        // order 2 drops terms <= 2 dmax^5 / R^4.
        const double rmin = fmax(Rc - a.dmax, 1e-3);
        const double d2 = a.dmax * a.dmax;
        const double rad = 6.283185307179586 * a.k_rev;
        const double e1 = rad * 0.625 * d2 * d2 / (rmin * rmin * rmin);
        const double e2 = rad * 2.0 * d2 * d2 * a.dmax / (rmin * rmin * rmin * rmin);
        const int ord = e1 <= a.order_tol ? 1 : (e2 <= a.order_tol ? 2 : 3);
        if (ord > 1) atomicMax(lds_order, ord);
      } else if constexpr (PREC == kF64) {
        double* pd = reinterpret_cast<double*>(pr);
        pd[0] = ax;
        pd[1] = ay;
        pd[2] = az;
        pd[3] = a.r0 + start * a.dr;  // range of window sample 0
      } else {
        pr[0] = (float)ax;
        pr[1] = (float)ay;
        pr[2] = (float)az;
        pr[3] = (float)(a.r0 + start * a.dr);
      }
    }
    __syncthreads();
    // Range-expansion order of this stage: read by every thread here, reset by
    // thread 0 after the next barrier, before any setup of the next stage.
    const int order = __builtin_amdgcn_readfirstlane(*lds_order);
    // ---- stage the range windows ----
    for (int i = tid; i < cnt * W; i += kThreads) {
      const int t = i / W, k = i - t * W;
      const int g = lds_start[t] + k;
      float2 v = make_float2(0.0f, 0.0f);
      if ((unsigned)g < (unsigned)a.n_samples)
        v = load_sample<DT>(a.prof, (size_t)(m0 + t) * a.n_samples + g);
      lds_raw[i] = v;
    }
    __syncthreads();
    if (tid == 0) *lds_order = 1;
    // ---- sinc8: interpolant coefficients for every (pulse, position) ----
    if constexpr (INTERP == kSinc8) {
      for (int i = tid; i < cnt * NPOS; i += kThreads) {
        const int t = i / NPOS, j = i - t * NPOS;
        const float2* r = lds_raw + t * W + j;
        float sr[kSincTaps], si[kSincTaps];
#pragma unroll
        for (int k = 0; k < kSincTaps; ++k) {
          const float2 v = r[k];
          sr[k] = v.x;
          si[k] = v.y;
        }
        float4* d = lds_coef + (size_t)i * 4;
#pragma unroll
        for (int k = 0; k < kSincCoef; k += 2) {
          float a0 = 0.0f, b0 = 0.0f, a1 = 0.0f, b1 = 0.0f;
#pragma unroll
          for (int tp = 0; tp < kSincTaps; ++tp) {
            const float c0 = a.sinc_poly[k * kSincTaps + tp];
            const float c1 = a.sinc_poly[(k + 1) * kSincTaps + tp];
            a0 = fmaf(c0, sr[tp], a0);
            b0 = fmaf(c0, si[tp], b0);
            a1 = fmaf(c1, sr[tp], a1);
            b1 = fmaf(c1, si[tp], b1);
          }
          d[k / 2] = make_float4(a0, b0, a1, b1);
        }
      }
      __syncthreads();
    }
    // ---- back-project the staged pulses ----
    if constexpr (PREC == kRef) {
      const float4* par = reinterpret_cast<const float4*>(lds_par);
      if (order == 1)
        stage_ref<INTERP, PPL, 1>(par, lds_rc2, lds_raw, lds_coef, W, cnt, fdi, fdj, dd, inv_dr, k_rev, lim, acc_r, acc_i);
      else if (order == 2)
        stage_ref<INTERP, PPL, 2>(par, lds_rc2, lds_raw, lds_coef, W, cnt, fdi, fdj, dd, inv_dr, k_rev, lim, acc_r, acc_i);
      else
        stage_ref<INTERP, PPL, 3>(par, lds_rc2, lds_raw, lds_coef, W, cnt, fdi, fdj, dd, inv_dr, k_rev, lim, acc_r, acc_i);
    } else {
      (void)order;
      for (int t = 0; t < cnt; ++t) {
        if constexpr (PREC == kF64) {
          const double* pd = reinterpret_cast<const double*>(lds_par + 8 * t);
          const double ax = pd[0], ay = pd[1], az = pd[2], rs = pd[3];
#pragma unroll
          for (int n = 0; n < PPL; ++n) {
            const double dx = px[n] - ax, dy = py[n] - ay, dz = pz[n] - az;
            const double R = sqrt(dx * dx + dy * dy + dz * dz);
            const float bin = (float)((R - rs) * a.inv_dr);
            double ph = a.k_rev * R;
            ph -= floor(ph);
            float vr, vi;
            interp<INTERP>(lds_raw, lds_coef, t, W, bin, lim, vr, vi);
            rotate_acc(vr, vi, (float)ph, acc_r[n], acc_i[n]);
          }
        } else {
          const float4 p0 = reinterpret_cast<const float4*>(lds_par)[2 * t];
#pragma unroll
          for (int n = 0; n < PPL; ++n) {
            const float dx = (float)px[n] - p0.x, dy = (float)py[n] - p0.y,
                        dz = (float)pz[n] - p0.z;
            const float R = sqrtf(dx * dx + dy * dy + dz * dz);
            const float bin = (R - p0.w) * inv_dr;
            const float ph = __builtin_amdgcn_fractf(R * k_rev);
            float vr, vi;
            interp<INTERP>(lds_raw, lds_coef, t, W, bin, lim, vr, vi);
            rotate_acc(vr, vi, ph, acc_r[n], acc_i[n]);
          }
        }
      }
    }
    __syncthreads();
  }
  // ---- write the tile ----
  const int x = x0 + lx;
  if (x < a.nx) {
#pragma unroll
    for (int n = 0; n < PPL; ++n) {
      const int y = y0 + wv * PPL + n;
      if (y < a.ny) {
        float2* o = a.img + (size_t)y * a.nx + x;
        float2 r = make_float2(acc_r[n], acc_i[n]);
        if (a.accumulate) {
          const float2 prev = *o;
          r.x += prev.x;
          r.y += prev.y;
        }
        *o = r;
      }
    }
  }
}

template <int INTERP, int PREC, int DT, int PPL>
__global__ __launch_bounds__(kThreads) void bp_plain_kernel(BpArgs a) {
  extern __shared__ __align__(16) char lds[];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  bp_tile<INTERP, PREC, DT, PPL>(a, tile, lds);
}

// Persistent batched kernel: workgroups claim (job, tile) items from a device
// counter until the list is exhausted; every wave of every workgroup reaches
// the exit once the counter passes n_items.
template <int INTERP, int PREC, int DT, int PPL>
__global__ __launch_bounds__(kThreads) void bp_batch_kernel(const JobDesc* __restrict__ jobs,
                                                            const BatchHeader* __restrict__ hdr,
                                                            unsigned long long* counter) {
  extern __shared__ __align__(16) char lds[];
  __shared__ long long item_sh;
  const int64_t n_items = hdr->n_items;
  const int32_t n_jobs = hdr->n_jobs;
  for (;;) {
    if (threadIdx.x == 0) item_sh = (long long)atomicAdd(counter, 1ull);
    __syncthreads();
    const long long item = item_sh;
    __syncthreads();
    if (item >= n_items) break;
    // job = last j with jobs[j].item_begin <= item (binary search, uniform)
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].item_begin <= item) lo = mid; else hi = mid - 1;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);  // uniform: descriptor reads become scalar loads
    bp_tile<INTERP, PREC, DT, PPL>(jobs[lo].a, (int)(item - jobs[lo].item_begin), lds);
  }
}

__global__ void reset_counter_kernel(unsigned long long* counter) { *counter = 0ull; }

// ---- dispatch over the template parameters ----
template <int INTERP, int PREC, int DT, int PPL>
hipError_t launch_plain_t(const BpArgs& a, hipStream_t s) {
  const int nb = a.tiles_x * a.tiles_y;
  const size_t shm = lds_layout(INTERP, a.win, a.pulses_per_stage).total;
  hipLaunchKernelGGL((bp_plain_kernel<INTERP, PREC, DT, PPL>), dim3(nb), dim3(kThreads), shm, s, a);
  return hipGetLastError();
}

template <int INTERP, int PREC, int DT, int PPL>
hipError_t launch_batch_t(const JobDesc* jobs, const BatchHeader* hdr, unsigned long long* counter,
                          size_t shm, int n_blocks, hipStream_t s) {
  hipLaunchKernelGGL((bp_batch_kernel<INTERP, PREC, DT, PPL>), dim3(n_blocks), dim3(kThreads),
                     shm, s, jobs, hdr, counter);
  return hipGetLastError();
}

template <int INTERP, int PREC, int DT, int PPL>
int occupancy_t(size_t shm) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &nb, reinterpret_cast<const void*>(&bp_batch_kernel<INTERP, PREC, DT, PPL>), kThreads,
          shm) != hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

// F is invoked with the compile-time parameters matching the runtime
// (interp, precision, dtype, ppl); combinations that are not instantiated
// return hipErrorInvalidValue.
template <typename F>
hipError_t dispatch(int interp, int prec, int dt, int ppl, const F& f) {
#define SARBP_CASE_PPL(I, P, D)                                  \
  switch (ppl) {                                                 \
    case 2: return f.template run<I, P, D, 2>();                 \
    case 4: return f.template run<I, P, D, 4>();                 \
    case 8: return f.template run<I, P, D, 8>();                 \
    default: return hipErrorInvalidValue;                        \
  }
#define SARBP_CASE_DT(I, P) \
  if (dt == 0) { SARBP_CASE_PPL(I, P, 0) } else { SARBP_CASE_PPL(I, P, 1) }
#define SARBP_CASE_PREC(I)                \
  switch (prec) {                         \
    case kRef: { SARBP_CASE_DT(I, kRef) } \
    case kF64: { SARBP_CASE_DT(I, kF64) } \
    case kF32: { SARBP_CASE_DT(I, kF32) } \
    default: return hipErrorInvalidValue; \
  }
  switch (interp) {
    case kNearest: { SARBP_CASE_PREC(kNearest) }
    case kLinear: { SARBP_CASE_PREC(kLinear) }
    case kSinc8: { SARBP_CASE_PREC(kSinc8) }
    default: return hipErrorInvalidValue;
  }
#undef SARBP_CASE_PREC
#undef SARBP_CASE_DT
#undef SARBP_CASE_PPL
}

struct PlainLauncher {
  const BpArgs& a;
  hipStream_t s;
  template <int I, int P, int D, int L>
  hipError_t run() const { return launch_plain_t<I, P, D, L>(a, s); }
};

struct BatchLauncher {
  const JobDesc* jobs;
  const BatchHeader* hdr;
  unsigned long long* counter;
  size_t shm;
  int n_blocks;
  hipStream_t s;
  template <int I, int P, int D, int L>
  hipError_t run() const {
    return launch_batch_t<I, P, D, L>(jobs, hdr, counter, shm, n_blocks, s);
  }
};

struct OccupancyQuery {
  size_t shm;
  int* out;
  template <int I, int P, int D, int L>
  hipError_t run() const {
    *out = occupancy_t<I, P, D, L>(shm);
    return hipSuccess;
  }
};

}  // namespace

hipError_t launch_plain(const BpArgs& a, int interp, int precision, int dtype, int ppl,
                        hipStream_t stream) {
  return dispatch(interp, precision, dtype, ppl, PlainLauncher{a, stream});
}

hipError_t launch_batch(const JobDesc* d_jobs, const BatchHeader* d_hdr,
                        unsigned long long* d_counter, int interp, int precision, int dtype,
                        int ppl, size_t shm, int n_blocks, hipStream_t stream) {
  return dispatch(interp, precision, dtype, ppl,
                  BatchLauncher{d_jobs, d_hdr, d_counter, shm, n_blocks, stream});
}

hipError_t reset_counter(unsigned long long* d_counter, hipStream_t stream) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, d_counter);
  return hipGetLastError();
}

int batch_blocks_per_cu(int interp, int precision, int dtype, int ppl, size_t shm) {
  int nb = 1;
  (void)dispatch(interp, precision, dtype, ppl, OccupancyQuery{shm, &nb});
  return nb;
}

}  // namespace sarbp
