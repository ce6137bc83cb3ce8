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
// window; the workgroup then copies each pulse's window of `win` samples from
// HBM into LDS with coalesced loads (zero outside the profile), and every lane
// back-projects its pixels against the staged pulses.  The pulse index is
// wavefront-uniform, so the per-pulse record is an LDS broadcast and the
// sample gathers hit a short, contiguous window.
//
// Precision (SARBP_PREC_REF).  With c the tile's reference pixel, e the unit
// vector from antenna to c, Rc = |c - antenna| and d = p - c:
//   s = e.d,  pp = |d|^2 - s^2 >= 0,
//   R = Rc + s + q,  q = pp / (2 (Rc + s)) + O(pp^2 / Rc^3)
// Rc, e and the phase of the reference, frac(2 fc Rc / c), are FP64 per
// (tile, pulse).  s, pp and q are small (metres) and FP32 carries them to
// ~1e-7 m.  The linear phase term k.s = (k e.ux) di + (k e.uy) dj (k = 2fc/c,
// di, dj integer pixel offsets) is evaluated with the per-step phase
// increments reduced modulo one revolution in FP64, which is exact for
// integer di, dj; FP32 then only ever holds a few revolutions.
//
// sin/cos: the hardware v_sin_f32 / v_cos_f32 take their argument in
// revolutions, which is what the phase already is.
#include "sarbp_internal.h"

namespace sarbp {
namespace {

enum { kNearest = 0, kLinear = 1, kSinc8 = 2 };
enum { kRef = 0, kF64 = 1, kF32 = 2 };

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

// Interpolates the staged window w (win samples) at window-relative bin `bin`.
// The tap index is clamped as an integer (v_cvt_i32_f32 maps NaN to 0), so no
// input, however malformed, addresses LDS outside the window.
template <int INTERP>
__device__ __forceinline__ void interp(const float2* __restrict__ w, float bin,
                                       int lim_hi, const float4* __restrict__ tab,
                                       float& vr, float& vi) {
  if constexpr (INTERP == kNearest) {
    const int i = min(max((int)(bin + 0.5f), 0), lim_hi);
    const float2 s = w[i];
    vr = s.x;
    vi = s.y;
  } else if constexpr (INTERP == kLinear) {
    const int i0 = min(max((int)bin, 0), lim_hi);
    const float f = __builtin_amdgcn_fractf(bin);
    const float2 s0 = w[i0], s1 = w[i0 + 1];
    vr = fmaf(f, s1.x - s0.x, s0.x);
    vi = fmaf(f, s1.y - s0.y, s0.y);
  } else {
    const int i0 = min(max((int)bin, 3), lim_hi);
    const float f = __builtin_amdgcn_fractf(bin);
    const float x = f * (float)kSincQ;
    const int fi = min(max((int)x, 0), kSincQ - 1);
    const float a = x - (float)fi;
    const float4* tr = tab + 4 * fi;
    const float2* s = w + (i0 - 3);
    vr = 0.0f;
    vi = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float4 t = tr[k];
      const float w0 = fmaf(t.y, a, t.x), w1 = fmaf(t.w, a, t.z);
      const float2 s0 = s[2 * k], s1 = s[2 * k + 1];
      vr = fmaf(w0, s0.x, vr);
      vi = fmaf(w0, s0.y, vi);
      vr = fmaf(w1, s1.x, vr);
      vi = fmaf(w1, s1.y, vi);
    }
  }
}

// Largest first-tap index that keeps every tap inside a window of `win`.
template <int INTERP>
__host__ __device__ constexpr int bin_hi(int win) {
  return INTERP == kSinc8 ? win - 5 : (INTERP == kLinear ? win - 2 : win - 1);
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
// Per-pulse LDS record: {alpha, beta, pa, pb | c0, g, 2g^2, bin(ref)}; the
// order-3 path also reads 2 Rc from the start-array neighbour (rc2).
template <int INTERP, int PPL, int ORD>
__device__ __forceinline__ void stage_ref(const float4* __restrict__ par, const float* __restrict__ rc2,
                                          const float2* __restrict__ win, int W, int cnt,
                                          const float4* __restrict__ tab, float fdi,
                                          const float* fdj, const float* dd, float inv_dr,
                                          float k_rev, int lim, float* acc_r, float* acc_i) {
  for (int t = 0; t < cnt; ++t) {
    const float2* w = win + t * W;
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
      interp<INTERP>(w, bin, lim, tab, vr, vi);
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

  float2* lds_win = reinterpret_cast<float2*>(lds);
  float* lds_par = reinterpret_cast<float*>(lds + (size_t)P * W * 8);
  int* lds_start = reinterpret_cast<int*>(lds + (size_t)P * W * 8 + (size_t)P * kParBytes);
  float* lds_rc2 = reinterpret_cast<float*>(lds_start + P);
  int* lds_order = reinterpret_cast<int*>(lds_rc2 + P);
  const float4* lds_tab = reinterpret_cast<const float4*>(lds + lds_fixed_bytes(W, P));

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
    // ---- stage the range windows ----
    for (int i = tid; i < cnt * W; i += kThreads) {
      const int t = i / W, k = i - t * W;
      const int g = lds_start[t] + k;
      float2 v = make_float2(0.0f, 0.0f);
      if ((unsigned)g < (unsigned)a.n_samples)
        v = load_sample<DT>(a.prof, (size_t)(m0 + t) * a.n_samples + g);
      lds_win[i] = v;
    }
    const int order = __builtin_amdgcn_readfirstlane(*lds_order);
    __syncthreads();
    // ---- back-project the staged pulses ----
    if constexpr (PREC == kRef) {
      if (tid == 0) *lds_order = 1;  // all reads of this stage's order are done
      const float4* par = reinterpret_cast<const float4*>(lds_par);
      if (order == 1)
        stage_ref<INTERP, PPL, 1>(par, lds_rc2, lds_win, W, cnt, lds_tab, fdi, fdj, dd, inv_dr, k_rev, lim, acc_r, acc_i);
      else if (order == 2)
        stage_ref<INTERP, PPL, 2>(par, lds_rc2, lds_win, W, cnt, lds_tab, fdi, fdj, dd, inv_dr, k_rev, lim, acc_r, acc_i);
      else
        stage_ref<INTERP, PPL, 3>(par, lds_rc2, lds_win, W, cnt, lds_tab, fdi, fdj, dd, inv_dr, k_rev, lim, acc_r, acc_i);
    } else {
      (void)order;
      for (int t = 0; t < cnt; ++t) {
        const float2* w = lds_win + t * W;
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
            interp<INTERP>(w, bin, lim, lds_tab, vr, vi);
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
            interp<INTERP>(w, bin, lim, lds_tab, vr, vi);
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

template <int INTERP>
__device__ __forceinline__ void load_table(const BpArgs& a, char* lds) {
  if constexpr (INTERP == kSinc8) {
    float4* t = reinterpret_cast<float4*>(lds + lds_fixed_bytes(a.win, a.pulses_per_stage));
    for (int i = threadIdx.x; i < 4 * kSincQ; i += kThreads) t[i] = a.sinc_tab[i];
  }
}

template <int INTERP, int PREC, int DT, int PPL>
__global__ __launch_bounds__(kThreads) void bp_plain_kernel(BpArgs a) {
  extern __shared__ __align__(16) char lds[];
  load_table<INTERP>(a, lds);
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  bp_tile<INTERP, PREC, DT, PPL>(a, tile, lds);
}

template <int INTERP, int PREC, int DT, int PPL>
__global__ __launch_bounds__(kThreads) void bp_batch_kernel(const JobDesc* __restrict__ jobs,
                                                            int32_t n_jobs, int64_t n_items,
                                                            unsigned long long* counter) {
  extern __shared__ __align__(16) char lds[];
  __shared__ long long item_sh;
  bool table_loaded = false;
  for (;;) {
    if (threadIdx.x == 0) item_sh = (long long)atomicAdd(counter, 1ull);
    __syncthreads();
    const long long item = item_sh;
    __syncthreads();
    if (item >= n_items) break;  // every wave of every block reaches this exit
    // job = last j with jobs[j].item_begin <= item (binary search, uniform)
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].item_begin <= item) lo = mid; else hi = mid - 1;
    }
    const BpArgs a = jobs[lo].a;
    if (!table_loaded) {
      load_table<INTERP>(a, lds);
      table_loaded = true;
    }
    bp_tile<INTERP, PREC, DT, PPL>(a, (int)(item - jobs[lo].item_begin), lds);
  }
}

template <int INTERP, int PREC, int DT, int PPL>
hipError_t launch_plain_t(const BpArgs& a, hipStream_t s) {
  const int nb = a.tiles_x * a.tiles_y;
  const size_t shm = lds_bytes(INTERP, a.win, a.pulses_per_stage);
  hipLaunchKernelGGL((bp_plain_kernel<INTERP, PREC, DT, PPL>), dim3(nb), dim3(kThreads), shm, s, a);
  return hipGetLastError();
}

template <int INTERP, int PREC, int DT, int PPL>
hipError_t launch_batch_t(const JobDesc* jobs, int32_t n_jobs, int64_t n_items,
                          unsigned long long* counter, size_t shm, int n_blocks,
                          hipStream_t s) {
  hipLaunchKernelGGL((bp_batch_kernel<INTERP, PREC, DT, PPL>), dim3(n_blocks), dim3(kThreads),
                     shm, s, jobs, n_jobs, n_items, counter);
  return hipGetLastError();
}

// Dispatch tables over the template parameters.
template <int INTERP, int PREC, int DT>
hipError_t plain_ppl(const BpArgs& a, int ppl, hipStream_t s) {
  switch (ppl) {
    case 2: return launch_plain_t<INTERP, PREC, DT, 2>(a, s);
    case 4: return launch_plain_t<INTERP, PREC, DT, 4>(a, s);
    case 8: return launch_plain_t<INTERP, PREC, DT, 8>(a, s);
    default: return hipErrorInvalidValue;
  }
}
template <int INTERP, int PREC>
hipError_t plain_dt(const BpArgs& a, int dt, int ppl, hipStream_t s) {
  return dt == 0 ? plain_ppl<INTERP, PREC, 0>(a, ppl, s) : plain_ppl<INTERP, PREC, 1>(a, ppl, s);
}
template <int INTERP>
hipError_t plain_prec(const BpArgs& a, int prec, int dt, int ppl, hipStream_t s) {
  switch (prec) {
    case kRef: return plain_dt<INTERP, kRef>(a, dt, ppl, s);
    case kF64: return plain_dt<INTERP, kF64>(a, dt, ppl, s);
    case kF32: return plain_dt<INTERP, kF32>(a, dt, ppl, s);
    default: return hipErrorInvalidValue;
  }
}

template <int INTERP, int PREC, int DT>
hipError_t batch_ppl(const JobDesc* j, int32_t n, int64_t ni, unsigned long long* c,
                     size_t shm, int nb, int ppl, hipStream_t s) {
  switch (ppl) {
    case 2: return launch_batch_t<INTERP, PREC, DT, 2>(j, n, ni, c, shm, nb, s);
    case 4: return launch_batch_t<INTERP, PREC, DT, 4>(j, n, ni, c, shm, nb, s);
    case 8: return launch_batch_t<INTERP, PREC, DT, 8>(j, n, ni, c, shm, nb, s);
    default: return hipErrorInvalidValue;
  }
}
template <int INTERP, int PREC>
hipError_t batch_dt(const JobDesc* j, int32_t n, int64_t ni, unsigned long long* c, size_t shm,
                    int nb, int dt, int ppl, hipStream_t s) {
  return dt == 0 ? batch_ppl<INTERP, PREC, 0>(j, n, ni, c, shm, nb, ppl, s)
                 : batch_ppl<INTERP, PREC, 1>(j, n, ni, c, shm, nb, ppl, s);
}
template <int INTERP>
hipError_t batch_prec(const JobDesc* j, int32_t n, int64_t ni, unsigned long long* c, size_t shm,
                      int nb, int prec, int dt, int ppl, hipStream_t s) {
  switch (prec) {
    case kRef: return batch_dt<INTERP, kRef>(j, n, ni, c, shm, nb, dt, ppl, s);
    case kF64: return batch_dt<INTERP, kF64>(j, n, ni, c, shm, nb, dt, ppl, s);
    case kF32: return batch_dt<INTERP, kF32>(j, n, ni, c, shm, nb, dt, ppl, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

size_t lds_bytes(int interp, int win, int pulses_per_stage) {
  size_t b = lds_fixed_bytes(win, pulses_per_stage);
  if (interp == kSinc8) b += (size_t)kSincQ * 64;
  return b;
}

hipError_t launch_plain(const BpArgs& a, int interp, int precision, int dtype, int ppl,
                        hipStream_t stream) {
  switch (interp) {
    case kNearest: return plain_prec<kNearest>(a, precision, dtype, ppl, stream);
    case kLinear: return plain_prec<kLinear>(a, precision, dtype, ppl, stream);
    case kSinc8: return plain_prec<kSinc8>(a, precision, dtype, ppl, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_batch(const JobDesc* d_jobs, int32_t n_jobs, int64_t n_items,
                        unsigned long long* d_counter, int interp, int precision, int dtype,
                        int ppl, int pulses_per_stage, int win, int n_blocks,
                        hipStream_t stream) {
  const size_t shm = lds_bytes(interp, win, pulses_per_stage);
  switch (interp) {
    case kNearest: return batch_prec<kNearest>(d_jobs, n_jobs, n_items, d_counter, shm, n_blocks, precision, dtype, ppl, stream);
    case kLinear: return batch_prec<kLinear>(d_jobs, n_jobs, n_items, d_counter, shm, n_blocks, precision, dtype, ppl, stream);
    case kSinc8: return batch_prec<kSinc8>(d_jobs, n_jobs, n_items, d_counter, shm, n_blocks, precision, dtype, ppl, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace sarbp
