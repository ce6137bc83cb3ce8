// sarbp_fft.h -- in-LDS power-of-two FFTs for one workgroup: Stockham passes
// of radix 8 (n/8 threads) or 16 (n/16 threads), after one smaller-radix pass
// when log2 n is not a multiple of 3 or 4; used by range compression
// (sarbp_range.hip) and by the omega-K transforms (sarbp_wk.hip).  fft_regs
// keeps the points in registers (8 or 16 per thread) and sends only the inner
// passes through LDS.  Not part of the public ABI.
//
// Element i of the transform lives at LDS slot fft::slot<LG>(i): one pad word
// per 16 for n <= 4096 (buffer fft::lds_words<LG>() complex64 words), an XOR
// swizzle of bits 1-3 by bits 4-6 for n >= 8192 (exactly n words: 64 KiB at
// 8192 points, 128 KiB at 16384).  Either turns the power-of-two strides of
// the Stockham writes into ones whose 16 lanes per LDS access hit distinct
// banks.
#pragma once

#include <hip/hip_runtime.h>

namespace sarbp {
namespace fft {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

template <int LG>
__device__ __forceinline__ int slot(int i) {
  if constexpr (LG <= 12)
    return i + (i >> 4);
  else
    return i ^ (((i >> 4) & 7) << 1);
}
template <int LG>
constexpr int lds_words() { return LG <= 12 ? (1 << LG) + (1 << LG) / 16 : 1 << LG; }

// R-point DFT of x in registers, in place, natural output order.  INV: the
// conjugate transform (unscaled inverse).
template <int R, bool INV>
__device__ __forceinline__ void dft_r(float2* x) {
  // multiply by -j (forward) / +j (inverse)
  auto mj = [](float2 v) { return INV ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x); };
  if constexpr (R == 2) {
    const float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const float2 t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]), t2 = cadd(x[1], x[3]), t3 = mj(csub(x[1], x[3]));
    x[0] = cadd(t0, t2);
    x[1] = cadd(t1, t3);
    x[2] = csub(t0, t2);
    x[3] = csub(t1, t3);
  } else if constexpr (R == 16) {
    // 16 = 4 x 4: A[n1][k1] = DFT4_n2(x[n1 + 4 n2]) W16^(n1 k1), then
    // y[k1 + 4 k2] = DFT4_n1(A[n1][k1]).
    constexpr float c1 = 0.92387953251128676f, s1 = 0.38268343236508977f, h = 0.70710678118654752f;
    float2 a[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
      a[n1][0] = x[n1];
      a[n1][1] = x[n1 + 4];
      a[n1][2] = x[n1 + 8];
      a[n1][3] = x[n1 + 12];
      dft_r<4, INV>(a[n1]);
    }
    // W16^e = (cos(2 pi e/16), -+sin(2 pi e/16)) for e = n1 k1 in {1,2,3,4,6,9}.
    const float sg = INV ? 1.0f : -1.0f;
    const float2 w1 = make_float2(c1, sg * s1), w2 = make_float2(h, sg * h), w3 = make_float2(s1, sg * c1);
    const float2 w6 = make_float2(-h, sg * h), w9 = make_float2(-c1, -sg * s1);
    a[1][1] = cmul(a[1][1], w1);
    a[1][2] = cmul(a[1][2], w2);
    a[1][3] = cmul(a[1][3], w3);
    a[2][1] = cmul(a[2][1], w2);
    a[2][2] = mj(a[2][2]);  // W16^4
    a[2][3] = cmul(a[2][3], w6);
    a[3][1] = cmul(a[3][1], w3);
    a[3][2] = cmul(a[3][2], w6);
    a[3][3] = cmul(a[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 col[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      dft_r<4, INV>(col);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) x[k1 + 4 * k2] = col[k2];
    }
  } else {
    // 8 = 4 x 2: u_k = x_k + x_{k+4}, v_k = W8^k (x_k - x_{k+4}); the even
    // outputs are DFT4(u), the odd ones DFT4(v).
    constexpr float s = 0.70710678118654752f;
    float2 u[4], v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      u[k] = cadd(x[k], x[k + 4]);
      v[k] = csub(x[k], x[k + 4]);
    }
    // W8 = (1 - j)/sqrt2 forward, (1 + j)/sqrt2 inverse; W8^2 = -j / +j.
    v[1] = INV ? make_float2(s * (v[1].x - v[1].y), s * (v[1].x + v[1].y))
               : make_float2(s * (v[1].x + v[1].y), s * (v[1].y - v[1].x));
    v[2] = mj(v[2]);
    v[3] = INV ? make_float2(-s * (v[3].x + v[3].y), s * (v[3].x - v[3].y))
               : make_float2(s * (v[3].y - v[3].x), -s * (v[3].x + v[3].y));
    dft_r<4, INV>(u);
    dft_r<4, INV>(v);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      x[2 * k] = u[k];
      x[2 * k + 1] = v[k];
    }
  }
}

// Twiddles and R-point DFT of butterfly b of a Stockham pass over NF = 2^LG
// points after sub-transforms of size ns = 2^lns: input r is multiplied by
// W^{r k}, k = b mod ns, of size R ns (tw[t] = exp(-j 2 pi t / NF); powers by
// repeated products).
template <int LG, int R, bool INV>
__device__ __forceinline__ void twiddle_dft(float2* y, const float2* __restrict__ tw, int b, int lns) {
  constexpr int LGR = R == 16 ? 4 : (R == 8 ? 3 : (R == 4 ? 2 : 1));
  if (lns > 0) {
    const int k = b & ((1 << lns) - 1);
    float2 w[R];
    w[1] = tw[k << (LG - lns - LGR)];
    if (INV) w[1].y = -w[1].y;
    // Powers by few products: w2 = w1^2, w3 = w2 w1, w4 = w2^2,
    // w5..w7 = w4 w1..w3.
    if constexpr (R >= 4) {
      w[2] = cmul(w[1], w[1]);
      w[3] = cmul(w[2], w[1]);
    }
    if constexpr (R >= 8) {
      w[4] = cmul(w[2], w[2]);
      w[5] = cmul(w[4], w[1]);
      w[6] = cmul(w[4], w[2]);
      w[7] = cmul(w[4], w[3]);
    }
    if constexpr (R == 16) {
      // w8 = w4^2, w9..w15 = w8 w1..w7: at most four products deep.
      w[8] = cmul(w[4], w[4]);
#pragma unroll
      for (int q = 1; q < 8; ++q) w[8 + q] = cmul(w[8], w[q]);
    }
#pragma unroll
    for (int r = 1; r < R; ++r) y[r] = cmul(y[r], w[r]);
  }
  dft_r<R, INV>(y);
}

// One radix-R Stockham pass over NF = 2^LG points in LDS after sub-transforms
// of size ns = 2^lns: butterfly b (k = b mod ns) reads b + r NF/R, multiplies
// input r by W^{r k} of size R ns (twiddle_dft), and writes output r to
// (b - k) R + k + r ns.  All reads complete (barrier) before any write.
template <int LG, int T, int R, bool INV>
__device__ __forceinline__ void stockham_pass(float2* buf, const float2* __restrict__ tw, int lns,
                                              int tid = threadIdx.x) {
  constexpr int NF = 1 << LG;
  constexpr int kT = T;
  constexpr int NB = NF / R;
  constexpr int PER = (NB + kT - 1) / kT;
  constexpr int LGR = R == 16 ? 4 : (R == 8 ? 3 : (R == 4 ? 2 : 1));
  const int ns = 1 << lns;
  float2 y[PER][R];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * kT;
    if (NB % kT == 0 || b < NB) {
#pragma unroll
      for (int r = 0; r < R; ++r) y[p][r] = buf[slot<LG>(b + r * NB)];
      twiddle_dft<LG, R, INV>(y[p], tw, b, lns);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * kT;
    if (NB % kT == 0 || b < NB) {
      const int k = b & (ns - 1);
      const int dst = ((b - k) << LGR) + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[slot<LG>(dst + r * ns)] = y[p][r];
    }
  }
  __syncthreads();
}

// FFT of NF = 2^LG points held in registers by T = NF/PT threads, PT = 8 or
// 16 points each: on entry x[e] is point tid + e T, on return x[e] is output
// tid + e T.  The first pass (radix R0 = 2^(LG mod log2 PT), or PT when that
// is 1; butterfly b = tid + p T reads b + r NF/R0, i.e. x[p + r PT/R0]) and the
// last radix-PT pass (butterfly tid writes tid + r T) touch only the thread's
// own points, so only the passes between go through LDS.  buf
// (lds_words<LG>() words) is scratch: every thread must be done reading it
// when the call starts, and a barrier must separate the call's last reads of
// buf from the next writes.
template <int LG, bool INV, int PT = 16>
__device__ __forceinline__ void fft_regs(float2* buf, const float2* __restrict__ tw, float2 (&x)[PT]) {
  static_assert(PT == 8 || PT == 16, "8 or 16 points per thread");
  constexpr int LGP = PT == 16 ? 4 : 3;
  constexpr int NF = 1 << LG, T = NF / PT;
  constexpr int LGR0 = LG % LGP == 0 ? LGP : LG % LGP;
  constexpr int R0 = 1 << LGR0, PER0 = PT / R0;
  static_assert(LG >= 2 * LGP, "two passes at least");
  // The thread index goes through an empty asm per call, so that the LDS
  // addresses, which depend on it alone, are recomputed in each transform
  // instead of being kept live (and spilled) from one transform to the next.
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
#pragma unroll
  for (int p = 0; p < PER0; ++p) {
    float2 y[R0];
#pragma unroll
    for (int r = 0; r < R0; ++r) y[r] = x[p + r * PER0];
    dft_r<R0, INV>(y);
    const int dst = (tid + p * T) << LGR0;
#pragma unroll
    for (int r = 0; r < R0; ++r) buf[slot<LG>(dst + r)] = y[r];
  }
  __syncthreads();
  int lns = LGR0;
#pragma unroll
  for (int pass = 0; pass < (LG - LGR0) / LGP - 1; ++pass) {
    stockham_pass<LG, T, PT, INV>(buf, tw, lns, tid);
    lns += LGP;
  }
#pragma unroll
  for (int r = 0; r < PT; ++r) x[r] = buf[slot<LG>(tid + r * T)];
  twiddle_dft<LG, PT, INV>(x, tw, tid, lns);
}

// In-place FFT of NF = 2^LG points held in LDS at slots slot<LG>(i) by a
// workgroup of T threads: a radix 2 or 4 pass when LG is not a multiple of 3,
// then radix-8 passes.  tw[t] = exp(-j 2 pi t / NF).  INV uses the conjugate
// twiddles (unscaled inverse).  Ends with a barrier.
template <int LG, int T, bool INV>
__device__ __forceinline__ void fft_lds(float2* buf, const float2* __restrict__ tw) {
  int lns = 0;
  if constexpr (T * 16 == (1 << LG)) {
    // 16 points per thread: a radix 2, 4 or 8 pass when LG is not a multiple
    // of 4, then radix-16 passes.
    if constexpr (LG % 4 == 1) {
      stockham_pass<LG, T, 2, INV>(buf, tw, 0);
      lns = 1;
    } else if constexpr (LG % 4 == 2) {
      stockham_pass<LG, T, 4, INV>(buf, tw, 0);
      lns = 2;
    } else if constexpr (LG % 4 == 3) {
      stockham_pass<LG, T, 8, INV>(buf, tw, 0);
      lns = 3;
    }
    if constexpr (LG >= 4) {
#pragma unroll
      for (int pass = 0; pass < LG / 4; ++pass) {
        stockham_pass<LG, T, 16, INV>(buf, tw, lns);
        lns += 4;
      }
    }
    return;
  }
  if constexpr (LG % 3 == 1) {
    stockham_pass<LG, T, 2, INV>(buf, tw, 0);
    lns = 1;
  } else if constexpr (LG % 3 == 2) {
    stockham_pass<LG, T, 4, INV>(buf, tw, 0);
    lns = 2;
  }
  if constexpr (LG >= 3) {
#pragma unroll
    for (int pass = 0; pass < LG / 3; ++pass) {
      stockham_pass<LG, T, 8, INV>(buf, tw, lns);
      lns += 3;
    }
  }
}

}  // namespace fft
}  // namespace sarbp
