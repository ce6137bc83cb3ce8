// sarbp_diag.hip -- calibration kernels for the roofline analysis of the BP
// kernels: issue-rate probes of the gfx950 vector ALU.  Each thread runs
// `iters` rounds of 8 independent FP32 chains unrolled 16 deep (128 VALU
// instructions per round), so the issue rate is not limited by latency; the
// caller times the launch with HIP events and reads the SQ counters with
// rocprofv3.  Kinds:
//   0  128 v_fma_f32 per round
//   1  128 v_sin_f32 per round (the hardware sine, argument in revolutions)
//   2  120 v_fma_f32 + 8 v_sin_f32 per round (1 : 15)
//   3  112 v_fma_f32 + 16 v_sin_f32 per round (1 : 7)
//   4  128 v_fmac_f32_e32 (VOP2 encoding: the addend is the destination)
//   5  128 v_mul_f32_e32 (VOP2)
//   6  128 v_add_f32_e32 (VOP2)
//   7  128 v_fract_f32_e32 (VOP1)
//   8  128 v_med3_f32 (VOP3)
//   9  64 v_fma_f32 + 64 v_fmac_f32_e32, alternating
// Kinds 10-16 interleave 64 v_fma_f32 on 4 chains with 64 instructions of one
// other class on 4 independent chains (the partner written as inline asm
// where the compiler would fold a chain of it):
//  10  v_add_u32 (integer, VOP2)         14  v_med3_f32 (VOP3)
//  11  v_mad_u32_u24 (integer, VOP3)     15  v_mul_f32 (VOP2)
//  12  v_cvt_f32_i32 (VOP1)              16  v_mov_b32 (VOP1)
//  13  v_fract_f32 (VOP1)                 17  v_add_f32 (VOP2)
//  18  v_lshrrev_b32 (VOP2)              19  v_and_b32 (VOP2)
//  20  v_or_b32 (VOP2)                   21  v_min_u32 (VOP2)
//  22  v_cndmask_b32 (VOP2)              23  v_lshl_add_u32 (VOP3)
//  24  v_add3_u32 (VOP3)                 25  v_cvt_i32_f32 (VOP1)
//  26  v_max_f32 (VOP2)                  27  v_floor_f32 (VOP1)
//  28  v_sub_u32 (VOP2)                  29  v_bfe_u32 (VOP3)
//  30  v_ldexp_f32 (VOP2)                31  v_ashrrev_i32 (VOP2)
// The BP pixel loop issues 2 transcendentals per 28.4 VALU instructions, between
// kinds 2 and 3, and mixes VOP3 and VOP1/VOP2 encodings.
//
// sarbp_rotation_probe evaluates the BP kernels' own phase rotation
// (sin_rev / cos_rev / rotate_acc of sarbp_rotation.h) on given phases, so a
// test can hold it against FP64 sin(2 pi x) and cos(2 pi x).
#include "sarbp.h"
#include "sarbp_host.h"
#include "sarbp_rotation.h"

namespace sarbp {
namespace {

constexpr int kProbeThreads = 256;

template <int KIND>
__global__ __launch_bounds__(kProbeThreads) void valu_probe_kernel(float* __restrict__ out, int iters,
                                                                  float a, float b) {
  float x[8], y[8];
  int iv[4];
  const int ic = (int)threadIdx.x | 1;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    x[k] = (float)(threadIdx.x + 7 * k) * 1e-3f;
    y[k] = (float)(threadIdx.x + 3 * k) * 1e-4f;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) iv[k] = (int)threadIdx.x + k;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if constexpr (KIND <= 3) {
          const bool trans = KIND == 1 || (KIND == 2 && u == 0) || (KIND == 3 && (u == 0 || u == 8));
          x[k] = trans ? __builtin_amdgcn_sinf(x[k]) : fmaf(x[k], a, b);
        } else if constexpr (KIND == 4) {
          x[k] = fmaf(a, y[k], x[k]);
        } else if constexpr (KIND == 5) {
          x[k] = a * x[k];
        } else if constexpr (KIND == 6) {
          x[k] = x[k] + b;
        } else if constexpr (KIND == 7) {
          x[k] = __builtin_amdgcn_fractf(x[k]);
        } else if constexpr (KIND == 8) {
          x[k] = __builtin_amdgcn_fmed3f(x[k], a, b);
        } else if constexpr (KIND == 9) {
          x[k] = (u & 1) ? fmaf(a, y[k], x[k]) : fmaf(x[k], a, b);
        } else if (k < 4) {
          x[k] = fmaf(x[k], a, b);
        } else if constexpr (KIND == 10) {
          asm volatile("v_add_u32 %0, %0, %1" : "+v"(iv[k - 4]) : "v"(ic));
        } else if constexpr (KIND == 11) {
          asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(iv[k - 4]) : "v"(ic), "v"(ic));
        } else if constexpr (KIND == 12) {
          asm volatile("v_cvt_f32_i32 %0, %1" : "=v"(x[k]) : "v"(iv[k - 4]));
        } else if constexpr (KIND == 13) {
          x[k] = __builtin_amdgcn_fractf(x[k]);
        } else if constexpr (KIND == 14) {
          x[k] = __builtin_amdgcn_fmed3f(x[k], a, b);
        } else if constexpr (KIND == 15) {
          x[k] = a * x[k];
        } else if constexpr (KIND == 16) {
          asm volatile("v_mov_b32 %0, %1" : "=v"(x[k]) : "v"(y[k]));
        } else if constexpr (KIND == 17) {
          asm volatile("v_add_f32 %0, %0, %1" : "+v"(x[k]) : "v"(y[k]));
        } else if constexpr (KIND == 18) {
          asm volatile("v_lshrrev_b32 %0, 1, %0" : "+v"(iv[k - 4]));
        } else if constexpr (KIND == 19) {
          asm volatile("v_and_b32 %0, %0, %1" : "+v"(iv[k - 4]) : "v"(ic));
        } else if constexpr (KIND == 20) {
          asm volatile("v_or_b32 %0, %0, %1" : "+v"(iv[k - 4]) : "v"(ic));
        } else if constexpr (KIND == 21) {
          asm volatile("v_min_u32 %0, %0, %1" : "+v"(iv[k - 4]) : "v"(ic));
        } else if constexpr (KIND == 22) {
          asm volatile("v_cndmask_b32 %0, %0, %1, vcc" : "+v"(iv[k - 4]) : "v"(ic));
        } else if constexpr (KIND == 23) {
          asm volatile("v_lshl_add_u32 %0, %0, 1, %1" : "+v"(iv[k - 4]) : "v"(ic));
        } else if constexpr (KIND == 24) {
          asm volatile("v_add3_u32 %0, %0, %1, %1" : "+v"(iv[k - 4]) : "v"(ic));
        } else if constexpr (KIND == 25) {
          asm volatile("v_cvt_i32_f32 %0, %1" : "=v"(iv[k - 4]) : "v"(y[k]));
        } else if constexpr (KIND == 26) {
          asm volatile("v_max_f32 %0, %0, %1" : "+v"(x[k]) : "v"(y[k]));
        } else if constexpr (KIND == 27) {
          asm volatile("v_floor_f32 %0, %1" : "=v"(x[k]) : "v"(y[k]));
        } else if constexpr (KIND == 28) {
          asm volatile("v_sub_u32 %0, %0, %1" : "+v"(iv[k - 4]) : "v"(ic));
        } else if constexpr (KIND == 29) {
          asm volatile("v_bfe_u32 %0, %0, 1, 20" : "+v"(iv[k - 4]));
        } else if constexpr (KIND == 30) {
          asm volatile("v_ldexp_f32 %0, %0, %1" : "+v"(x[k]) : "v"(ic));
        } else {
          asm volatile("v_ashrrev_i32 %0, 1, %0" : "+v"(iv[k - 4]));
        }
      }
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += x[k];
#pragma unroll
  for (int k = 0; k < 4; ++k) s += (float)iv[k];
  out[(size_t)blockIdx.x * kProbeThreads + threadIdx.x] = s;
}

}  // namespace
}  // namespace sarbp

using namespace sarbp;

// out[2 i], out[2 i + 1] = the rotation of the sample 1 + 0j by exp(+j 2 pi
// ph[i]) as the BP kernels accumulate it: (cos, sin) of ph[i] revolutions.
__global__ void rotation_probe_kernel(const float* __restrict__ ph, float* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float ar = 0.0f, ai = 0.0f;
  sarbp::rotate_acc(1.0f, 0.0f, ph[i], ar, ai);
  out[2 * i] = ar;
  out[2 * i + 1] = ai;
}

extern "C" sarbp_status sarbp_rotation_probe(const float* d_phase, float* d_out, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && (!d_phase || !d_out))) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL buffer or n < 0");
  if (n == 0) return SARBP_OK;
  if (n > ((int64_t)1 << 31) * kProbeThreads / 2)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "n = %lld phases: too many", (long long)n);
  const unsigned blocks = (unsigned)((n + kProbeThreads - 1) / kProbeThreads);
  hipLaunchKernelGGL(rotation_probe_kernel, dim3(blocks), dim3(kProbeThreads), 0, (hipStream_t)stream, d_phase,
                     d_out, n);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SARBP_OK : hip_fail(e, "rotation probe launch");
}

extern "C" sarbp_status sarbp_valu_probe(int32_t kind, int32_t n_blocks, int32_t iters, float* d_out,
                                         void* stream) {
  if (kind < 0 || kind > 31) return fail(SARBP_ERR_INVALID_ARGUMENT, "probe kind %d", kind);
  if (n_blocks < 1 || iters < 1) return fail(SARBP_ERR_INVALID_ARGUMENT, "n_blocks and iters must be >= 1");
  if (!d_out) return fail(SARBP_ERR_INVALID_ARGUMENT, "d_out is NULL");
  hipStream_t s = (hipStream_t)stream;
  // a, b: the FMA chains converge to 1 (no denormals, no overflow)
  const float a = 0.999f, b = 1e-3f;
#define SARBP_PROBE(K) \
  case K: hipLaunchKernelGGL(valu_probe_kernel<K>, dim3(n_blocks), dim3(kProbeThreads), 0, s, d_out, iters, a, b); break;
  switch (kind) {
    SARBP_PROBE(0) SARBP_PROBE(1) SARBP_PROBE(2) SARBP_PROBE(3) SARBP_PROBE(4) SARBP_PROBE(5)
    SARBP_PROBE(6) SARBP_PROBE(7) SARBP_PROBE(8) SARBP_PROBE(9) SARBP_PROBE(10) SARBP_PROBE(11)
    SARBP_PROBE(12) SARBP_PROBE(13) SARBP_PROBE(14) SARBP_PROBE(15) SARBP_PROBE(16) SARBP_PROBE(17)
    SARBP_PROBE(18) SARBP_PROBE(19) SARBP_PROBE(20) SARBP_PROBE(21) SARBP_PROBE(22) SARBP_PROBE(23)
    SARBP_PROBE(24) SARBP_PROBE(25) SARBP_PROBE(26) SARBP_PROBE(27) SARBP_PROBE(28) SARBP_PROBE(29)
    SARBP_PROBE(30) SARBP_PROBE(31)
  }
#undef SARBP_PROBE
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SARBP_OK : hip_fail(e, "VALU probe launch");
}
