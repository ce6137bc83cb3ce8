// sarbp_rotation.h -- the phase rotation of the back-projection sum, shared by
// the BP kernels and the diagnostic kernel that measures its accuracy.
//
// The phase is carried in revolutions (frac(2 R / lambda) and its FP32
// differentials), which is the argument unit of the hardware v_sin_f32 /
// v_cos_f32: no 2 pi scaling and no software range reduction.  The kernels
// keep |phase| below 128 revolutions (the instructions' domain is +-256).
// Replaces the sincosf of code/SAR_BP/CUDA/BPA_CUDA_kernels.cu:67.
#pragma once

#include <hip/hip_runtime.h>

namespace sarbp {

__device__ __forceinline__ float sin_rev(float x) { return __builtin_amdgcn_sinf(x); }
__device__ __forceinline__ float cos_rev(float x) { return __builtin_amdgcn_cosf(x); }

// MAC of one interpolated sample rotated by exp(+j 2 pi ph).
__device__ __forceinline__ void rotate_acc(float vr, float vi, float ph, float& ar, float& ai) {
  const float sn = sin_rev(ph), cs = cos_rev(ph);
  ar = fmaf(vr, cs, ar);
  ar = fmaf(-vi, sn, ar);
  ai = fmaf(vr, sn, ai);
  ai = fmaf(vi, cs, ai);
}

}  // namespace sarbp
