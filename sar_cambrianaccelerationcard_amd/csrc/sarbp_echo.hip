// sarbp_echo.hip -- point-target raw-echo generator on the device: the
// reference's generateEchoKernel / launchGenerateEcho
// (code/SAR_BP/CUDA/BPA_CUDA_kernels.cu:3-45, :79-89) behind the C ABI.
//
// Sample (i, j) of Srnm [Mslow][Nfast] is the sum over targets k of
// sigma_k exp(j (pi Kr t^2 - 4 pi R / lambda)) for the k inside the pulse
// (0 < t < Tr) and the beam (|Az| < La/2), Az = ta[i] vr - x_k,
// R = sqrt(Az^2 + y_k^2 + H^2), t = tr[j] - 2 R / c.  Range, gate and phase
// are FP64, in that order with one rounding per operation (no contraction),
// so the gate decisions are those of an FP64 evaluation of the formula; the
// phase is reduced to revolutions in FP64 and its sine and cosine are the
// hardware ones (which take revolutions).  The reference evaluates all of it
// in FP32, where 4 pi R / lambda ~ 1e6 rad carries ~0.1 rad of error.
//
// One thread per sample, grid-stride over Mslow * Nfast in 64-bit indices;
// the target list is read uniformly by every lane of a wave.
#include "sarbp.h"
#include "sarbp_host.h"

#include <cmath>

namespace sarbp {
namespace {

struct EchoArgs {
  float2* out;
  const float* tg;  // [ntarget][3]: azimuth, ground range, sigma
  const float* ta;  // [mslow]
  const float* tr;  // [nfast]
  long long total;  // mslow * nfast
  int nfast, ntarget;
  double c, lambda, kr, tw, vr, h, la_half;
};

__global__ __launch_bounds__(256) void echo_kernel(EchoArgs a) {
#pragma clang fp contract(off)
  const long long step = (long long)gridDim.x * blockDim.x;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < a.total; idx += step) {
    const int i = (int)(idx / a.nfast);
    const int j = (int)(idx - (long long)i * a.nfast);
    const double ta = a.ta[i], tr = a.tr[j];
    float sx = 0.0f, sy = 0.0f;
    for (int k = 0; k < a.ntarget; ++k) {
      const double x = a.tg[3 * k], y = a.tg[3 * k + 1];
      const float sigma = a.tg[3 * k + 2];
      const double az = ta * a.vr - x;
      const double r = sqrt(az * az + y * y + a.h * a.h);
      const double t = tr - 2.0 * r / a.c;
      if (t > 0.0 && t < a.tw && fabs(az) < a.la_half) {
        double rev = 0.5 * a.kr * t * t - 2.0 * r / a.lambda;
        rev -= rint(rev);
        const float ph = (float)rev;
        sx += sigma * __builtin_amdgcn_cosf(ph);
        sy += sigma * __builtin_amdgcn_sinf(ph);
      }
    }
    a.out[idx] = make_float2(sx, sy);
  }
}

}  // namespace
}  // namespace sarbp

extern "C" {

sarbp_status sarbp_launch_generate_echo(sarbp_cf32* Srnm, const float* Ptarget, const float* ta, const float* tr,
                                        sarbp_sar_params p, int32_t Mslow, int32_t Nfast) {
  using namespace sarbp;
  if (Mslow < 0 || Nfast < 0 || p.Ntarget < 0)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "Mslow %d, Nfast %d, Ntarget %d: negative", Mslow, Nfast, p.Ntarget);
  const long long total = (long long)Mslow * Nfast;
  if (total == 0) return SARBP_OK;
  if (!Srnm || !ta || !tr || (p.Ntarget > 0 && !Ptarget))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "Srnm, Ptarget, ta or tr is NULL");
  if (!(std::isfinite(p.c) && p.c > 0.0f && std::isfinite(p.lambda) && p.lambda != 0.0f && std::isfinite(p.Kr) &&
        std::isfinite(p.Tr) && std::isfinite(p.vr) && std::isfinite(p.H) && std::isfinite(p.La)))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "SARParams c, lambda, Kr, Tr, vr, H or La invalid");
  EchoArgs a{};
  a.out = reinterpret_cast<float2*>(Srnm);
  a.tg = Ptarget;
  a.ta = ta;
  a.tr = tr;
  a.total = total;
  a.nfast = Nfast;
  a.ntarget = p.Ntarget;
  a.c = p.c;
  a.lambda = p.lambda;
  a.kr = p.Kr;
  a.tw = p.Tr;
  a.vr = p.vr;
  a.h = p.H;
  a.la_half = (double)p.La / 2.0;
  const long long blocks = (total + 255) / 256;
  const unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
  hipLaunchKernelGGL(echo_kernel, dim3(grid), dim3(256), 0, nullptr, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "echo_kernel launch");
  // Synchronous, as the original (BPA_CUDA_kernels.cu:88).
  if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
  return SARBP_OK;
}

}  // extern "C"
