// sarbp_kernels_linear.hip -- the BP kernel templates of sarbp_kernels.inc for
// the linear interpolator (one translation unit per interpolator).
#define SARBP_TU_INTERP kLinear
#define SARBP_TU_NAME(f) f##_linear
#include "sarbp_kernels.inc"
