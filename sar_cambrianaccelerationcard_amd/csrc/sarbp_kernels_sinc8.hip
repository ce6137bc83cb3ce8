// sarbp_kernels_sinc8.hip -- the BP kernel templates of sarbp_kernels.inc for
// the sinc8 interpolator (one translation unit per interpolator).
#define SARBP_TU_INTERP kSinc8
#define SARBP_TU_NAME(f) f##_sinc8
#include "sarbp_kernels.inc"
