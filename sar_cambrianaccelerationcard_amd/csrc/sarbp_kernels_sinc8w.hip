// sarbp_kernels_sinc8w.hip -- the BP kernel templates of sarbp_kernels.inc for
// the whole-bin sinc8 interpolator (one translation unit per interpolator).
#define SARBP_TU_INTERP kSinc8W
#define SARBP_TU_NAME(f) f##_sinc8w
#include "sarbp_kernels.inc"
