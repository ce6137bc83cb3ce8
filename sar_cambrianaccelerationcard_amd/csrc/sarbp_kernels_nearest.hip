// sarbp_kernels_nearest.hip -- the BP kernel templates of sarbp_kernels.inc for
// the nearest interpolator (one translation unit per interpolator).
#define SARBP_TU_INTERP kNearest
#define SARBP_TU_NAME(f) f##_nearest
#include "sarbp_kernels.inc"
