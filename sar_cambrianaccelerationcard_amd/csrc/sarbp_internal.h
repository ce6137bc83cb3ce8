// sarbp_internal.h -- launch descriptors shared by the host planner
// (sarbp_api.hip) and the gfx950 BP kernels (sarbp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace sarbp {

// One image tile is 64 PX pixels along the grid's x axis by 4 PY along y: a
// workgroup of 4 wave64s, each lane accumulating PX columns (stride 64) x PY
// rows.
constexpr int kTileX = 64;
constexpr int kWaves = 4;
constexpr int kThreads = kTileX * kWaves;
// Sinc8 comes in two kernels.  kSinc8 (half bins): the weight of tap t
// (sample i0 - 3 + t) at fractional bin f, in half h = floor(2 f) of the bin,
// is the degree-5 polynomial sum_k C[h][k][t] e^k of e = 2 f - h - 1/2.
// kSinc8W (whole bins): the degree-7 polynomial sum_k C[k][t] e^k of
// e = f - 1/2.  Half bins cost a pixel 10 FMA instead of 14 and a window
// position 96 FMA instead of 64 of coefficient precompute (mirror sums and
// differences of the taps, the polynomials being even or odd in e about the
// interval's centre); the planner picks per launch.
constexpr int kSincTaps = 8;
constexpr int kSincHalves = 2;
constexpr int kSincCoef = 6;
constexpr int kSincCoefW = 8;
// Per-pulse parameter record in LDS (12 floats or 4 doubles).
constexpr int kParBytes = 48;

enum { kNearest = 0, kLinear = 1, kSinc8 = 2, kSinc8W = 3 };

// Complex coefficients D (8 B each) of the interpolant per window position;
// also the number of 4-float rows of the kernel's tap-polynomial table.
__host__ __device__ constexpr int sinc_rows(int interp) {
  return interp == kSinc8 ? kSincHalves * kSincCoef : (interp == kSinc8W ? kSincCoefW : 0);
}
// LDS stride of one window position's coefficient block, in 16-B units.
// Half bins: 96 B (two 48-B half blocks; 48-B blocks of consecutive half
// positions fall on distinct banks).  Whole bins: the 64-B block padded to
// 80 B, so 8 consecutive positions start on distinct 16-B bank groups; at
// 64 B, positions 2 apart share banks and a coarse grid's lanes, spread
// over many positions, lost ~23% of the CU cycles to bank conflicts.
__host__ __device__ constexpr int coef_stride16(int interp) {
  return interp == kSinc8 ? kSincHalves * kSincCoef / 2 : (interp == kSinc8W ? kSincCoefW / 2 + 1 : 0);
}
enum { kRef = 0, kF64 = 1, kF32 = 2, kDF = 3 };

// Everything one plain launch needs; passed by value as the kernel argument.
struct BpArgs {
  const void* prof;   // [n_pulses][n_samples] c64 or bf16x2
  const double* pos;  // [n_pulses][3]
  float2* img;        // [ny][nx]
  double origin[3], ux[3], uy[3];
  double r0, dr, inv_dr, k_rev;
  double dmax;       // max |pixel - tile reference pixel| over a tile, m
  double order_tol;  // truncation tolerance of the range expansion, rad
  double s_quant;    // kDF: 2^e, the inverse quantum of the direction cosines
                     // e.ux, e.uy, so that their products with integer pixel
                     // offsets, and the sums of those, are exact in FP32
  float inv_dr_f, k_rev_f;
  float k_rev_lo;    // k_rev - k_rev_f (kDF's double-float k)
  int32_t n_pulses, n_samples, nx, ny;
  int32_t tiles_x, tiles_y;
  int32_t win;   // LDS window samples per pulse
  int32_t half;  // window start = floor(ref bin) - half
  int32_t pulses_per_stage;
  int32_t pulse_begin, pulse_end;  // pulses [begin, end) are summed
  int32_t accumulate;
  int32_t kernel_interp;  // kNearest, kLinear, kSinc8 or kSinc8W
  int32_t kernel_prec;    // kRef, kF64, kF32 or kDF: the arithmetic launched
  // tap-polynomial rows of 4 (sinc8_kernel_rows): kSinc8 [6][P, Q] of the
  // mirror sums and differences, kSinc8W [8] (k even: sums, odd: differences)
  float sinc_poly[kSincHalves * kSincCoef * 4];
};

// Batched (persistent) launch: one work item = one tile of one job.
struct JobDesc {
  BpArgs a;            // per-job launch descriptor (tiles_x/tiles_y of the job)
  int64_t item_begin;  // first global work-item index of this job
};

// Size of the job table a batched launch drains (device memory, so one
// captured graph serves tables of any length).
struct BatchHeader {
  int64_t n_items;
  int32_t n_jobs;
  int32_t reserved;
};

// Pulses whose per-(tile, pulse) setup is computed together: the largest
// multiple of the stage size P up to 64 (one wave's lanes), so the FP64 setup
// runs once per SB / P stages.
__host__ __device__ inline int setup_batch(int P) { return P >= 64 ? P : P * (64 / P); }

// LDS layout of one workgroup (byte offsets), SB = setup_batch(P):
//   raw   [P][win] float2         staged range windows (nearest, linear)
//   coef  [P][win-7] blocks       sinc8 only: the R = sinc_rows polynomial
//                                 coefficients (float2) of the interpolant at
//                                 each window position ([2][6] half bins or
//                                 [8] + 16 B of padding), coef_stride16 x 16 B
//   par   [SB] 48-B pulse records
//   rng   [SB] int2: first staged sample, first sinc8 position of each pulse
//   start [SB] int window starts
//   stage [SB/P][3] int per stage: the set of its pulses' range-expansion
//                   orders (bit o), staged samples and sinc8 positions per
//                   pulse
//   poly  [R][4] float  sinc8 only: the tap-polynomial rows (read as
//                       broadcasts)
struct LdsLayout {
  size_t raw, coef, par, rng, start, stage, poly, total;
};

__host__ __device__ inline LdsLayout lds_layout(int interp, int win, int P) {
  LdsLayout L;
  const int SB = setup_batch(P);
  size_t o = 0;
  L.raw = o;
  if (sinc_rows(interp) == 0) o += (size_t)P * win * 8;  // sinc8 reads its taps from the profile
  L.coef = o;
  o += (size_t)P * (win - 7) * coef_stride16(interp) * 16;
  L.par = o;
  o += (size_t)SB * kParBytes;
  L.rng = o;
  o += (size_t)SB * 8;
  L.start = o;
  o += (size_t)SB * 4;
  L.stage = o;
  o += (size_t)(SB / P) * 12;
  o = (o + 15) & ~(size_t)15;
  L.poly = o;
  o += (size_t)sinc_rows(interp) * 16;
  L.total = (o + 15) & ~(size_t)15;
  return L;
}

// Tile shapes (PX columns x PY rows per lane) with compiled kernels: every
// shape for SARBP_PREC_REF, the PX = 1 shapes for the F64, F32 and DF modes.
inline bool tile_shape_supported(int precision, int px, int py) {
  const bool x1 = px == 1 && (py == 2 || py == 4 || py == 8);
  if (precision != kRef) return x1;
  return x1 || (px == 2 && (py == 2 || py == 4 || py == 8)) || (px == 4 && (py == 2 || py == 4));
}

// Per-interpolator entry points (sarbp_kernels_<interp>.hip).
#define SARBP_DECLARE_TU(I)                                                                   \
  hipError_t launch_plain_##I(const BpArgs& a, int precision, int dtype, int px, int py,      \
                              hipStream_t stream);                                            \
  hipError_t launch_batch_##I(const JobDesc* d_jobs, const BatchHeader* d_hdr,               \
                              unsigned long long* d_counter, int precision, int dtype,        \
                              int px, int py, size_t shm, int n_blocks, hipStream_t stream);  \
  int blocks_per_cu_##I(int precision, int dtype, int px, int py, size_t shm);
SARBP_DECLARE_TU(nearest)
SARBP_DECLARE_TU(linear)
SARBP_DECLARE_TU(sinc8)
SARBP_DECLARE_TU(sinc8w)
#undef SARBP_DECLARE_TU

// Launches the plain tile kernel.  Returns hipSuccess or the launch error.
hipError_t launch_plain(const BpArgs& a, int interp, int precision, int dtype,
                        int px, int py, hipStream_t stream);

// Launches the persistent batched kernel over the job table described by
// *d_hdr (device memory); counter = device work counter (zeroed on the
// stream before the launch); shm = LDS bytes >= every job's layout.
hipError_t launch_batch(const JobDesc* d_jobs, const BatchHeader* d_hdr,
                        unsigned long long* d_counter, int interp,
                        int precision, int dtype, int px, int py, size_t shm,
                        int n_blocks, hipStream_t stream);

// Zeroes the batch work counter with a one-thread kernel (a kernel node, not a
// memset node, so the sequence can be captured into a graph).
hipError_t reset_counter(unsigned long long* d_counter, hipStream_t stream);

// Workgroups of the batched kernel that fit one CU at the given LDS size.
int batch_blocks_per_cu(int interp, int precision, int dtype, int px, int py, size_t shm);

}  // namespace sarbp
