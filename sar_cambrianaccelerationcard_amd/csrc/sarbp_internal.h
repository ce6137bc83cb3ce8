// sarbp_internal.h -- launch descriptors shared by the host planner
// (sarbp_api.hip) and the gfx950 BP kernels (sarbp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sarbp {

// One image tile is 64 pixels along the grid's x axis (one wave64, lane = x)
// by 4 * PPL pixels along y (4 waves, PPL rows per lane).
constexpr int kTileX = 64;
constexpr int kWaves = 4;
constexpr int kThreads = kTileX * kWaves;
// Sinc8 weight table: kSincQ intervals over the fractional bin [0, 1); row i
// holds, for the 8 taps t = 0..7 (sample i0 - 3 + t), the pair
// (K(i/Q + 3 - t), K((i+1)/Q + 3 - t) - K(i/Q + 3 - t)).
constexpr int kSincQ = 512;
// Per-pulse parameter record in LDS (8 floats or 4 doubles).
constexpr int kParBytes = 32;

// Everything one plain launch needs; passed by value as the kernel argument.
struct BpArgs {
  const void* prof;        // [n_pulses][n_samples] c64 or bf16x2
  const double* pos;       // [n_pulses][3]
  float2* img;             // [ny][nx]
  const float4* sinc_tab;  // [kSincQ][4] float4 = 8 taps x (K, dK)
  double origin[3], ux[3], uy[3];
  double r0, dr, inv_dr, k_rev;
  float inv_dr_f, k_rev_f;
  int32_t n_pulses, n_samples, nx, ny;
  int32_t tiles_x, tiles_y;
  int32_t win;        // LDS window samples per pulse
  int32_t half;       // window start = floor(ref bin) - half
  int32_t pulses_per_stage;
  int32_t pulse_begin, pulse_end;  // pulses [begin, end) are summed
  int32_t accumulate;
  double dmax;       // max |pixel - tile reference pixel| over a tile, m
  double order_tol;  // truncation tolerance of the range expansion, rad
};

// Batched (persistent) launch: one work item = one tile of one job.
struct JobDesc {
  BpArgs a;              // per-job launch descriptor (tiles_x/tiles_y of the job)
  int64_t item_begin;    // first global work-item index of this job
};

// LDS layout of one workgroup: [P][win] float2 windows | [P] 32-B pulse
// records | [P] int window starts | [P] float 2Rc | int stage order | pad to
// 16 B | sinc table (sinc8 only).
__host__ __device__ inline size_t lds_fixed_bytes(int win, int pulses_per_stage) {
  size_t b = (size_t)pulses_per_stage * win * 8 + (size_t)pulses_per_stage * kParBytes +
             (size_t)pulses_per_stage * 8 + 4;
  return (b + 15) & ~(size_t)15;
}
size_t lds_bytes(int interp, int win, int pulses_per_stage);

// Launches the plain tile kernel.  Returns hipSuccess or the launch error.
hipError_t launch_plain(const BpArgs& a, int interp, int precision, int dtype,
                        int ppl, hipStream_t stream);

// Launches the persistent batched kernel over n_jobs descriptors in device
// memory; n_items = total tiles; counter = device int64 work counter (zeroed
// by the caller on the stream before the launch).
hipError_t launch_batch(const JobDesc* d_jobs, int32_t n_jobs, int64_t n_items,
                        unsigned long long* d_counter, int interp,
                        int precision, int dtype, int ppl, int pulses_per_stage,
                        int win, int n_blocks, hipStream_t stream);

}  // namespace sarbp
