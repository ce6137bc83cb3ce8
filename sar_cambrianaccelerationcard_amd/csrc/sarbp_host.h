// sarbp_host.h -- host-side planning helpers and the graph-replayed batched
// launcher shared by the C ABI (sarbp_api.hip) and the server
// (sarbp_server.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "sarbp.h"
#include "sarbp_internal.h"

namespace sarbp {

constexpr size_t kLdsBudget = 64 * 1024;  // dynamic LDS per workgroup
constexpr size_t kCuLds = 160 * 1024;     // LDS of one gfx950 CU
constexpr size_t kBatchStaticLds = 1024;  // bound on the batched kernel's static LDS

// Records a detail message for sarbp_last_error_detail and returns st.
sarbp_status fail(sarbp_status st, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
sarbp_status hip_fail(hipError_t e, const char* what);
const char* last_detail();

// Makes `dev` current for the scope and restores the caller's device after.
struct DeviceGuard {
  int prev = -1;
  hipError_t err = hipSuccess;
  explicit DeviceGuard(int dev) {
    if ((err = hipGetDevice(&prev)) == hipSuccess && prev != dev) err = hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
  DeviceGuard(const DeviceGuard&) = delete;
  DeviceGuard& operator=(const DeviceGuard&) = delete;
};

// ---- planning (sarbp_plan.cpp; no HIP runtime calls) ----
bool finite_all(const double* v, int n);
double norm3(const double* v);
// Kaiser-sinc tap polynomials of the sinc8 kernels (half bins, whole bins).
void sinc8_poly(double beta, float* out);
void sinc8_poly_whole(double beta, float* out);
// Their mirror-sum / mirror-difference rows for a kernel (kSinc8: 48 floats,
// kSinc8W: 32); nothing for the other interpolators.
void sinc8_kernel_rows(int kinterp, double beta, float* out);
sarbp_status validate(const sarbp_grid* g, const sarbp_params* p);
// Size class of a job (0 .. SARBP_SIZE_CLASSES - 1; sarbp_size_class).
int size_class(const sarbp_grid* g, const sarbp_params* p, const double* h_pos);
// Built-in tuning of an interpolator for a size class and precision mode.
// Built-in tuning of a size class; `positions`: the plan knows the antenna
// positions (without them a tile's window covers a disc of its radius, and
// other tunings win).
sarbp_tuning default_tuning(int interp, int cls, int precision, bool positions = true);
// SARBP_ERR_INVALID_ARGUMENT unless the shape is compiled and 1 <= P <= 256.
sarbp_status check_tuning(int precision, const sarbp_tuning* t);
// A tuning fitted to one grid and the arithmetic (kRef, kF64, kF32, kDF) its
// kernels run.
struct Fitted {
  sarbp_tuning t;
  int prec;
};
// The tuning with its tile shape reduced, if needed, to the largest compiled
// shape within it whose tiles fit the REF tile-radius limit on this grid
// (h_pos, if given, bounds the range from below and relaxes the limit).  A
// REF job that no shape fits runs the double-float differential (kDF, no
// radius limit) with one column per lane.
Fitted fit_tuning(sarbp_tuning t, int precision, const sarbp_grid* g, const sarbp_params* p,
                  const double* h_pos);
// Workgroups per CU of the batched BP kernel of interpolator kinterp,
// arithmetic kprec, profile dtype and tile shape px x py with `shm` bytes of
// dynamic LDS (registers, LDS and its allocation granularity counted), as the
// device library's registered source (the HIP occupancy query) reports it; 0
// if no source is registered (host-only builds).
using OccupancyFn = int (*)(int kinterp, int kprec, int dtype, int px, int py, size_t shm);
void set_occupancy_source(OccupancyFn fn);
int kernel_occupancy(int kinterp, int kprec, int dtype, int px, int py, size_t shm);
// Fills a launch descriptor (device pointers left NULL) for one image, with
// kernel arithmetic kprec.  The tuning's pulses per stage is an upper bound:
// while one pulse fewer per stage lets more workgroups share a CU
// (kernel_occupancy), the stage is made one pulse shallower, down to 4 pulses
// for a plain launch and to 1 for a batched one (`batched`).
sarbp_status plan_args(const double* h_pos, const sarbp_grid* g, const sarbp_params* p,
                       const sarbp_tuning& tune, int kprec, BpArgs* a, sarbp_plan_info* info,
                       bool batched = false);
// Launches the plain kernel of a planned descriptor (its kernel_interp and
// kernel_prec, the dtype of p).
sarbp_status launch(const BpArgs& a, const sarbp_params* p, const sarbp_tuning& t, hipStream_t s);

// Launch group of a planned job in a batched launch.  The jobs of a group
// share one persistent kernel (tile shape, interpolation kernel, arithmetic)
// and its LDS size, the largest of theirs; occ (workgroups per CU that the
// job's LDS allows) keeps jobs of different occupancy apart, so one
// LDS-hungry job does not lower the occupancy of all the others.
struct GroupKey {
  int px, py, kinterp, kprec, occ;
};
GroupKey group_key(const BpArgs& a, const sarbp_tuning& t, int dtype);
// Assigns keys[0..n) to groups in first-seen order (group indices to out).
// Keys that differ only in occ share a group when keeping them apart would
// take more than max_groups groups.  Returns the number of groups, or -1 if
// that still exceeds max_groups.  Allocates nothing.
int assign_groups(const GroupKey* keys, int n, int max_groups, int* out);

// Batched persistent launch replayed from a captured hipGraph.  The graph is
// [counter reset kernel -> persistent BP kernel] and reads the job table and
// its header from fixed device buffers, which each run rewrites (async H2D
// from pinned staging) on the caller's stream before replaying the graph.
// One graph is captured per (interp, precision, dtype, tile shape, LDS size
// class); buffers grow on demand (which drops the cached graphs).  Not
// thread-safe: callers serialise.
class BatchExec {
 public:
  BatchExec() = default;
  ~BatchExec();
  BatchExec(const BatchExec&) = delete;
  BatchExec& operator=(const BatchExec&) = delete;

  // Enqueues one batched launch over `jobs` (planned descriptors with device
  // pointers set; item_begin filled in here) on stream s.
  sarbp_status run(std::vector<JobDesc>& jobs, int interp, int precision, int dtype,
                   const sarbp_tuning& tune, int n_cu, hipStream_t s);
  // Blocks until the last run has finished on the device.
  sarbp_status wait();
  // Event recorded after the last run (NULL before the first run).
  hipEvent_t done_event() const { return done_; }
  // Number of graphs captured so far (for diagnostics and tests).
  int graphs_captured() const { return captured_; }

 private:
  sarbp_status reserve(size_t n_jobs);
  void drop_graphs();

  JobDesc* d_jobs_ = nullptr;
  JobDesc* h_jobs_ = nullptr;
  BatchHeader* d_hdr_ = nullptr;
  BatchHeader* h_hdr_ = nullptr;
  unsigned long long* d_counter_ = nullptr;
  size_t cap_ = 0;
  hipEvent_t done_ = nullptr;
  std::map<std::tuple<int, int, int, int, int, size_t>, hipGraphExec_t> graphs_;
  int captured_ = 0;
};

// Iterative radix-2 FFT, FP64, forward (exp(-j ...)), size a power of two.
void fft_host(std::vector<std::complex<double>>& a);

// Device tables of range compression (twiddles, matched filter, phase ramp),
// one set per (FFT size, upsampling, chirp) seen by a context.
struct RangeCache;
void range_cache_destroy(RangeCache* rc);

// hipFFT plans per (na, nr) and the axis buffers of the omega-K entry points.
struct WkCache;
void wk_cache_destroy(WkCache* wc);

}  // namespace sarbp

struct sarbp_context {
  int device = 0;
  int n_cu = 256;
  std::mutex mu;
  std::map<int, sarbp_tuning> tuning;
  // Graph-replayed batched launches, one executor per tile shape of a call
  // (jobs whose grids need different shapes run as separate launches).
  sarbp::BatchExec* batch[8] = {};
  void* scratch = nullptr;               // this device's address of the
                                         // allocation-free batch's scratch
  sarbp::RangeCache* range = nullptr;    // range-compression tables
  sarbp::WkCache* wk = nullptr;          // omega-K transform plans
};
