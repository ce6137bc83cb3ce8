// sarbp_batch.hip -- BatchExec: the batched persistent BP launch, captured once
// per configuration into a hipGraph and replayed over a device job table.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "sarbp_host.h"

namespace sarbp {

namespace {
__global__ void reset_counter_kernel(unsigned long long* counter) { *counter = 0ull; }
}  // namespace

hipError_t reset_counter(unsigned long long* d_counter, hipStream_t stream) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, d_counter);
  return hipGetLastError();
}

hipError_t launch_plain(const BpArgs& a, int interp, int precision, int dtype, int px, int py,
                        hipStream_t stream) {
  switch (interp) {
    case kNearest: return launch_plain_nearest(a, precision, dtype, px, py, stream);
    case kLinear: return launch_plain_linear(a, precision, dtype, px, py, stream);
    case kSinc8: return launch_plain_sinc8(a, precision, dtype, px, py, stream);
    case kSinc8W: return launch_plain_sinc8w(a, precision, dtype, px, py, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_batch(const JobDesc* d_jobs, const BatchHeader* d_hdr, unsigned long long* d_counter,
                        int interp, int precision, int dtype, int px, int py, size_t shm, int n_blocks,
                        hipStream_t stream) {
  switch (interp) {
    case kNearest:
      return launch_batch_nearest(d_jobs, d_hdr, d_counter, precision, dtype, px, py, shm, n_blocks, stream);
    case kLinear:
      return launch_batch_linear(d_jobs, d_hdr, d_counter, precision, dtype, px, py, shm, n_blocks, stream);
    case kSinc8:
      return launch_batch_sinc8(d_jobs, d_hdr, d_counter, precision, dtype, px, py, shm, n_blocks, stream);
    case kSinc8W:
      return launch_batch_sinc8w(d_jobs, d_hdr, d_counter, precision, dtype, px, py, shm, n_blocks, stream);
    default: return hipErrorInvalidValue;
  }
}

int batch_blocks_per_cu(int interp, int precision, int dtype, int px, int py, size_t shm) {
  switch (interp) {
    case kNearest: return blocks_per_cu_nearest(precision, dtype, px, py, shm);
    case kLinear: return blocks_per_cu_linear(precision, dtype, px, py, shm);
    case kSinc8: return blocks_per_cu_sinc8(precision, dtype, px, py, shm);
    case kSinc8W: return blocks_per_cu_sinc8w(precision, dtype, px, py, shm);
    default: return 1;
  }
}

namespace {
// Held across every graph capture of a batched launch, and across the
// occupancy queries of planning: a query can load a kernel's code object,
// which must not overlap another thread's capture (a server plans jobs on its
// client and admission threads while its scheduler thread captures).
std::mutex g_capture_mu;

// The occupancy source of planning (set_occupancy_source): the HIP query for
// the batched kernel, cached per kernel and LDS size (a plan asks for a few
// stage depths; every gfx950 device gives the same answer).
int cached_occupancy(int kinterp, int kprec, int dtype, int px, int py, size_t shm) {
  static std::map<std::tuple<int, int, int, int, int, size_t>, int> cache;
  const auto key = std::make_tuple(kinterp, kprec, dtype, px, py, shm);
  std::lock_guard<std::mutex> lk(g_capture_mu);
  const auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const int n = batch_blocks_per_cu(kinterp, kprec, dtype, px, py, shm);
  cache.emplace(key, n);
  return n;
}
const bool g_occupancy_registered = (set_occupancy_source(&cached_occupancy), true);
}  // namespace

BatchExec::~BatchExec() {
  if (done_) {
    (void)hipEventSynchronize(done_);
    (void)hipEventDestroy(done_);
  }
  drop_graphs();
  if (d_jobs_) (void)hipFree(d_jobs_);
  if (h_jobs_) (void)hipHostFree(h_jobs_);
  if (d_hdr_) (void)hipFree(d_hdr_);
  if (h_hdr_) (void)hipHostFree(h_hdr_);
  if (d_counter_) (void)hipFree(d_counter_);
}

void BatchExec::drop_graphs() {
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
}

sarbp_status BatchExec::wait() {
  if (!done_) return SARBP_OK;
  hipError_t e = hipEventSynchronize(done_);
  return e == hipSuccess ? SARBP_OK : hip_fail(e, "hipEventSynchronize(batch)");
}

sarbp_status BatchExec::reserve(size_t n) {
  hipError_t e;
  if (!d_hdr_) {
    if ((e = hipMalloc(&d_hdr_, sizeof(BatchHeader))) != hipSuccess) return hip_fail(e, "hipMalloc(header)");
    if ((e = hipHostMalloc(&h_hdr_, sizeof(BatchHeader), hipHostMallocDefault)) != hipSuccess)
      return hip_fail(e, "hipHostMalloc(header)");
    if ((e = hipMalloc(&d_counter_, sizeof(unsigned long long))) != hipSuccess)
      return hip_fail(e, "hipMalloc(counter)");
  }
  if (n <= cap_) return SARBP_OK;
  // The graphs hold the old table pointer.
  drop_graphs();
  if (d_jobs_) (void)hipFree(d_jobs_);
  if (h_jobs_) (void)hipHostFree(h_jobs_);
  d_jobs_ = nullptr;
  h_jobs_ = nullptr;
  cap_ = 0;
  const size_t cap = std::max<size_t>(64, n + n / 2);
  if ((e = hipMalloc(&d_jobs_, sizeof(JobDesc) * cap)) != hipSuccess) return hip_fail(e, "hipMalloc(jobs)");
  if ((e = hipHostMalloc(&h_jobs_, sizeof(JobDesc) * cap, hipHostMallocDefault)) != hipSuccess)
    return hip_fail(e, "hipHostMalloc(jobs)");
  cap_ = cap;
  return SARBP_OK;
}

sarbp_status BatchExec::run(std::vector<JobDesc>& jobs, int interp, int precision, int dtype,
                            const sarbp_tuning& tune, int n_cu, hipStream_t s) {
  const int px = tune.cols_per_lane, py = tune.rows_per_lane;
  if (jobs.empty()) return SARBP_OK;
  hipError_t e;
  // The pinned staging copies of the previous run must have been consumed.
  sarbp_status st = wait();
  if (st != SARBP_OK) return st;
  if (!done_ && (e = hipEventCreateWithFlags(&done_, hipEventDisableTiming)) != hipSuccess)
    return hip_fail(e, "hipEventCreate(batch)");
  if ((st = reserve(jobs.size())) != SARBP_OK) return st;
  int64_t items = 0;
  size_t shm = 0;
  for (JobDesc& d : jobs) {
    d.item_begin = items;
    items += (int64_t)d.a.tiles_x * d.a.tiles_y;
    shm = std::max(shm, lds_layout(interp, d.a.win, d.a.pulses_per_stage).total);
  }
  if (items == 0) return SARBP_OK;
  // LDS size class: the largest size (256-B steps) that keeps as many
  // workgroups per CU as the group's largest job needs, so the class costs no
  // occupancy and at most ~16 graphs are captured per kernel.  The estimate
  // bounds the static LDS from above; the occupancy query confirms it, else
  // the job's own size is the class.
  const int per_cu_job = cached_occupancy(interp, precision, dtype, px, py, shm);
  const size_t per_cu_fit = std::max<size_t>(1, kCuLds / (shm + kBatchStaticLds));
  size_t shm_class = std::min(kLdsBudget, std::max(shm, (kCuLds / per_cu_fit - kBatchStaticLds) & ~(size_t)255));
  if (shm_class != shm && cached_occupancy(interp, precision, dtype, px, py, shm_class) < per_cu_job)
    shm_class = shm;
  std::memcpy(h_jobs_, jobs.data(), sizeof(JobDesc) * jobs.size());
  h_hdr_->n_items = items;
  h_hdr_->n_jobs = (int32_t)jobs.size();
  h_hdr_->reserved = 0;
  if ((e = hipMemcpyAsync(d_jobs_, h_jobs_, sizeof(JobDesc) * jobs.size(), hipMemcpyHostToDevice, s)) !=
      hipSuccess)
    return hip_fail(e, "hipMemcpyAsync(jobs)");
  if ((e = hipMemcpyAsync(d_hdr_, h_hdr_, sizeof(BatchHeader), hipMemcpyHostToDevice, s)) != hipSuccess)
    return hip_fail(e, "hipMemcpyAsync(header)");

  const auto key = std::make_tuple(interp, precision, dtype, px, py, shm_class);
  auto it = graphs_.find(key);
  if (it == graphs_.end()) {
    // The capture stream lives only for the capture: a process's streams share
    // a few hardware queues, and a long-lived idle stream shifts the queue
    // assignment of streams created later (a server's copy stream could land
    // on its compute stream's queue and wait behind every kernel).
    hipStream_t cs = nullptr;
    if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess)
      return hip_fail(e, "hipStreamCreate(capture)");
    const int per_cu = cached_occupancy(interp, precision, dtype, px, py, shm_class);
    const int n_blocks = std::max(1, n_cu * per_cu);
    std::lock_guard<std::mutex> cap(g_capture_mu);
    if ((e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
      (void)hipStreamDestroy(cs);
      return hip_fail(e, "hipStreamBeginCapture");
    }
    hipError_t e1 = reset_counter(d_counter_, cs);
    hipError_t e2 = e1 == hipSuccess ? launch_batch(d_jobs_, d_hdr_, d_counter_, interp, precision, dtype, px,
                                                    py, shm_class, n_blocks, cs)
                                     : e1;
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(e2, "batched launch capture");
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    it = graphs_.emplace(key, ge).first;
    ++captured_;
  }
  if ((e = hipGraphLaunch(it->second, s)) != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  if ((e = hipEventRecord(done_, s)) != hipSuccess) return hip_fail(e, "hipEventRecord(batch)");
  return SARBP_OK;
}

}  // namespace sarbp
