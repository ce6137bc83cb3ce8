// sarbp_server.hip -- the device side of the continuous-batching server: the
// HIP executor the server core (sarbp_server_core.cpp) runs against, and the
// two C ABI calls that need the device (sarbp_server_create,
// sarbp_server_submit_raw).  Each of the server's threads has its own stream
// at the default priority: admission copies and range-compresses on one,
// epochs run as graph-replayed batched launches on another (two sets of
// executors: two epochs may be in flight), completion copies images back on
// the third.
#include <memory>
#include <new>

#include "sarbp_host.h"
#include "sarbp_server_core.h"

namespace sarbp {
namespace {

class HipServerExec final : public ServerExec {
 public:
  explicit HipServerExec(int device) : device_(device) {}
  ~HipServerExec() override {
    DeviceGuard dg(device_);
    if (s_compute_) (void)hipStreamSynchronize(s_compute_);
    for (auto& row : exec_)
      for (auto& b : row) b.reset();
    if (arena_) (void)hipFree(arena_);
    if (s_admit_) (void)hipStreamDestroy(s_admit_);
    if (s_compute_) (void)hipStreamDestroy(s_compute_);
    if (s_complete_) (void)hipStreamDestroy(s_complete_);
    sarbp_destroy(rctx_);
  }

  // Streams, the range-compression context and the CU count, on the device.
  sarbp_status init() {
    DeviceGuard dg(device_);
    if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device_) == hipSuccess && cu > 0)
      n_cu_ = cu;
    sarbp_status st = sarbp_create(&rctx_, device_);
    if (st != SARBP_OK) return st;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&s_admit_, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&s_compute_, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&s_complete_, hipStreamNonBlocking)) != hipSuccess)
      return hip_fail(e, "hipStreamCreate");
    return SARBP_OK;
  }

  int n_cu() const override { return n_cu_; }

  sarbp_status alloc_arena(size_t bytes, char** base) override {
    DeviceGuard dg(device_);
    if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
    const hipError_t e = hipMalloc(&arena_, bytes);
    if (e != hipSuccess) {
      arena_ = nullptr;
      return hip_fail(e, "hipMalloc(server arena)");
    }
    *base = static_cast<char*>(arena_);
    return SARBP_OK;
  }

  void bind_thread() override { (void)hipSetDevice(device_); }

  sarbp_status upload(void* dst, const void* src, size_t n) override {
    const hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s_admit_);
    return e == hipSuccess ? SARBP_OK : hip_fail(e, "server H2D");
  }
  sarbp_status compress(const sarbp_cf32* d_raw, int32_t n_pulses, const sarbp_range_params* rp,
                        void* d_out) override {
    return sarbp_range_compress(rctx_, d_raw, n_pulses, rp, d_out, s_admit_);
  }
  sarbp_status admit_wait() override {
    const hipError_t e = hipStreamSynchronize(s_admit_);
    return e == hipSuccess ? SARBP_OK : hip_fail(e, "server H2D");
  }

  sarbp_status epoch_begin(EpochMark* m) override {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    m->begin = e0;
    m->end = e1;
    if (e == hipSuccess) e = hipEventRecord(e0, s_compute_);
    return e == hipSuccess ? SARBP_OK : hip_fail(e, "epoch events");
  }
  sarbp_status run_group(int slot, int g, std::vector<JobDesc>& desc, int kinterp, int kprec, int dtype,
                         const sarbp_tuning& t) override {
    if (slot < 0 || slot > 1 || g < 0 || g >= kServerGroups)
      return fail(SARBP_ERR_INVALID_ARGUMENT, "executor slot %d, group %d", slot, g);
    std::unique_ptr<BatchExec>& b = exec_[slot][g];
    if (!b) b.reset(new (std::nothrow) BatchExec);
    if (!b) return fail(SARBP_ERR_OUT_OF_MEMORY, "batch executor");
    return b->run(desc, kinterp, kprec, dtype, t, n_cu_, s_compute_);
  }
  sarbp_status epoch_end(EpochMark* m) override {
    const hipError_t e = hipEventRecord(static_cast<hipEvent_t>(m->end), s_compute_);
    return e == hipSuccess ? SARBP_OK : hip_fail(e, "hipEventRecord");
  }
  int epoch_done(const EpochMark& m) override {
    const hipError_t e = hipEventQuery(static_cast<hipEvent_t>(m.end));
    return e == hipSuccess ? 1 : (e == hipErrorNotReady ? 0 : -1);
  }
  sarbp_status epoch_wait(const EpochMark& m, float* ms) override {
    *ms = 0.0f;
    const hipError_t e = hipEventSynchronize(static_cast<hipEvent_t>(m.end));
    if (e != hipSuccess) return hip_fail(e, "epoch");
    (void)hipEventElapsedTime(ms, static_cast<hipEvent_t>(m.begin), static_cast<hipEvent_t>(m.end));
    return SARBP_OK;
  }
  void epoch_release(EpochMark* m) override {
    if (m->begin) (void)hipEventDestroy(static_cast<hipEvent_t>(m->begin));
    if (m->end) (void)hipEventDestroy(static_cast<hipEvent_t>(m->end));
    m->begin = m->end = nullptr;
  }
  void compute_wait() override {
    DeviceGuard dg(device_);
    (void)hipStreamSynchronize(s_compute_);
  }

  sarbp_status download(void* dst, const void* src, size_t n) override {
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s_complete_);
    if (e == hipSuccess) e = hipStreamSynchronize(s_complete_);
    return e == hipSuccess ? SARBP_OK : hip_fail(e, "server D2H");
  }

 private:
  int device_;
  int n_cu_ = 256;
  void* arena_ = nullptr;
  hipStream_t s_admit_ = nullptr, s_compute_ = nullptr, s_complete_ = nullptr;
  std::unique_ptr<BatchExec> exec_[2][kServerGroups];  // [epoch parity][group]
  sarbp_context* rctx_ = nullptr;  // range-compression tables of raw-echo jobs
};

}  // namespace
}  // namespace sarbp

using namespace sarbp;

extern "C" {

sarbp_status sarbp_server_create(const sarbp_server_config* cfg, sarbp_server** out) {
  if (!cfg || !out) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
  if (cfg->device < 0 || cfg->device >= n) return fail(SARBP_ERR_INVALID_ARGUMENT, "device %d", cfg->device);
  std::unique_ptr<HipServerExec> ex(new (std::nothrow) HipServerExec(cfg->device));
  if (!ex) return fail(SARBP_ERR_OUT_OF_MEMORY, "server executor");
  sarbp_status st = ex->init();
  if (st != SARBP_OK) return st;
  return server_create(cfg, std::move(ex), out);
}

sarbp_status sarbp_server_submit_raw(sarbp_server* s, const sarbp_raw_request* raw, uint64_t* job_id) {
  if (!s || !raw || !job_id) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  sarbp_raw_request r = *raw;
  r.range.dtype = server_dtype(s);
  double r0 = 0.0, dr = 0.0;
  sarbp_status st = sarbp_range_info(&r.range, nullptr, &r0, &dr);
  if (st != SARBP_OK) return st;
  sarbp_request req{};
  req.h_profiles = nullptr;
  req.h_pos = r.h_pos;
  req.grid = r.grid;
  req.params = r.params;
  req.params.n_samples = r.range.n_out;
  req.params.r0 = r0;
  req.params.dr = dr;
  req.h_image = r.h_image;
  return server_submit(s, &req, &r, job_id);
}

}  // extern "C"
