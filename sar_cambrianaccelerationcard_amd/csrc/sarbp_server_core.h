// sarbp_server_core.h -- the continuous-batching server's host logic and the
// device interface it runs against.  The core (sarbp_server_core.cpp) holds
// the three threads, the job table, the arena bookkeeping, epoch scheduling,
// the watchdog and the C ABI calls on a server; it makes no HIP runtime call.
// The library's executor (sarbp_server.hip) runs the device side on a HIP
// device; the CPU tests run the same core against a host-memory stand-in
// under thread and address sanitizers.  Not part of the public ABI.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

#include "sarbp.h"
#include "sarbp_host.h"

namespace sarbp {

// Markers that bracket one epoch's device work (executor-defined handles).
struct EpochMark {
  void* begin = nullptr;
  void* end = nullptr;
};

// What a server's threads need from the device.  Each group of methods is
// called from one server thread only (the admission, scheduler or
// completion thread; bind_thread from each of them first), alloc_arena and
// n_cu from the creating thread before the threads start, and the
// destructor after they have been joined.
class ServerExec {
 public:
  virtual ~ServerExec() = default;
  virtual int n_cu() const = 0;
  // The arena: `bytes` of device memory, released by the destructor.
  virtual sarbp_status alloc_arena(size_t bytes, char** base) = 0;
  // Makes the server's device current on the calling thread.
  virtual void bind_thread() = 0;

  // Admission: enqueue host-to-device copies and the range compression of a
  // raw-echo job's echoes (d_raw -> d_out), then wait for all of them.
  virtual sarbp_status upload(void* dst, const void* src, size_t n) = 0;
  virtual sarbp_status compress(const sarbp_cf32* d_raw, int32_t n_pulses, const sarbp_range_params* rp,
                                void* d_out) = 0;
  virtual sarbp_status admit_wait() = 0;

  // Scheduler: an epoch is epoch_begin, one batched launch per job group
  // (slot = epoch parity: the launches of two consecutive epochs may be in
  // flight; g = group index), epoch_end.
  virtual sarbp_status epoch_begin(EpochMark* m) = 0;
  virtual sarbp_status run_group(int slot, int g, std::vector<JobDesc>& desc, int kinterp, int kprec, int dtype,
                                 const sarbp_tuning& t) = 0;
  virtual sarbp_status epoch_end(EpochMark* m) = 0;
  // 1 if the epoch's device work has finished, 0 if not yet, < 0 on an error
  // (epoch_wait then reports it).
  virtual int epoch_done(const EpochMark& m) = 0;
  // Waits for the epoch; *ms = its device time between the markers.
  virtual sarbp_status epoch_wait(const EpochMark& m, float* ms) = 0;
  virtual void epoch_release(EpochMark* m) = 0;
  // Waits until the scheduler's queued device work has finished.
  virtual void compute_wait() = 0;

  // Completion: a device-to-host copy, waited for.
  virtual sarbp_status download(void* dst, const void* src, size_t n) = 0;
};

// Launch groups of one epoch's chunks; at most this many.
constexpr int kServerGroups = 8;

// Creates a server over an executor (sarbp_server_create passes the HIP
// one): validates cfg, takes the tunings, allocates the arena, starts the
// threads.
sarbp_status server_create(const sarbp_server_config* cfg, std::unique_ptr<ServerExec> exec,
                           sarbp_server** out);
// The profile storage (sarbp_dtype) a server was created for.
int server_dtype(const sarbp_server* s);
// Queues one job; raw is NULL for a profile request, or holds the raw echoes
// and their range compression (req->params already describe the profiles
// the compression makes).
sarbp_status server_submit(sarbp_server* s, const sarbp_request* req, const sarbp_raw_request* raw,
                           uint64_t* job_id);

}  // namespace sarbp
