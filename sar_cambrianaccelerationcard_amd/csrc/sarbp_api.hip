// sarbp_api.hip -- host side of the C ABI declared in include/sarbp.h:
// contexts and their tuning caches, the plain / planned / batched launches,
// the allocation-free batch, the autotuner and the reference-compatible
// entry points (the planning they share is sarbp_plan.cpp).  Nothing here
// exits or throws across the ABI.
#include "sarbp.h"
#include "sarbp_host.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

namespace sarbp {

sarbp_status hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? SARBP_ERR_OUT_OF_MEMORY : SARBP_ERR_DEVICE,
              "%s: %s", what, hipGetErrorString(e));
}

// Tuning cache key: (interp, precision, dtype, size class, whether the call
// has host positions); class kAllClasses holds a tuning set for every class
// at once (sarbp_set_tuning), with and without positions.  Without positions
// a tile's window covers the disc of its radius, so the two kinds of calls
// keep separate per-class entries.
constexpr int kAllClasses = SARBP_SIZE_CLASSES;
int tuning_key(int interp, int prec, int dtype, int cls, bool positions = true) {
  return ((((interp * 10 + prec) * 10 + dtype) * 10 + cls) << 1) | (positions || cls == kAllClasses ? 0 : 1);
}

// ---- device scratch of the allocation-free batched launch ----------------
// Layout: kNoAllocGroups slots of kNoAllocSlot bytes (one per tile-shape
// group: its BatchHeader at offset 0 and its work counter at offset 64),
// then the groups' job tables back to back (at most kNoAllocJobs JobDescs).
constexpr int kNoAllocJobs = 256;
constexpr int kNoAllocGroups = 8;
constexpr size_t kNoAllocSlot = 256;
constexpr size_t kNoAllocCounter = 64;
constexpr size_t kNoAllocBytes = kNoAllocGroups * kNoAllocSlot + kNoAllocJobs * sizeof(JobDesc);
constexpr int kMaxDevices = 64;

}  // namespace sarbp

// File-scope device array (each GPU holds its own copy at its own address,
// for the life of the process).  Host code never uses this name as an
// address: sarbp_create looks up each device's copy with hipGetSymbolAddress
// and keeps that pointer in the context.
static __device__ __align__(256) unsigned char scratch[sarbp::kNoAllocBytes];

namespace sarbp {

std::mutex g_scratch_mu;               // guards g_scratch_dev and the host staging
void* g_scratch_dev[kMaxDevices] = {};  // each device's address of `scratch`
alignas(256) unsigned char g_noalloc_host[kNoAllocBytes];  // host image of the scratch
JobDesc g_noalloc_plan[kNoAllocJobs];  // planned descriptors, submission order

// Looks up `scratch` on `device` (current) once per device; *out = its address.
sarbp_status scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return SARBP_OK;  // entry point then reports it
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_dev[device] = p;
  }
  *out = g_scratch_dev[device];
  return SARBP_OK;
}

}  // namespace sarbp

using namespace sarbp;

struct sarbp_plan {
  sarbp_context* ctx;
  BpArgs args;
  int interp, precision, dtype;
  sarbp_tuning tune;
  sarbp_plan_info info;
};

namespace sarbp {

// The context's tuning of a class, else its all-classes tuning, else the
// built-in default (`positions`: for a plan that knows the antenna positions).
sarbp_tuning get_tuning(sarbp_context* ctx, int interp, int prec, int dtype, int cls, bool positions = true) {
  auto it = ctx->tuning.find(tuning_key(interp, prec, dtype, cls, positions));
  if (it != ctx->tuning.end()) return it->second;
  it = ctx->tuning.find(tuning_key(interp, prec, dtype, kAllClasses));
  if (it != ctx->tuning.end()) return it->second;
  return default_tuning(interp, cls, prec, positions);
}

sarbp_status launch(const BpArgs& a, const sarbp_params* p, const sarbp_tuning& t, hipStream_t s) {
  if (a.nx == 0 || a.ny == 0) return SARBP_OK;
  hipError_t e = launch_plain(a, a.kernel_interp, a.kernel_prec, p->dtype, t.cols_per_lane, t.rows_per_lane, s);
  if (e != hipSuccess) return hip_fail(e, "BP kernel launch");
  return SARBP_OK;
}

}  // namespace sarbp

extern "C" {

const char* sarbp_strerror(sarbp_status s) {
  switch (s) {
    case SARBP_OK: return "ok";
    case SARBP_ERR_INVALID_ARGUMENT: return "invalid argument";
    case SARBP_ERR_OUT_OF_MEMORY: return "out of memory";
    case SARBP_ERR_OVER_BUDGET: return "over HBM budget";
    case SARBP_ERR_DEVICE: return "device error";
    case SARBP_ERR_UNSUPPORTED: return "unsupported";
    case SARBP_ERR_NOT_READY: return "not ready";
    case SARBP_ERR_NOT_FOUND: return "not found";
    case SARBP_ERR_CANCELLED: return "cancelled";
    case SARBP_ERR_TIMEOUT: return "timeout";
    case SARBP_ERR_SHUTDOWN: return "shutting down";
  }
  return "unknown status";
}

int32_t sarbp_abi_version(void) { return SARBP_ABI_VERSION; }

const char* sarbp_last_error_detail(void) { return last_detail(); }

sarbp_status sarbp_create(sarbp_context** out, int32_t device) {
  if (!out) return fail(SARBP_ERR_INVALID_ARGUMENT, "ctx out-pointer is NULL");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
  if (device < 0 || device >= n) return fail(SARBP_ERR_INVALID_ARGUMENT, "device %d of %d", device, n);
  sarbp_context* c = new (std::nothrow) sarbp_context;
  if (!c) return fail(SARBP_ERR_OUT_OF_MEMORY, "context");
  c->device = device;
  int cu = 0;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0)
    c->n_cu = cu;
  {
    DeviceGuard dg(device);
    sarbp_status st = dg.err != hipSuccess ? hip_fail(dg.err, "hipSetDevice")
                                           : scratch_address(device, &c->scratch);
    if (st != SARBP_OK) {
      delete c;
      return st;
    }
  }
  *out = c;
  return SARBP_OK;
}

void sarbp_destroy(sarbp_context* c) {
  if (!c) return;
  {
    DeviceGuard dg(c->device);
    for (BatchExec* b : c->batch) delete b;
    range_cache_destroy(c->range);
    wk_cache_destroy(c->wk);
  }
  delete c;
}

sarbp_status sarbp_set_tuning(sarbp_context* ctx, int32_t interp, int32_t precision, int32_t dtype,
                              const sarbp_tuning* t) {
  if (!ctx || !t) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  if (interp < 0 || interp > 2 || precision < 0 || precision > 3 || dtype < 0 || dtype > 1)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "bad (interp, precision, dtype)");
  std::lock_guard<std::mutex> lk(ctx->mu);
  for (int c = 0; c <= kAllClasses; ++c) {
    ctx->tuning.erase(tuning_key(interp, precision, dtype, c, true));
    ctx->tuning.erase(tuning_key(interp, precision, dtype, c, false));
  }
  if (t->rows_per_lane == 0 && t->cols_per_lane == 0 && t->pulses_per_stage == 0) return SARBP_OK;
  sarbp_status st = check_tuning(precision, t);
  if (st != SARBP_OK) return st;
  ctx->tuning[tuning_key(interp, precision, dtype, kAllClasses)] = *t;
  return SARBP_OK;
}

int32_t sarbp_size_class(const sarbp_grid* grid, const sarbp_params* params, const double* h_pos) {
  if (validate(grid, params) != SARBP_OK) return -1;
  if (h_pos && !finite_all(h_pos, 3 * params->n_pulses)) return -1;
  return size_class(grid, params, h_pos);
}

sarbp_status sarbp_set_class_tuning(sarbp_context* ctx, int32_t interp, int32_t precision, int32_t dtype,
                                    int32_t size_cls, const sarbp_tuning* t) {
  if (!ctx || !t) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  if (interp < 0 || interp > 2 || precision < 0 || precision > 3 || dtype < 0 || dtype > 1 || size_cls < 0 ||
      size_cls >= SARBP_SIZE_CLASSES)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "bad (interp, precision, dtype, size class)");
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (t->rows_per_lane == 0 && t->cols_per_lane == 0 && t->pulses_per_stage == 0) {
    ctx->tuning.erase(tuning_key(interp, precision, dtype, size_cls));
    return SARBP_OK;
  }
  sarbp_status st = check_tuning(precision, t);
  if (st != SARBP_OK) return st;
  ctx->tuning[tuning_key(interp, precision, dtype, size_cls)] = *t;
  return SARBP_OK;
}

sarbp_status sarbp_get_class_tuning(sarbp_context* ctx, int32_t interp, int32_t precision, int32_t dtype,
                                    int32_t size_cls, sarbp_tuning* t) {
  if (!ctx || !t) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  if (size_cls < 0 || size_cls >= SARBP_SIZE_CLASSES) return fail(SARBP_ERR_INVALID_ARGUMENT, "size class %d", size_cls);
  std::lock_guard<std::mutex> lk(ctx->mu);
  *t = get_tuning(ctx, interp, precision, dtype, size_cls);
  return SARBP_OK;
}

sarbp_status sarbp_get_tuning(sarbp_context* ctx, int32_t interp, int32_t precision, int32_t dtype,
                              sarbp_tuning* t) {
  if (!ctx || !t) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  *t = get_tuning(ctx, interp, precision, dtype, 0);
  return SARBP_OK;
}

sarbp_status sarbp_backproject(sarbp_context* ctx, const void* d_profiles, const double* d_pos,
                               const sarbp_grid* grid, const sarbp_params* params,
                               sarbp_cf32* d_image, void* stream) {
  if (!ctx) return fail(SARBP_ERR_INVALID_ARGUMENT, "ctx is NULL");
  sarbp_status st = validate(grid, params);
  if (st != SARBP_OK) return st;
  if ((params->n_pulses > 0 && (!d_profiles || !d_pos)) || (!d_image && grid->nx > 0 && grid->ny > 0))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "device pointer is NULL");
  std::lock_guard<std::mutex> lk(ctx->mu);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  const Fitted fit =
      fit_tuning(get_tuning(ctx, params->interp, params->precision, params->dtype, size_class(grid, params, nullptr),
                            false),
                 params->precision, grid, params, nullptr);
  BpArgs a;
  st = plan_args(nullptr, grid, params, fit.t, fit.prec, &a, nullptr);
  if (st != SARBP_OK) return st;
  a.prof = d_profiles;
  a.pos = d_pos;
  a.img = reinterpret_cast<float2*>(d_image);
  return launch(a, params, fit.t, (hipStream_t)stream);
}

sarbp_status sarbp_backproject_sync(sarbp_context* ctx, const void* d_profiles, const double* d_pos,
                                    const sarbp_grid* grid, const sarbp_params* params,
                                    sarbp_cf32* d_image, void* stream) {
  sarbp_status st = sarbp_backproject(ctx, d_profiles, d_pos, grid, params, d_image, stream);
  if (st != SARBP_OK) return st;
  DeviceGuard dg(ctx->device);
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  return SARBP_OK;
}

sarbp_status sarbp_plan_create(sarbp_context* ctx, const double* h_pos, const sarbp_grid* grid,
                               const sarbp_params* params, sarbp_plan** out) {
  if (!ctx || !out) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  sarbp_status st = validate(grid, params);
  if (st != SARBP_OK) return st;
  if (h_pos && !finite_all(h_pos, 3 * params->n_pulses))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "positions must be finite");
  sarbp_plan* pl = new (std::nothrow) sarbp_plan;
  if (!pl) return fail(SARBP_ERR_OUT_OF_MEMORY, "plan");
  std::lock_guard<std::mutex> lk(ctx->mu);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) {
    delete pl;
    return hip_fail(dg.err, "hipSetDevice");
  }
  const Fitted fit =
      fit_tuning(get_tuning(ctx, params->interp, params->precision, params->dtype, size_class(grid, params, h_pos),
                            h_pos != nullptr),
                 params->precision, grid, params, h_pos);
  const sarbp_tuning& tune = fit.t;
  std::memset(&pl->info, 0, sizeof pl->info);
  st = plan_args(h_pos, grid, params, tune, fit.prec, &pl->args, &pl->info);
  if (st != SARBP_OK) {
    delete pl;
    return st;
  }
  pl->ctx = ctx;
  pl->interp = params->interp;
  pl->precision = params->precision;
  pl->dtype = params->dtype;
  pl->tune = tune;
  *out = pl;
  return SARBP_OK;
}

sarbp_status sarbp_plan_execute(sarbp_plan* pl, const void* d_profiles, const double* d_pos,
                                sarbp_cf32* d_image, void* stream) {
  if (!pl) return fail(SARBP_ERR_INVALID_ARGUMENT, "plan is NULL");
  if ((pl->args.n_pulses > 0 && (!d_profiles || !d_pos)) || (!d_image && pl->args.nx > 0 && pl->args.ny > 0))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "device pointer is NULL");
  DeviceGuard dg(pl->ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  BpArgs a = pl->args;
  a.prof = d_profiles;
  a.pos = d_pos;
  a.img = reinterpret_cast<float2*>(d_image);
  sarbp_params p;
  std::memset(&p, 0, sizeof p);
  p.interp = pl->interp;
  p.precision = pl->precision;
  p.dtype = pl->dtype;
  return launch(a, &p, pl->tune, (hipStream_t)stream);
}

sarbp_status sarbp_plan_get_info(const sarbp_plan* pl, sarbp_plan_info* info) {
  if (!pl || !info) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  *info = pl->info;
  return SARBP_OK;
}

void sarbp_plan_destroy(sarbp_plan* pl) { delete pl; }

}  // extern "C"

namespace sarbp {

// Checks a batch's jobs: each valid, all sharing job 0's interp, precision and
// dtype, device pointers present where there is work.
sarbp_status validate_batch(const sarbp_job* jobs, int32_t n_jobs) {
  const sarbp_params& p0 = jobs[0].params;
  for (int32_t j = 0; j < n_jobs; ++j) {
    const sarbp_job& jb = jobs[j];
    sarbp_status st = validate(&jb.grid, &jb.params);
    if (st != SARBP_OK) return fail(st, "job %d: %s", j, last_detail());
    if (jb.params.interp != p0.interp || jb.params.precision != p0.precision || jb.params.dtype != p0.dtype)
      return fail(SARBP_ERR_INVALID_ARGUMENT, "job %d: interp, precision and dtype must match job 0", j);
    if ((jb.params.n_pulses > 0 && (!jb.d_profiles || !jb.d_pos)) ||
        (!jb.d_image && jb.grid.nx > 0 && jb.grid.ny > 0))
      return fail(SARBP_ERR_INVALID_ARGUMENT, "job %d: device pointer is NULL", j);
  }
  return SARBP_OK;
}

}  // namespace sarbp

extern "C" {

sarbp_status sarbp_backproject_batch(sarbp_context* ctx, const sarbp_job* jobs, int32_t n_jobs,
                                     void* stream) {
  if (!ctx || (!jobs && n_jobs > 0) || n_jobs < 0)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL context or job list");
  if (n_jobs == 0) return SARBP_OK;
  const sarbp_params& p0 = jobs[0].params;
  sarbp_status vst = validate_batch(jobs, n_jobs);
  if (vst != SARBP_OK) return vst;
  std::lock_guard<std::mutex> lk(ctx->mu);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  // Largest per-tile cost first (pulses per tile), ties by submission order:
  // the persistent grid drains items in this order.
  std::vector<int32_t> order((size_t)n_jobs);
  for (int32_t j = 0; j < n_jobs; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
    return jobs[x].params.n_pulses > jobs[y].params.n_pulses;
  });
  // Each job runs with the tile shape a plain launch of it would use (the
  // tuned shape, reduced to fit its grid), so its image is bitwise the plain
  // launch's; jobs that share a shape share one persistent launch.
  struct Group {
    sarbp_tuning tune;
    int kinterp, kprec;
    std::vector<JobDesc> desc;
  };
  std::vector<JobDesc> planned;
  std::vector<sarbp_tuning> tunes;
  std::vector<GroupKey> keys;
  for (int32_t j : order) {
    const sarbp_job& jb = jobs[j];
    if (jb.grid.nx == 0 || jb.grid.ny == 0) continue;
    const Fitted f = fit_tuning(get_tuning(ctx, p0.interp, p0.precision, p0.dtype, size_class(&jb.grid, &jb.params, jb.h_pos),
                                           jb.h_pos != nullptr),
                                p0.precision, &jb.grid, &jb.params, jb.h_pos);
    JobDesc d;
    sarbp_status st = plan_args(jb.h_pos, &jb.grid, &jb.params, f.t, f.prec, &d.a, nullptr, true);
    if (st != SARBP_OK) return fail(st, "job %d: %s", j, last_detail());
    d.a.prof = jb.d_profiles;
    d.a.pos = jb.d_pos;
    d.a.img = reinterpret_cast<float2*>(jb.d_image);
    d.item_begin = 0;
    planned.push_back(d);
    tunes.push_back(f.t);
    keys.push_back(group_key(d.a, f.t, p0.dtype));
  }
  const int max_groups = (int)(sizeof ctx->batch / sizeof ctx->batch[0]);
  std::vector<int> gid(planned.size());
  const int n_groups = assign_groups(keys.data(), (int)keys.size(), max_groups, gid.data());
  if (n_groups < 0) return fail(SARBP_ERR_UNSUPPORTED, "more than %d tile shapes in one batch", max_groups);
  std::vector<Group> groups((size_t)n_groups);
  for (size_t i = 0; i < planned.size(); ++i) {
    Group& G = groups[gid[i]];
    if (G.desc.empty()) G = Group{tunes[i], keys[i].kinterp, keys[i].kprec, {}};
    G.desc.push_back(planned[i]);
  }
  for (size_t g = 0; g < groups.size(); ++g) {
    if (!ctx->batch[g] && !(ctx->batch[g] = new (std::nothrow) BatchExec))
      return fail(SARBP_ERR_OUT_OF_MEMORY, "batch executor");
    const sarbp_status st = ctx->batch[g]->run(groups[g].desc, groups[g].kinterp, groups[g].kprec, p0.dtype,
                                               groups[g].tune, ctx->n_cu, (hipStream_t)stream);
    if (st != SARBP_OK) return st;
  }
  return SARBP_OK;
}

sarbp_status sarbp_backproject_batch_noalloc(sarbp_context* ctx, const sarbp_job* jobs, int32_t n_jobs,
                                             void* stream) {
  if (!ctx || (!jobs && n_jobs > 0) || n_jobs < 0)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL context or job list");
  if (n_jobs > kNoAllocJobs)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "%d jobs: the allocation-free batch takes at most %d", n_jobs,
                kNoAllocJobs);
  if (n_jobs == 0) return SARBP_OK;
  const sarbp_params& p0 = jobs[0].params;
  sarbp_status st = validate_batch(jobs, n_jobs);
  if (st != SARBP_OK) return st;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (!ctx->scratch)
    return fail(SARBP_ERR_UNSUPPORTED, "device %d has no scratch address (devices 0..%d)", ctx->device,
                kMaxDevices - 1);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  // The host image of the scratch is shared by every context: one call at a time.
  std::lock_guard<std::mutex> sk(g_scratch_mu);
  // Largest pulse count first, ties in submission order (insertion sort: no
  // temporary buffer), as sarbp_backproject_batch orders its work list.
  int32_t order[kNoAllocJobs];
  for (int32_t j = 0; j < n_jobs; ++j) {
    int32_t k = j;
    while (k > 0 && jobs[order[k - 1]].params.n_pulses < jobs[j].params.n_pulses) {
      order[k] = order[k - 1];
      --k;
    }
    order[k] = j;
  }
  // Plan every job with the tile shape its plain launch would use and put it
  // in the group of its (tile shape, sinc8 kernel).
  struct Group {
    sarbp_tuning tune;
    int kinterp, kprec, n;
    size_t shm;
    int64_t items;
  } groups[kNoAllocGroups];
  int n_planned = 0;
  int group_of[kNoAllocJobs];
  GroupKey keys[kNoAllocJobs];
  sarbp_tuning tunes[kNoAllocJobs];
  for (int32_t r = 0; r < n_jobs; ++r) {
    const sarbp_job& jb = jobs[order[r]];
    if (jb.grid.nx == 0 || jb.grid.ny == 0) continue;
    const Fitted f = fit_tuning(get_tuning(ctx, p0.interp, p0.precision, p0.dtype, size_class(&jb.grid, &jb.params, jb.h_pos),
                                           jb.h_pos != nullptr),
                                p0.precision, &jb.grid, &jb.params, jb.h_pos);
    JobDesc& d = g_noalloc_plan[n_planned];
    st = plan_args(jb.h_pos, &jb.grid, &jb.params, f.t, f.prec, &d.a, nullptr, true);
    if (st != SARBP_OK) return fail(st, "job %d: %s", order[r], last_detail());
    d.a.prof = jb.d_profiles;
    d.a.pos = jb.d_pos;
    d.a.img = reinterpret_cast<float2*>(jb.d_image);
    tunes[n_planned] = f.t;
    keys[n_planned] = group_key(d.a, f.t, p0.dtype);
    ++n_planned;
  }
  if (n_planned == 0) return SARBP_OK;
  const int n_groups = assign_groups(keys, n_planned, kNoAllocGroups, group_of);
  if (n_groups < 0) return fail(SARBP_ERR_UNSUPPORTED, "more than %d tile shapes in one batch", kNoAllocGroups);
  for (int g = 0; g < n_groups; ++g) groups[g] = Group{sarbp_tuning{}, 0, 0, 0, 0, 0};
  for (int i = 0; i < n_planned; ++i) {
    Group& G = groups[group_of[i]];
    if (G.n == 0) {
      G.tune = tunes[i];
      G.kinterp = keys[i].kinterp;
      G.kprec = keys[i].kprec;
    }
    JobDesc& d = g_noalloc_plan[i];
    d.item_begin = G.items;
    G.items += (int64_t)d.a.tiles_x * d.a.tiles_y;
    G.shm = std::max(G.shm, lds_layout(G.kinterp, d.a.win, d.a.pulses_per_stage).total);
    ++G.n;
  }
  // Host image of the scratch: headers (counters zero), then each group's
  // table in planning order.  One copy on the stream writes all of it, so
  // the counters start at zero for this call's launches.
  std::memset(g_noalloc_host, 0, kNoAllocGroups * kNoAllocSlot);
  size_t table_off[kNoAllocGroups];
  size_t off = kNoAllocGroups * kNoAllocSlot;
  for (int g = 0; g < n_groups; ++g) {
    BatchHeader h;
    h.n_items = groups[g].items;
    h.n_jobs = groups[g].n;
    h.reserved = 0;
    std::memcpy(g_noalloc_host + g * kNoAllocSlot, &h, sizeof h);
    table_off[g] = off;
    off += (size_t)groups[g].n * sizeof(JobDesc);
  }
  int filled[kNoAllocGroups] = {};
  for (int i = 0; i < n_planned; ++i) {
    const int g = group_of[i];
    std::memcpy(g_noalloc_host + table_off[g] + (size_t)filled[g]++ * sizeof(JobDesc), &g_noalloc_plan[i],
                sizeof(JobDesc));
  }
  hipStream_t s = (hipStream_t)stream;
  char* dev = static_cast<char*>(ctx->scratch);
  // Pageable source: the call returns once the bytes are staged, so the host
  // image may be rewritten by the next call.
  hipError_t e = hipMemcpyAsync(dev, g_noalloc_host, off, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(scratch)");
  for (int g = 0; g < n_groups; ++g) {
    const Group& G = groups[g];
    const int per_cu = batch_blocks_per_cu(G.kinterp, G.kprec, p0.dtype, G.tune.cols_per_lane,
                                           G.tune.rows_per_lane, G.shm);
    e = launch_batch(reinterpret_cast<const JobDesc*>(dev + table_off[g]),
                     reinterpret_cast<const BatchHeader*>(dev + g * kNoAllocSlot),
                     reinterpret_cast<unsigned long long*>(dev + g * kNoAllocSlot + kNoAllocCounter),
                     G.kinterp, G.kprec, p0.dtype, G.tune.cols_per_lane, G.tune.rows_per_lane, G.shm,
                     std::max(1, ctx->n_cu * per_cu), s);
    if (e != hipSuccess) return hip_fail(e, "allocation-free batched launch");
  }
  return SARBP_OK;
}

sarbp_status sarbp_autotune(sarbp_context* ctx, const void* d_profiles, const double* d_pos,
                            const double* h_pos, const sarbp_grid* grid, const sarbp_params* params,
                            sarbp_cf32* d_scratch, size_t scratch_pixels, int32_t max_crop_px,
                            void* stream, sarbp_autotune_result* out) {
  if (!ctx || !out) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  sarbp_status st = validate(grid, params);
  if (st != SARBP_OK) return st;
  if (params->n_pulses < 1 || !d_profiles || !d_pos || !d_scratch)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "autotune needs pulses and device buffers");
  if (max_crop_px < kTileX) return fail(SARBP_ERR_INVALID_ARGUMENT, "max_crop_px < %d", kTileX);
  // The crop must hold many waves of workgroups, or small-tile configurations
  // win on the last wave's tail alone (a 4096 x 1024 crop, ~3 waves, already
  // mis-ranks 64x32 against 64x16 tiles): keep up to 4096 columns and as many
  // rows as a max_crop_px^2 pixel budget allows.
  sarbp_grid crop = *grid;
  const int64_t budget = (int64_t)max_crop_px * max_crop_px;
  crop.nx = std::min(grid->nx, std::max(max_crop_px, 4096));
  crop.ny = (int32_t)std::min<int64_t>(grid->ny, std::max<int64_t>(1, budget / std::max(crop.nx, 1)));
  if (crop.nx < 1 || crop.ny < 1) return fail(SARBP_ERR_INVALID_ARGUMENT, "empty image");
  if ((size_t)crop.nx * crop.ny > scratch_pixels)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "scratch holds %zu pixels, crop needs %d", scratch_pixels,
                crop.nx * crop.ny);
  const int ox = (grid->nx - crop.nx) / 2, oy = (grid->ny - crop.ny) / 2;
  for (int k = 0; k < 3; ++k) crop.origin[k] = grid->origin[k] + ox * grid->ux[k] + oy * grid->uy[k];
  sarbp_params p = *params;
  p.accumulate = 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  DeviceGuard dg(ctx->device);
  if (dg.err != hipSuccess) return hip_fail(dg.err, "hipSetDevice");
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e;
  if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return hip_fail(e, "hipEventCreate");
  }
  auto time_one = [&](const sarbp_tuning& t, double* ms, int* eff_p) -> sarbp_status {
    BpArgs a;
    sarbp_status r = plan_args(h_pos, &crop, &p, t, p.precision, &a, nullptr);
    if (r != SARBP_OK) return r;
    *eff_p = a.pulses_per_stage;
    a.prof = d_profiles;
    a.pos = d_pos;
    a.img = reinterpret_cast<float2*>(d_scratch);
    if ((r = launch(a, &p, t, s)) != SARBP_OK) return r;  // warm-up
    hipError_t he;
    if ((he = hipEventRecord(e0, s)) != hipSuccess) return hip_fail(he, "hipEventRecord");
    const int reps = 2;
    for (int i = 0; i < reps; ++i)
      if ((r = launch(a, &p, t, s)) != SARBP_OK) return r;
    if ((he = hipEventRecord(e1, s)) != hipSuccess || (he = hipEventSynchronize(e1)) != hipSuccess)
      return hip_fail(he, "hipEventSynchronize");
    float f = 0.0f;
    if ((he = hipEventElapsedTime(&f, e0, e1)) != hipSuccess) return hip_fail(he, "hipEventElapsedTime");
    *ms = f / reps;
    return SARBP_OK;
  };
  // default first (built-in, ignoring any stored tuning)
  const int cls = size_class(grid, params, h_pos);
  const sarbp_tuning def = default_tuning(p.interp, cls, p.precision, h_pos != nullptr);
  int eff = 0;
  double def_ms = INFINITY;
  st = time_one(def, &def_ms, &eff);
  if (st == SARBP_ERR_UNSUPPORTED) st = SARBP_OK;  // default shape too large for this grid
  sarbp_tuning best = def;
  best.pulses_per_stage = eff;
  double best_ms = def_ms;
  int n = std::isfinite(def_ms) ? 1 : 0;
  std::vector<std::tuple<int, int, int>> seen;
  seen.emplace_back(def.cols_per_lane, def.rows_per_lane, eff);
  const int shapes[][2] = {{1, 2}, {1, 4}, {1, 8}, {2, 2}, {2, 4}, {2, 8}, {4, 2}, {4, 4}};
  const int stages[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64};
  for (const auto& sh : shapes) {
    if (st != SARBP_OK) break;
    if (!tile_shape_supported(p.precision, sh[0], sh[1])) continue;
    for (int P : stages) {
      sarbp_tuning t;
      t.cols_per_lane = sh[0];
      t.rows_per_lane = sh[1];
      t.pulses_per_stage = P;
      BpArgs probe;
      if (plan_args(h_pos, &crop, &p, t, p.precision, &probe, nullptr) != SARBP_OK) continue;
      // fewer than 4 pulses per stage only pay where 4 would leave fewer
      // than 4 workgroups per CU (40 KiB of LDS each)
      if (P < 4 && lds_layout(probe.kernel_interp, probe.win, 4).total <= 40 * 1024) continue;
      const auto key = std::make_tuple(sh[0], sh[1], probe.pulses_per_stage);
      if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
      seen.push_back(key);
      double ms = 0.0;
      if ((st = time_one(t, &ms, &eff)) != SARBP_OK) break;
      ++n;
      // a candidate replaces the built-in default only when it is more than
      // 2% faster: near-ties on one crop do not carry to the other jobs of
      // the class (a two-column tile that ties on a 4096-pulse image loses
      // 37% under an 8192-pulse aperture)
      if (ms < best_ms && ms < 0.98 * def_ms) {
        best_ms = ms;
        best = t;
        best.pulses_per_stage = eff;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != SARBP_OK) return st;
  ctx->tuning[tuning_key(p.interp, p.precision, p.dtype, cls, h_pos != nullptr)] = best;
  out->best = best;
  out->best_ms = best_ms;
  out->default_ms = def_ms;
  out->n_candidates = n;
  out->crop_nx = crop.nx;
  out->crop_ny = crop.ny;
  return SARBP_OK;
}

sarbp_status sarbp_launch_backprojection(sarbp_cf32* Srnm, sarbp_cf32* Image, sarbp_sar_params sp,
                                         int32_t Na, int32_t Nr) {
  if (!Srnm || !Image) return fail(SARBP_ERR_INVALID_ARGUMENT, "Srnm or Image is NULL");
  if (Na < 1 || Nr < 1 || sp.Mslow < 1 || sp.Nfast < 1)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "Na, Nr, Mslow and Nfast must be >= 1");
  if (!(sp.c > 0 && sp.fc > 0 && sp.Fr > 0 && sp.vr > 0 && std::isfinite(sp.La)))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "SARParams c, fc, Fr, vr must be > 0");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  // One default context per device, created on first use and kept for the
  // process (the reference keeps no state between calls; this caches tables).
  static std::mutex mu;
  static std::map<int, sarbp_context*> ctxs;
  sarbp_context* ctx = nullptr;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = ctxs.find(dev);
    if (it == ctxs.end()) {
      sarbp_status st = sarbp_create(&ctx, dev);
      if (st != SARBP_OK) return st;
      ctxs[dev] = ctx;
    } else {
      ctx = it->second;
    }
  }
  const int M = sp.Mslow;
  std::vector<double> pos((size_t)3 * M);
  const double ta0 = -(double)sp.La / (2.0 * sp.vr), ta1 = (double)sp.La / (2.0 * sp.vr);
  for (int m = 0; m < M; ++m) {
    const double ta = M > 1 ? ta0 + (ta1 - ta0) * m / (M - 1) : 0.0;
    pos[3 * (size_t)m] = (double)sp.vr * ta;
    pos[3 * (size_t)m + 1] = 0.0;
    pos[3 * (size_t)m + 2] = sp.H;
  }
  double* d_pos = nullptr;
  if ((e = hipMalloc(&d_pos, pos.size() * sizeof(double))) != hipSuccess) return hip_fail(e, "hipMalloc(pos)");
  if ((e = hipMemcpy(d_pos, pos.data(), pos.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) {
    (void)hipFree(d_pos);
    return hip_fail(e, "hipMemcpy(pos)");
  }
  // The track runs along x, so image rows (azimuth, i) step along x and
  // columns (ground range, j) along y: pixel (j, i) = (Az0 + i - Na/2,
  // Rg0 + j - Nr/2, 0).
  sarbp_grid g;
  std::memset(&g, 0, sizeof g);
  g.origin[0] = (double)sp.Az0 - (double)(Na / 2);
  g.origin[1] = (double)sp.Rg0 - (double)(Nr / 2);
  g.ux[1] = 1.0;
  g.uy[0] = 1.0;
  g.nx = Nr;
  g.ny = Na;
  sarbp_params p;
  std::memset(&p, 0, sizeof p);
  p.c = sp.c;
  p.fc = sp.fc;
  p.r0 = 0.0;
  p.dr = (double)sp.c / (2.0 * sp.Fr);
  p.kaiser_beta = 0.0;
  p.n_pulses = M;
  p.n_samples = sp.Nfast;
  p.interp = SARBP_INTERP_NEAREST;
  p.precision = SARBP_PREC_REF;
  p.dtype = SARBP_DTYPE_C64;
  p.accumulate = 0;
  sarbp_status st = sarbp_backproject_sync(ctx, Srnm, d_pos, &g, &p, Image, nullptr);
  (void)hipFree(d_pos);
  return st;
}

sarbp_status sarbp_sinc8_coefficients(double beta, float* out) {
  if (!out) return fail(SARBP_ERR_INVALID_ARGUMENT, "out is NULL");
  if (!(std::isfinite(beta) && beta >= 0 && beta <= 50)) return fail(SARBP_ERR_INVALID_ARGUMENT, "beta");
  sinc8_poly(beta, out);
  return SARBP_OK;
}

sarbp_status sarbp_sinc8_coefficients_whole(double beta, float* out) {
  if (!out) return fail(SARBP_ERR_INVALID_ARGUMENT, "out is NULL");
  if (!(std::isfinite(beta) && beta >= 0 && beta <= 50)) return fail(SARBP_ERR_INVALID_ARGUMENT, "beta");
  sinc8_poly_whole(beta, out);
  return SARBP_OK;
}

}  // extern "C"
