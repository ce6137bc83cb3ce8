// sarbp_server_core.cpp -- continuous-batching BP server (see include/sarbp.h,
// "serving"), host logic only: the device work goes through a ServerExec
// (sarbp_server_core.h).  Three host threads per server:
//   admission   FIFO queue -> device arena (profiles, positions, image) + H2D
//   scheduler   epochs: one batched launch per job group in which every
//               resident job back-projects its next pulse chunk; two epochs in
//               flight (double-buffered job tables)
//   completion  D2H of finished images, arena release, job bookkeeping
// No HIP runtime call is made here.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "sarbp_server_core.h"

namespace sarbp {
namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t n) { return (n + kAlign - 1) & ~(kAlign - 1); }

// First-fit sub-allocator over one device allocation (the executor's),
// coalescing on free.
class Arena {
 public:
  sarbp_status init(size_t bytes, ServerExec& ex) {
    size_ = bytes & ~(kAlign - 1);
    if (size_ == 0) return fail(SARBP_ERR_INVALID_ARGUMENT, "hbm_budget_bytes too small");
    sarbp_status st = ex.alloc_arena(size_, &base_);
    if (st != SARBP_OK) {
      base_ = nullptr;
      return st;
    }
    free_[0] = size_;
    return SARBP_OK;
  }
  bool alloc(size_t n, size_t* off) {
    n = align_up(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const size_t rest = it->second - n, at = it->first + n;
      free_.erase(it);
      if (rest) free_[at] = rest;
      in_use_ += n;
      peak_ = std::max(peak_, in_use_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = align_up(n);
    in_use_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  char* base() const { return base_; }
  size_t size() const { return size_; }
  size_t in_use() const { return in_use_; }
  size_t peak() const { return peak_; }

 private:
  char* base_ = nullptr;
  size_t size_ = 0, in_use_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

struct Job {
  uint64_t id = 0;
  sarbp_request req{};
  BpArgs args{};
  sarbp_tuning tune{};  // the server's tuning, tile shape fitted to this job's grid
  // Raw-echo jobs: range-compressed in HBM at admission; the raw echoes sit
  // after the image in the job's arena block.
  bool raw = false;
  const sarbp_cf32* h_raw = nullptr;
  sarbp_range_params rp{};
  size_t prof_bytes = 0, pos_bytes = 0, img_bytes = 0, raw_bytes = 0, total = 0;
  size_t off = 0;
  bool has_mem = false;
  int32_t state = SARBP_JOB_QUEUED;
  sarbp_status status = SARBP_OK;
  char detail[192] = {};  // the failing call's detail text (sarbp_server_wait reports it)
  int32_t next_pulse = 0;
  bool started = false;
  int inflight = 0;
  bool cancel = false;
  int32_t epochs = 0;
  double t_submit = 0, t_admit = -1, t_first = -1, t_done = -1;
  double work = 0;
};

// Keeps the calling thread's last error detail (the call that failed the job).
void set_detail(Job& j) { std::snprintf(j.detail, sizeof j.detail, "%s", last_detail()); }

struct Epoch {
  std::vector<std::pair<uint64_t, int32_t>> parts;  // (job, pulse_end of its chunk)
  EpochMark mark;
  double t_enqueue = 0;
};

bool terminal(int32_t s) {
  return s == SARBP_JOB_DONE || s == SARBP_JOB_FAILED || s == SARBP_JOB_CANCELLED;
}

}  // namespace
}  // namespace sarbp

using namespace sarbp;

struct sarbp_server {
  sarbp_server_config cfg{};
  sarbp_tuning tune[SARBP_SIZE_CLASSES]{};  // per size class; guarded by mu
  int n_cu = 256;
  Arena arena;
  std::mutex mu;
  std::condition_variable cv_admit, cv_sched, cv_complete, cv_done;
  std::deque<uint64_t> queue;       // awaiting admission, arrival order
  std::vector<uint64_t> active;     // resident, admission order
  std::deque<uint64_t> finishing;   // device work done, awaiting D2H
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  uint64_t next_id = 1;
  bool stop = false;
  int admitting = 0;                // jobs between queue pop and residency
  std::thread th_admit, th_sched, th_complete;
  std::unique_ptr<ServerExec> exec;  // the device side
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  int64_t n_submitted = 0, n_completed = 0, n_failed = 0, n_cancelled = 0, n_epochs = 0;
  double work_done = 0, gpu_busy = 0;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  void admit_loop();
  void sched_loop();
  void complete_loop();
  void finish_epoch(Epoch& e);
  void release_job(Job& j) {  // caller holds mu
    if (j.has_mem) {
      arena.release(j.off, j.total);
      j.has_mem = false;
      cv_admit.notify_all();  // room for queued jobs
      cv_done.notify_all();   // sarbp_server_drain waits for the memory too
    }
  }
  void remove_active(uint64_t id) {
    active.erase(std::remove(active.begin(), active.end(), id), active.end());
  }
};

void sarbp_server::admit_loop() {
  exec->bind_thread();
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    Job* j = nullptr;
    cv_admit.wait(lk, [&] {
      if (stop) return true;
      while (!queue.empty() && terminal(jobs[queue.front()]->state)) queue.pop_front();
      if (queue.empty()) return false;
      if (cfg.max_active_jobs > 0 && (int)active.size() + admitting >= cfg.max_active_jobs) return false;
      Job* h = jobs[queue.front()].get();
      return arena.alloc(h->total, &h->off) ? (h->has_mem = true) : false;
    });
    if (stop) return;
    j = jobs[queue.front()].get();
    queue.pop_front();
    ++admitting;
    char* base = arena.base() + j->off;
    const sarbp_request req = j->req;
    const size_t pb = j->prof_bytes, qb = j->pos_bytes, ib = j->img_bytes;
    const bool raw = j->raw;
    const sarbp_cf32* h_raw = j->h_raw;
    const sarbp_range_params rp = j->rp;
    const size_t rb = j->raw_bytes;
    lk.unlock();
    sarbp_status st = SARBP_OK;
    if (raw) {
      sarbp_cf32* d_raw = reinterpret_cast<sarbp_cf32*>(base + align_up(pb) + align_up(qb) + align_up(ib));
      if (rb) st = exec->upload(d_raw, h_raw, rb);
      if (st == SARBP_OK && rb) st = exec->compress(d_raw, req.params.n_pulses, &rp, base);
    } else if (pb) {
      st = exec->upload(base, req.h_profiles, pb);
    }
    if (st == SARBP_OK && qb) st = exec->upload(base + align_up(pb), req.h_pos, qb);
    if (st == SARBP_OK) st = exec->admit_wait();
    lk.lock();
    --admitting;
    if (st != SARBP_OK) {
      j->state = SARBP_JOB_FAILED;
      j->status = st;
      set_detail(*j);
      j->t_done = now();
      ++n_failed;
      release_job(*j);
      cv_done.notify_all();
      continue;
    }
    if (j->cancel) {
      j->state = SARBP_JOB_CANCELLED;
      j->t_done = now();
      ++n_cancelled;
      release_job(*j);
      cv_done.notify_all();
      continue;
    }
    j->args.prof = base;
    j->args.pos = reinterpret_cast<const double*>(base + align_up(pb));
    j->args.img = reinterpret_cast<float2*>(base + align_up(pb) + align_up(qb));
    j->state = SARBP_JOB_RESIDENT;
    j->t_admit = now();
    active.push_back(j->id);
    cv_sched.notify_all();
  }
}

void sarbp_server::finish_epoch(Epoch& ep) {
  // Wait (unlocked) for the epoch; the watchdog only reports -- a kernel runs
  // to completion and its memory is released only after it has.
  bool late = false;
  if (cfg.epoch_watchdog_s > 0) {
    while (exec->epoch_done(ep.mark) == 0) {
      if (now() - ep.t_enqueue > cfg.epoch_watchdog_s) {
        late = true;
        break;
      }
      std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
  }
  float ms = 0.0f;
  const sarbp_status est = exec->epoch_wait(ep.mark, &ms);
  std::lock_guard<std::mutex> lk(mu);
  gpu_busy += ms / 1e3;
  for (auto& pr : ep.parts) {
    Job& j = *jobs[pr.first];
    --j.inflight;
    if (!terminal(j.state) && (est != SARBP_OK || late)) {
      j.state = SARBP_JOB_FAILED;
      j.status = est != SARBP_OK ? est
                                 : fail(SARBP_ERR_TIMEOUT, "epoch exceeded the %.3g s watchdog", cfg.epoch_watchdog_s);
      set_detail(j);
      j.t_done = now();
      ++n_failed;
      cv_done.notify_all();
    }
    if (j.inflight > 0) continue;  // a later chunk of this job is still in flight
    if (j.state == SARBP_JOB_FAILED) {
      remove_active(j.id);
      release_job(j);
    } else if (j.cancel) {
      remove_active(j.id);
      j.state = SARBP_JOB_CANCELLED;
      j.t_done = now();
      ++n_cancelled;
      release_job(j);
      cv_done.notify_all();
    } else if (j.next_pulse >= j.args.n_pulses && j.started) {
      remove_active(j.id);
      finishing.push_back(j.id);
      cv_complete.notify_all();
    }
  }
  exec->epoch_release(&ep.mark);
  cv_sched.notify_all();
}

void sarbp_server::sched_loop() {
  exec->bind_thread();
  std::deque<Epoch> inflight;
  int k = 0;
  for (;;) {
    // chunks grouped by tile shape, kernel and occupancy: one batched launch per group
    struct Group {
      sarbp_tuning tune;
      int kinterp, kprec;
      std::vector<JobDesc> desc;
    };
    std::vector<Group> groups;
    std::vector<JobDesc> chunks;
    std::vector<sarbp_tuning> tunes;
    std::vector<GroupKey> keys;
    bool over_groups = false;
    Epoch ep;
    {
      std::unique_lock<std::mutex> lk(mu);
      auto ready = [&] {
        for (uint64_t id : active) {
          const Job& j = *jobs[id];
          if (!j.cancel && !terminal(j.state) && (j.next_pulse < j.args.n_pulses || !j.started))
            return true;
        }
        return false;
      };
      if (inflight.empty()) {
        cv_sched.wait(lk, [&] { return stop || ready(); });
        if (stop) break;
      } else if (!ready()) {
        lk.unlock();
        finish_epoch(inflight.front());
        inflight.pop_front();
        continue;
      }
      const double t = now();
      const int32_t chunk = cfg.pulses_per_epoch > 0 ? cfg.pulses_per_epoch : INT32_MAX;
      for (uint64_t id : active) {
        Job& j = *jobs[id];
        // One chunk per job per epoch.  A job may have chunks in both epochs in
        // flight: they run in order on the compute stream, each adding into
        // the image the previous one wrote.
        if (j.cancel || terminal(j.state)) continue;
        if (j.next_pulse >= j.args.n_pulses && j.started) continue;
        JobDesc d;
        d.a = j.args;
        d.a.pulse_begin = j.next_pulse;
        d.a.pulse_end = (int32_t)std::min<int64_t>(j.args.n_pulses, (int64_t)j.next_pulse + chunk);
        d.a.accumulate = j.next_pulse > 0 ? 1 : 0;
        d.item_begin = 0;
        chunks.push_back(d);
        tunes.push_back(j.tune);
        keys.push_back(group_key(d.a, j.tune, cfg.dtype));
        j.next_pulse = d.a.pulse_end;
        j.started = true;
        ++j.inflight;
        ++j.epochs;
        if (j.state == SARBP_JOB_RESIDENT) {
          j.state = SARBP_JOB_RUNNING;
          j.t_first = t;
        }
        ep.parts.emplace_back(j.id, d.a.pulse_end);
      }
      ++n_epochs;
      std::vector<int> gid(chunks.size());
      const int ng = assign_groups(keys.data(), (int)keys.size(), kServerGroups, gid.data());
      groups.resize(ng < 0 ? 0 : (size_t)ng);
      for (size_t i = 0; ng >= 0 && i < chunks.size(); ++i) {
        Group& G = groups[gid[i]];
        if (G.desc.empty()) G = Group{tunes[i], keys[i].kinterp, keys[i].kprec, {}};
        G.desc.push_back(chunks[i]);
      }
      if (ng < 0) over_groups = true;
    }
    if (inflight.size() >= 2) {
      finish_epoch(inflight.front());
      inflight.pop_front();
    }
    ep.t_enqueue = now();
    sarbp_status st = exec->epoch_begin(&ep.mark);
    if (st == SARBP_OK && over_groups)
      st = fail(SARBP_ERR_UNSUPPORTED, "more than %d tile shapes in an epoch", kServerGroups);
    for (size_t g = 0; st == SARBP_OK && g < groups.size(); ++g)
      st = exec->run_group(k & 1, (int)g, groups[g].desc, groups[g].kinterp, groups[g].kprec, cfg.dtype,
                           groups[g].tune);
    if (st == SARBP_OK) st = exec->epoch_end(&ep.mark);
    if (st != SARBP_OK) {
      // Nothing was enqueued after the begin marker (or a launch failed):
      // fail the parts once the device is idle.
      exec->compute_wait();
      std::lock_guard<std::mutex> lk(mu);
      for (auto& pr : ep.parts) {
        Job& j = *jobs[pr.first];
        --j.inflight;
        j.state = SARBP_JOB_FAILED;
        j.status = st;
        set_detail(j);
        j.t_done = now();
        ++n_failed;
        remove_active(j.id);
        release_job(j);
      }
      exec->epoch_release(&ep.mark);
      cv_done.notify_all();
      continue;
    }
    inflight.push_back(ep);
    ++k;
  }
  while (!inflight.empty()) {
    finish_epoch(inflight.front());
    inflight.pop_front();
  }
}

void sarbp_server::complete_loop() {
  exec->bind_thread();
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_complete.wait(lk, [&] { return !finishing.empty() || stop; });
    if (finishing.empty()) {
      if (stop) return;
      continue;
    }
    Job* j = jobs[finishing.front()].get();
    finishing.pop_front();
    const void* src = j->args.img;
    void* dst = j->req.h_image;
    const size_t n = j->img_bytes;
    lk.unlock();
    const sarbp_status st = exec->download(dst, src, n);
    lk.lock();
    if (st != SARBP_OK) {
      j->state = SARBP_JOB_FAILED;
      j->status = st;
      set_detail(*j);
      ++n_failed;
    } else {
      j->state = SARBP_JOB_DONE;
      ++n_completed;
      work_done += j->work;
    }
    j->t_done = now();
    release_job(*j);
    cv_done.notify_all();
  }
}

namespace {

void fill_info(const Job& j, sarbp_job_info* info) {
  if (!info) return;
  std::memset(info, 0, sizeof *info);
  info->state = j.state;
  info->status = j.status;
  info->epochs = j.epochs;
  info->submit_s = j.t_submit;
  info->admit_s = j.t_admit;
  info->first_epoch_s = j.t_first;
  info->done_s = j.t_done;
  info->work = j.work;
}

}  // namespace

sarbp_status sarbp::server_create(const sarbp_server_config* cfg, std::unique_ptr<ServerExec> exec,
                                  sarbp_server** out) {
  if (!cfg || !out || !exec) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (cfg->interp < 0 || cfg->interp > 2 || cfg->precision < 0 || cfg->precision > 3 || cfg->dtype < 0 ||
      cfg->dtype > 1)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "bad (interp, precision, dtype)");
  if (cfg->max_active_jobs < 0 || cfg->pulses_per_epoch < 0 || cfg->epoch_watchdog_s < 0)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "negative limit");
  std::unique_ptr<sarbp_server> s(new (std::nothrow) sarbp_server);
  if (!s) return fail(SARBP_ERR_OUT_OF_MEMORY, "server");
  s->cfg = *cfg;
  for (int c = 0; c < SARBP_SIZE_CLASSES; ++c) {
    sarbp_tuning& t = s->tune[c];
    t = default_tuning(cfg->interp, c, cfg->precision);
    if (cfg->rows_per_lane) t.rows_per_lane = cfg->rows_per_lane;
    if (cfg->cols_per_lane) t.cols_per_lane = cfg->cols_per_lane;
    if (cfg->pulses_per_stage) t.pulses_per_stage = cfg->pulses_per_stage;
    sarbp_status st = check_tuning(cfg->precision, &t);
    if (st != SARBP_OK) return st;
  }
  s->exec = std::move(exec);
  s->n_cu = s->exec->n_cu();
  sarbp_status st = s->arena.init(cfg->hbm_budget_bytes, *s->exec);
  if (st != SARBP_OK) return st;
  sarbp_server* p = s.get();
  try {
    p->th_admit = std::thread([p] { p->admit_loop(); });
    p->th_sched = std::thread([p] { p->sched_loop(); });
    p->th_complete = std::thread([p] { p->complete_loop(); });
  } catch (...) {
    {
      std::lock_guard<std::mutex> lk(p->mu);
      p->stop = true;
    }
    p->cv_admit.notify_all();
    p->cv_sched.notify_all();
    p->cv_complete.notify_all();
    if (p->th_admit.joinable()) p->th_admit.join();
    if (p->th_sched.joinable()) p->th_sched.join();
    if (p->th_complete.joinable()) p->th_complete.join();
    return fail(SARBP_ERR_OUT_OF_MEMORY, "server threads");
  }
  *out = s.release();
  return SARBP_OK;
}

int sarbp::server_dtype(const sarbp_server* s) { return s->cfg.dtype; }

sarbp_status sarbp::server_submit(sarbp_server* s, const sarbp_request* req, const sarbp_raw_request* raw,
                                  uint64_t* job_id) {
  if (!s || !req || !job_id) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  const sarbp_params& p = req->params;
  sarbp_status st = validate(&req->grid, &p);
  if (st != SARBP_OK) return st;
  if (p.interp != s->cfg.interp || p.precision != s->cfg.precision || p.dtype != s->cfg.dtype)
    return fail(SARBP_ERR_INVALID_ARGUMENT, "request interp/precision/dtype differ from the server's");
  if (p.accumulate) return fail(SARBP_ERR_INVALID_ARGUMENT, "served jobs cannot accumulate into a host image");
  const size_t npix = (size_t)req->grid.nx * req->grid.ny;
  if ((p.n_pulses > 0 && ((!raw && !req->h_profiles) || (raw && !raw->h_raw) || !req->h_pos)) ||
      (npix > 0 && !req->h_image))
    return fail(SARBP_ERR_INVALID_ARGUMENT, "host pointer is NULL");
  std::unique_ptr<Job> j(new (std::nothrow) Job);
  if (!j) return fail(SARBP_ERR_OUT_OF_MEMORY, "job");
  j->req = *req;
  sarbp_tuning base;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    base = s->tune[size_class(&req->grid, &p, req->h_pos)];
  }
  const Fitted fit = fit_tuning(base, p.precision, &req->grid, &p, req->h_pos);
  j->tune = fit.t;
  if ((st = plan_args(req->h_pos, &req->grid, &p, j->tune, fit.prec, &j->args, nullptr, true)) != SARBP_OK)
    return st;
  j->prof_bytes = (size_t)p.n_pulses * p.n_samples * (p.dtype == SARBP_DTYPE_C64 ? 8 : 4);
  j->pos_bytes = (size_t)p.n_pulses * 24;
  j->img_bytes = npix * 8;
  if (raw) {
    j->raw = true;
    j->h_raw = raw->h_raw;
    j->rp = raw->range;
    j->raw_bytes = (size_t)p.n_pulses * raw->range.n_raw * 8;
  }
  j->total = align_up(j->prof_bytes) + align_up(j->pos_bytes) + align_up(j->img_bytes) + align_up(j->raw_bytes);
  j->work = (double)npix * p.n_pulses;
  if (j->total > s->arena.size())
    return fail(SARBP_ERR_OVER_BUDGET, "job needs %zu bytes, the arena holds %zu", j->total, s->arena.size());
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return fail(SARBP_ERR_SHUTDOWN, "server is shutting down");
  j->id = s->next_id++;
  j->t_submit = s->now();
  *job_id = j->id;
  ++s->n_submitted;
  if (npix == 0) {  // nothing to image
    j->state = SARBP_JOB_DONE;
    j->t_done = j->t_submit;
    ++s->n_completed;
    s->jobs.emplace(j->id, std::move(j));
    s->cv_done.notify_all();
    return SARBP_OK;
  }
  const uint64_t id = j->id;
  s->jobs.emplace(id, std::move(j));
  s->queue.push_back(id);
  s->cv_admit.notify_all();
  return SARBP_OK;
}

extern "C" {

sarbp_status sarbp_server_set_class_tuning(sarbp_server* s, int32_t size_cls, const sarbp_tuning* t) {
  if (!s || !t) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  if (size_cls < 0 || size_cls >= SARBP_SIZE_CLASSES) return fail(SARBP_ERR_INVALID_ARGUMENT, "size class %d", size_cls);
  sarbp_tuning v = *t;
  if (v.rows_per_lane == 0 && v.cols_per_lane == 0 && v.pulses_per_stage == 0)
    v = default_tuning(s->cfg.interp, size_cls, s->cfg.precision);
  sarbp_status st = check_tuning(s->cfg.precision, &v);
  if (st != SARBP_OK) return st;
  std::lock_guard<std::mutex> lk(s->mu);
  s->tune[size_cls] = v;
  return SARBP_OK;
}

sarbp_status sarbp_server_submit(sarbp_server* s, const sarbp_request* req, uint64_t* job_id) {
  return server_submit(s, req, nullptr, job_id);
}

sarbp_status sarbp_server_poll(sarbp_server* s, uint64_t id, sarbp_job_info* info) {
  if (!s) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return fail(SARBP_ERR_NOT_FOUND, "job %llu", (unsigned long long)id);
  fill_info(*it->second, info);
  return terminal(it->second->state) ? SARBP_OK : SARBP_ERR_NOT_READY;
}

sarbp_status sarbp_server_wait(sarbp_server* s, uint64_t id, double timeout_s, sarbp_job_info* info) {
  if (!s) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return fail(SARBP_ERR_NOT_FOUND, "job %llu", (unsigned long long)id);
  Job* j = it->second.get();
  auto pred = [&] { return terminal(j->state); };
  if (timeout_s < 0) {
    s->cv_done.wait(lk, pred);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
    fill_info(*j, info);
    return fail(SARBP_ERR_TIMEOUT, "job %llu not finished after %.3g s", (unsigned long long)id, timeout_s);
  }
  fill_info(*j, info);
  if (j->state == SARBP_JOB_FAILED) return fail(j->status, "job %llu: %s", (unsigned long long)id, j->detail);
  if (j->state == SARBP_JOB_CANCELLED) return fail(SARBP_ERR_CANCELLED, "job %llu cancelled", (unsigned long long)id);
  return SARBP_OK;
}

sarbp_status sarbp_server_cancel(sarbp_server* s, uint64_t id) {
  if (!s) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return fail(SARBP_ERR_NOT_FOUND, "job %llu", (unsigned long long)id);
  Job& j = *it->second;
  if (terminal(j.state)) return SARBP_OK;
  j.cancel = true;
  if (j.state == SARBP_JOB_QUEUED && !j.has_mem) {  // still waiting for admission
    j.state = SARBP_JOB_CANCELLED;
    j.t_done = s->now();
    ++s->n_cancelled;
    s->cv_done.notify_all();
    s->cv_admit.notify_all();
  } else if (j.inflight == 0 && (j.state == SARBP_JOB_RESIDENT || j.state == SARBP_JOB_RUNNING)) {
    s->remove_active(j.id);
    j.state = SARBP_JOB_CANCELLED;
    j.t_done = s->now();
    ++s->n_cancelled;
    s->release_job(j);
    s->cv_done.notify_all();
  }  // else: the scheduler or admission thread finishes the cancellation
  return SARBP_OK;
}

sarbp_status sarbp_server_release(sarbp_server* s, uint64_t id) {
  if (!s) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return fail(SARBP_ERR_NOT_FOUND, "job %llu", (unsigned long long)id);
  if (!terminal(it->second->state) || it->second->has_mem || it->second->inflight > 0)
    return fail(SARBP_ERR_NOT_READY, "job %llu still running", (unsigned long long)id);
  s->queue.erase(std::remove(s->queue.begin(), s->queue.end(), id), s->queue.end());
  s->jobs.erase(it);
  return SARBP_OK;
}

sarbp_status sarbp_server_get_stats(sarbp_server* s, sarbp_server_stats* out) {
  if (!s || !out) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof *out);
  out->submitted = s->n_submitted;
  out->completed = s->n_completed;
  out->failed = s->n_failed;
  out->cancelled = s->n_cancelled;
  out->epochs = s->n_epochs;
  int32_t q = 0;
  for (uint64_t id : s->queue)
    if (!terminal(s->jobs[id]->state)) ++q;
  out->queued = q + s->admitting;
  out->active = (int32_t)s->active.size() + (int32_t)s->finishing.size();
  out->work_done = s->work_done;
  out->gpu_busy_s = s->gpu_busy;
  out->uptime_s = s->now();
  out->hbm_in_use = s->arena.in_use();
  out->hbm_peak = s->arena.peak();
  return SARBP_OK;
}

sarbp_status sarbp_server_drain(sarbp_server* s, double timeout_s) {
  if (!s) return fail(SARBP_ERR_INVALID_ARGUMENT, "NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto idle = [&] {
    for (auto& kv : s->jobs)
      if (!terminal(kv.second->state) || kv.second->has_mem) return false;
    return true;
  };
  if (timeout_s < 0) {
    s->cv_done.wait(lk, idle);
    return SARBP_OK;
  }
  if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), idle))
    return fail(SARBP_ERR_TIMEOUT, "server not drained after %.3g s", timeout_s);
  return SARBP_OK;
}

void sarbp_server_destroy(sarbp_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& kv : s->jobs) {
      Job& j = *kv.second;
      if (terminal(j.state)) continue;
      j.cancel = true;
      if (j.state == SARBP_JOB_QUEUED && !j.has_mem) {
        j.state = SARBP_JOB_CANCELLED;
        ++s->n_cancelled;
      }
    }
    s->stop = true;
  }
  s->cv_admit.notify_all();
  s->cv_sched.notify_all();
  s->cv_complete.notify_all();
  if (s->th_admit.joinable()) s->th_admit.join();
  if (s->th_sched.joinable()) s->th_sched.join();  // drains in-flight epochs
  if (s->th_complete.joinable()) s->th_complete.join();
  s->exec->compute_wait();
  delete s;  // the executor releases the arena and its device resources
}

}  // extern "C"
