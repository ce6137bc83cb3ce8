// server_stress.cpp -- the continuous-batching server's host logic
// (csrc/sarbp_server_core.cpp) run against a CPU stand-in for the device, so
// its three threads, the job table, the arena and the watchdog can be run
// under ThreadSanitizer and AddressSanitizer + UndefinedBehaviorSanitizer on
// a machine without a GPU (tests/test_host_sanitizers.py builds and runs it).
//
// StubExec keeps the arena in host memory and runs a "compute stream": a
// worker thread executing epoch markers and batched launches in order.  A
// launch adds, for every job chunk, the chunk's pulse count into each pixel
// of the job's image (real part), so a finished image must hold n_pulses in
// every pixel -- a check of the chunking, the accumulation flags and the
// D2H copy.  It also checks every launch descriptor it is given (pulse range,
// image inside the arena, one kernel per group).
//
// Scenarios: many client threads submitting, cancelling, waiting and
// releasing at once; admission backpressure; over-budget rejection; the
// epoch watchdog; cancellation of running jobs; raw-echo submissions;
// shutdown with work in flight; argument errors.  Exit status 0 when every
// scenario passes.
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "sarbp.h"
#include "sarbp_server_core.h"

using namespace sarbp;

namespace {

std::atomic<int> g_failures{0};

#define CHECK(cond, ...)                                      \
  do {                                                        \
    if (!(cond)) {                                            \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);                      \
      std::fprintf(stderr, "\n");                             \
      g_failures.fetch_add(1);                                \
    }                                                         \
  } while (0)

struct Marker {
  bool done = false;
  std::chrono::steady_clock::time_point t;
};

class StubExec final : public ServerExec {
 public:
  explicit StubExec(int delay_ms = 0) : delay_ms_(delay_ms) {
    dev_ = std::thread([this] { device_loop(); });
  }
  ~StubExec() override {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    dev_.join();
    std::free(arena_);
  }
  void set_delay_ms(int ms) { delay_ms_.store(ms); }
  int violations() const { return violations_.load(); }
  int launches() const { return launches_.load(); }

  int n_cu() const override { return 4; }
  sarbp_status alloc_arena(size_t bytes, char** base) override {
    arena_ = static_cast<char*>(std::calloc(bytes, 1));
    if (!arena_) return SARBP_ERR_OUT_OF_MEMORY;
    arena_bytes_ = bytes;
    *base = arena_;
    return SARBP_OK;
  }
  void bind_thread() override {}

  sarbp_status upload(void* dst, const void* src, size_t n) override {
    if (!in_arena(dst, n)) violations_.fetch_add(1);
    std::memcpy(dst, src, n);
    return SARBP_OK;
  }
  sarbp_status compress(const sarbp_cf32* d_raw, int32_t n_pulses, const sarbp_range_params* rp,
                        void* d_out) override {
    const size_t n = (size_t)n_pulses * rp->n_out * sizeof(sarbp_cf32);
    if (!in_arena(d_raw, (size_t)n_pulses * rp->n_raw * sizeof(sarbp_cf32)) || !in_arena(d_out, n))
      violations_.fetch_add(1);
    std::memset(d_out, 0, n);
    return SARBP_OK;
  }
  sarbp_status admit_wait() override { return SARBP_OK; }

  sarbp_status epoch_begin(EpochMark* m) override {
    Marker* b = new Marker;
    Marker* e = new Marker;
    m->begin = b;
    m->end = e;
    enqueue([this, b] { mark(b); });
    return SARBP_OK;
  }
  sarbp_status run_group(int slot, int g, std::vector<JobDesc>& desc, int kinterp, int kprec, int dtype,
                         const sarbp_tuning& t) override {
    if (slot < 0 || slot > 1 || g < 0 || g >= kServerGroups) violations_.fetch_add(1);
    for (const JobDesc& d : desc) {
      const BpArgs& a = d.a;
      const bool ok = a.kernel_interp == kinterp && a.kernel_prec == kprec && a.pulse_begin >= 0 &&
                      a.pulse_begin < a.pulse_end && a.pulse_end <= a.n_pulses &&
                      a.accumulate == (a.pulse_begin > 0 ? 1 : 0) && t.rows_per_lane > 0 && dtype >= 0 &&
                      in_arena(a.img, (size_t)a.nx * a.ny * sizeof(sarbp_cf32));
      if (!ok) violations_.fetch_add(1);
    }
    std::vector<JobDesc> work(desc);
    launches_.fetch_add(1);
    enqueue([work] {
      for (const JobDesc& d : work) {
        float* img = reinterpret_cast<float*>(d.a.img);
        const float add = (float)(d.a.pulse_end - d.a.pulse_begin);
        const size_t n = (size_t)d.a.nx * d.a.ny;
        for (size_t i = 0; i < n; ++i) {
          img[2 * i] = (d.a.accumulate ? img[2 * i] : 0.0f) + add;
          img[2 * i + 1] = 0.0f;
        }
      }
    });
    return SARBP_OK;
  }
  sarbp_status epoch_end(EpochMark* m) override {
    Marker* e = static_cast<Marker*>(m->end);
    enqueue([this, e] {
      const int ms = delay_ms_.load();
      if (ms > 0) std::this_thread::sleep_for(std::chrono::milliseconds(ms));
      mark(e);
    });
    return SARBP_OK;
  }
  int epoch_done(const EpochMark& m) override {
    std::lock_guard<std::mutex> lk(m_);
    return static_cast<Marker*>(m.end)->done ? 1 : 0;
  }
  sarbp_status epoch_wait(const EpochMark& m, float* ms) override {
    std::unique_lock<std::mutex> lk(m_);
    Marker* b = static_cast<Marker*>(m.begin);
    Marker* e = static_cast<Marker*>(m.end);
    cv_done_.wait(lk, [&] { return e->done; });
    *ms = std::chrono::duration<float, std::milli>(e->t - b->t).count();
    return SARBP_OK;
  }
  void epoch_release(EpochMark* m) override {
    delete static_cast<Marker*>(m->begin);
    delete static_cast<Marker*>(m->end);
    m->begin = m->end = nullptr;
  }
  void compute_wait() override {
    Marker f;
    enqueue([this, &f] { mark(&f); });
    std::unique_lock<std::mutex> lk(m_);
    cv_done_.wait(lk, [&] { return f.done; });
  }
  sarbp_status download(void* dst, const void* src, size_t n) override {
    if (!in_arena(src, n)) violations_.fetch_add(1);
    std::memcpy(dst, src, n);
    return SARBP_OK;
  }

 private:
  bool in_arena(const void* p, size_t n) const {
    const char* c = static_cast<const char*>(p);
    return c >= arena_ && c + n <= arena_ + arena_bytes_;
  }
  void enqueue(std::function<void()> f) {
    {
      std::lock_guard<std::mutex> lk(m_);
      q_.push_back(std::move(f));
    }
    cv_.notify_all();
  }
  void mark(Marker* k) {
    std::lock_guard<std::mutex> lk(m_);
    k->t = std::chrono::steady_clock::now();
    k->done = true;
    cv_done_.notify_all();
  }
  void device_loop() {
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
      if (q_.empty()) return;
      std::function<void()> f = std::move(q_.front());
      q_.pop_front();
      lk.unlock();
      f();
      lk.lock();
    }
  }

  char* arena_ = nullptr;
  size_t arena_bytes_ = 0;
  std::mutex m_;
  std::condition_variable cv_, cv_done_;
  std::deque<std::function<void()>> q_;
  bool stop_ = false;
  std::thread dev_;
  std::atomic<int> delay_ms_;
  std::atomic<int> violations_{0};
  std::atomic<int> launches_{0};
};

// A small X-band job 5 km from a straight track at 3 km altitude.
struct ClientJob {
  std::vector<sarbp_cf32> prof;
  std::vector<double> pos;
  std::vector<sarbp_cf32> img;
  sarbp_request req{};
};

void make_job(ClientJob& j, int nx, int ny, int n_pulses, int n_samples, double spacing, int interp) {
  j.prof.assign((size_t)n_pulses * n_samples, sarbp_cf32{0.5f, -0.25f});
  j.pos.resize((size_t)3 * n_pulses);
  for (int m = 0; m < n_pulses; ++m) {
    j.pos[3 * m] = -0.15 * n_pulses / 2 + 0.15 * m;
    j.pos[3 * m + 1] = 0.0;
    j.pos[3 * m + 2] = 3000.0;
  }
  j.img.assign((size_t)nx * ny, sarbp_cf32{-1.0f, -1.0f});
  sarbp_request& r = j.req;
  r.h_profiles = j.prof.data();
  r.h_pos = j.pos.data();
  r.h_image = j.img.data();
  r.grid = sarbp_grid{{-0.5 * nx * spacing, 4000.0 - 0.5 * ny * spacing, 0.0}, {spacing, 0, 0}, {0, spacing, 0}, nx, ny};
  r.params.c = 299792458.0;
  r.params.fc = 9.6e9;
  r.params.dr = 0.1875;
  r.params.r0 = 5000.0 - 0.5 * n_samples * r.params.dr;
  r.params.kaiser_beta = 6.0;
  r.params.n_pulses = n_pulses;
  r.params.n_samples = n_samples;
  r.params.interp = interp;
  r.params.precision = SARBP_PREC_REF;
  r.params.dtype = SARBP_DTYPE_C64;
  r.params.accumulate = 0;
}

bool image_holds(const ClientJob& j) {
  for (const sarbp_cf32& v : j.img)
    if (v.re != (float)j.req.params.n_pulses || v.im != 0.0f) return false;
  return true;
}

sarbp_server_config config(size_t budget, int pulses_per_epoch, int max_active, double watchdog) {
  sarbp_server_config c{};
  c.device = 0;
  c.interp = SARBP_INTERP_SINC8;
  c.precision = SARBP_PREC_REF;
  c.dtype = SARBP_DTYPE_C64;
  c.hbm_budget_bytes = budget;
  c.max_active_jobs = max_active;
  c.pulses_per_epoch = pulses_per_epoch;
  c.epoch_watchdog_s = watchdog;
  return c;
}

sarbp_server* create(const sarbp_server_config& c, StubExec** ex_out) {
  StubExec* ex = new StubExec;
  *ex_out = ex;
  sarbp_server* s = nullptr;
  const sarbp_status st = server_create(&c, std::unique_ptr<ServerExec>(ex), &s);
  CHECK(st == SARBP_OK, "server_create: %s", last_detail());
  return s;
}

// Waits and drains must be woken by the server, not end at their timeouts:
// every call below has 60 s and must return within kPrompt.
constexpr double kPrompt = 10.0;

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

sarbp_status drain(sarbp_server* s, const char* what) {
  const auto t0 = std::chrono::steady_clock::now();
  const sarbp_status st = sarbp_server_drain(s, 60.0);
  CHECK(seconds_since(t0) < kPrompt, "%s: drain returned after %.1f s (a missed wakeup)", what, seconds_since(t0));
  return st;
}

sarbp_status wait(sarbp_server* s, uint64_t id, sarbp_job_info* info) {
  const auto t0 = std::chrono::steady_clock::now();
  const sarbp_status st = sarbp_server_wait(s, id, 60.0, info);
  CHECK(seconds_since(t0) < kPrompt, "job %llu: wait returned after %.1f s (a missed wakeup)",
        (unsigned long long)id, seconds_since(t0));
  return st;
}

void check_drained(sarbp_server* s, StubExec* ex, const char* what) {
  CHECK(drain(s, what) == SARBP_OK, "%s: drain", what);
  sarbp_server_stats st{};
  CHECK(sarbp_server_get_stats(s, &st) == SARBP_OK, "%s: stats", what);
  CHECK(st.submitted == st.completed + st.failed + st.cancelled, "%s: %lld submitted, %lld + %lld + %lld", what,
        (long long)st.submitted, (long long)st.completed, (long long)st.failed, (long long)st.cancelled);
  CHECK(st.hbm_in_use == 0, "%s: %zu bytes still in the arena", what, st.hbm_in_use);
  CHECK(st.queued == 0 && st.active == 0, "%s: queued %d active %d", what, st.queued, st.active);
  CHECK(ex->violations() == 0, "%s: %d bad launch descriptors or copies", what, ex->violations());
}

// Client threads submitting, cancelling, polling, waiting and releasing at once.
void scenario_concurrent() {
  StubExec* ex = nullptr;
  sarbp_server* s = create(config(24u << 20, 48, 6, 0.0), &ex);
  const int n_threads = 4, per_thread = 40;
  std::atomic<int> done{0}, cancelled{0};
  std::vector<std::thread> th;
  for (int t = 0; t < n_threads; ++t) {
    th.emplace_back([&, t] {
      std::mt19937 rng(1234 + t);
      std::vector<std::unique_ptr<ClientJob>> mine;
      std::vector<uint64_t> ids;
      for (int k = 0; k < per_thread; ++k) {
        auto j = std::make_unique<ClientJob>();
        const int side = 8 + (int)(rng() % 120);
        const double sp = 0.25 + 0.25 * (rng() % 4);
        make_job(*j, side, side + (int)(rng() % 9), 1 + (int)(rng() % 200), 64 + (int)(rng() % 448), sp,
                 SARBP_INTERP_SINC8);
        uint64_t id = 0;
        const sarbp_status st = sarbp_server_submit(s, &j->req, &id);
        CHECK(st == SARBP_OK, "submit: %s", last_detail());
        if (st != SARBP_OK) continue;
        if (rng() % 8 == 0) CHECK(sarbp_server_cancel(s, id) == SARBP_OK, "cancel");
        sarbp_job_info info{};
        (void)sarbp_server_poll(s, id, &info);
        ids.push_back(id);
        mine.push_back(std::move(j));
      }
      for (size_t k = 0; k < ids.size(); ++k) {
        sarbp_job_info info{};
        const sarbp_status st = wait(s, ids[k], &info);
        if (info.state == SARBP_JOB_DONE) {
          CHECK(st == SARBP_OK && image_holds(*mine[k]), "job %llu: image or status wrong",
                (unsigned long long)ids[k]);
          done.fetch_add(1);
        } else {
          CHECK(info.state == SARBP_JOB_CANCELLED && st == SARBP_ERR_CANCELLED, "job %llu: state %d status %d",
                (unsigned long long)ids[k], info.state, st);
          cancelled.fetch_add(1);
        }
        CHECK(sarbp_server_release(s, ids[k]) == SARBP_OK, "release");
      }
    });
  }
  for (auto& t : th) t.join();
  CHECK(done.load() + cancelled.load() == n_threads * per_thread, "%d done + %d cancelled", done.load(),
        cancelled.load());
  CHECK(done.load() > 0 && ex->launches() > 0, "no work ran");
  check_drained(s, ex, "concurrent");
  sarbp_server_destroy(s);
  std::printf("concurrent: %d done, %d cancelled\n", done.load(), cancelled.load());
}

// An arena that holds one job at a time: every job still completes, in turn;
// one that can never fit is rejected at submission.
void scenario_backpressure_and_budget() {
  StubExec* ex = nullptr;
  ClientJob probe;
  make_job(probe, 96, 96, 100, 256, 0.5, SARBP_INTERP_SINC8);
  const size_t one = 100 * 256 * 8 + 100 * 24 + 96 * 96 * 8 + 3 * 256;
  sarbp_server* s = create(config(one + one / 2, 32, 0, 0.0), &ex);
  std::vector<std::unique_ptr<ClientJob>> js;
  std::vector<uint64_t> ids;
  for (int k = 0; k < 8; ++k) {
    js.push_back(std::make_unique<ClientJob>());
    make_job(*js.back(), 96, 96, 100, 256, 0.5, SARBP_INTERP_SINC8);
    uint64_t id = 0;
    CHECK(sarbp_server_submit(s, &js.back()->req, &id) == SARBP_OK, "submit %d", k);
    ids.push_back(id);
  }
  ClientJob big;
  make_job(big, 512, 512, 400, 2048, 0.5, SARBP_INTERP_SINC8);
  uint64_t bid = 0;
  CHECK(sarbp_server_submit(s, &big.req, &bid) == SARBP_ERR_OVER_BUDGET, "over-budget job accepted");
  for (size_t k = 0; k < ids.size(); ++k) {
    sarbp_job_info info{};
    CHECK(wait(s, ids[k], &info) == SARBP_OK && image_holds(*js[k]), "job %zu", k);
  }
  sarbp_server_stats st{};
  (void)sarbp_server_get_stats(s, &st);
  CHECK(st.hbm_peak <= one + one / 2, "arena peak %zu above budget", st.hbm_peak);
  check_drained(s, ex, "backpressure");
  sarbp_server_destroy(s);
  std::printf("backpressure: 8 jobs through a one-job arena, over-budget rejected\n");
}

// Epochs slower than the watchdog fail their jobs with SARBP_ERR_TIMEOUT and
// release their memory once the device work has finished; later jobs run.
void scenario_watchdog() {
  StubExec* ex = nullptr;
  sarbp_server* s = create(config(8u << 20, 16, 0, 0.01), &ex);
  ex->set_delay_ms(60);
  ClientJob a;
  make_job(a, 32, 32, 64, 128, 0.5, SARBP_INTERP_SINC8);
  uint64_t ia = 0;
  CHECK(sarbp_server_submit(s, &a.req, &ia) == SARBP_OK, "submit");
  sarbp_job_info info{};
  const sarbp_status st = wait(s, ia, &info);
  CHECK(st == SARBP_ERR_TIMEOUT && info.state == SARBP_JOB_FAILED, "watchdog: status %d state %d", st,
        info.state);
  CHECK(drain(s, "watchdog") == SARBP_OK, "drain after the watchdog");
  ex->set_delay_ms(0);
  ClientJob b;
  make_job(b, 40, 24, 50, 128, 0.5, SARBP_INTERP_SINC8);
  uint64_t ib = 0;
  CHECK(sarbp_server_submit(s, &b.req, &ib) == SARBP_OK, "submit after the watchdog");
  CHECK(wait(s, ib, &info) == SARBP_OK && image_holds(b), "job after the watchdog");
  check_drained(s, ex, "watchdog");
  sarbp_server_destroy(s);
  std::printf("watchdog: slow epoch failed with a timeout, the next job completed\n");
}

// Cancelling a job between its epochs; cancelling queued jobs.
void scenario_cancel_running() {
  StubExec* ex = nullptr;
  sarbp_server* s = create(config(16u << 20, 8, 1, 0.0), &ex);
  ex->set_delay_ms(5);
  ClientJob a, b, c;
  make_job(a, 64, 64, 400, 256, 0.5, SARBP_INTERP_SINC8);  // 50 epochs
  make_job(b, 16, 16, 10, 128, 0.5, SARBP_INTERP_SINC8);
  make_job(c, 16, 16, 10, 128, 0.5, SARBP_INTERP_SINC8);
  uint64_t ia = 0, ib = 0, ic = 0;
  CHECK(sarbp_server_submit(s, &a.req, &ia) == SARBP_OK, "submit a");
  CHECK(sarbp_server_submit(s, &b.req, &ib) == SARBP_OK, "submit b");  // waits: max_active_jobs = 1
  CHECK(sarbp_server_submit(s, &c.req, &ic) == SARBP_OK, "submit c");
  sarbp_job_info info{};
  for (int k = 0; k < 2000 && !(sarbp_server_poll(s, ia, &info), info.epochs >= 3); ++k)
    std::this_thread::sleep_for(std::chrono::milliseconds(1));
  CHECK(info.epochs >= 3, "job a did not start");
  CHECK(sarbp_server_cancel(s, ia) == SARBP_OK && sarbp_server_cancel(s, ib) == SARBP_OK, "cancel");
  CHECK(wait(s, ia, &info) == SARBP_ERR_CANCELLED && info.state == SARBP_JOB_CANCELLED,
        "running job not cancelled: state %d", info.state);
  CHECK(wait(s, ib, &info) == SARBP_ERR_CANCELLED, "queued job not cancelled");
  CHECK(wait(s, ic, &info) == SARBP_OK && image_holds(c), "job c");
  CHECK(sarbp_server_release(s, ia) == SARBP_OK, "release a");
  CHECK(sarbp_server_release(s, ia) == SARBP_ERR_NOT_FOUND, "released twice");
  check_drained(s, ex, "cancel");
  sarbp_server_destroy(s);
  std::printf("cancel: running and queued jobs cancelled, the next job completed\n");
}

// Raw-echo jobs go through the compression step of admission.
void scenario_raw() {
  StubExec* ex = nullptr;
  sarbp_server* s = create(config(8u << 20, 32, 0, 0.0), &ex);
  ClientJob j;
  make_job(j, 48, 40, 70, 512, 0.5, SARBP_INTERP_SINC8);
  std::vector<sarbp_cf32> raw((size_t)70 * 200, sarbp_cf32{1.0f, 0.0f});
  sarbp_raw_request r{};
  r.h_raw = raw.data();
  r.range.n_raw = 200;
  r.range.n_out = 512;
  r.range.upsample = 4;
  uint64_t id = 0;
  CHECK(server_submit(s, &j.req, &r, &id) == SARBP_OK, "raw submit: %s", last_detail());
  sarbp_job_info info{};
  CHECK(wait(s, id, &info) == SARBP_OK && image_holds(j), "raw job");
  check_drained(s, ex, "raw");
  sarbp_server_destroy(s);
  std::printf("raw: raw-echo job completed\n");
}

// Destroying a server with queued, resident and running jobs.
void scenario_shutdown() {
  for (int rep = 0; rep < 3; ++rep) {
    StubExec* ex = nullptr;
    sarbp_server* s = create(config(4u << 20, 16, 0, 0.0), &ex);
    ex->set_delay_ms(rep);
    std::vector<std::unique_ptr<ClientJob>> js;
    for (int k = 0; k < 30; ++k) {
      js.push_back(std::make_unique<ClientJob>());
      make_job(*js.back(), 40 + k, 30, 60 + 5 * k, 256, 0.5, SARBP_INTERP_SINC8);
      uint64_t id = 0;
      CHECK(sarbp_server_submit(s, &js.back()->req, &id) == SARBP_OK, "submit");
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(3 * rep));
    sarbp_server_destroy(s);  // must return with no thread or memory left behind
  }
  std::printf("shutdown: servers destroyed with work in flight\n");
}

// Argument errors are reported, not crashed on.
void scenario_errors() {
  StubExec* ex = nullptr;
  sarbp_server_config bad = config(1u << 20, 16, 0, 0.0);
  bad.interp = 9;
  sarbp_server* s = nullptr;
  CHECK(server_create(&bad, std::unique_ptr<ServerExec>(new StubExec), &s) == SARBP_ERR_INVALID_ARGUMENT,
        "bad interp accepted");
  bad = config(1u << 20, -1, 0, 0.0);
  CHECK(server_create(&bad, std::unique_ptr<ServerExec>(new StubExec), &s) == SARBP_ERR_INVALID_ARGUMENT,
        "negative limit accepted");
  s = create(config(1u << 20, 16, 0, 0.0), &ex);
  ClientJob j;
  make_job(j, 16, 16, 8, 64, 0.5, SARBP_INTERP_SINC8);
  uint64_t id = 0;
  j.req.params.interp = SARBP_INTERP_LINEAR;
  CHECK(sarbp_server_submit(s, &j.req, &id) == SARBP_ERR_INVALID_ARGUMENT, "interp mismatch accepted");
  j.req.params.interp = SARBP_INTERP_SINC8;
  j.req.params.accumulate = 1;
  CHECK(sarbp_server_submit(s, &j.req, &id) == SARBP_ERR_INVALID_ARGUMENT, "accumulate accepted");
  j.req.params.accumulate = 0;
  j.req.h_image = nullptr;
  CHECK(sarbp_server_submit(s, &j.req, &id) == SARBP_ERR_INVALID_ARGUMENT, "NULL image accepted");
  j.req.h_image = j.img.data();
  j.req.grid.ux[0] = std::nan("");
  CHECK(sarbp_server_submit(s, &j.req, &id) == SARBP_ERR_INVALID_ARGUMENT, "NaN grid accepted");
  CHECK(sarbp_server_poll(s, 999999, nullptr) == SARBP_ERR_NOT_FOUND, "unknown job");
  sarbp_tuning t{3, 1, 4};
  CHECK(sarbp_server_set_class_tuning(s, 0, &t) == SARBP_ERR_INVALID_ARGUMENT, "bad tile shape accepted");
  CHECK(sarbp_server_set_class_tuning(s, SARBP_SIZE_CLASSES, &t) == SARBP_ERR_INVALID_ARGUMENT, "bad class");
  check_drained(s, ex, "errors");
  sarbp_server_destroy(s);
  std::printf("errors: bad configurations and requests rejected\n");
}

}  // namespace

int main(int argc, char** argv) {
  const std::string only = argc > 1 ? argv[1] : "";
  const std::pair<const char*, void (*)()> all[] = {
      {"concurrent", scenario_concurrent}, {"backpressure", scenario_backpressure_and_budget},
      {"watchdog", scenario_watchdog},     {"cancel", scenario_cancel_running},
      {"raw", scenario_raw},               {"shutdown", scenario_shutdown},
      {"errors", scenario_errors}};
  for (const auto& sc : all)
    if (only.empty() || only == sc.first) sc.second();
  if (g_failures.load() != 0) {
    std::printf("%d checks failed\n", g_failures.load());
    return 1;
  }
  std::printf("all scenarios passed\n");
  return 0;
}
