// The serving loop: zoom-window requests against a device-resident survey,
// served by continuous batching.
//
// A request is the reference's zoomWindow (tests/utils/viewersim.nim:167-191)
// without the decode: margin = requiredMargin of the flow at the survey's dt,
// window = decodeWindow(shown, margin, samples), plan = planFor(flow, window
// width, samples) (viewersim.nim:53-62, compiled on first use), the plan on
// the window's rows, the shown columns kept.  The window's rows are read in
// place from the survey (a job's source pitch is the survey's), the cropped
// result goes to the server's device arena.
//
// One worker thread owns the GPU side.  Each turn it retires finished
// launches (an event per launch), then admits queued requests in arrival
// order while their results fit in the arena and the batch limits allow,
// and packs them into one job table per (threads per row, kernel form)
// class: one launch per class, one workgroup per row in heaviest-first
// order, the table copied from pinned memory on the launch's own stream (a
// class of one request passes its job by value).  With two launches in
// flight the next batch is built
// and queued while the previous one runs.  A row's arithmetic depends only
// on its plan and its samples, so a request's bits do not depend on what it
// shares a launch with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "pseg_plan.hpp"

using namespace pseg;

namespace {

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

std::string fmt(const char *f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char *f, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

// First-fit allocator over one device buffer (256-byte granules).
class Arena {
 public:
  void reset(char *base, size_t bytes) {
    base_ = base;
    free_.clear();
    if (bytes) free_[0] = bytes;
    used_ = peak_ = 0;
  }
  char *alloc(size_t n) {
    n = (std::max<size_t>(n, 1) + 255) / 256 * 256;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        used_ += n;
        peak_ = std::max(peak_, used_);
        sizes_[off] = n;
        return base_ + off;
      }
    }
    return nullptr;
  }
  void release(char *p) {
    const size_t off = (size_t)(p - base_);
    auto s = sizes_.find(off);
    if (s == sizes_.end()) return;
    size_t n = s->second;
    sizes_.erase(s);
    used_ -= n;
    size_t start = off;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && next->first == off + n) {
      n += next->second;
      free_.erase(next);
    }
    auto prev = free_.lower_bound(off);
    if (prev != free_.begin()) {
      --prev;
      if (prev->first + prev->second == off) {
        start = prev->first;
        n += prev->second;
        free_.erase(prev);
      }
    }
    free_[start] = n;
  }
  size_t peak() const { return peak_; }

 private:
  char *base_ = nullptr;
  std::map<size_t, size_t> free_;   // offset -> length
  std::map<size_t, size_t> sizes_;  // offset -> allocated length
  size_t used_ = 0, peak_ = 0;
};

struct Flow {
  std::vector<pseg_op> ops;
  int32_t margin;
};

enum State { kQueued, kRunning, kDone };

struct Request {
  int32_t flow, trace0, ntraces, shown_first, shown_last;
  int32_t first, window;
  State state = kQueued;
  pseg_status status = PSEG_OK;
  std::string error;
  char *out = nullptr;
  size_t out_bytes = 0;
  pseg_plan *plan = nullptr;
  int32_t batch_requests = 0, plan_compiled = 0;
  int64_t submit_ns = 0, start_ns = 0, done_ns = 0;
};

constexpr int kSlots = 2;        // launches in flight at most
constexpr int kMaxJobs = 1024;   // requests per launch at most

struct Slot {
  bool busy = false;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  Job *host_jobs = nullptr;      // pinned staging
  char *dev = nullptr;           // the job table
  std::vector<uint64_t> tickets;
};

}  // namespace

struct pseg_server {
  pseg_server_config cfg;
  int device = 0;
  size_t esz = 4;
  char *arena_mem = nullptr;
  Arena arena;
  // flows and the plan cache (keyed like planFor: (flow, cols, record)),
  // under plan_mu: the worker compiles while submitters and add_flow run
  std::mutex plan_mu;
  std::vector<Flow> flows;
  std::map<std::tuple<int, int, int>, pseg_plan *> plans;
  Slot slots[kSlots];
  std::mutex mu;
  std::condition_variable work_cv, done_cv;
  std::deque<uint64_t> queue;
  std::unordered_map<uint64_t, Request> reqs;
  uint64_t next_ticket = 1;
  bool stop = false;
  // bumped (under mu) by every submit and release: the worker's unlocked
  // poll of its launches also watches for work it could admit
  std::atomic<uint64_t> wake_seq{0};
  pseg_server_info info{};
  std::thread worker;

  void run();
  void retire_finished(std::unique_lock<std::mutex> &lk);
  bool launch_batch(std::unique_lock<std::mutex> &lk, int slot);
  pseg_plan *plan_for(int flow, int32_t cols, bool *compiled, std::string *err,
                      pseg_status *st);
};

void pseg_server::retire_finished(std::unique_lock<std::mutex> &) {
  for (Slot &s : slots) {
    if (!s.busy) continue;
    hipError_t q = hipEventQuery(s.done);
    if (q == hipErrorNotReady) continue;
    const int64_t t = now_ns();
    for (uint64_t tk : s.tickets) {
      auto it = reqs.find(tk);
      if (it == reqs.end()) continue;
      Request &r = it->second;
      r.state = kDone;
      r.done_ns = t;
      if (q != hipSuccess) {
        r.status = PSEG_E_HIP;
        r.error = fmt("launch: %s", hipGetErrorString(q));
      }
    }
    s.tickets.clear();
    s.busy = false;
    done_cv.notify_all();
  }
}

pseg_plan *pseg_server::plan_for(int flow, int32_t cols, bool *compiled, std::string *err,
                                 pseg_status *st) {
  auto key = std::make_tuple(flow, (int)cols, (int)cfg.samples);
  std::vector<pseg_op> ops;
  *compiled = false;
  {
    std::lock_guard<std::mutex> g(plan_mu);
    auto it = plans.find(key);
    if (it != plans.end()) return it->second;
    ops = flows[flow].ops;
  }
  pseg_plan *p = nullptr;
  *st = pseg_plan_compile(ops.data(), (int32_t)ops.size(), cfg.dtype, cfg.dt_ms, cols,
                          cfg.samples, &p);
  if (*st != PSEG_OK) {
    char buf[512];
    pseg_last_error(buf, sizeof buf);
    *err = buf;
    return nullptr;
  }
  std::lock_guard<std::mutex> g(plan_mu);
  plans[key] = p;
  *compiled = true;
  return p;
}

// Takes queued requests in arrival order into one launch on `slot`.
// Returns false when nothing was admitted.
bool pseg_server::launch_batch(std::unique_lock<std::mutex> &lk, int slot_index) {
  Slot &S = slots[slot_index];
  std::vector<uint64_t> take;
  int64_t rows = 0;
  const int max_req = std::min(std::max(cfg.max_batch_requests, 1), kMaxJobs);
  while (!queue.empty() && (int)take.size() < max_req) {
    Request &r = reqs[queue.front()];
    if (cfg.max_batch_rows > 0 && !take.empty() && rows + r.ntraces > cfg.max_batch_rows) break;
    char *out = arena.alloc(r.out_bytes);
    if (!out) break;  // admission: wait for results to be released
    r.out = out;
    r.state = kRunning;
    rows += r.ntraces;
    take.push_back(queue.front());
    queue.pop_front();
  }
  if (take.empty()) return false;
  info.arena_peak_bytes = (int64_t)arena.peak();
  // the plans (compiled outside the lock: it runs host FFTs and copies)
  std::vector<Request *> rq;
  for (uint64_t tk : take) rq.push_back(&reqs[tk]);
  lk.unlock();
  const int64_t t_start = now_ns();
  struct Item {
    Request *r;
    pseg_plan *plan;
  };
  std::vector<Item> ok;
  std::vector<std::tuple<Request *, pseg_status, std::string>> bad;
  int compiled_now = 0;
  for (Request *r : rq) {
    bool compiled = false;
    std::string err;
    pseg_status st = PSEG_OK;
    pseg_plan *p = plan_for(r->flow, r->window, &compiled, &err, &st);
    if (!p) {
      bad.emplace_back(r, st, err);
      continue;
    }
    r->plan = p;
    r->plan_compiled = compiled ? 1 : 0;
    compiled_now += compiled;
    ok.push_back({r, p});
  }
  // one launch per (threads per row, kernel form) class; heavier rows first
  std::map<std::pair<int, int>, std::vector<Item>> classes;  // (threads, kernel form)
  for (const Item &it : ok) classes[{it.plan->threads, it.plan->full}].push_back(it);
  size_t n_jobs = 0;
  std::vector<JobLaunch> launches;
  std::vector<int> launch_threads;
  for (auto &kv : classes) {
    std::vector<Item> &v = kv.second;
    std::stable_sort(v.begin(), v.end(), [](const Item &a, const Item &b) {
      return a.r->window > b.r->window;
    });
    JobLaunch L;
    memset(&L, 0, sizeof L);
    size_t lds = 1024;
    int32_t item0 = 0;
    const size_t first_job = n_jobs;
    // A class of one request runs as a Plan.apply does: its job by value,
    // with no job table to copy (what a lightly loaded server launches for
    // most requests).
    const bool by_value = v.size() == 1;
    for (const Item &it : v) {
      Request &r = *it.r;
      Job &J = by_value ? L.single : S.host_jobs[n_jobs++];
      memset(&J, 0, sizeof J);
      J.plan = it.plan->dev;
      J.src = (const char *)cfg.survey + ((size_t)r.trace0 * cfg.pitch + r.first) * esz;
      J.dst = r.out;
      J.num = cfg.trace_len + r.trace0;
      J.src_pitch = cfg.pitch;
      J.dst_pitch = r.shown_last - r.shown_first + 1;
      J.rows = r.ntraces;
      J.cols = r.window;
      J.first = r.first;
      J.keys = cfg.trace_keys ? cfg.trace_keys + 2 * (size_t)r.trace0 : nullptr;
      J.col0 = r.first;  // the window's column 0 is trace sample `first`
      J.crop0 = r.shown_first - r.first;
      J.crop1 = J.crop0 + (r.shown_last - r.shown_first + 1);
      J.item0 = item0;
      J.step_first = 0;
      J.step_last = it.plan->n_steps;
      item0 += r.ntraces;
      for (int s = 0; s < it.plan->n_steps; s++) lds = std::max(lds, it.plan->step_lds[s]);
      L.max_cols = std::max(L.max_cols, r.window);
    }
    L.jobs = by_value ? nullptr : (const Job *)(S.dev + first_job * sizeof(Job));
    L.n_jobs = by_value ? 0 : (int32_t)v.size();
    L.total = item0;
    L.lds_bytes = lds;
    const int threads = kv.first.first;
    L.full = kv.first.second;
    L.grid = L.total;  // one workgroup per row (no counter)
    launches.push_back(L);
    launch_threads.push_back(threads);
  }
  hipError_t e = hipSuccess;
  if (n_jobs > 0)
    e = hipMemcpyAsync(S.dev, S.host_jobs, n_jobs * sizeof(Job), hipMemcpyHostToDevice, S.stream);
  for (size_t i = 0; i < launches.size() && e == hipSuccess; i++) {
    // one workgroup per row (the dispatcher hands rows out as workgroups
    // finish, heaviest first): measured faster than a persistent grid that
    // takes rows from a counter, as the batch tuner finds for batches too
    e = launch_pipeline(cfg.dtype, launch_threads[i], launches[i], S.stream);
  }
  if (e == hipSuccess) e = hipEventRecord(S.done, S.stream);
  lk.lock();
  info.batches++;
  info.launches += (int64_t)launches.size();
  info.plans_compiled += compiled_now;
  for (auto &b : bad) {
    Request *r = std::get<0>(b);
    r->status = std::get<1>(b);
    r->error = std::get<2>(b);
    r->state = kDone;
    r->done_ns = now_ns();
    r->start_ns = t_start;
  }
  S.tickets.clear();
  for (const Item &it : ok) {
    it.r->start_ns = t_start;
    it.r->batch_requests = (int32_t)ok.size();
  }
  for (uint64_t tk : take) {
    Request &r = reqs[tk];
    if (r.state == kRunning) S.tickets.push_back(tk);
  }
  if (e != hipSuccess) {
    for (uint64_t tk : S.tickets) {
      Request &r = reqs[tk];
      r.state = kDone;
      r.status = PSEG_E_HIP;
      r.error = fmt("serving launch: %s", hipGetErrorString(e));
      r.done_ns = now_ns();
    }
    S.tickets.clear();
    S.busy = false;
  } else {
    S.busy = !S.tickets.empty();
    if (!S.busy) (void)hipEventSynchronize(S.done);
  }
  if (!bad.empty()) done_cv.notify_all();
  return true;
}

void pseg_server::run() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  const int inflight = std::min(std::max(cfg.max_inflight, 1), kSlots);
  for (;;) {
    retire_finished(lk);
    int busy = 0;
    for (int i = 0; i < inflight; i++) busy += slots[i].busy;
    if (stop && queue.empty() && busy == 0) break;
    bool launched = false;
    for (int i = 0; i < inflight && !queue.empty(); i++) {
      if (slots[i].busy) continue;
      launched = launch_batch(lk, i);
      break;
    }
    if (launched) continue;
    if (busy > 0) {
      // A launch is running: poll its event without the lock (slots are the
      // worker's alone) until one finishes, or new work arrives while a slot
      // is free.  A timed sleep would add the kernel's timer slack (about
      // 50 us) to every result; the spin lasts only while a launch is on the
      // device, and submitters and waiters never wait on it for the lock.
      const uint64_t seen = wake_seq.load(std::memory_order_acquire);
      lk.unlock();
      for (;;) {
        bool finished = false;
        for (int i = 0; i < inflight && !finished; i++)
          finished = slots[i].busy && hipEventQuery(slots[i].done) != hipErrorNotReady;
        if (finished) break;
        if (busy < inflight && wake_seq.load(std::memory_order_acquire) != seen) break;
        std::this_thread::yield();
      }
      lk.lock();
    } else {
      work_cv.wait_for(lk, std::chrono::milliseconds(50));
    }
  }
}

extern "C" {

pseg_status pseg_server_create(const pseg_server_config *cfg, pseg_server **out) {
  if (!out) return set_error(PSEG_E_INVALID, "out is NULL");
  *out = nullptr;
  if (!cfg) return set_error(PSEG_E_INVALID, "cfg is NULL");
  const pseg_server_config &c = *cfg;
  if (c.dtype != PSEG_F32 && c.dtype != PSEG_F64)
    return set_error(PSEG_E_INVALID, fmt("dtype must be PSEG_F32 or PSEG_F64, got %d", c.dtype));
  if (c.traces < 1 || c.samples < 1 || c.pitch < c.samples)
    return set_error(PSEG_E_INVALID, fmt("a survey needs traces >= 1, samples >= 1 and pitch >= "
                                         "samples, got %d, %d, %lld", c.traces, c.samples,
                                         (long long)c.pitch));
  if (!c.survey || !c.trace_len) return set_error(PSEG_E_INVALID, "NULL survey or trace_len");
  if (!(c.dt_ms > 0) || !std::isfinite(c.dt_ms))
    return set_error(PSEG_E_INVALID, fmt("dtMs must be finite and positive, got %g", c.dt_ms));
  if (c.arena_bytes < 256) return set_error(PSEG_E_INVALID, "arena_bytes must be >= 256");
  pseg_server *S = new pseg_server;
  S->cfg = c;
  S->esz = c.dtype == PSEG_F32 ? 4 : 8;
  hipError_t e = hipGetDevice(&S->device);
  if (e == hipSuccess) e = hipMalloc((void **)&S->arena_mem, c.arena_bytes);
  for (int i = 0; i < kSlots && e == hipSuccess; i++) {
    Slot &sl = S->slots[i];
    e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e == hipSuccess)
      e = hipHostMalloc((void **)&sl.host_jobs, kMaxJobs * sizeof(Job), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void **)&sl.dev, kMaxJobs * sizeof(Job));
  }
  if (e != hipSuccess) {
    pseg_status st = hip_error(e, "server set-up");
    for (Slot &sl : S->slots) {
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.host_jobs) (void)hipHostFree(sl.host_jobs);
      if (sl.dev) (void)hipFree(sl.dev);
    }
    if (S->arena_mem) (void)hipFree(S->arena_mem);
    delete S;
    return st;
  }
  S->arena.reset(S->arena_mem, c.arena_bytes);
  S->worker = std::thread([S] { S->run(); });
  *out = S;
  return PSEG_OK;
}

pseg_status pseg_server_add_flow(pseg_server *S, const pseg_op *ops, int32_t n_ops,
                                 int32_t *flow) {
  if (!S || !flow) return set_error(PSEG_E_INVALID, "NULL argument");
  if (n_ops < 0 || n_ops > PSEG_MAX_OPS || (n_ops > 0 && !ops))
    return set_error(PSEG_E_INVALID, fmt("a flow has 0 .. %d ops", PSEG_MAX_OPS));
  for (int i = 0; i < n_ops; i++) {
    if (ops[i].kind == PSEG_OP_NORMALIZE && ops[i].scope == 1)
      return set_error(PSEG_E_UNSUPPORTED,
                       fmt("op %d (normalize): a per-panel normalization is served by "
                           "pseg_plan_apply, not by the batching server", i));
    if (ops[i].kind == PSEG_OP_CALIBRATE || ops[i].kind == PSEG_OP_UNCALIBRATE)
      return set_error(PSEG_E_UNSUPPORTED,
                       fmt("op %d: a calibration is applied by pseg_plan_apply_calib, not by "
                           "the batching server", i));
  }
  int32_t m = 0;
  pseg_status st = pseg_required_margin(ops, n_ops, S->cfg.dt_ms, &m);
  if (st != PSEG_OK) return st;
  // compiling once at the survey's width checks every op before any request
  pseg_plan *probe = nullptr;
  st = pseg_plan_compile(ops, n_ops, S->cfg.dtype, S->cfg.dt_ms, S->cfg.samples, S->cfg.samples,
                         &probe);
  if (st != PSEG_OK) return st;
  std::lock_guard<std::mutex> g(S->plan_mu);
  Flow f;
  f.ops.assign(ops, ops + n_ops);
  f.margin = m;
  S->flows.push_back(f);
  *flow = (int32_t)S->flows.size() - 1;
  S->plans[std::make_tuple((int)*flow, (int)S->cfg.samples, (int)S->cfg.samples)] = probe;
  return PSEG_OK;
}

pseg_status pseg_server_submit(pseg_server *S, int32_t flow, int32_t trace0, int32_t ntraces,
                               int32_t shown_first, int32_t shown_last, uint64_t *ticket) {
  if (!S || !ticket) return set_error(PSEG_E_INVALID, "NULL argument");
  const pseg_server_config &c = S->cfg;
  if (ntraces < 1 || trace0 < 0 || (int64_t)trace0 + ntraces > c.traces)
    return set_error(PSEG_E_INVALID, fmt("traces %d .. %lld are outside the survey's 0 .. %d",
                                         trace0, (long long)trace0 + ntraces - 1, c.traces - 1));
  if (shown_first < 0 || shown_last < shown_first || shown_last >= c.samples)
    return set_error(PSEG_E_INVALID, fmt("samples %d .. %d are not a window of 0 .. %d",
                                         shown_first, shown_last, c.samples - 1));
  Request r;
  {
    std::lock_guard<std::mutex> g(S->plan_mu);
    if (flow < 0 || flow >= (int32_t)S->flows.size())
      return set_error(PSEG_E_INVALID, fmt("no flow %d", flow));
    r.flow = flow;
    int32_t first = 0, last = 0;
    pseg_decode_window(shown_first, shown_last, S->flows[flow].margin, c.samples, &first, &last);
    r.first = first;
    r.window = last - first + 1;
  }
  if (pipeline_threads_for(r.window) == 0)
    return set_error(PSEG_E_UNSUPPORTED, fmt("a window of %d samples is beyond this GPU path",
                                             r.window));
  r.trace0 = trace0;
  r.ntraces = ntraces;
  r.shown_first = shown_first;
  r.shown_last = shown_last;
  r.out_bytes = (size_t)ntraces * (size_t)(shown_last - shown_first + 1) * S->esz;
  if ((r.out_bytes + 255) / 256 * 256 > c.arena_bytes)
    return set_error(PSEG_E_WORKSPACE, fmt("a result of %zu bytes does not fit the server's "
                                           "arena of %zu", r.out_bytes, c.arena_bytes));
  r.submit_ns = now_ns();
  std::lock_guard<std::mutex> g(S->mu);
  if (S->stop) return set_error(PSEG_E_INVALID, "the server is stopping");
  const uint64_t t = S->next_ticket++;
  S->reqs.emplace(t, std::move(r));
  S->queue.push_back(t);
  S->info.requests++;
  S->wake_seq.fetch_add(1, std::memory_order_release);
  S->work_cv.notify_one();
  *ticket = t;
  return PSEG_OK;
}

pseg_status pseg_server_wait(pseg_server *S, uint64_t ticket, int32_t timeout_ms,
                             pseg_serve_result *out) {
  if (!S || !out) return set_error(PSEG_E_INVALID, "NULL argument");
  std::unique_lock<std::mutex> lk(S->mu);
  auto done = [&] {
    auto it = S->reqs.find(ticket);
    return it == S->reqs.end() || it->second.state == kDone;
  };
  if (timeout_ms < 0) S->done_cv.wait(lk, done);
  else S->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), done);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return set_error(PSEG_E_INVALID, fmt("no request %llu", (unsigned long long)ticket));
  const Request &r = it->second;
  if (r.state != kDone) return set_error(PSEG_E_BUSY, "the request is not done");
  memset(out, 0, sizeof *out);
  out->status = r.status;
  out->rows = r.ntraces;
  out->cols = r.shown_last - r.shown_first + 1;
  out->first = r.first;
  out->window = r.window;
  out->batch_requests = r.batch_requests;
  out->plan_compiled = r.plan_compiled;
  out->samples = r.status == PSEG_OK ? r.out : nullptr;
  out->generation = r.plan ? r.plan->generation : 0;
  out->submit_ns = r.submit_ns;
  out->start_ns = r.start_ns;
  out->done_ns = r.done_ns;
  if (r.status != PSEG_OK) set_error(r.status, r.error);
  return PSEG_OK;
}

pseg_status pseg_server_read(pseg_server *S, uint64_t ticket, void *dst, void *stream) {
  if (!S || !dst) return set_error(PSEG_E_INVALID, "NULL argument");
  const char *src = nullptr;
  size_t bytes = 0;
  {
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end() || it->second.state != kDone || it->second.status != PSEG_OK)
      return set_error(PSEG_E_INVALID, "no finished request with that ticket");
    src = it->second.out;
    bytes = it->second.out_bytes;
  }
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
  if (e != hipSuccess) return hip_error(e, "reading a result");
  return PSEG_OK;
}

pseg_status pseg_server_release(pseg_server *S, uint64_t ticket) {
  if (!S) return set_error(PSEG_E_INVALID, "NULL argument");
  std::lock_guard<std::mutex> g(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return set_error(PSEG_E_INVALID, "no request with that ticket");
  if (it->second.state != kDone) return set_error(PSEG_E_BUSY, "the request is not done");
  if (it->second.out) S->arena.release(it->second.out);
  S->reqs.erase(it);
  S->wake_seq.fetch_add(1, std::memory_order_release);
  S->work_cv.notify_one();  // room for queued requests
  return PSEG_OK;
}

pseg_status pseg_server_info_get(pseg_server *S, pseg_server_info *out) {
  if (!S || !out) return set_error(PSEG_E_INVALID, "NULL argument");
  std::lock_guard<std::mutex> g(S->mu);
  *out = S->info;
  out->arena_peak_bytes = (int64_t)S->arena.peak();
  return PSEG_OK;
}

pseg_status pseg_server_destroy(pseg_server *S) {
  if (!S) return PSEG_OK;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
    // requests that cannot be admitted (the arena is full of unreleased
    // results) would never run: fail them instead of waiting
    S->work_cv.notify_all();
  }
  // results still held by callers keep their arena room; queued requests
  // that do not fit are dropped by the worker's exit condition below
  {
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
      bool any_running = false;
      for (auto &kv : S->reqs) any_running |= kv.second.state == kRunning;
      if (!any_running) break;
      S->done_cv.wait_for(lk, std::chrono::milliseconds(5));
    }
    for (uint64_t tk : S->queue) {
      Request &r = S->reqs[tk];
      r.state = kDone;
      r.status = PSEG_E_INVALID;
      r.error = "the server stopped before the request ran";
    }
    S->queue.clear();
  }
  S->worker.join();
  for (auto &kv : S->plans) pseg_plan_destroy(kv.second);
  for (Slot &sl : S->slots) {
    (void)hipStreamSynchronize(sl.stream);
    (void)hipStreamDestroy(sl.stream);
    (void)hipEventDestroy(sl.done);
    (void)hipHostFree(sl.host_jobs);
    (void)hipFree(sl.dev);
  }
  (void)hipFree(S->arena_mem);
  delete S;
  return PSEG_OK;
}

}  // extern "C"
