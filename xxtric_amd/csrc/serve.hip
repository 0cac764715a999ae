// serve.hip -- continuous batching of arriving graphs (host runtime, C++).
//
// Requests (host CSRs) are queued by xt_serve_submit.  A worker thread owns
// `max_inflight` batch slots, each an xt_batch with its own HIP stream and
// pre-sized HBM arena.  Whenever a slot is free it admits queued requests in
// arrival order while their HBM footprint fits the budget and the slot's
// arena (below saturation, a request of kLargeRequest entries or more alone),
// uploads them and launches the persistent batch pipeline on the slot's
// stream; slots in flight run concurrently on the device.  Finished slots are
// read back, their requests completed (latency = completion - arrival) and
// their footprint released, which admits the next requests.  A large
// admission is copied into its slot's staging memory by a pool of host
// threads while the worker goes on serving the other slots.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "xt_internal.hpp"

namespace xt {
int batch_launch(xt_batch* b, int mode, uint32_t flags);
int batch_poll(xt_batch* b);
int batch_wait(xt_batch* b);
int batch_read(xt_batch* b, uint64_t* totals, int64_t n);
const char* check_host_graph(int64_t nv, int64_t ne, const int64_t* off, const int32_t* tails,
                             int64_t* off_out, int32_t* tails_out);
int batch_stage_reserve(xt_batch* b);
int batch_readback_reserve(xt_batch* b);
int batch_stage_host(xt_batch* b, int64_t nv, int64_t ne, const int64_t* off, const int32_t* tails);
const char* batch_stage_host_checked(xt_batch* b, int64_t nv, int64_t ne, const int64_t* off,
                                     const int32_t* tails, int* rc);
int batch_stage_layout(xt_batch* b, int64_t nv, int64_t ne, int64_t** off_dst, int32_t** tails_dst);
}  // namespace xt

using namespace xt;

namespace {

// HBM bytes one graph occupies in a slot arena: union + local CSR, vertex map,
// orientation and scan scratch (per vertex) and oriented ids (per entry).
constexpr int64_t kBytesPerVertex = 40;
constexpr int64_t kBytesPerEntry = 12;

enum ReqState : int { kQueued = 0, kRunning = 1, kDone = 2 };

struct Req {
  uint64_t ticket = 0;
  int64_t nv = 0, ne = 0;
  std::unique_ptr<int64_t[]> off;  // the request's copy (not zero-filled first)
  std::unique_ptr<int32_t[]> tails;
  const int64_t* b_off = nullptr;  // borrowed caller arrays (xt_serve_submit_borrowed)
  const int32_t* b_tails = nullptr;
  double arrival_us = 0, done_us = 0;
  uint64_t total = 0;
  int64_t footprint = 0;
  int state = kQueued;
  int err = XT_OK;  // a borrowed request that failed its check
};

// Host threads that copy admitted requests into a slot's pinned staging
// mirror in parallel, ranges of at most kCopyRange elements, checking borrowed
// requests in the same pass (offsets monotone, ids in [0, nv)).
constexpr int64_t kCopyRange = 1 << 22;
// Requests with at least this many directed entries get a batch of their own.
constexpr int64_t kLargeRequest = int64_t(1) << 22;

struct CopyTask {
  const int64_t* off = nullptr;  // offsets [lo, hi) (off != nullptr) or ids [lo, hi)
  const int32_t* tails = nullptr;
  int64_t* off_dst = nullptr;
  int32_t* tails_dst = nullptr;
  int64_t nv = 0, lo = 0, hi = 0;
  bool check = false;
  std::atomic<int>* bad = nullptr;
};

void run_copy(const CopyTask& t) {
  if (t.off) {
    int64_t down = 0;
    for (int64_t i = t.lo; i < t.hi; i++) {
      t.off_dst[i] = t.off[i];
      if (t.check && i < t.nv) down |= (int64_t)(t.off[i + 1] < t.off[i]);
    }
    if (down) t.bad->store(1, std::memory_order_relaxed);
  } else if (t.check) {
    uint32_t out = 0;
    const uint32_t n = (uint32_t)t.nv;
    for (int64_t e = t.lo; e < t.hi; e++) {
      const int32_t v = t.tails[e];
      t.tails_dst[e] = v;
      out |= (uint32_t)((uint32_t)v >= n);
    }
    if (out) t.bad->store(1, std::memory_order_relaxed);
  } else {
    std::memcpy(t.tails_dst + t.lo, t.tails + t.lo, (size_t)(t.hi - t.lo) * sizeof(int32_t));
  }
}

// A copy job: its tasks are handed out from one counter to the pool's
// threads (and to the worker when it helps); `done` counts finished tasks.
struct CopyJob {
  std::vector<CopyTask> tasks;
  std::atomic<size_t> next{0};
  std::atomic<size_t> done{0};
  bool finished() const { return done.load(std::memory_order_acquire) == tasks.size(); }
  // Runs unclaimed tasks in the calling thread; false once none is left.
  bool help() {
    const size_t i = next.fetch_add(1);
    if (i >= tasks.size()) return false;
    run_copy(tasks[i]);
    done.fetch_add(1, std::memory_order_release);
    return true;
  }
};

// Host threads that run copy jobs in submission order.  A job stays listed
// until all its tasks are claimed; its owner keeps it alive until finished().
class CopyPool {
 public:
  explicit CopyPool(int n) {
    for (int i = 0; i < n; i++) th_.emplace_back([this] { loop(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (std::thread& t : th_) t.join();
  }
  void submit(const std::shared_ptr<CopyJob>& job) {
    if (job->tasks.empty()) return;
    {
      std::lock_guard<std::mutex> lk(mu_);
      jobs_.push_back(job);
    }
    cv_.notify_all();
  }
  // Submits the job, helps with it, and returns when every task is done.
  void run(const std::shared_ptr<CopyJob>& job) {
    submit(job);
    while (job->help()) {
    }
    while (!job->finished()) std::this_thread::yield();
  }

 private:
  void loop() {
    while (true) {
      std::shared_ptr<CopyJob> job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
        if (stop_) return;
        job = jobs_.front();
      }
      if (!job->help()) {  // every task of this job is claimed: unlist it
        std::lock_guard<std::mutex> lk(mu_);
        if (!jobs_.empty() && jobs_.front() == job) jobs_.pop_front();
        continue;
      }
      while (job->help()) {
      }
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::shared_ptr<CopyJob>> jobs_;
  bool stop_ = false;
};

// A large request admitted alone (below saturation) with at least this many
// directed entries is copied into its slot's staging mirror by the pool while
// the worker goes on serving the other slots; every other admission is copied
// by the pool with the worker's help before its launch.
constexpr int64_t kAsyncStage = int64_t(1) << 20;

struct ServeSlot {
  xt_batch* b = nullptr;
  std::vector<Req*> reqs;
  int64_t V = 0, E = 0;
  bool inflight = false;
  // staging by the pool: the job, and per request its destination and check flag
  bool staging = false;
  std::shared_ptr<CopyJob> job;
  std::vector<int64_t*> odst;
  std::vector<int32_t*> tdst;
  std::unique_ptr<std::atomic<int>[]> bad;
};

}  // namespace

struct xt_server {
  xt_ctx* ctx = nullptr;
  xt_serve_config cfg = {};
  int64_t slot_v = 0, slot_e = 0, slot_g = 0;
  std::vector<ServeSlot> slots;
  std::mutex mu;
  std::condition_variable cv, cv_done;
  std::deque<Req*> queue;
  std::unordered_map<uint64_t, Req*> live;
  uint64_t next_ticket = 1;
  int64_t resident = 0, peak_resident = 0;
  std::vector<double> latencies;
  double first_arrival = -1, last_done = 0;
  int64_t submitted = 0, completed = 0, rejected = 0, batches = 0, isolated = 0;
  double t_admit = 0, t_stage = 0, t_launch = 0, t_complete = 0, t_device = 0;  // seconds (under mu)
  int worker_err = XT_OK;
  bool stop = false;
  bool paused = false;  // admission held (xt_serve_pause); running batches still finish
  std::unique_ptr<CopyPool> pool;
  std::thread worker;
  std::chrono::steady_clock::time_point t0;

  double now() const {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

void complete_slot(xt_server* s, ServeSlot& sl) {
  const double t_in = s->now();
  std::vector<uint64_t> tot(sl.reqs.size());
  int rc = batch_read(sl.b, tot.data(), (int64_t)tot.size());
  const double t = s->now();
  std::lock_guard<std::mutex> lk(s->mu);
  if (rc != XT_OK) s->worker_err = rc;
  xt_batch_timing bt = {};
  if (xt_batch_timing_get(sl.b, &bt) == XT_OK) s->t_device += bt.run_ms * 1e-3;
  for (size_t i = 0; i < sl.reqs.size(); i++) {
    Req* r = sl.reqs[i];
    r->total = tot[i];
    r->done_us = t;
    r->state = kDone;
    s->resident -= r->footprint;
    if (r->err != XT_OK) continue;  // failed its check: counted as rejected, no latency
    s->latencies.push_back(t - r->arrival_us);
    s->completed++;
    s->last_done = t;
  }
  sl.reqs.clear();
  sl.V = sl.E = 0;
  sl.inflight = false;
  (void)xt_batch_clear(sl.b);
  s->cv_done.notify_all();
  s->t_complete += (s->now() - t_in) * 1e-6;
}

// The slot's requests are in its staging mirror: requests that failed their
// check become harmless graphs (reported at poll), host copies are released,
// and the slot's batch is launched.
void launch_staged(xt_server* s, ServeSlot& sl) {
  for (size_t i = 0; i < sl.reqs.size(); i++) {
    Req* r = sl.reqs[i];
    if (sl.bad[i].load()) {
      // Keep the slot's layout: the graph becomes rows of vertex 0 ids
      // (valid, never reported) and the request fails at poll.
      for (int64_t v = 0; v < r->nv; v++) sl.odst[i][v] = 0;
      sl.odst[i][r->nv] = r->ne;
      if (r->ne > 0) std::memset(sl.tdst[i], 0, (size_t)r->ne * sizeof(int32_t));
      std::lock_guard<std::mutex> lk(s->mu);
      r->err = XT_ERR_INPUT;
      s->rejected++;
    }
    r->off.reset();  // host copies no longer needed
    r->tails.reset();
  }
  sl.staging = false;
  sl.job.reset();
  const double t_in = s->now();
  int rc = batch_launch(sl.b, XT_BATCH_PERSISTENT, XT_ORIENT_DEGREE);
  const double t_out = s->now();
  std::lock_guard<std::mutex> lk(s->mu);
  s->t_launch += (t_out - t_in) * 1e-6;
  if (rc != XT_OK) s->worker_err = rc;
  sl.inflight = true;
  s->batches++;
}

void worker_main(xt_server* s) {
  (void)hipSetDevice(s->ctx->device);
  while (true) {
    bool progressed = false;
    for (ServeSlot& sl : s->slots) {
      if (sl.inflight && batch_poll(sl.b) == 1) {
        complete_slot(s, sl);
        progressed = true;
      }
    }
    // Fill one free slot from the queue (arrival order), then launch it.
    ServeSlot* free_slot = nullptr;
    for (ServeSlot& sl : s->slots)
      if (!sl.inflight && !sl.staging) {
        free_slot = &sl;
        break;
      }
    bool held = false;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      held = s->paused && !s->stop;
    }
    if (free_slot && !held) {
      const double t_admit0 = s->now();
      std::vector<Req*> take;
      bool alone = false;  // one large request admitted by itself (below saturation)
      {
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->stop && s->queue.empty()) {
          // Background copies still read request arrays and write the slots'
          // staging memory: they end before the worker does.
          lk.unlock();
          for (ServeSlot& o : s->slots)
            if (o.staging)
              while (!o.job->finished())
                if (!o.job->help()) std::this_thread::yield();
          break;
        }
        if (!s->queue.empty() && s->cfg.max_wait_us > 0) {
          // batching window: let a short queue fill unless a slot's worth is waiting
          const double wait_until = s->queue.front()->arrival_us + s->cfg.max_wait_us;
          if ((int64_t)s->queue.size() < s->slot_g && s->now() < wait_until && !s->stop) {
            s->cv.wait_for(lk, std::chrono::microseconds(20));
            continue;
          }
        }
        int64_t V = free_slot->V, E = free_slot->E, G = (int64_t)free_slot->reqs.size();
        // Below saturation (less than an eighth of a slot's graphs queued) a
        // large request is counted in a batch of its own, so small graphs never
        // wait for it inside one batch: they go on through the other slots
        // while it runs.  Near saturation everything is batched together, for
        // throughput.
        const bool isolate = (int64_t)s->queue.size() * 8 < s->slot_g;
        while (!s->queue.empty()) {
          Req* r = s->queue.front();
          const bool large = isolate && r->ne >= kLargeRequest;
          if (large && (G > 0 || !take.empty())) break;
          if (G + 1 > s->slot_g || V + r->nv > s->slot_v || E + r->ne > s->slot_e) break;
          if (s->resident + r->footprint > s->cfg.hbm_budget_bytes) break;  // wait for memory
          s->queue.pop_front();
          s->resident += r->footprint;
          s->peak_resident = std::max(s->peak_resident, s->resident);
          r->state = kRunning;
          take.push_back(r);
          V += r->nv;
          E += r->ne;
          G++;
          if (large) {
            s->isolated++;
            alone = true;
            break;
          }
        }
      }
      // Lay the admitted requests out in the slot's pinned mirror, then copy
      // (and check the borrowed ones) with the copy pool: inline for small
      // admissions, in the background for large ones.
      ServeSlot& sl = *free_slot;
      sl.job = std::make_shared<CopyJob>();
      sl.odst.assign(take.size(), nullptr);
      sl.tdst.assign(take.size(), nullptr);
      sl.bad.reset(new std::atomic<int>[take.size()]);
      int64_t entries = 0;
      for (size_t i = 0; i < take.size(); i++) {
        Req* r = take[i];
        sl.bad[i].store(0);
        int rc = batch_stage_layout(sl.b, r->nv, r->ne, &sl.odst[i], &sl.tdst[i]);
        if (rc != XT_OK) {
          // Admission keeps a slot's graphs within its arena, so this is an
          // internal failure: this request and the rest of the admitted ones
          // end with an error and give their footprint back.
          std::lock_guard<std::mutex> lk(s->mu);
          const double t = s->now();
          for (size_t j = i; j < take.size(); j++) {
            Req* q = take[j];
            q->err = XT_ERR_INTERNAL;
            q->state = kDone;
            q->done_us = t;
            q->off.reset();
            q->tails.reset();
            s->resident -= q->footprint;
            s->rejected++;
          }
          take.resize(i);
          s->cv_done.notify_all();
          break;
        }
        sl.reqs.push_back(r);
        sl.V += r->nv;
        sl.E += r->ne;
        entries += r->ne;
        const bool borrowed = r->b_off != nullptr;
        const int64_t* off = borrowed ? r->b_off : r->off.get();
        const int32_t* tails = borrowed ? r->b_tails : r->tails.get();
        if (borrowed && (off[0] != 0 || off[r->nv] != r->ne)) sl.bad[i].store(1);
        for (int64_t lo = 0; lo <= r->nv; lo += kCopyRange)
          sl.job->tasks.push_back(CopyTask{off, nullptr, sl.odst[i], nullptr, r->nv, lo,
                                           std::min(r->nv + 1, lo + kCopyRange), borrowed, &sl.bad[i]});
        for (int64_t lo = 0; lo < r->ne; lo += kCopyRange)
          sl.job->tasks.push_back(CopyTask{nullptr, tails, nullptr, sl.tdst[i], r->nv, lo,
                                           std::min(r->ne, lo + kCopyRange), borrowed, &sl.bad[i]});
      }
      const double t_admit1 = s->now();
      {
        std::lock_guard<std::mutex> lk(s->mu);
        s->t_admit += (t_admit1 - t_admit0) * 1e-6;
      }
      if (!sl.reqs.empty()) {
        if (alone && entries >= kAsyncStage) {
          // Below saturation: the large request is copied in the background
          // while the worker keeps launching small graphs in the other slots.
          s->pool->submit(sl.job);
          sl.staging = true;
        } else {
          // Near saturation everything is batched, and the worker helps the
          // pool copy, for throughput.
          s->pool->run(sl.job);
          {
            std::lock_guard<std::mutex> lk(s->mu);
            s->t_stage += (s->now() - t_admit1) * 1e-6;
          }
          launch_staged(s, sl);
        }
        progressed = true;
      } else {
        sl.job.reset();
      }
    }
    // Slots whose background staging has finished are launched.
    for (ServeSlot& sl : s->slots)
      if (sl.staging && sl.job->finished()) {
        launch_staged(s, sl);
        progressed = true;
      }
    if (!progressed) {
      std::unique_lock<std::mutex> lk(s->mu);
      bool any_inflight = false;
      for (ServeSlot& sl : s->slots) any_inflight |= sl.inflight || sl.staging;
      if (s->stop && s->queue.empty() && !any_inflight) break;
      s->cv.wait_for(lk, std::chrono::microseconds(any_inflight ? 20 : 200));
    }
  }
}

}  // namespace

extern "C" {

int xt_serve_create(xt_ctx* ctx, const xt_serve_config* cfg, xt_server** out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !out) return XT_ERR_ARG;
    *out = nullptr;
    xt_server* s = new xt_server();
    s->ctx = ctx;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.hbm_budget_bytes <= 0) s->cfg.hbm_budget_bytes = 8LL << 30;
    if (s->cfg.max_batch_graphs <= 0) s->cfg.max_batch_graphs = 4096;
    if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 2;
    if (s->cfg.max_queue <= 0) s->cfg.max_queue = 1 << 20;
    if (s->cfg.max_wait_us < 0) s->cfg.max_wait_us = 0;
    // Each slot's arena gets an equal share of the budget.
    const int64_t share = s->cfg.hbm_budget_bytes / s->cfg.max_inflight;
    s->slot_e = std::max<int64_t>(1024, share / (2 * kBytesPerEntry + kBytesPerVertex / 2));
    s->slot_v = std::max<int64_t>(256, s->slot_e / 4);
    s->slot_g = s->cfg.max_batch_graphs;
    s->slots.resize(s->cfg.max_inflight);
    for (ServeSlot& sl : s->slots) {
      int rc = xt_batch_create(ctx, s->slot_v, s->slot_e, s->slot_g, &sl.b);
      if (rc == XT_OK) rc = batch_stage_reserve(sl.b);  // pinned staging now, not on the first request
      if (rc == XT_OK) rc = batch_readback_reserve(sl.b);  // results come back with each launch
      if (rc != XT_OK) {
        for (ServeSlot& o : s->slots)
          if (o.b) xt_batch_destroy(o.b);
        delete s;
        return rc;
      }
    }
    const unsigned hc = std::thread::hardware_concurrency();
    s->pool.reset(new CopyPool((int)std::min(7u, std::max(1u, hc / 2))));
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread(worker_main, s);
    *out = s;
    return XT_OK;
  });
}

int xt_serve_submit(xt_server* s, int64_t nv, int64_t ne, const int64_t* offsets, const int32_t* tails,
                    double arrival_us, uint64_t* ticket) {
  return xt_guarded([&]() -> int {
    if (!s || !ticket || nv < 1 || ne < 0 || !offsets || (ne > 0 && !tails)) return XT_ERR_ARG;
    const int64_t fp = kBytesPerVertex * nv + kBytesPerEntry * ne;
    auto reject = [&](int code, const char* msg) {
      {
        std::lock_guard<std::mutex> lk(s->mu);
        s->submitted++;
        s->rejected++;
      }
      return set_err(s->ctx, code, msg);  // outside the server lock
    };
    if (nv > s->slot_v || ne > s->slot_e || fp > s->cfg.hbm_budget_bytes)
      return reject(XT_ERR_CAPACITY, "request rejected: larger than a batch slot, the HBM budget or the queue");
    // Checked and copied in one pass, in the caller's thread, outside the server lock.
    Req* r = new Req();
    r->nv = nv;
    r->ne = ne;
    r->off.reset(new int64_t[nv + 1]);
    r->tails.reset(new int32_t[ne > 0 ? ne : 1]);
    if (const char* bad = check_host_graph(nv, ne, offsets, tails, r->off.get(), r->tails.get())) {
      delete r;
      return reject(XT_ERR_INPUT, bad);
    }
    std::unique_lock<std::mutex> lk(s->mu);
    if ((int64_t)s->queue.size() >= s->cfg.max_queue) {
      lk.unlock();
      delete r;
      return reject(XT_ERR_CAPACITY, "request rejected: larger than a batch slot, the HBM budget or the queue");
    }
    s->submitted++;
    r->ticket = s->next_ticket++;
    r->arrival_us = arrival_us >= 0 ? arrival_us : s->now();
    r->footprint = fp;
    if (s->first_arrival < 0 || r->arrival_us < s->first_arrival) s->first_arrival = r->arrival_us;
    s->live[r->ticket] = r;
    s->queue.push_back(r);
    *ticket = r->ticket;
    s->cv.notify_all();
    return XT_OK;
  });
}

int xt_serve_submit_borrowed(xt_server* s, int64_t nv, int64_t ne, const int64_t* offsets,
                             const int32_t* tails, double arrival_us, uint64_t* ticket) {
  return xt_guarded([&]() -> int {
    if (!s || !ticket || nv < 1 || ne < 0 || !offsets || (ne > 0 && !tails)) return XT_ERR_ARG;
    const int64_t fp = kBytesPerVertex * nv + kBytesPerEntry * ne;
    std::unique_lock<std::mutex> lk(s->mu);
    s->submitted++;
    if (nv > s->slot_v || ne > s->slot_e || fp > s->cfg.hbm_budget_bytes ||
        (int64_t)s->queue.size() >= s->cfg.max_queue) {
      s->rejected++;
      lk.unlock();
      return set_err(s->ctx, XT_ERR_CAPACITY,
                     "request rejected: larger than a batch slot, the HBM budget or the queue");
    }
    Req* r = new Req();
    r->nv = nv;
    r->ne = ne;
    r->b_off = offsets;
    r->b_tails = tails;
    r->ticket = s->next_ticket++;
    r->arrival_us = arrival_us >= 0 ? arrival_us : s->now();
    r->footprint = fp;
    if (s->first_arrival < 0 || r->arrival_us < s->first_arrival) s->first_arrival = r->arrival_us;
    s->live[r->ticket] = r;
    s->queue.push_back(r);
    *ticket = r->ticket;
    s->cv.notify_all();
    return XT_OK;
  });
}

int xt_serve_poll(xt_server* s, uint64_t ticket, uint64_t* total, double* latency_us) {
  return xt_guarded([&]() -> int {
    if (!s) return XT_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return XT_ERR_NOTFOUND;
    Req* r = it->second;
    // A request that has not finished stays live (and a borrowed one keeps
    // being read) whatever the server-wide state; its own result, once there,
    // is reported and the ticket retired.
    if (r->state != kDone) return s->worker_err != XT_OK ? s->worker_err : XT_PENDING;
    if (r->err != XT_OK) {  // failed its check, or could not be staged
      const int err = r->err;
      s->live.erase(it);
      delete r;
      return set_err(s->ctx, err, err == XT_ERR_INPUT
                                      ? "request rejected: adjacency out of range or offsets not monotone"
                                      : "request failed: it could not be staged into its batch slot");
    }
    if (total) *total = r->total;
    if (latency_us) *latency_us = r->done_us - r->arrival_us;
    s->live.erase(it);
    delete r;
    return XT_OK;
  });
}

int xt_serve_ticket_state(xt_server* s, uint64_t ticket) {
  return xt_guarded([&]() -> int {
    if (!s) return XT_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return XT_ERR_NOTFOUND;
    return it->second->state == kDone ? XT_SERVE_DONE : it->second->state == kRunning ? XT_SERVE_RUNNING
                                                                                        : XT_SERVE_QUEUED;
  });
}

int xt_serve_pause(xt_server* s, int paused) {
  return xt_guarded([&]() -> int {
    if (!s) return XT_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->paused = paused != 0;
    s->cv.notify_all();
    return XT_OK;
  });
}

int xt_serve_replay(xt_server* s, int64_t npool, const int64_t* nv, const int64_t* ne,
                    const int64_t* const* offsets, const int32_t* const* tails, const uint64_t* expected,
                    int64_t nreq, const int64_t* order, const double* arrival_us, double timeout_s,
                    int64_t* mismatches, double* latency_us) {
  return xt_guarded([&]() -> int {
    if (!s || npool < 1 || !nv || !ne || !offsets || !tails || nreq < 0 || (nreq > 0 && (!order || !arrival_us)))
      return XT_ERR_ARG;
    for (int64_t i = 0; i < nreq; i++)
      if (order[i] < 0 || order[i] >= npool) return set_err(s->ctx, XT_ERR_ARG, "replay: pool index out of range");
    std::vector<uint64_t> tk((size_t)nreq, 0);
    // Each request is submitted (borrowed) at its scheduled arrival time on
    // the server clock, which is also the arrival its latency is measured from.
    for (int64_t i = 0; i < nreq; i++) {
      while (s->now() < arrival_us[i]) std::this_thread::yield();
      const int64_t k = order[i];
      const int rc = xt_serve_submit_borrowed(s, nv[k], ne[k], offsets[k], tails[k], arrival_us[i], &tk[i]);
      if (rc != XT_OK && rc != XT_ERR_CAPACITY) return rc;  // rejected requests keep ticket 0
    }
    XT_TRY(xt_serve_drain(s, timeout_s));
    int64_t bad = 0;
    for (int64_t i = 0; i < nreq; i++) {
      if (tk[i] == 0) {
        if (latency_us) latency_us[i] = -1.0;
        continue;
      }
      uint64_t tot = 0;
      double lat = 0;
      const int rc = xt_serve_poll(s, tk[i], &tot, &lat);
      if (rc != XT_OK) return rc;
      if (expected && tot != expected[order[i]]) bad++;
      if (latency_us) latency_us[i] = lat;
    }
    if (mismatches) *mismatches = bad;
    return XT_OK;
  });
}

int xt_serve_drain(xt_server* s, double timeout_s) {
  return xt_guarded([&]() -> int {
    if (!s) return XT_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
      return s->completed + s->rejected >= s->submitted || s->worker_err != XT_OK;
  });
  if (s->worker_err != XT_OK) return s->worker_err;
  return ok ? XT_OK : XT_ERR_STATE;
  });
}

int xt_serve_stats_get(xt_server* s, xt_serve_stats* out) {
  return xt_guarded([&]() -> int {
    if (!s || !out) return XT_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    xt_serve_stats st = {};
    st.submitted = s->submitted;
    st.completed = s->completed;
    st.rejected = s->rejected;
    st.batches = s->batches;
    st.peak_resident_bytes = s->peak_resident;
    st.isolated_batches = s->isolated;
    st.admit_s = s->t_admit;
    st.stage_s = s->t_stage;
    st.launch_s = s->t_launch;
    st.complete_s = s->t_complete;
    st.device_s = s->t_device;
    if (!s->latencies.empty()) {
      std::vector<double> v = s->latencies;
      std::sort(v.begin(), v.end());
      auto q = [&](double p) {
        size_t k = (size_t)std::ceil(p * (double)v.size()) - 1;
        return v[std::min(k, v.size() - 1)];
      };
      st.p50_us = q(0.50);
      st.p99_us = q(0.99);
      double sum = 0;
      for (double x : v) sum += x;
      st.mean_us = sum / (double)v.size();
      st.max_us = v.back();
      const double span = s->last_done - s->first_arrival;
      st.graphs_per_s = span > 0 ? (double)s->completed / (span * 1e-6) : 0.0;
    }
    *out = st;
    return XT_OK;
  });
}

double xt_serve_now_us(const xt_server* s) { return s ? s->now() : 0.0; }

int xt_serve_destroy(xt_server* s) {
  return xt_guarded([&]() -> int {
    if (!s) return XT_ERR_ARG;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
      s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    for (ServeSlot& sl : s->slots) {
      if (sl.b) {
        (void)batch_wait(sl.b);
        xt_batch_destroy(sl.b);
      }
    }
    for (auto& kv : s->live) delete kv.second;
    delete s;
    return XT_OK;
  });
}

}  // extern "C"
