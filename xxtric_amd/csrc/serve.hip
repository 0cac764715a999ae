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
// their footprint released, which admits the next requests.
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

// A run hands out task indices from one counter; a thread takes part in a run
// only after registering (under the lock, while the run's task list is
// current), and run() returns once every task is done and every registered
// thread has left, so no thread keeps a run's task list past it.
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
  // Runs every task (the calling thread helps) and returns when all are done.
  void run(const std::vector<CopyTask>& tasks) {
    if (tasks.empty()) return;
    {
      std::lock_guard<std::mutex> lk(mu_);
      tasks_ = &tasks;
      next_.store(0);
      left_ = tasks.size();
      active_ = 1;  // this thread
      gen_++;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return left_ == 0 && active_ == 0; });
    tasks_ = nullptr;
  }

 private:
  void work() {  // the caller is registered in active_
    const std::vector<CopyTask>& ts = *tasks_;
    size_t finished = 0;
    for (size_t i = next_.fetch_add(1); i < ts.size(); i = next_.fetch_add(1)) {
      run_copy(ts[i]);
      finished++;
    }
    std::lock_guard<std::mutex> lk(mu_);
    left_ -= finished;
    active_--;
    if (left_ == 0 && active_ == 0) done_cv_.notify_all();
  }
  void loop() {
    uint64_t seen = 0;
    while (true) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || (gen_ != seen && tasks_ != nullptr); });
        if (stop_) return;
        seen = gen_;
        active_++;
      }
      work();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::vector<CopyTask>* tasks_ = nullptr;
  std::atomic<size_t> next_{0};
  size_t left_ = 0;
  int active_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

struct ServeSlot {
  xt_batch* b = nullptr;
  std::vector<Req*> reqs;
  int64_t V = 0, E = 0;
  bool inflight = false;
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
  int64_t submitted = 0, completed = 0, rejected = 0, batches = 0;
  int worker_err = XT_OK;
  bool stop = false;
  std::unique_ptr<CopyPool> pool;
  std::thread worker;
  std::chrono::steady_clock::time_point t0;

  double now() const {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

void complete_slot(xt_server* s, ServeSlot& sl) {
  std::vector<uint64_t> tot(sl.reqs.size());
  int rc = batch_read(sl.b, tot.data(), (int64_t)tot.size());
  const double t = s->now();
  std::lock_guard<std::mutex> lk(s->mu);
  if (rc != XT_OK) s->worker_err = rc;
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
      if (!sl.inflight) {
        free_slot = &sl;
        break;
      }
    if (free_slot) {
      std::vector<Req*> take;
      {
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->stop && s->queue.empty()) break;
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
          if (large) break;
        }
      }
      // Lay the admitted requests out in the slot's pinned mirror, then copy
      // (and check the borrowed ones) with the copy pool.
      std::vector<CopyTask> tasks;
      std::vector<int64_t*> odst(take.size());
      std::vector<int32_t*> tdst(take.size());
      std::unique_ptr<std::atomic<int>[]> bad(new std::atomic<int>[take.size()]);
      for (size_t i = 0; i < take.size(); i++) {
        Req* r = take[i];
        bad[i].store(0);
        int rc = batch_stage_layout(free_slot->b, r->nv, r->ne, &odst[i], &tdst[i]);
        if (rc != XT_OK) {
          std::lock_guard<std::mutex> lk(s->mu);
          s->worker_err = rc;
          break;
        }
        free_slot->reqs.push_back(r);
        free_slot->V += r->nv;
        free_slot->E += r->ne;
        const bool borrowed = r->b_off != nullptr;
        const int64_t* off = borrowed ? r->b_off : r->off.get();
        const int32_t* tails = borrowed ? r->b_tails : r->tails.get();
        if (borrowed && (off[0] != 0 || off[r->nv] != r->ne)) bad[i].store(1);
        for (int64_t lo = 0; lo <= r->nv; lo += kCopyRange)
          tasks.push_back(CopyTask{off, nullptr, odst[i], nullptr, r->nv, lo, std::min(r->nv + 1, lo + kCopyRange),
                                   borrowed, &bad[i]});
        for (int64_t lo = 0; lo < r->ne; lo += kCopyRange)
          tasks.push_back(CopyTask{nullptr, tails, nullptr, tdst[i], r->nv, lo, std::min(r->ne, lo + kCopyRange),
                                   borrowed, &bad[i]});
      }
      s->pool->run(tasks);
      for (size_t i = 0; i < free_slot->reqs.size() && i < take.size(); i++) {
        Req* r = take[i];
        if (bad[i].load()) {
          // Keep the slot's layout: the graph becomes rows of vertex 0 ids
          // (valid, never reported) and the request fails at poll.
          for (int64_t v = 0; v < r->nv; v++) odst[i][v] = 0;
          odst[i][r->nv] = r->ne;
          if (r->ne > 0) std::memset(tdst[i], 0, (size_t)r->ne * sizeof(int32_t));
          std::lock_guard<std::mutex> lk(s->mu);
          r->err = XT_ERR_INPUT;
          s->rejected++;
        }
        r->off.reset();  // host copies no longer needed
        r->tails.reset();
      }
      if (!free_slot->reqs.empty()) {
        int rc = batch_launch(free_slot->b, XT_BATCH_PERSISTENT, XT_ORIENT_DEGREE);
        std::lock_guard<std::mutex> lk(s->mu);
        if (rc != XT_OK) s->worker_err = rc;
        free_slot->inflight = true;
        s->batches++;
        progressed = true;
      }
    }
    if (!progressed) {
      std::unique_lock<std::mutex> lk(s->mu);
      bool any_inflight = false;
      for (ServeSlot& sl : s->slots) any_inflight |= sl.inflight;
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
    if (s->worker_err != XT_OK) return s->worker_err;
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return XT_ERR_NOTFOUND;
    Req* r = it->second;
    if (r->state != kDone) return XT_PENDING;
    if (r->err != XT_OK) {  // a borrowed request that failed its check
      const int err = r->err;
      s->live.erase(it);
      delete r;
      return set_err(s->ctx, err, "request rejected: adjacency out of range or offsets not monotone");
    }
    if (total) *total = r->total;
    if (latency_us) *latency_us = r->done_us - r->arrival_us;
    s->live.erase(it);
    delete r;
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
