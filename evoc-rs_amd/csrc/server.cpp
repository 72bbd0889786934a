// Serving loop: continuous batching of knn_graph requests on one MI355X.
//
// Requests arrive through evoc_mi355x_submit at any time.  A fixed set of
// slots (one device context = one HIP stream + stream-ordered memory pool
// each, all at the default priority) takes requests off the queue as soon as a
// slot is free and the request's device memory fits the budget the running
// requests leave.  Requests on different slots run concurrently on the device,
// so a small request that arrives while a large one is scoring starts at once
// instead of waiting behind it.
//
// Admission is FIFO with bounded bypass: if the oldest request does not fit,
// younger requests that fit may start, but only MAX_BYPASS times before the
// queue holds everything back for the oldest one.  A request gets its full
// working set when that fits, otherwise at least its minimum (the exact search
// then works in smaller query slabs); the context's budget is set to the grant.
//
// Small requests (at most `batch_rows` rows, exact mode, f32 or int8 rows,
// k + margin <= 64) queue apart: the next free slot takes every queued small
// request of the oldest one's element kind that fits the batch workspace and
// runs them as one batch on its own stream (batch.cpp: one persistent launch
// per pass over all of them, replayed from a captured hipGraph), before any
// regular request.  Small requests that arrive meanwhile form the next batch.
// The workspace is reserved from the budget when the server starts.  Batches
// run on the slots' streams rather than a stream of their own: with a few
// hardware queues per process, one more stream would share a queue with a
// slot and wait behind that slot's kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "evoc_host.h"
#include "evoc_mi355x.h"

namespace {

constexpr int MAX_BYPASS = 16;

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t ticket = 0;
  evoc_knn_input in{};
  evoc_knn_options opts{};
  int64_t tau[3] = {0, 0, 0};
  std::vector<int64_t> trees;
  evoc_knn_output* out = nullptr;
  uint64_t need_full = 0, need_min = 0;
  int bypassed = 0;
  bool done = false;
  bool batched = false;  // in the batch lane
  BatchJob job;          // the batch lane's view of the request
  int status = EVOC_ERR_PENDING;
  std::string error;
  Clock::time_point t_submit;
};

}  // namespace

struct evoc_mi355x_server {
  int device = 0;
  uint64_t budget = 0;
  uint64_t granted = 0;
  int64_t running = 0;
  std::vector<evoc_mi355x_ctx*> slots;
  std::vector<std::thread> workers;
  mutable std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_ticket = 1;
  bool stopping = false;
  // requests beyond one that are queued or running: running searches read it
  // at every query slab and shrink their slabs while it is non-zero
  std::atomic<int> contention{0};
  // batch lane
  BatchEngine* batch = nullptr;
  int64_t batch_rows = 0;  // 0: no batching
  std::deque<Request*> bqueue;
  int64_t batch_running = 0;
  bool batch_busy = false;  // a slot is running a batch (one at a time: one workspace)
  int64_t n_batches = 0, n_batched = 0, max_batch = 0;

  void publish() {
    contention.store(
        (int)std::max<int64_t>(0, (int64_t)(queue.size() + bqueue.size()) + running + batch_running - 1));
  }

  // Takes every queued small request of the oldest one's kind that fits the
  // batch workspace, runs them as one batch on `ctx`'s stream, and completes
  // them; a request whose buffers overflowed goes back to the front of the
  // regular queue (mu held on entry and on return).
  void run_batch(evoc_mi355x_ctx* ctx, std::unique_lock<std::mutex>& lk) {
    const int kind = bqueue.front()->job.kind;
    const size_t room = batch_workspace_bytes(batch);
    size_t used = 0;
    std::vector<Request*> take;
    for (auto it = bqueue.begin(); it != bqueue.end() && (int)take.size() < batch_max_requests(batch);) {
      Request* r = *it;
      if (r->job.kind == kind && used + r->job.bytes <= room) {
        used += r->job.bytes;
        take.push_back(r);
        it = bqueue.erase(it);
      } else {
        ++it;
      }
    }
    batch_busy = true;
    batch_running = (int64_t)take.size();
    publish();
    lk.unlock();
    const Clock::time_point t_start = Clock::now();
    std::vector<BatchJob*> jobs;
    for (Request* r : take) jobs.push_back(&r->job);
    double ms = 0.0;
    batch_run(batch, ctx, jobs, &ms);
    lk.lock();
    batch_busy = false;
    batch_running = 0;
    n_batches++;
    n_batched += (int64_t)take.size();
    max_batch = std::max<int64_t>(max_batch, (int64_t)take.size());
    for (Request* r : take) {
      r->out->queue_seconds = std::chrono::duration<double>(t_start - r->t_submit).count();
      if (r->job.overflow) {
        // the single-request search resolves overflowing buffers: run it alone, next
        r->batched = false;
        queue.push_front(r);
        continue;
      }
      r->status = r->job.status;
      r->error = r->job.error;
      r->done = true;
    }
    publish();
    cv_done.notify_all();
    cv_work.notify_all();
  }

  // next request to start and its grant, or nullptr (mu held)
  Request* pick(uint64_t& grant) {
    const uint64_t avail = budget > granted ? budget - granted : 0;
    for (size_t i = 0; i < queue.size(); i++) {
      Request* r = queue[i];
      if (r->need_min <= avail) {
        grant = std::min(r->need_full, avail);
        // passing the oldest request counts against it
        if (i > 0) queue.front()->bypassed++;
        queue.erase(queue.begin() + (std::ptrdiff_t)i);
        return r;
      }
      if (i == 0 && r->bypassed >= MAX_BYPASS) return nullptr;  // hold the line for it
    }
    return nullptr;
  }

  void work(evoc_mi355x_ctx* ctx) {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      Request* r = nullptr;
      uint64_t grant = 0;
      bool batch_turn = false;
      // small requests first: a batch of them is short and holds no grant
      cv_work.wait(lk, [&] {
        if (stopping) return true;
        if (batch && !batch_busy && !bqueue.empty()) return batch_turn = true;
        return (r = pick(grant)) != nullptr;
      });
      if (stopping) break;
      if (batch_turn) {
        run_batch(ctx, lk);
        continue;
      }
      granted += grant;
      running++;
      publish();
      lk.unlock();
      const double queued = std::chrono::duration<double>(Clock::now() - r->t_submit).count();
      ctx->budget = grant;
      // (a request handed back by the batch lane has waited there as well)
      const double waited = r->out->queue_seconds;
      const int rc = evoc_mi355x_knn_graph_seeded(ctx, &r->in, &r->opts, r->tau,
                                                  r->trees.empty() ? nullptr : r->trees.data(),
                                                  (int64_t)(r->trees.size() / 3), r->out);
      r->out->queue_seconds = std::max(queued, waited);
      // a slot's pool keeps at most a share of the budget cached between
      // requests, so cached blocks and granted bytes together stay bounded
      (void)hipMemPoolTrimTo(ctx->pool, budget / (4 * slots.size()));
      lk.lock();
      granted -= grant;
      running--;
      publish();
      r->status = rc;
      if (rc != EVOC_OK) r->error = ctx->last_error;
      r->done = true;
      cv_done.notify_all();
      cv_work.notify_all();  // memory came back: queued requests may fit now
    }
  }
};

extern "C" {

int evoc_mi355x_server_create(int32_t device, int32_t n_slots, uint64_t hbm_budget_bytes,
                              evoc_mi355x_server** out) {
  if (!out) return EVOC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (n_slots <= 0) n_slots = 4;
  if (n_slots > 32) return EVOC_ERR_INVALID_ARGUMENT;
  auto* s = new (std::nothrow) evoc_mi355x_server;
  if (!s) return EVOC_ERR_OUT_OF_BUDGET;
  s->device = device;
  for (int i = 0; i < n_slots; i++) {
    evoc_mi355x_ctx* c = nullptr;
    const int rc = evoc_mi355x_create(device, 0, &c);
    if (rc != EVOC_OK) {
      for (auto* x : s->slots) evoc_mi355x_destroy(x);
      delete s;
      return rc;
    }
    s->slots.push_back(c);
  }
  // the budget is shared by all slots.  Between requests each slot's pool keeps
  // up to budget / (4 slots) bytes cached (see work()), so granted and cached
  // bytes together reach 1.25 budgets: by default the budget is 90% of what is
  // free now divided by 1.25, so both fit the device
  s->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)((double)s->slots[0]->budget / 1.25);
  for (auto* c : s->slots) {
    c->serving = true;
    c->contention = &s->contention;
    const int rc = enable_staging(c, 32u << 20);
    if (rc != EVOC_OK) {
      for (auto* x : s->slots) evoc_mi355x_destroy(x);
      delete s;
      return rc;
    }
  }
  // the batch lane: a workspace of min(4 GiB, budget / 8) reserved from the
  // budget, unless that is under 256 MiB; EVOC_MI355X_BATCH_ROWS sets the size
  // limit (0: no batching), EVOC_MI355X_BATCH_GRAPH=0 launches the passes
  // instead of replaying their captured graph
  {
    const char* er = std::getenv("EVOC_MI355X_BATCH_ROWS");
    const int64_t rows = er ? std::atoll(er) : 16384;
    const char* eg = std::getenv("EVOC_MI355X_BATCH_GRAPH");
    const bool graph = !(eg && eg[0] == '0');
    const size_t ws = std::min<size_t>((size_t)4 << 30, (size_t)(s->budget / 8));
    if (rows > 0 && ws >= ((size_t)256 << 20)) {
      s->batch = batch_engine_create(device, ws, 64, graph);
      if (s->batch) {
        s->batch_rows = rows;
        s->budget -= ws + ((size_t)64 << 20);
      }
    }
  }
  for (auto* c : s->slots) s->workers.emplace_back([s, c] { s->work(c); });
  *out = s;
  return EVOC_OK;
}

void evoc_mi355x_server_destroy(evoc_mi355x_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    for (auto* q : {&s->queue, &s->bqueue}) {
      for (Request* r : *q) {
        r->status = EVOC_ERR_CANCELLED;
        r->error = "server destroyed before the request started";
        r->done = true;
      }
      q->clear();
    }
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->workers) t.join();
  batch_engine_destroy(s->batch);
  for (auto* c : s->slots) evoc_mi355x_destroy(c);
  delete s;
}

int evoc_mi355x_server_request_bytes(const evoc_mi355x_server* s, const evoc_knn_input* in,
                                     const evoc_knn_options* opts, int32_t out_location, uint64_t* full_bytes,
                                     uint64_t* min_bytes) {
  if (!s || !full_bytes || !min_bytes) return EVOC_ERR_INVALID_ARGUMENT;
  const int rc = evoc_mi355x_request_bytes(s->slots[0], in, opts, out_location, full_bytes, min_bytes);
  if (rc != EVOC_OK) return rc;
  // DevBuf rounding and small buffers of the search: keep 1% + 4 MiB headroom
  *full_bytes += *full_bytes / 100 + (4u << 20);
  *min_bytes += *min_bytes / 100 + (4u << 20);
  return EVOC_OK;
}

int evoc_mi355x_submit(evoc_mi355x_server* s, const evoc_knn_input* in, const evoc_knn_options* opts,
                       const int64_t tau_state[3], const int64_t* tree_states, int64_t n_tree_rows,
                       evoc_knn_output* out, uint64_t* ticket) {
  if (!s || !in || !opts || !tau_state || !out || !ticket) return EVOC_ERR_INVALID_ARGUMENT;
  if (opts->n_neighbors <= 0) return EVOC_ERR_INVALID_NEIGHBORS;
  if (in->n <= 0) return EVOC_ERR_EMPTY_DATA;
  if (in->d <= 0 || in->ld < in->d || !in->data || in->dtype < EVOC_DTYPE_F32 || in->dtype > EVOC_DTYPE_U8)
    return EVOC_ERR_INVALID_ARGUMENT;
  auto r = std::make_unique<Request>();
  r->in = *in;
  r->opts = *opts;
  std::memcpy(r->tau, tau_state, sizeof(r->tau));
  if (tree_states && n_tree_rows > 0) r->trees.assign(tree_states, tree_states + 3 * n_tree_rows);
  r->out = out;
  const int rc = evoc_mi355x_server_request_bytes(s, in, opts, out->location, &r->need_full, &r->need_min);
  if (rc != EVOC_OK) return rc;
  if (r->need_min > s->budget) return EVOC_ERR_OUT_OF_BUDGET;
  out->queue_seconds = 0.0;
  r->job.in = *in;
  r->job.opts = *opts;
  r->job.out = out;
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stopping) return EVOC_ERR_CANCELLED;
  r->batched = s->batch && s->batch_rows > 0 && batch_fit(s->batch, &r->job, s->batch_rows);
  r->ticket = s->next_ticket++;
  *ticket = r->ticket;
  (r->batched ? s->bqueue : s->queue).push_back(r.get());
  s->reqs.emplace(r->ticket, std::move(r));
  s->publish();
  s->cv_work.notify_one();
  return EVOC_OK;
}

int evoc_mi355x_poll(evoc_mi355x_server* s, uint64_t ticket) {
  if (!s) return EVOC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return EVOC_ERR_INVALID_ARGUMENT;
  return it->second->done ? it->second->status : EVOC_ERR_PENDING;
}

int evoc_mi355x_wait(evoc_mi355x_server* s, uint64_t ticket, double timeout_seconds, char* err,
                     int64_t err_len) {
  if (!s) return EVOC_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return EVOC_ERR_INVALID_ARGUMENT;
  Request* r = it->second.get();
  auto ready = [r] { return r->done; };
  if (timeout_seconds < 0) {
    s->cv_done.wait(lk, ready);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_seconds), ready)) {
    return EVOC_ERR_PENDING;
  }
  const int status = r->status;
  if (err && err_len > 0) {
    const size_t m = std::min<size_t>((size_t)err_len - 1, r->error.size());
    std::memcpy(err, r->error.data(), m);
    err[m] = 0;
  }
  s->reqs.erase(it);
  return status;
}

void evoc_mi355x_server_stats(const evoc_mi355x_server* s, int64_t* queued, int64_t* running,
                              uint64_t* granted_bytes, uint64_t* budget_bytes) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(s->mu);
  if (queued) *queued = (int64_t)(s->queue.size() + s->bqueue.size());
  if (running) *running = s->running + s->batch_running;
  if (granted_bytes) *granted_bytes = s->granted;
  if (budget_bytes) *budget_bytes = s->budget;
}

}  // extern "C"

extern "C" int evoc_mi355x_server_set_batching(evoc_mi355x_server* s, int64_t max_rows) {
  if (!s || max_rows < 0) return EVOC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  if (max_rows > 0 && !s->batch) return EVOC_ERR_UNSUPPORTED;
  s->batch_rows = max_rows;
  return EVOC_OK;
}

extern "C" void evoc_mi355x_server_batch_stats(const evoc_mi355x_server* s, int64_t* batches, int64_t* requests,
                                               int64_t* largest, int64_t* graph_replays) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(s->mu);
  if (batches) *batches = s->n_batches;
  if (requests) *requests = s->n_batched;
  if (largest) *largest = s->max_batch;
  // (read after the batch worker has published its batch under the lock)
  if (graph_replays) *graph_replays = s->batch ? batch_graph_replays(s->batch) : 0;
}
