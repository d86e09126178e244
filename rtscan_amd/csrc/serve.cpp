// serve.cpp -- request loop over resident tables: HBM-budgeted admission of result bitmaps
// and continuous batching of the admitted requests across several streams.
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "admission.h"
#include "rts_serve.h"

namespace {

using Clock = std::chrono::steady_clock;

inline uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count();
}

// The result pool: one device allocation; which ranges are free is the admission's
// business (host only).
struct DeviceRegion {
  char *base = nullptr;
  bool init(size_t bytes) { return hipMalloc((void **)&base, bytes) == hipSuccess; }
  void destroy() {
    if (base) (void)hipFree(base);
    base = nullptr;
  }
  BITS *at(size_t off) const { return reinterpret_cast<BITS *>(base + off); }
};

// Completion tracking of one rts_server_replay call: its requests count here when they are
// done or failed, so a replay returns only when none of them can still touch its arrays.
struct ReplayState {
  uint64_t finished = 0;
  hipError_t error = hipSuccess;
};

struct Table {
  std::vector<const CODE *> columns;
  std::vector<const BinDex *> indexes;
  int64_t n = 0;
  size_t bitmap_bytes = 0;
};

struct Request {
  int table = 0;
  rts_query_spec spec{};
  uint64_t arrival_ns = 0, done_ns = 0;
  int state = RTS_REQ_QUEUED;
  size_t off = SIZE_MAX, bytes = 0;
  unsigned long long count = 0;
  // replay mode: write the outcome here and free the bitmap at completion
  bool auto_release = false;
  uint64_t *lat_out = nullptr;
  unsigned long long *count_out = nullptr;
  ReplayState *replay = nullptr;
};

struct Worker {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<rts_batch_plan *> plans;  // per table
  unsigned long long *dcounts = nullptr;
  unsigned long long *hcounts = nullptr;  // pinned
  std::vector<uint64_t> batch;            // tickets in flight
  bool busy = false;
};

}  // namespace

struct rts_server {
  rts_server_config cfg{};
  std::vector<Table> tables;
  DeviceRegion region;
  rts::Admission adm;
  std::vector<Worker> workers;
  std::unordered_map<uint64_t, Request> reqs;
  uint64_t next_ticket = 1;
  mutable std::mutex mu;
  std::condition_variable cv_loop, cv_done;
  std::thread loop;
  std::atomic<bool> running{false};
  std::atomic<uint64_t> submitted{0}, completed{0}, batches{0};
  int max_batch_seen = 0;
  hipError_t loop_error = hipSuccess;
  int device = 0;
};

namespace {

void complete_batch(rts_server *s, Worker &w, uint64_t t_done) {
  for (size_t i = 0; i < w.batch.size(); i++) {
    auto it = s->reqs.find(w.batch[i]);
    if (it == s->reqs.end()) continue;
    Request &r = it->second;
    r.count = w.hcounts[i];
    r.done_ns = t_done;
    r.state = RTS_REQ_DONE;
    if (r.auto_release) {
      if (r.lat_out) *r.lat_out = t_done - r.arrival_ns;
      if (r.count_out) *r.count_out = r.count;
      if (r.replay) r.replay->finished++;
      s->adm.release(r.off, r.bytes);
      s->reqs.erase(it);
    }
    s->completed++;
  }
  w.batch.clear();
  w.busy = false;
}

// Requests whose batch failed: their bitmaps return to the pool once the worker's stream
// has drained (a kernel launched before the failure may still write them), replayed ones
// are counted and dropped (so a replay ends), the others stay for rts_server_wait to report.
void fail_requests(rts_server *s, Worker &w, const std::vector<uint64_t> &tickets, hipError_t e) {
  (void)hipStreamSynchronize(w.stream);
  s->loop_error = e;
  for (uint64_t t : tickets) {
    auto it = s->reqs.find(t);
    if (it == s->reqs.end()) continue;
    Request &r = it->second;
    r.state = RTS_REQ_FAILED;
    if (r.off != SIZE_MAX) s->adm.release(r.off, r.bytes);
    r.off = SIZE_MAX;
    if (r.auto_release) {
      if (r.replay) {
        r.replay->finished++;
        r.replay->error = e;
      }
      s->reqs.erase(it);
    }
    s->completed++;
  }
  w.batch.clear();
  w.busy = false;
  s->cv_done.notify_all();
}

// Admit the oldest waiting request's table's requests (in arrival order) into one batch on
// worker w.  Returns true when a batch was launched.
bool launch_batch(rts_server *s, Worker &w) {
  if (!s->adm.waiting()) return false;
  const int table = s->adm.front_table();
  const std::vector<rts::Admission::Admitted> got =
      s->adm.admit(s->cfg.max_batch, [s](int t) { return s->tables[t].bitmap_bytes; });
  if (got.empty()) return false;
  const Table &tb = s->tables[table];
  std::vector<uint64_t> admitted;
  std::vector<rts_query_spec> specs;
  std::vector<BITS *> outs;
  for (const auto &a : got) {
    Request &r = s->reqs.at(a.ticket);
    r.off = a.offset;
    r.bytes = tb.bitmap_bytes;
    r.state = RTS_REQ_RUNNING;
    admitted.push_back(a.ticket);
    specs.push_back(r.spec);
    outs.push_back(s->region.at(a.offset));
  }
  rts_batch_plan *plan = w.plans[table];
  hipError_t e = rts_batch_plan_set_queries_out(plan, specs.data(), outs.data(), (int)specs.size(), w.stream);
  if (e == hipSuccess) e = rts_batch_plan_launch(plan, nullptr, 0, w.dcounts, w.stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(w.hcounts, w.dcounts, specs.size() * sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, w.stream);
  if (e == hipSuccess) e = hipEventRecord(w.done, w.stream);
  if (e != hipSuccess) {
    fail_requests(s, w, admitted, e);
    return false;
  }
  w.batch = admitted;
  w.busy = true;
  s->batches++;
  s->max_batch_seen = std::max(s->max_batch_seen, (int)admitted.size());
  return true;
}

void loop_main(rts_server *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lock(s->mu);
  while (s->running) {
    bool progress = false;
    for (Worker &w : s->workers) {
      if (!w.busy) continue;
      const hipError_t q = hipEventQuery(w.done);
      if (q == hipSuccess) {
        complete_batch(s, w, now_ns());
        progress = true;
      } else if (q != hipErrorNotReady) {  // the batch's work failed: fail its requests
        const std::vector<uint64_t> tickets = w.batch;
        fail_requests(s, w, tickets, q);
        progress = true;
      }
    }
    if (progress) s->cv_done.notify_all();
    for (Worker &w : s->workers)
      if (!w.busy && launch_batch(s, w)) progress = true;
    bool any_busy = false;
    for (Worker &w : s->workers) any_busy = any_busy || w.busy;
    if (progress) continue;
    if (any_busy) {
      // completions are polled: release the lock briefly
      lock.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(5));
      lock.lock();
    } else {
      s->cv_loop.wait_for(lock, std::chrono::milliseconds(2));
    }
  }
}

}  // namespace

extern "C" {

void rts_server_default_config(rts_server_config *cfg) {
  cfg->max_batch = 256;
  cfg->streams = 3;
  cfg->pool_bytes = (size_t)8 << 30;
  cfg->queue_capacity = 1 << 20;
  rts_batch_default_config(&cfg->batch);
}

hipError_t rts_server_create(rts_server **out, const rts_server_config *cfg) {
  if (!out) return hipErrorInvalidValue;
  auto *s = new rts_server();
  if (cfg) s->cfg = *cfg; else rts_server_default_config(&s->cfg);
  if (s->cfg.max_batch < 1 || s->cfg.streams < 1 || s->cfg.queue_capacity < 1) {
    delete s;
    return hipErrorInvalidValue;
  }
  (void)hipGetDevice(&s->device);
  const size_t pool = s->cfg.pool_bytes / rts::PoolAllocator::kGranule * rts::PoolAllocator::kGranule;
  if (pool == 0 || !s->region.init(pool)) {
    delete s;
    return hipErrorOutOfMemory;
  }
  s->adm.reset(pool);
  s->workers.resize(s->cfg.streams);
  for (Worker &w : s->workers) {
    hipError_t e = hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&w.done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc((void **)&w.dcounts, s->cfg.max_batch * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc((void **)&w.hcounts, s->cfg.max_batch * sizeof(unsigned long long), 0);
    if (e != hipSuccess) {
      rts_server_destroy(s);
      return e;
    }
  }
  *out = s;
  return hipSuccess;
}

hipError_t rts_server_add_table(rts_server *s, const CODE *const *columns, const BinDex *const *indexes,
                                int ncols, int64_t n, int *table_id) {
  if (!s || !columns || ncols < 1 || n < 1 || s->running) return hipErrorInvalidValue;
  Table t;
  t.columns.assign(columns, columns + ncols);
  t.indexes.assign(ncols, nullptr);
  for (int c = 0; indexes && c < ncols; c++) t.indexes[c] = indexes[c];
  t.n = n;
  t.bitmap_bytes = (size_t)((n + 31) / 32) * 4;
  for (size_t wi = 0; wi < s->workers.size(); wi++) {
    rts_batch_plan *p = nullptr;
    hipError_t e = rts_batch_plan_create(&p, t.columns.data(), t.indexes.data(), ncols, n,
                                         s->cfg.max_batch, &s->cfg.batch, s->workers[wi].stream);
    if (e != hipSuccess) {  // keep every worker's plan list aligned with the table list
      for (size_t k = 0; k < wi; k++) {
        rts_batch_plan_free(s->workers[k].plans.back());
        s->workers[k].plans.pop_back();
      }
      return e;
    }
    s->workers[wi].plans.push_back(p);
  }
  s->tables.push_back(std::move(t));
  if (table_id) *table_id = (int)s->tables.size() - 1;
  return hipSuccess;
}

hipError_t rts_server_start(rts_server *s) {
  if (!s || s->running) return hipErrorInvalidValue;
  s->running = true;
  s->loop = std::thread(loop_main, s);
  return hipSuccess;
}

static hipError_t submit_at(rts_server *s, int table, const rts_query_spec *q, uint64_t arrival,
                            uint64_t *lat_out, unsigned long long *count_out, ReplayState *replay,
                            uint64_t *ticket) {
  if (!s || !q || table < 0 || table >= (int)s->tables.size() || q->ncols < 1 ||
      q->ncols > RTS_BATCH_MAX_COLS)
    return hipErrorInvalidValue;
  for (int c = 0; c < q->ncols; c++)
    if (q->column[c] < 0 || q->column[c] >= (int)s->tables[table].columns.size()) return hipErrorInvalidValue;
  if (s->tables[table].bitmap_bytes > s->adm.pool().capacity()) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(s->mu);
  if ((int)s->adm.waiting() >= s->cfg.queue_capacity) return hipErrorNotReady;
  const uint64_t t = s->next_ticket++;
  Request &r = s->reqs[t];
  r.table = table;
  r.spec = *q;
  r.arrival_ns = arrival;
  r.auto_release = replay != nullptr;
  r.lat_out = lat_out;
  r.count_out = count_out;
  r.replay = replay;
  s->adm.enqueue(t, table);
  s->submitted++;
  if (ticket) *ticket = t;
  s->cv_loop.notify_one();
  return hipSuccess;
}

hipError_t rts_server_submit(rts_server *s, int table_id, const rts_query_spec *query, uint64_t *ticket) {
  return submit_at(s, table_id, query, now_ns(), nullptr, nullptr, nullptr, ticket);
}

int rts_server_wait(rts_server *s, uint64_t ticket, int timeout_ms, const BITS **bitmap,
                    unsigned long long *count, uint64_t *latency_ns) {
  if (!s) return RTS_REQ_FAILED;
  std::unique_lock<std::mutex> lock(s->mu);
  auto done = [&] {
    auto it = s->reqs.find(ticket);
    return it == s->reqs.end() || it->second.state >= RTS_REQ_DONE;
  };
  if (timeout_ms < 0)
    s->cv_done.wait(lock, done);
  else
    s->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), done);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return RTS_REQ_FAILED;
  const Request &r = it->second;
  if (r.state == RTS_REQ_DONE) {
    if (bitmap) *bitmap = s->region.at(r.off);
    if (count) *count = r.count;
    if (latency_ns) *latency_ns = r.done_ns - r.arrival_ns;
  }
  return r.state;
}

void rts_server_release(rts_server *s, uint64_t ticket) {
  if (!s) return;
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end() || it->second.state == RTS_REQ_RUNNING) return;
  if (it->second.off != SIZE_MAX) s->adm.release(it->second.off, it->second.bytes);
  if (it->second.state == RTS_REQ_QUEUED) s->adm.remove(ticket);
  s->reqs.erase(it);
  s->cv_loop.notify_one();
}

void rts_server_stats(const rts_server *s, uint64_t *submitted, uint64_t *completed, uint64_t *batches,
                      int *max_batch_seen, size_t *pool_in_use, size_t *pool_peak) {
  if (!s) return;
  std::lock_guard<std::mutex> lock(s->mu);
  if (submitted) *submitted = s->submitted;
  if (completed) *completed = s->completed;
  if (batches) *batches = s->batches;
  if (max_batch_seen) *max_batch_seen = s->max_batch_seen;
  if (pool_in_use) *pool_in_use = s->adm.pool().in_use();
  if (pool_peak) *pool_peak = s->adm.pool().peak();
}

hipError_t rts_server_stop(rts_server *s) {
  if (!s) return hipErrorInvalidValue;
  if (s->running) {
    {
      std::lock_guard<std::mutex> lock(s->mu);
      s->running = false;
    }
    s->cv_loop.notify_all();
    if (s->loop.joinable()) s->loop.join();
    s->cv_done.notify_all();
  }
  for (Worker &w : s->workers)
    if (w.stream) (void)hipStreamSynchronize(w.stream);
  return s->loop_error;
}

void rts_server_destroy(rts_server *s) {
  if (!s) return;
  (void)rts_server_stop(s);
  for (Worker &w : s->workers) {
    for (rts_batch_plan *p : w.plans) rts_batch_plan_free(p);
    if (w.dcounts) (void)hipFree(w.dcounts);
    if (w.hcounts) (void)hipHostFree(w.hcounts);
    if (w.done) (void)hipEventDestroy(w.done);
    if (w.stream) (void)hipStreamDestroy(w.stream);
  }
  s->region.destroy();
  delete s;
}

hipError_t rts_server_replay(rts_server *s, const int *table_ids, const rts_query_spec *queries,
                             const uint64_t *arrival_ns, int nq, uint64_t *latency_ns,
                             unsigned long long *counts, double *elapsed_s) {
  if (!s || !s->running || nq < 0 || (nq && (!table_ids || !queries || !arrival_ns || !latency_ns)))
    return hipErrorInvalidValue;
  ReplayState rs;  // this replay's requests only: an earlier failure does not end it early
  const uint64_t t0 = now_ns();
  hipError_t err = hipSuccess;
  int sent = 0;  // requests actually submitted: the wait below counts only those
  std::thread submitter([&] {
    for (int i = 0; i < nq; i++) {
      const uint64_t due = t0 + arrival_ns[i];
      uint64_t t = now_ns();
      if (due > t + 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(due - t - 100000));
      while (now_ns() < due) {
      }
      hipError_t e;
      while ((e = submit_at(s, table_ids[i], &queries[i], due, &latency_ns[i], counts ? &counts[i] : nullptr,
                            &rs, nullptr)) == hipErrorNotReady)
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      if (e != hipSuccess) {
        err = e;
        return;
      }
      sent++;
    }
  });
  submitter.join();
  {
    // every submitted request must be done or failed: until then the loop may still write
    // latency_ns / counts, which belong to the caller
    std::unique_lock<std::mutex> lock(s->mu);
    s->cv_done.wait(lock, [&] { return rs.finished >= (uint64_t)sent || !s->running; });
    if (rs.finished < (uint64_t)sent) {
      // the server was stopped under this replay: its loop no longer runs, so drop the
      // replay's remaining requests (their batches drained first) so that nothing refers
      // to this call's state after it returns
      for (Worker &w : s->workers) (void)hipStreamSynchronize(w.stream);
      for (auto it = s->reqs.begin(); it != s->reqs.end();) {
        Request &r = it->second;
        if (r.replay != &rs) {
          ++it;
          continue;
        }
        if (r.state == RTS_REQ_QUEUED) s->adm.remove(it->first);
        if (r.off != SIZE_MAX) s->adm.release(r.off, r.bytes);
        it = s->reqs.erase(it);
      }
      if (rs.error == hipSuccess) rs.error = hipErrorNotReady;
    }
  }
  if (elapsed_s) *elapsed_s = (now_ns() - t0) * 1e-9;
  return err != hipSuccess ? err : rs.error;
}

}  // extern "C"
