// server.hip — a serving loop over one device (dfr_server_*): queries are
// submitted from any thread and answered through tickets.
//
// Each lane is a worker thread with its own context (stream, scratch pool).
// A lane that is free takes the queries waiting at the head of the queue,
// as many as fit both the batch size and the device-memory budget that is
// not held by queries in flight on other lanes, and runs them as one batch
// (dfr_groupby_batch: one persistent launch). Queries that arrive while
// batches run wait only until a lane frees: with two or more lanes, their
// batches overlap on the device (continuous batching). A query's result is
// what dfr_groupby gives for it alone, whatever batch it ran in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/dfrocm.h"
#include "groupby.h"
#include "runtime.h"

namespace {

// A submitted query with its own copies of the arrays dfr_query points to.
struct Pending {
  int64_t ticket = 0;
  int64_t bytes = 0;  // admission estimate
  int64_t t_submit = 0;
  std::vector<dfr_column *> keys, columns;
  std::vector<dfr_agg> aggs;
  dfr_query q = {};
  void bind() {
    q.keys = keys.data();
    q.columns = columns.data();
    q.aggs = aggs.data();
  }
};

struct Done {
  int status = DFR_OK;
  dfr_result *result = nullptr;
  char *error = nullptr;
  int64_t latency_ns = 0;  // submit to completion
};

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Device bytes a query holds while it runs in a batch: its group table,
// keys and output columns (the batch path keeps nothing else per query).
int64_t estimate_bytes(dfr_context *ctx, const dfr_query *q) {
  Compiled C;
  char *msg = nullptr;
  dfr::Tuning t = ctx->tune;
  t.agg_mode = dfr::AGG_ATOMIC;
  int st = compile_query(ctx, q, C, &msg);
  if (st == DFR_OK) choose_layout(ctx, C, t.hash_slots, 0, &t);
  free(msg);
  if (st) return 0;  // the batch reports the error; it holds nothing
  const dfr::DevPlan &P = C.plan;
  int64_t b = (int64_t)(P.cap * (uint64_t)P.words * 8);
  if (P.mode == dfr::MODE_HASH) b += (int64_t)P.cap * 8;
  const int64_t per_out = 8 + (int64_t)(C.key_dtypes.size() + C.agg_dtypes.size()) * 9;
  return b + (int64_t)P.cap * per_out;
}

}  // namespace

struct dfr_server {
  dfr_context *owner = nullptr;  // the columns' context
  std::vector<dfr_context *> lanes;
  std::vector<std::thread> workers;
  int max_batch = 256;
  int64_t budget = 0;
  std::mutex mu;
  std::condition_variable work_cv, done_cv;
  std::deque<Pending *> queue;
  std::unordered_map<int64_t, Done> done;
  std::unordered_set<int64_t> outstanding;  // submitted, not yet handed over
  int64_t next_ticket = 1;
  int64_t in_flight_bytes = 0;
  int busy_lanes = 0;
  bool stop = false;
  dfr_server_stats stats = {};

  void worker(int lane);
};

void dfr_server::worker(int lane) {
  dfr_context *ctx = lanes[lane];
  hipSetDevice(ctx->device);
  std::vector<Pending *> batch;
  for (;;) {
    int64_t held = 0;
    {
      std::unique_lock<std::mutex> lk(mu);
      work_cv.wait(lk, [&] { return stop || !queue.empty(); });
      if (queue.empty()) return;  // stopping, nothing left
      // admit from the head while the batch and the budget allow; a lane
      // with nothing in flight anywhere always takes at least one query
      while (!queue.empty() && (int)batch.size() < max_batch) {
        Pending *p = queue.front();
        const bool fits = budget <= 0 || in_flight_bytes + held + p->bytes <= budget;
        if (!fits && !(batch.empty() && in_flight_bytes == 0)) break;
        batch.push_back(p);
        held += p->bytes;
        queue.pop_front();
      }
      if (batch.empty()) {
        // the head waits for memory held by other lanes
        done_cv.wait(lk, [&] { return in_flight_bytes == 0; });
        continue;
      }
      in_flight_bytes += held;
      ++busy_lanes;
      stats.batches++;
      stats.max_batch = std::max<int64_t>(stats.max_batch, (int64_t)batch.size());
      stats.concurrent_max = std::max<int64_t>(stats.concurrent_max, busy_lanes);
    }
    const int n = (int)batch.size();
    std::vector<dfr_result *> res(n, nullptr);
    std::vector<int32_t> st(n, DFR_OK);
    std::vector<char *> errs(n, nullptr);
    char *berr = nullptr;
    int bst = DFR_OK;
    // ordered queries, and a batch of one, take the one-query path; the
    // rest run as one batch
    std::vector<dfr_query> qs;
    std::vector<int> at;
    for (int i = 0; i < n; ++i) {
      if (n == 1 || batch[i]->q.maintain_order) {
        st[i] = dfr_groupby(ctx, &batch[i]->q, &res[i], &errs[i]);
      } else {
        qs.push_back(batch[i]->q);
        at.push_back(i);
      }
    }
    if (!qs.empty()) {
      const int m = (int)qs.size();
      std::vector<dfr_result *> r2(m, nullptr);
      std::vector<int32_t> s2(m, DFR_OK);
      std::vector<char *> e2(m, nullptr);
      bst = dfr_groupby_batch(ctx, qs.data(), m, r2.data(), s2.data(), e2.data(), &berr);
      for (int k = 0; k < m; ++k) {
        res[at[k]] = r2[k];
        st[at[k]] = bst == DFR_OK ? s2[k] : bst;
        errs[at[k]] = bst == DFR_OK ? e2[k] : (berr ? strdup(berr) : nullptr);
      }
    }
    // results keep their lane's context alive past dfr_server_destroy
    for (int i = 0; i < n; ++i) context_hold(res[i]);
    const int64_t t_done = now_ns();
    {
      std::lock_guard<std::mutex> lk(mu);
      for (int i = 0; i < n; ++i) {
        Done d;
        d.latency_ns = t_done - batch[i]->t_submit;
        d.status = st[i];
        d.result = res[i];
        d.error = errs[i];
        done[batch[i]->ticket] = d;
        delete batch[i];
      }
      stats.completed += n;
      in_flight_bytes -= held;
      --busy_lanes;
    }
    free(berr);
    batch.clear();
    done_cv.notify_all();
    work_cv.notify_all();
  }
}

extern "C" int dfr_server_create(dfr_context *ctx, int32_t lanes, int32_t max_batch, int64_t budget_bytes,
                                 dfr_server **out, char **error) {
  if (!ctx || !out) return dfr_fail(error, DFR_INVALID, "null argument");
  if (lanes < 1 || lanes > 16) return dfr_fail(error, DFR_INVALID, "lanes must be 1..16");
  if (max_batch < 1) return dfr_fail(error, DFR_INVALID, "max_batch must be positive");
  dfr_server *s = new dfr_server();
  s->owner = ctx;
  s->max_batch = max_batch;
  s->budget = budget_bytes;
  for (int i = 0; i < lanes; ++i) {
    dfr_context *c = nullptr;
    int st = dfr_context_create(ctx->device, 0, &c, error);
    if (st) {
      for (dfr_context *l : s->lanes) dfr_context_destroy(l);
      delete s;
      return st;
    }
    c->tune = ctx->tune;
    c->tune.jit_wait = 0;  // a query never waits for a runtime compile (seconds)
    s->lanes.push_back(c);
  }
  for (int i = 0; i < lanes; ++i) s->workers.emplace_back([s, i] { s->worker(i); });
  *out = s;
  return DFR_OK;
}

extern "C" int dfr_server_submit(dfr_server *s, const dfr_query *q, int64_t *ticket, char **error) {
  if (!s || !q || !ticket) return dfr_fail(error, DFR_INVALID, "null argument");
  if (q->n_keys < 0 || q->n_keys > DFR_MAX_KEYS || q->n_columns < 0 || q->n_columns > DFR_MAX_COLUMNS ||
      q->n_aggs < 0 || q->n_aggs > 64)
    return dfr_fail(error, DFR_UNSUPPORTED, "query shape outside the device path");
  Pending *p = new Pending();
  p->keys.assign(q->keys, q->keys + q->n_keys);
  p->columns.assign(q->columns, q->columns + q->n_columns);
  p->aggs.assign(q->aggs, q->aggs + q->n_aggs);
  p->q = *q;
  p->bind();
  p->t_submit = now_ns();
  p->bytes = estimate_bytes(s->owner, &p->q);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      delete p;
      return dfr_fail(error, DFR_INVALID, "server is stopping");
    }
    p->ticket = s->next_ticket++;
    *ticket = p->ticket;
    s->outstanding.insert(p->ticket);
    s->queue.push_back(p);
    s->stats.submitted++;
  }
  s->work_cv.notify_one();
  return DFR_OK;
}

extern "C" int dfr_server_wait(dfr_server *s, int64_t ticket, int64_t timeout_ms, dfr_result **out,
                               int64_t *latency_ns, char **error) {
  if (!s || !out) return dfr_fail(error, DFR_INVALID, "null argument");
  *out = nullptr;
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->outstanding.count(ticket))
    return dfr_fail(error, DFR_INVALID, "ticket %lld is unknown or already answered", (long long)ticket);
  auto ready = [&] { return s->done.count(ticket) > 0; };
  if (timeout_ms < 0) {
    s->done_cv.wait(lk, ready);
  } else if (!s->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
    return dfr_fail(error, DFR_PENDING, "ticket %lld is not done", (long long)ticket);
  }
  Done d = s->done[ticket];
  s->done.erase(ticket);
  s->outstanding.erase(ticket);
  lk.unlock();
  *out = d.result;
  if (latency_ns) *latency_ns = d.latency_ns;
  if (d.status != DFR_OK) {
    if (error) *error = d.error;
    else free(d.error);
    return d.status;
  }
  free(d.error);
  return DFR_OK;
}

extern "C" int dfr_server_stats_get(dfr_server *s, dfr_server_stats *out) {
  if (!s || !out) return DFR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->queued = (int64_t)s->queue.size();
  out->in_flight_bytes = s->in_flight_bytes;
  return DFR_OK;
}

extern "C" void dfr_server_destroy(dfr_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->work_cv.notify_all();
  s->done_cv.notify_all();
  for (std::thread &t : s->workers) t.join();
  // results never waited for are released; a lane whose results are
  // still held by callers is destroyed when the last of them is released
  for (auto &kv : s->done) {
    dfr_result_release(kv.second.result);
    free(kv.second.error);
  }
  for (dfr_context *c : s->lanes) context_retire(c);
  delete s;
}
