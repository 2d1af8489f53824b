// Serving loop over device-resident tables: queries are submitted from any
// host thread, a scheduler thread forms batches of the queries that are
// waiting (continuous batching: whatever arrived while earlier batches ran),
// admits each batch against an HBM budget, and runs up to max_inflight
// batches at once, each on its own stream, so that their device work
// overlaps. A batch is planned and executed by the batch engine
// (mgb_batch_*); with batching off every query runs alone on the plain path,
// in arrival order (the per-query baseline).
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "groupby_host.hpp"

using namespace mgb;

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0, Clock::time_point t1)
{
  return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

struct Table {
  std::vector<mgb_column> cols;
  int64_t bytes = 0;
};

struct Query {
  int64_t ticket = 0;
  int32_t table = 0;
  int32_t null_policy = 0;
  std::vector<mgb_column> keys;   // the table's columns, resolved at submission
  std::vector<mgb_column> vals;
  std::string group_key;          // table and key columns: the cache key of its group-count estimate
  std::vector<std::vector<mgb_aggregation>> aggs;
  std::string signature;  // queries of one batch share it
  // a second-pass or sort-path aggregation (VARIANCE, STD, M2, ARGMIN,
  // ARGMAX, MEDIAN, QUANTILE, NUNIQUE, NTH_ELEMENT) or null keys grouped under
  // INCLUDE: batches do not serve it, so the query runs alone on the plain path
  bool plain_only = false;
  Clock::time_point arrival;
};

// A batch held until every one of its results has been taken or released.
struct HeldBatch {
  mgb_batch* batch = nullptr;
  int64_t bytes = 0;
  int remaining = 0;
  ~HeldBatch() { mgb_batch_destroy(batch); }
};

struct Done {
  mgb_status status = MGB_SUCCESS;
  std::string error;
  int64_t groups = 0;
  double latency_ms = 0;
  std::shared_ptr<HeldBatch> batch;  // batched result, or
  int32_t index = 0;                 // the query's index in it
  std::vector<mgb_out_column> keys, results;  // plain-path result (owned)
  int64_t bytes = 0;
};

// Device memory of batches and results: chunks taken from the device once
// and kept until the server is destroyed, carved first-fit in 256-byte
// units. A block comes back only when nothing on the device uses it any
// more (a batch is released after its work and the copies of its results
// have completed), so it can be handed out again at once.
class Arena {
 public:
  explicit Arena(int64_t limit) : limit_(limit) {}
  ~Arena()
  {
    for (auto& c : chunks_) (void)hipFree(c.base);
  }
  // One chunk of `bytes`, taken now.
  bool reserve(size_t bytes)
  {
    std::lock_guard<std::mutex> lk(mu_);
    return add_chunk(align_up(bytes, kUnit));
  }
  void* allocate(size_t bytes)
  {
    const size_t n = align_up(std::max<size_t>(bytes, 1), kUnit);
    std::lock_guard<std::mutex> lk(mu_);
    if (void* p = take(n)) return p;
    // a new chunk: the request, and at least the memory held so far (up to
    // 8 GiB), so that chunks are few
    size_t sz = std::max(align_up(n, kGrow), std::min<size_t>(total_, 8ull << 30));
    if ((int64_t)(total_ + sz) > limit_) sz = n;
    // near the limit: chunks with nothing in use go back first (hipFree
    // waits for the device; only under memory pressure)
    if ((int64_t)(total_ + sz) > limit_) drop_idle_chunks();
    if ((int64_t)(total_ + sz) > limit_ || !add_chunk(sz)) return nullptr;
    return take(n);
  }
  void deallocate(void* p)
  {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) return;
    const size_t n = it->second;
    live_.erase(it);
    for (auto& c : chunks_) {
      uint8_t* u = static_cast<uint8_t*>(p);
      if (u < c.base || u >= c.base + c.size) continue;
      size_t off = (size_t)(u - c.base), len = n;
      auto next = c.free.lower_bound(off);
      if (next != c.free.end() && off + len == next->first) {
        len += next->second;
        next = c.free.erase(next);
      }
      if (next != c.free.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
          prev->second += len;
          return;
        }
      }
      c.free.emplace(off, len);
      return;
    }
  }
  int64_t total()
  {
    std::lock_guard<std::mutex> lk(mu_);
    return (int64_t)total_;
  }
  bool in_use()
  {
    std::lock_guard<std::mutex> lk(mu_);
    return !live_.empty();
  }
  void set_limit(int64_t limit)
  {
    std::lock_guard<std::mutex> lk(mu_);
    limit_ = limit;
  }

 private:
  static constexpr size_t kUnit = 256, kGrow = 256ull << 20;
  struct Chunk {
    uint8_t* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free;  // offset -> length
  };
  bool add_chunk(size_t sz)
  {
    void* p = nullptr;
    if (sz == 0 || hipMalloc(&p, sz) != hipSuccess) return false;
    Chunk c;
    c.base = static_cast<uint8_t*>(p);
    c.size = sz;
    c.free.emplace(0, sz);
    chunks_.push_back(std::move(c));
    total_ += sz;
    return true;
  }
  void drop_idle_chunks()
  {
    for (size_t i = 0; i < chunks_.size();) {
      Chunk& c = chunks_[i];
      if (c.free.size() == 1 && c.free.begin()->second == c.size) {
        (void)hipFree(c.base);
        total_ -= c.size;
        chunks_.erase(chunks_.begin() + (std::ptrdiff_t)i);
      } else {
        ++i;
      }
    }
  }
  void* take(size_t n)
  {
    for (auto& c : chunks_)
      for (auto it = c.free.begin(); it != c.free.end(); ++it) {
        if (it->second < n) continue;
        const size_t off = it->first, len = it->second;
        c.free.erase(it);
        if (len > n) c.free.emplace(off + n, len - n);
        void* p = c.base + off;
        live_.emplace(p, n);
        return p;
      }
    return nullptr;
  }
  std::mutex mu_;
  std::vector<Chunk> chunks_;
  std::map<void*, size_t> live_;
  size_t total_ = 0;
  int64_t limit_;
};

void* arena_alloc(void* ctx, size_t bytes, void*) { return static_cast<Arena*>(ctx)->allocate(bytes); }
void arena_free(void* ctx, void* p, size_t, void*) { static_cast<Arena*>(ctx)->deallocate(p); }

struct Slot {
  hipStream_t stream = nullptr;
  bool busy = false;
  std::atomic<bool> finished{false};
  std::shared_ptr<HeldBatch> batch;
  std::vector<Query> queries;
  struct mgb_server* server = nullptr;
  // temporaries of the plain-path calls run on this slot's stream: carved
  // from the server's arena (so they count against the HBM budget) and
  // returned to it only after the slot has waited for its stream
  Arena* arena = nullptr;
  std::vector<void*> deferred;
  mgb_allocator temps{};
  void flush_deferred()
  {
    for (void* p : deferred) arena->deallocate(p);
    deferred.clear();
  }
};

void* slot_temp_alloc(void* ctx, size_t bytes, void*) { return static_cast<Slot*>(ctx)->arena->allocate(bytes); }
void slot_temp_free(void* ctx, void* p, size_t, void*) { static_cast<Slot*>(ctx)->deferred.push_back(p); }

// Released after the query's work and its copies have completed.
void free_out(Arena* arena, std::vector<mgb_out_column>& cols)
{
  for (auto& c : cols) {
    arena->deallocate(c.data);
    arena->deallocate(c.null_mask);
    arena->deallocate(c.child_data);
    arena->deallocate(c.child_null_mask);
    c.data = nullptr;
    c.null_mask = nullptr;
    c.child_data = nullptr;
    c.child_null_mask = nullptr;
  }
}

}  // namespace

struct mgb_server {
  mgb_server_options opt{};
  std::mutex mu;
  std::condition_variable cv;        // scheduler wake-ups
  std::condition_variable done_cv;   // waiters
  bool stop = false;
  std::vector<Table> tables;
  int64_t resident_bytes = 0;
  int64_t held_bytes = 0;            // batches and plain results not yet taken
  std::deque<Query> pending;
  std::map<int64_t, Done> done;
  int64_t next_ticket = 1;
  std::vector<std::unique_ptr<Slot>> slots;
  std::thread scheduler;
  mgb_server_stats stats{};
  std::vector<double> batch_sizes;
  std::map<std::string, int64_t> estimates;  // distinct key rows per (table, key columns); tables are immutable
  bool wait_memory = false;                   // the next query waits for memory to be released
  std::unique_ptr<Arena> arena;               // memory of batches and plain-path results
  mgb_allocator alloc{};                      // over `arena`

  void run();
  void retire(Slot& s);
  bool launch(Slot& s);
  void fail_query(const Query& q, mgb_status st, const std::string& msg);
};

namespace {

// Host callback after a slot's batch (a runtime thread). It takes the
// server's lock only to order its wake-up after the scheduler's predicate
// check; no HIP call is ever made under that lock, so nothing holding it can
// be waiting for this callback.
void slot_done(void* arg)
{
  Slot* s = static_cast<Slot*>(arg);
  s->finished.store(true);
  { std::lock_guard<std::mutex> lk(s->server->mu); }
  s->server->cv.notify_all();
}

std::string signature_of(const Query& q)
{
  // tables with the same column types and null-ness batch together
  std::string sig = std::to_string(q.null_policy) + "|";
  for (const mgb_column& k : q.keys)
    sig += "k" + std::to_string(k.type_id) + (k.null_mask && k.null_count ? "n" : "") + ",";
  for (size_t r = 0; r < q.vals.size(); ++r) {
    const mgb_column& c = q.vals[r];
    sig += "v" + std::to_string(c.type_id) + (c.null_mask && c.null_count ? "n" : "") + ":";
    for (auto& a : q.aggs[r])
      sig += std::to_string(a.kind) + "." + std::to_string(a.ddof) + "." + std::to_string(a.null_policy) + "." +
             std::to_string(a.interpolation) + "." + std::to_string(a.quantile) + ",";
  }
  return sig;
}

}  // namespace

void mgb_server::fail_query(const Query& q, mgb_status st, const std::string& msg)
{
  Done d;
  d.status = st;
  d.error = msg;
  d.latency_ms = ms_since(q.arrival, Clock::now());
  done[q.ticket] = std::move(d);
  ++stats.failed;
}

// Completed slot -> results (the scheduler thread, lock not held).
void mgb_server::retire(Slot& s)
{
  const auto t0 = Clock::now();
  std::vector<Done> out(s.queries.size());
  if (s.batch) {
    mgb_status st = mgb_batch_finish(s.batch->batch, s.stream);
    std::string err = st == MGB_SUCCESS ? "" : mgb_last_error_message();
    mgb_batch_stats bs{};
    (void)mgb_batch_get_stats(s.batch->batch, &bs);
    const auto now = Clock::now();
    for (size_t i = 0; i < s.queries.size(); ++i) {
      out[i].status = st;
      out[i].error = err;
      out[i].latency_ms = ms_since(s.queries[i].arrival, now);
      out[i].batch = s.batch;
      out[i].index = (int32_t)i;
      if (st == MGB_SUCCESS) (void)mgb_batch_result(s.batch->batch, (int32_t)i, nullptr, nullptr, &out[i].groups);
    }
    std::lock_guard<std::mutex> lk(mu);
    stats.reruns += bs.reruns;
  }
  std::shared_ptr<HeldBatch> last = std::move(s.batch);  // released below, outside the lock
  std::lock_guard<std::mutex> lk(mu);
  for (size_t i = 0; i < s.queries.size(); ++i) {
    if (!last) continue;
    done[s.queries[i].ticket] = std::move(out[i]);
    ++stats.completed;
  }
  s.queries.clear();
  s.busy = false;
  s.finished.store(false);
  stats.host_finish_ms += ms_since(t0, Clock::now());
  done_cv.notify_all();
}

// Forms and starts one batch on an idle slot; false when nothing could start.
bool mgb_server::launch(Slot& s)
{
  std::vector<Query> take;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (pending.empty()) return false;
    const std::string sig = pending.front().signature;
    const int limit = (opt.batching && !pending.front().plain_only) ? std::max(1, opt.max_batch_queries) : 1;
    for (auto it = pending.begin(); it != pending.end() && (int)take.size() < limit;) {
      if (it->signature == sig) {
        take.push_back(std::move(*it));
        it = pending.erase(it);
      } else {
        ++it;
      }
    }
  }
  if (!opt.batching || take[0].plain_only) {
    // the per-query baseline (and the queries batches do not serve): the
    // plain path, synchronously on this slot's stream
    Query& q = take[0];
    std::vector<mgb_column>& keys = q.keys;
    std::vector<mgb_request> reqs(q.vals.size());
    int nout = 0;
    for (size_t r = 0; r < q.vals.size(); ++r) {
      reqs[r].values = q.vals[r];
      reqs[r].aggs = q.aggs[r].data();
      reqs[r].num_aggs = (int32_t)q.aggs[r].size();
      nout += reqs[r].num_aggs;
    }
    Done d;
    d.keys.assign(keys.size(), mgb_out_column{});
    d.results.assign(nout, mgb_out_column{});
    mgb_allocator call_alloc = alloc;
    call_alloc.temporaries = &s.temps;
    d.status = mgb_groupby_aggregate(keys.data(), (int32_t)keys.size(), q.null_policy, reqs.data(),
                                     (int32_t)reqs.size(), d.keys.data(), d.results.data(), &d.groups, s.stream,
                                     &call_alloc);
    // the call returns with its last writes still queued on the stream: the
    // result is complete (and may be read from any stream) after this wait,
    // and the call's temporaries may go back to the arena
    if (hipStreamSynchronize(s.stream) != hipSuccess && d.status == MGB_SUCCESS)
      d.status = fail(MGB_ERROR_HIP, "stream sync failed");
    s.flush_deferred();
    if (d.status == MGB_ERROR_ALLOC) {
      // no memory for its temporaries now: wait for batches to release theirs
      // unless nothing else holds any
      std::lock_guard<std::mutex> lk(mu);
      if (arena->in_use()) {
        ++stats.deferred;
        pending.push_front(std::move(q));
        wait_memory = true;
        return false;
      }
    }
    if (d.status != MGB_SUCCESS) d.error = mgb_last_error_message();
    d.latency_ms = ms_since(q.arrival, Clock::now());
    for (auto& c : d.keys) d.bytes += (int64_t)(c.data_bytes + c.mask_bytes);
    for (auto& c : d.results)
      d.bytes += (int64_t)(c.data_bytes + c.mask_bytes + c.child_data_bytes + c.child_mask_bytes);
    std::lock_guard<std::mutex> lk(mu);
    held_bytes += d.bytes;
    ++stats.completed;
    ++stats.batches;
    batch_sizes.push_back(1);
    done[q.ticket] = std::move(d);
    done_cv.notify_all();
    return true;
  }
  // continuous batching: plan the waiting queries as one batch; halve it
  // until it fits the HBM budget next to what is resident and held
  bool retried = false;
  while (!take.empty()) {
    const int Q = (int)take.size();
    const int nk = (int)take[0].keys.size(), nr = (int)take[0].vals.size();
    std::vector<mgb_column> keys((size_t)Q * nk);
    std::vector<mgb_request> reqs((size_t)Q * nr);
    for (int q = 0; q < Q; ++q) {
      for (int k = 0; k < nk; ++k) keys[(size_t)q * nk + k] = take[q].keys[k];
      for (int r = 0; r < nr; ++r) {
        reqs[(size_t)q * nr + r].values = take[q].vals[r];
        reqs[(size_t)q * nr + r].aggs = take[q].aggs[r].data();
        reqs[(size_t)q * nr + r].num_aggs = (int32_t)take[q].aggs[r].size();
      }
    }
    // group counts of resident tables: estimated once per table and key
    // columns (one synchronisation, the first time), then planned from the
    // cache with no device round trip
    std::vector<int64_t> est(Q);
    for (int q = 0; q < Q; ++q) {
      auto it = estimates.find(take[q].group_key);  // only this thread touches the cache
      if (it == estimates.end()) {
        int64_t e = 0;
        if (mgb_estimate_groups(take[q].keys.data(), (int32_t)take[q].keys.size(), take[q].null_policy, s.stream,
                                &s.temps, &e) != MGB_SUCCESS)
          e = take[q].keys.empty() ? 1 : take[q].keys[0].size;
        (void)hipStreamSynchronize(s.stream);
        s.flush_deferred();
        it = estimates.emplace(take[q].group_key, e).first;
      }
      est[q] = it->second;
    }
    // a served batch runs once: plain launches (capturing a graph for a
    // single execution only adds its capture and instantiation)
    mgb_batch_options bo{};
    bo.force_mode = -1;
    bo.use_graph = 0;
    bo.direct_item_rows = 0;
    bo.estimate_scale = 1.0;
    bo.estimates = est.data();
    bo.alloc = &alloc;
    mgb_batch* b = nullptr;
    const auto t_plan = Clock::now();
    mgb_status st = mgb_batch_create(keys.data(), nk, take[0].null_policy, reqs.data(), nr, Q, &bo, s.stream, &b);
    const double plan_ms = ms_since(t_plan, Clock::now());
    {
      std::lock_guard<std::mutex> lk(mu);
      stats.host_plan_ms += plan_ms;
    }
    if (st != MGB_SUCCESS && st != MGB_ERROR_ALLOC) {
      const std::string msg = mgb_last_error_message();
      std::lock_guard<std::mutex> lk(mu);
      for (auto& q : take) fail_query(q, st, msg);
      done_cv.notify_all();
      return true;
    }
    // no memory for the batch now (the chunks are taken and may not grow):
    // like a batch over the budget
    const bool no_memory = st == MGB_ERROR_ALLOC;
    mgb_batch_stats bs{};
    if (!no_memory) (void)mgb_batch_get_stats(b, &bs);
    bool fits = false;
    if (!no_memory) {
      std::lock_guard<std::mutex> lk(mu);
      fits = resident_bytes + held_bytes + bs.device_bytes <= opt.hbm_budget_bytes;
      if (fits) held_bytes += bs.device_bytes;
    }
    if (!fits) {
      if (b) mgb_batch_destroy(b);
      std::lock_guard<std::mutex> lk(mu);
      ++stats.deferred;
      if (Q == 1) {
        // fails only when it could not fit even with nothing else held
        if (no_memory ? arena->in_use() : resident_bytes + bs.device_bytes <= opt.hbm_budget_bytes) {
          pending.push_front(std::move(take[0]));  // wait for memory to be released
          wait_memory = true;
          return false;
        }
        if (no_memory && !retried) {
          // the last block in use may have come back since the attempt
          retried = true;
          continue;
        }
        fail_query(take[0], MGB_ERROR_ALLOC,
                   no_memory ? "no device memory for the query"
                             : "query does not fit the HBM budget (" + std::to_string(bs.device_bytes) +
                                   " bytes next to " + std::to_string(resident_bytes) + " resident, budget " +
                                   std::to_string(opt.hbm_budget_bytes) + ")");
        done_cv.notify_all();
        return true;
      }
      // put the second half back at the front, in order
      for (int q = Q - 1; q >= Q / 2; --q) pending.push_front(std::move(take[q]));
      take.resize(Q / 2);
      continue;
    }
    st = mgb_batch_execute(b, s.stream);
    if (st == MGB_SUCCESS && hipLaunchHostFunc(s.stream, slot_done, &s) != hipSuccess) st = MGB_ERROR_HIP;
    if (st != MGB_SUCCESS) {
      const std::string msg = mgb_last_error_message();
      mgb_batch_destroy(b);
      std::lock_guard<std::mutex> lk(mu);
      held_bytes -= bs.device_bytes;
      for (auto& q : take) fail_query(q, st, msg);
      done_cv.notify_all();
      return true;
    }
    auto hb = std::make_shared<HeldBatch>();
    hb->batch = b;
    hb->bytes = bs.device_bytes;
    hb->remaining = Q;
    std::lock_guard<std::mutex> lk(mu);
    s.batch = hb;
    s.queries = std::move(take);
    s.busy = true;
    ++stats.batches;
    batch_sizes.push_back(Q);
    stats.peak_held_bytes = std::max(stats.peak_held_bytes, held_bytes);
    return true;
  }
  return false;
}

void mgb_server::run()
{
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait_for(lk, std::chrono::milliseconds(1), [&] {
        if (stop) return true;
        for (auto& s : slots)
          if (s->busy && s->finished.load()) return true;
        bool idle = false;
        for (auto& s : slots) idle = idle || !s->busy;
        return idle && !pending.empty() && !wait_memory;
      });
      if (stop) {
        bool busy = false;
        for (auto& s : slots) busy = busy || s->busy;
        if (!busy) return;
      }
    }
    for (auto& s : slots)
      if (s->busy && s->finished.load()) retire(*s);
    for (auto& s : slots) {
      if (s->busy) continue;
      if (!launch(*s)) break;
    }
  }
}

extern "C" mgb_status mgb_server_create(const mgb_server_options* options, mgb_server** out)
{
  if (!out) return fail(MGB_ERROR_LOGIC, "out is NULL");
  *out = nullptr;
  std::unique_ptr<mgb_server> s(new mgb_server());
  if (options) s->opt = *options;
  if (s->opt.hbm_budget_bytes <= 0) s->opt.hbm_budget_bytes = 200ll << 30;
  if (s->opt.max_inflight <= 0) s->opt.max_inflight = 2;
  if (s->opt.max_batch_queries <= 0) s->opt.max_batch_queries = 256;
  if (!s->opt.batching) s->opt.max_inflight = 1;
  s->arena = std::make_unique<Arena>(s->opt.hbm_budget_bytes);
  s->alloc = mgb_allocator{arena_alloc, arena_free, s->arena.get()};
  if (s->opt.reserve_bytes > 0 &&
      (s->opt.reserve_bytes > s->opt.hbm_budget_bytes || !s->arena->reserve((size_t)s->opt.reserve_bytes)))
    return fail(MGB_ERROR_ALLOC, "the reserve does not fit the budget or the device");
  for (int i = 0; i < s->opt.max_inflight; ++i) {
    auto slot = std::make_unique<Slot>();
    slot->server = s.get();
    slot->arena = s->arena.get();
    slot->temps = mgb_allocator{slot_temp_alloc, slot_temp_free, slot.get(), nullptr, nullptr};
    MGB_HIP(hipStreamCreateWithFlags(&slot->stream, hipStreamNonBlocking));
    s->slots.push_back(std::move(slot));
  }
  mgb_server* raw = s.get();
  s->scheduler = std::thread([raw] { raw->run(); });
  *out = s.release();
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_server_add_table(mgb_server* s, const mgb_column* columns, int32_t num_columns,
                                           int32_t* table_id)
{
  if (!s || !columns || num_columns <= 0 || !table_id) return fail(MGB_ERROR_LOGIC, "bad table");
  Table t;
  t.cols.assign(columns, columns + num_columns);
  for (auto& c : t.cols) {
    if (c.size != t.cols[0].size) return fail(MGB_ERROR_LOGIC, "table columns differ in size");
    t.bytes += c.size * (int64_t)type_width(c.type_id) + (c.null_mask ? bitmask_bytes(c.size) : 0);
  }
  std::lock_guard<std::mutex> lk(s->mu);
  // the server's own device memory counts against the budget too
  if (s->resident_bytes + t.bytes + s->arena->total() > s->opt.hbm_budget_bytes)
    return fail(MGB_ERROR_ALLOC, "table does not fit the HBM budget");
  s->resident_bytes += t.bytes;
  s->arena->set_limit(s->opt.hbm_budget_bytes - s->resident_bytes);
  *table_id = (int32_t)s->tables.size();
  s->tables.push_back(std::move(t));
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_server_submit(mgb_server* s, int32_t table_id, const int32_t* key_columns, int32_t num_keys,
                                        int32_t null_policy, const mgb_server_request* requests,
                                        int32_t num_requests, int64_t* ticket)
{
  if (!s || !ticket || !key_columns || num_keys <= 0 || num_requests < 0 || (num_requests && !requests))
    return fail(MGB_ERROR_LOGIC, "bad query");
  Query q;
  q.arrival = Clock::now();
  q.null_policy = null_policy;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (table_id < 0 || table_id >= (int32_t)s->tables.size()) return fail(MGB_ERROR_LOGIC, "unknown table");
    const int ncol = (int)s->tables[table_id].cols.size();
    for (int k = 0; k < num_keys; ++k)
      if (key_columns[k] < 0 || key_columns[k] >= ncol) return fail(MGB_ERROR_LOGIC, "key column out of range");
    for (int r = 0; r < num_requests; ++r)
      if (requests[r].column < 0 || requests[r].column >= ncol || requests[r].num_aggs < 0 ||
          (requests[r].num_aggs && !requests[r].aggs))
        return fail(MGB_ERROR_LOGIC, "request out of range");
    q.table = table_id;
    const Table& t = s->tables[table_id];
    q.group_key = std::to_string(table_id) + "|" + std::to_string(null_policy) + "|";
    for (int k = 0; k < num_keys; ++k) {
      q.keys.push_back(t.cols[key_columns[k]]);
      q.group_key += std::to_string(key_columns[k]) + ",";
    }
    for (int r = 0; r < num_requests; ++r) {
      q.vals.push_back(t.cols[requests[r].column]);
      q.aggs.emplace_back(requests[r].aggs, requests[r].aggs + requests[r].num_aggs);
      for (auto& a : q.aggs.back())
        q.plain_only |= a.kind == MGB_AGG_VARIANCE || a.kind == MGB_AGG_STD || a.kind == MGB_AGG_M2 ||
                        a.kind == MGB_AGG_ARGMIN || a.kind == MGB_AGG_ARGMAX || a.kind == MGB_AGG_MEDIAN ||
                        a.kind == MGB_AGG_QUANTILE || a.kind == MGB_AGG_NUNIQUE || a.kind == MGB_AGG_NTH_ELEMENT ||
                        a.kind == MGB_AGG_COLLECT_LIST || a.kind == MGB_AGG_COLLECT_SET;
    }
    // null keys grouped as a key (null_policy INCLUDE): plain path too
    if (null_policy == MGB_NULL_INCLUDE)
      for (auto& k : q.keys) q.plain_only |= k.null_mask != nullptr && k.null_count != 0;
    q.signature = signature_of(q);
    q.ticket = s->next_ticket++;
    *ticket = q.ticket;
    s->pending.push_back(std::move(q));
    ++s->stats.submitted;
  }
  s->cv.notify_all();
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_server_wait(mgb_server* s, int64_t ticket, double timeout_ms, int64_t* num_groups,
                                      double* latency_ms)
{
  if (!s) return fail(MGB_ERROR_LOGIC, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  auto ready = [&] { return s->done.count(ticket) != 0; };
  if (timeout_ms < 0) s->done_cv.wait(lk, ready);
  else if (!s->done_cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
    return fail(MGB_ERROR_LOGIC, "timed out waiting for ticket " + std::to_string(ticket));
  const Done& d = s->done[ticket];
  if (num_groups) *num_groups = d.groups;
  if (latency_ms) *latency_ms = d.latency_ms;
  if (d.status != MGB_SUCCESS) return fail(d.status, d.error);
  return MGB_SUCCESS;
}

namespace {

// What a released ticket held. Freed with no lock held: hipFree waits for
// the device, and so for the host callbacks of batches still in flight.
struct Released {
  mgb_server* server = nullptr;
  std::shared_ptr<HeldBatch> batch;
  std::vector<mgb_out_column> keys, results;
  ~Released()
  {
    batch.reset();  // the last reference destroys the batch
    free_out(server->arena.get(), keys);
    free_out(server->arena.get(), results);
    // the memory is back: a query waiting for it may start
    {
      std::lock_guard<std::mutex> lk(server->mu);
      server->wait_memory = false;
    }
    server->cv.notify_all();
  }
};

// Takes a finished ticket out of the map (caller holds s->mu); the memory
// accounting changes at once, the memory itself is freed by the caller.
std::unique_ptr<Released> release_locked(mgb_server* s, std::map<int64_t, Done>::iterator it)
{
  auto r = std::make_unique<Released>();
  r->server = s;
  Done& d = it->second;
  if (d.batch) {
    if (--d.batch->remaining == 0) s->held_bytes -= d.batch->bytes;
    r->batch = std::move(d.batch);
  } else {
    s->held_bytes -= d.bytes;
    r->keys = std::move(d.keys);
    r->results = std::move(d.results);
  }
  s->done.erase(it);
  s->cv.notify_all();  // the budget may admit a waiting batch
  return r;
}

}  // namespace

extern "C" mgb_status mgb_server_take_result(mgb_server* s, int64_t ticket, mgb_out_column* out_keys,
                                             mgb_out_column* out_results, int64_t* num_groups, void* stream,
                                             const mgb_allocator* alloc)
{
  if (!s || !alloc) return fail(MGB_ERROR_LOGIC, "NULL argument");
  std::shared_ptr<HeldBatch> hb;
  int32_t index = 0;
  std::vector<mgb_out_column> own_keys, own_results;
  std::unique_ptr<Released> freed;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->done.find(ticket);
    if (it == s->done.end()) return fail(MGB_ERROR_LOGIC, "ticket not finished or already taken");
    if (it->second.status != MGB_SUCCESS) {
      const mgb_status st = it->second.status;
      const std::string err = it->second.error;
      freed = release_locked(s, it);
      lk.unlock();
      freed.reset();
      return fail(st, err);
    }
    hb = it->second.batch;
    index = it->second.index;
    if (!hb) {
      own_keys = it->second.keys;
      own_results = it->second.results;
      if (num_groups) *num_groups = it->second.groups;
    }
  }
  mgb_status st = MGB_SUCCESS;
  if (hb) {
    st = mgb_batch_copy_result(hb->batch, index, out_keys, out_results, num_groups, stream, alloc);
  } else {
    // plain-path result: copy into the caller's allocations
    hipStream_t hs = static_cast<hipStream_t>(stream);
    auto copy = [&](const mgb_out_column& src, mgb_out_column& dst) -> mgb_status {
      dst = src;
      dst.data = nullptr;
      dst.null_mask = nullptr;
      dst.child_data = nullptr;
      dst.child_null_mask = nullptr;
      if (src.child_data) {
        dst.child_data = alloc->allocate(alloc->ctx, src.child_data_bytes, stream);
        if (!dst.child_data) return fail(MGB_ERROR_ALLOC, "allocation of a list result's values failed");
        MGB_HIP(hipMemcpyAsync(dst.child_data, src.child_data, src.child_data_bytes, hipMemcpyDeviceToDevice, hs));
      }
      if (src.child_null_mask) {
        dst.child_null_mask = static_cast<uint32_t*>(alloc->allocate(alloc->ctx, src.child_mask_bytes, stream));
        if (!dst.child_null_mask) return fail(MGB_ERROR_ALLOC, "allocation of a list result's null mask failed");
        MGB_HIP(hipMemcpyAsync(dst.child_null_mask, src.child_null_mask, src.child_mask_bytes,
                               hipMemcpyDeviceToDevice, hs));
      }
      if (src.data) {
        dst.data = alloc->allocate(alloc->ctx, src.data_bytes, stream);
        if (!dst.data) return fail(MGB_ERROR_ALLOC, "allocation of a result column failed");
        MGB_HIP(hipMemcpyAsync(dst.data, src.data, src.data_bytes, hipMemcpyDeviceToDevice, hs));
      }
      if (src.null_mask) {
        dst.null_mask = static_cast<uint32_t*>(alloc->allocate(alloc->ctx, src.mask_bytes, stream));
        if (!dst.null_mask) return fail(MGB_ERROR_ALLOC, "allocation of a result null mask failed");
        MGB_HIP(hipMemcpyAsync(dst.null_mask, src.null_mask, src.mask_bytes, hipMemcpyDeviceToDevice, hs));
      }
      return MGB_SUCCESS;
    };
    for (size_t c = 0; c < own_keys.size() && st == MGB_SUCCESS; ++c) st = copy(own_keys[c], out_keys[c]);
    for (size_t c = 0; c < own_results.size() && st == MGB_SUCCESS; ++c) st = copy(own_results[c], out_results[c]);
    if (st == MGB_SUCCESS && hipStreamSynchronize(hs) != hipSuccess) st = fail(MGB_ERROR_HIP, "stream sync failed");
  }
  hb.reset();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->done.find(ticket);
    if (it != s->done.end()) freed = release_locked(s, it);
  }
  freed.reset();
  return st;
}

extern "C" mgb_status mgb_server_release(mgb_server* s, int64_t ticket)
{
  if (!s) return fail(MGB_ERROR_LOGIC, "server is NULL");
  std::unique_ptr<Released> freed;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->done.find(ticket);
    if (it == s->done.end()) return fail(MGB_ERROR_LOGIC, "ticket not finished or already taken");
    freed = release_locked(s, it);
  }
  freed.reset();
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_server_get_stats(mgb_server* s, mgb_server_stats* out)
{
  if (!s || !out) return fail(MGB_ERROR_LOGIC, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->resident_bytes = s->resident_bytes;
  out->held_bytes = s->held_bytes;
  out->pending = (int64_t)s->pending.size();
  out->device_bytes = s->arena->total();
  double sum = 0;
  for (double b : s->batch_sizes) sum += b;
  out->mean_batch_queries = s->batch_sizes.empty() ? 0.0 : sum / (double)s->batch_sizes.size();
  return MGB_SUCCESS;
}

extern "C" void mgb_server_destroy(mgb_server* s)
{
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->scheduler.joinable()) s->scheduler.join();
  std::vector<std::unique_ptr<Released>> freed;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    while (!s->done.empty()) freed.push_back(release_locked(s, s->done.begin()));
  }
  freed.clear();
  for (auto& slot : s->slots) {
    if (slot->stream) {
      (void)hipStreamSynchronize(slot->stream);
      slot->flush_deferred();
      // the host staging the plain path keeps for the stream goes with it
      (void)mgb_release_workspace(slot->stream);
      (void)hipStreamDestroy(slot->stream);
    }
  }
  delete s;
}
