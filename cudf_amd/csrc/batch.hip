// Batched groupby-aggregate: many independent queries that share one schema
// (key types, value types, aggregations) run as one captured HIP graph.
//
// Every query gets one of three executions, picked per query at plan time
// from a sampled distinct-count estimate and a cost model:
//
//   * LDS direct: its rows are cut into items; one workgroup aggregates an
//     item in an LDS table sized for the query (a power of two of slots, not
//     the whole LDS), then folds the groups into the query's global table.
//   * global: its rows are cut into items whose rows update the query's
//     global table with device atomics (small queries with many groups, where
//     an LDS table would not reduce anything).
//   * partitioned: its rows are hash partitioned (histogram, scan, tiled
//     scatter) into 2^k partitions whose groups fit one LDS table; each
//     partition is one item that writes its groups as output rows.
//
// The graph is: clear counters and fill the global tables; histogram of the
// partitioned queries' chunks; scan; scatter; one persistent kernel that
// takes items from a queue (largest estimated cost first); finalize of the
// global tables into output rows; validity masks. Outputs of all queries are
// concatenated columns; query q owns rows [out_base[q], out_base[q] +
// out_cap[q]) of them. No host synchronisation happens inside the graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "groupby_host.hpp"
#include "lds_agg.hpp"
#include "partition.hpp"

namespace mgb {

enum BatchMode : int32_t { BM_DIRECT = 0, BM_GLOBAL = 1, BM_PARTITIONED = 2 };
enum BatchItemKind : int32_t { BI_DIRECT = 0, BI_GLOBAL = 1, BI_PART = 2 };

// Per-query device descriptor. The kernels take the descriptor arrays as
// const __restrict__ arguments: read-only for the whole launch, so that their
// uniform fields are fetched with scalar loads (through the constant cache)
// rather than with a vector load per access.
struct BatchQuery {
  AggParams ap;        // the query's input columns and ops; table / capacity / counters: its global table
  AggParams rec_ap;    // partitioned: the same ops over the partitioned records (same global table)
  PartParams pp;       // partitioned: histogram / scatter of its chunks
  int32_t mode;
  int32_t slots;       // LDS slots of its direct items
  int32_t cap;         // groups a direct item's LDS table may hold
  int32_t pad;
  int64_t hist_base;   // partitioned: first entry of its [P][nchunks] histogram
  int64_t out_base;    // first output row of its groups
  unsigned long long* out_count;  // groups written so far
};

struct BatchItem {
  int32_t query;
  int32_t kind;
  int32_t part;        // BI_PART: partition
  int32_t nsub;        // BI_PART: the partition is cut into nsub items ...
  int64_t begin, end;  // BI_DIRECT / BI_GLOBAL: rows; BI_PART: begin = which of them
};

struct BatchChunk {
  int32_t query;
  int32_t chunk;
};

// A stretch of one query's global table (records [first, first + count)),
// the unit of work of the fill and finalize kernels.
struct BatchSeg {
  int32_t query;
  int32_t pad;
  int64_t first;
  int64_t count;
};

// records per stretch of the fill / finalize kernels: between one and eight
// workgroup passes, so that a batch of one large table still spreads over
// every CU
constexpr int64_t kSegMin = 1024, kSegMax = 8192;

struct BatchParams {
  const BatchQuery* queries;
  int32_t num_queries;
  int32_t words;                  // LDS record words
  const BatchItem* items;
  int64_t num_items;
  unsigned long long* queue;      // next item of the persistent kernel
  const BatchChunk* chunks;       // histogram / scatter work of partitioned queries
  int64_t num_chunks;
  const uint32_t* offsets;        // scanned histogram (record positions)
  const uint64_t* tmpl;           // identity record
  const OutTable* out;            // concatenated outputs
  int32_t part_slots;             // LDS slots of a partition item
  int32_t part_cap;
  uint32_t* hist;                 // [hist_entries + 1] partition counts; the last entry stays 0
  int64_t hist_entries;
  const BatchSeg* segs;           // the global tables, cut into stretches
  int64_t num_segs;
};

// ---------------------------------------------------------------- kernels

constexpr int kSampleRows = 8192;   // rows sampled per query for the distinct-count estimate
constexpr int kSampleSlots = 16384; // LDS hash set of the sample (64 KB + 64 KB)

// Distinct row hashes among kSampleRows evenly spaced rows of each query:
// est[q] = {distinct, seen once, seen twice, sampled}.
__global__ __launch_bounds__(kBlock) void batch_sample_kernel(const BatchQuery* queries, unsigned long long* est)
{
  __shared__ uint32_t hs[kSampleSlots];
  __shared__ uint32_t cnt[kSampleSlots];
  __shared__ unsigned long long tot[3];
  const BatchQuery& Q = queries[blockIdx.x];
  const AggParams& p = Q.ap;
  for (int i = threadIdx.x; i < kSampleSlots; i += kBlock) {
    hs[i] = 0xffffffffu;
    cnt[i] = 0;
  }
  if (threadIdx.x < 3) tot[threadIdx.x] = 0;
  __syncthreads();
  const int64_t R = p.num_rows;
  const int64_t S = R < kSampleRows ? R : kSampleRows;
  for (int64_t i = threadIdx.x; i < S; i += kBlock) {
    const int64_t row = S == R ? i : i * (R / S) + (i * (R % S)) / S;  // i * R / S without overflow
    if (!key_row_valid(p, row)) continue;
    uint32_t h = row_key(p, row).hash;
    if (h == 0xffffffffu) h = 0xfffffffeu;
    uint32_t slot = h & (kSampleSlots - 1);
    for (;;) {
      const uint32_t prev = atomicCAS(&hs[slot], 0xffffffffu, h);
      if (prev == 0xffffffffu || prev == h) {
        atomicAdd(&cnt[slot], 1u);
        break;
      }
      slot = (slot + 1) & (kSampleSlots - 1);
    }
  }
  __syncthreads();
  unsigned long long d = 0, f1 = 0, f2 = 0;
  for (int i = threadIdx.x; i < kSampleSlots; i += kBlock) {
    const uint32_t c = cnt[i];
    d += c ? 1 : 0;
    f1 += c == 1 ? 1 : 0;
    f2 += c == 2 ? 1 : 0;
  }
  atomicAdd(&tot[0], d);
  atomicAdd(&tot[1], f1);
  atomicAdd(&tot[2], f2);
  __syncthreads();
  if (threadIdx.x < 3) est[(int64_t)blockIdx.x * 4 + threadIdx.x] = tot[threadIdx.x];
  if (threadIdx.x == 3) est[(int64_t)blockIdx.x * 4 + 3] = (unsigned long long)S;
}

// Identity records into every query's global table (one workgroup per
// stretch of a table, one thread per record).
__global__ __launch_bounds__(kBlock) void batch_fill_kernel(const BatchParams bp, const BatchQuery* __restrict__ queries,
                                                          const BatchSeg* __restrict__ segs, uint64_t special_hdr)
{
  if (blockIdx.x == 0 && threadIdx.x == 0 && bp.hist) bp.hist[bp.hist_entries] = 0u;
  const BatchSeg sg = segs[blockIdx.x];
  const AggParams& p = queries[sg.query].ap;
  const int words = (int)(p.stride / 8);
  global_t<uint64_t>* t = to_global(reinterpret_cast<uint64_t*>(p.table));
  for (int64_t rec = sg.first + threadIdx.x; rec < sg.first + sg.count; rec += kBlock) {
    global_t<uint64_t>* r = t + rec * words;
    r[0] = rec == p.capacity ? special_hdr : bp.tmpl[0];
    for (int w = 1; w < words; ++w) r[w] = bp.tmpl[w];
  }
}

__global__ __launch_bounds__(kBlock) void batch_hist_kernel(const BatchQuery* __restrict__ queries,
                                                          const BatchChunk* __restrict__ chunks)
{
  extern __shared__ uint32_t counts[];
  const BatchChunk c = chunks[blockIdx.x];
  hist_chunk(queries[c.query].pp, c.chunk, counts);
}

__global__ __launch_bounds__(kLdsBlock) void batch_scatter_kernel(const BatchQuery* __restrict__ queries,
                                                                 const BatchChunk* __restrict__ chunks)
{
  extern __shared__ __align__(16) uint8_t smem[];
  const BatchChunk c = chunks[blockIdx.x];
  scatter_chunk(queries[c.query].pp, c.chunk, smem);
}

// Rows [begin, end) into p's global table with device atomics, one row per
// thread. New groups are counted in `inserted` (LDS, zero on entry and on
// return) and added to the table's group counter every 32 rows of a thread
// and at the end: the budget is tested against the counter as last added to.
__device__ __forceinline__ void global_rows(const AggParams& p, int64_t begin, int64_t end, uint32_t* inserted)
{
  uint32_t iter = 0;
  for (int64_t row = begin + threadIdx.x; row < end; row += kLdsBlock) {
    if ((++iter & 31u) == 0u) {
      add_inserted(p, atomicExch(inserted, 0u));
      // the overflow flag every 32 rows rather than at every probe
      if (__hip_atomic_load(&p.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    }
    if (!key_row_valid(p, row)) continue;
    uint8_t* g = global_find_or_insert(p, row_key(p, row), row, false, inserted, true);
    if (g) apply_row(p, g, row);
  }
  __syncthreads();
  if (threadIdx.x == 0) add_inserted(p, atomicExch(inserted, 0u));
}

// Persistent workgroups take items from the queue until it is empty.
__global__ __launch_bounds__(kLdsBlock) void batch_items_kernel(const BatchParams bp, const BatchQuery* __restrict__ queries,
                                                               const BatchItem* __restrict__ items)
{
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ int64_t next;
  __shared__ uint32_t inserted;
  if (threadIdx.x == 0) inserted = 0u;
  for (;;) {
    if (threadIdx.x == 0) next = (int64_t)std::min<unsigned long long>(atomicAdd(bp.queue, 1ull), 0x7fffffffull);
    __syncthreads();
    // the same for every lane: scalar loads of the item and query descriptors
    const int64_t it = (int64_t)__builtin_amdgcn_readfirstlane((int)next);
    __syncthreads();
    if (it >= bp.num_items) break;
    const BatchItem item = items[it];
    const BatchQuery& Q = queries[item.query];
    if (item.kind == BI_GLOBAL) {
      global_rows(Q.ap, item.begin, item.end, &inserted);
    } else {
      // one LDS aggregation for both kinds (a single inlined copy of the row loop)
      const bool part = item.kind == BI_PART;
      const AggParams& p = part ? Q.rec_ap : Q.ap;
      int64_t begin = item.begin, end = item.end;
      if (part) {
        const int64_t nchunks = Q.pp.nblocks;
        const int64_t ps = bp.offsets[Q.hist_base + (int64_t)item.part * nchunks];
        const int64_t pe = bp.offsets[Q.hist_base + (int64_t)(item.part + 1) * nchunks];
        begin = ps + (pe - ps) * item.begin / item.nsub;
        end = ps + (pe - ps) * (item.begin + 1) / item.nsub;
      }
      const LdsTable t = lds_table_at(lds, part ? bp.part_slots : Q.slots, bp.words, part ? bp.part_cap : Q.cap);
      lds_table_init(t, bp.tmpl, bp.words);
      lds_aggregate_rows(p, t, lds, begin, end, part ? Q.pp.log2p : 0, !part);
      // a partition aggregated whole writes its groups; parts of one fold
      // them into the query's table, as does an item whose groups overflowed
      if (part && item.nsub == 1 && !*t.overflow) lds_write_groups(p, t, Q.out_count, Q.out_base, *bp.out);
      else lds_fold_to_global(p, t, part ? (int)item.begin : 0, part ? item.nsub : 1);
    }
    __syncthreads();  // the next item reuses the LDS
  }
}

// Groups of every query's global table -> output rows (one workgroup per
// stretch of a table; one output-counter atomic per workgroup pass).
__global__ __launch_bounds__(kBlock) void batch_finalize_kernel(const BatchParams bp, const BatchQuery* __restrict__ queries,
                                                              const BatchSeg* __restrict__ segs)
{
  __shared__ uint32_t scan[kBlock];
  __shared__ unsigned long long base;
  const BatchSeg sg = segs[blockIdx.x];
  const BatchQuery& Q = queries[sg.query];
  // a partitioned query's table holds spilled groups whose representative
  // rows are partitioned records
  const AggParams& p = Q.mode == BM_PARTITIONED ? Q.rec_ap : Q.ap;
  const int64_t end = sg.first + sg.count;
  for (int64_t i0 = sg.first; i0 < end; i0 += kBlock) {
    const int64_t i = i0 + threadIdx.x;
    bool occ = false;
    if (i < end) {
      const uint64_t hdr = *reinterpret_cast<const uint64_t*>(p.table + i * (int64_t)p.stride);
      if (i == p.capacity) occ = p.key_mode == KEY_INT && hdr == 1ull;
      else occ = hdr != (p.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef);
    }
    const unsigned long long ballot = __ballot(occ);
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) scan[wave] = (uint32_t)__popcll(ballot);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t total = 0;
      for (int w = 0; w < kBlock / 64; ++w) {
        const uint32_t c = scan[w];
        scan[w] = total;
        total += c;
      }
      base = total ? atomicAdd(Q.out_count, (unsigned long long)total) : 0ull;
    }
    __syncthreads();
    if (occ) {
      const uint32_t rank = scan[wave] + (uint32_t)__popcll(ballot & ((1ull << lane) - 1ull));
      const int64_t pos = Q.out_base + (int64_t)base + rank;
      const uint8_t* rec = p.table + i * (int64_t)p.stride;
      finalize_keys(*reinterpret_cast<const uint64_t*>(rec), i == p.capacity, p.key_mode, p.keys, p.num_keys, pos,
                    *bp.out);
      finalize_record(rec, pos, *bp.out);
    }
    __syncthreads();
  }
}

// dst bits [0, n) = src bits [src_off, src_off + n); nulls counted into
// *null_count. dst has `words` words (the tail past n is cleared).
__global__ __launch_bounds__(kBlock) void copy_bits_kernel(const uint32_t* src, int64_t src_off, int64_t n,
                                                           uint32_t* dst, int64_t words,
                                                           unsigned long long* null_count)
{
  const int64_t w = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  uint32_t v = 0;
  if (w < words) {
    for (int b = 0; b < 32; ++b) {
      const int64_t i = w * 32 + b;
      if (i < n && bit_valid(src, src_off + i)) v |= 1u << b;
    }
    dst[w] = v;
  }
  const int64_t valid_here = w * 32 < n ? std::min<int64_t>(32, n - w * 32) : 0;
  const unsigned long long nulls = (unsigned long long)(valid_here - __popc(v));
  if (nulls) atomicAdd(null_count, nulls);
}

}  // namespace mgb

// ---------------------------------------------------------------- host

using namespace mgb;

namespace {

// Cost model of the three executions, in seconds of the whole GPU (rates
// of the whole chip; an item's fixed cost occupies one of the CUs). Rates
// are the ones measured on MI355X for the plain path's kernels: LDS direct
// aggregation ~180 G rows/s, partitioned ~20-30 G rows/s for one or two
// value columns, device atomics ~23.6 G/s; an LDS item costs a few
// microseconds of one CU to set up and write out.
struct CostModel {
  double lds_row = 6e-12;          // per row aggregated in LDS (direct)
  double part_row = 35e-12;        // per row hash partitioned and aggregated
  double atomic = 42e-12;          // per device atomic
  double item = 4e-6;              // per LDS item, on one CU
  double slot_init = 0.5e-9;       // per LDS slot initialised / scanned, on one CU
};

struct QueryPlan {
  int mode = BM_DIRECT;
  int64_t rows = 0;
  int64_t est = 0;        // distinct-count estimate (upper side)
  int slots = 0;          // direct: LDS slots
  int cap = 0;            // direct: LDS groups
  int64_t capacity = 0;   // global table records
  int64_t item_rows = 0;  // direct / global items
  int log2p = 0;          // partitioned
  int nsub = 1;           // partitioned: items per partition
  int64_t nchunks = 0;
  int64_t chunk_rows = 0;
  int64_t hist_base = 0;
  int64_t rec_base = 0;   // partitioned: first record (informational)
  int64_t out_base = 0;
  int64_t out_cap = 0;
  double cost = 0;
};

int64_t next_pow2(int64_t x)
{
  int64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

}  // namespace

namespace {

// Reusable pinned host blocks (never returned to the driver: hipHostFree
// waits for the whole device). A batch stages its descriptors in one block
// and uploads them with one asynchronous copy, and reads its counters back
// through it, so planning and finishing a batch never wait for other work on
// the device.
struct PinnedPool {
  std::mutex mu;
  std::multimap<size_t, void*> free;
} g_pinned;

void* pinned_get(size_t bytes, size_t* got)
{
  size_t sz = 4096;
  while (sz < bytes) sz <<= 1;
  {
    std::lock_guard<std::mutex> lk(g_pinned.mu);
    auto it = g_pinned.free.lower_bound(sz);
    if (it != g_pinned.free.end() && it->first <= 4 * sz) {
      void* p = it->second;
      *got = it->first;
      g_pinned.free.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess) return nullptr;
  *got = sz;
  return p;
}

void pinned_put(void* p, size_t bytes)
{
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_pinned.mu);
  g_pinned.free.emplace(bytes, p);
}

}  // namespace

struct mgb_batch {
  int Q = 0;
  int num_keys = 0, num_requests = 0, num_outs = 0;
  int32_t null_policy = 0;
  Plan plan;
  std::vector<mgb_column> keys;              // [Q][num_keys]
  std::vector<mgb_column> values;            // [Q][num_requests]
  std::vector<std::vector<mgb_aggregation>> aggs;  // [Q * num_requests]
  std::vector<QueryPlan> qp;
  int64_t total_out = 0;
  // device memory
  void* dev = nullptr;
  BatchParams bp{};
  unsigned long long* counters = nullptr;    // [Q][4]: groups, overflow, out_count, -
  unsigned long long* null_counts = nullptr;
  std::vector<uint8_t*> key_out, out_data, out_valid;
  std::vector<uint32_t*> out_mask;
  void* pinned = nullptr;                    // staged descriptors, then the counters read back
  size_t pinned_bytes = 0;
  unsigned long long* host_counters = nullptr;  // in `pinned`
  hipStream_t alloc_stream = nullptr;        // stream-ordered allocation and release of `dev`
  size_t dev_bytes = 0;
  bool own_alloc = false;                    // `dev` and re-run results from the caller's allocator
  mgb_allocator alloc{};
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = true;
  // results after mgb_batch_finish
  std::vector<int64_t> groups;
  std::vector<int> fallback;                 // 1: re-run with the plain path
  std::vector<std::vector<mgb_out_column>> fb_keys, fb_results;
  mgb_batch_stats stats{};
  // launch geometry
  size_t hist_lds = 0, scatter_lds = 0, items_lds = 0;
  int64_t num_chunks = 0, num_items = 0, hist_entries = 0, num_segs = 0;
  int items_grid = 0;
  uint32_t* bsum = nullptr;  // block sums of the histogram scan
  int64_t nbh = 0;

  ~mgb_batch()
  {
    release_fallbacks();
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
    // stream-ordered release: no device-wide synchronisation (a batch may be
    // destroyed while other batches run)
    if (dev) {
      if (own_alloc) alloc.deallocate(alloc.ctx, dev, dev_bytes, alloc_stream);
      else (void)hipFreeAsync(dev, alloc_stream);
    }
    pinned_put(pinned, pinned_bytes);
  }
  void release_fallbacks()
  {
    for (auto* v : {&fb_keys, &fb_results})
      for (auto& cols : *v)
        for (auto& c : cols) {
          if (own_alloc) {
            if (c.data) alloc.deallocate(alloc.ctx, c.data, c.data_bytes, alloc_stream);
            if (c.null_mask) alloc.deallocate(alloc.ctx, c.null_mask, c.mask_bytes, alloc_stream);
          } else {
            if (c.data) (void)hipFree(c.data);
            if (c.null_mask) (void)hipFree(c.null_mask);
          }
          c.data = nullptr;
          c.null_mask = nullptr;
        }
  }
};

namespace {

void* fb_alloc(void*, size_t bytes, void*)
{
  void* p = nullptr;
  return hipMalloc(&p, bytes) == hipSuccess ? p : nullptr;
}
void fb_free(void*, void* p, size_t, void*) { (void)hipFree(p); }

// Temporaries of the plain-path re-runs: from the batch's allocator, given
// back only after the stream has been waited for (the re-runs' last kernels
// may still read them when a call returns, and the allocator may hand a
// returned block to another stream at once).
struct DeferredTemps {
  const mgb_allocator* base;
  std::vector<std::pair<void*, size_t>> pending;
  static void* allocate(void* ctx, size_t bytes, void* stream)
  {
    const DeferredTemps* d = static_cast<const DeferredTemps*>(ctx);
    return d->base->allocate(d->base->ctx, bytes, stream);
  }
  static void deallocate(void* ctx, void* p, size_t bytes, void*)
  {
    static_cast<DeferredTemps*>(ctx)->pending.push_back({p, bytes});
  }
  void flush(hipStream_t stream)
  {
    for (auto& b : pending) base->deallocate(base->ctx, b.first, b.second, (void*)stream);
    pending.clear();
  }
};

// AggParams of one query's columns with the batch plan's ops and record.
mgb_status query_params(const mgb_batch& b, int q, AggParams& ap)
{
  std::vector<mgb_request> reqs(b.num_requests);
  for (int r = 0; r < b.num_requests; ++r) {
    reqs[r].values = b.values[(size_t)q * b.num_requests + r];
    reqs[r].aggs = b.aggs[(size_t)q * b.num_requests + r].data();
    reqs[r].num_aggs = (int32_t)b.aggs[(size_t)q * b.num_requests + r].size();
  }
  Plan plan;
  mgb_status st = build_plan(&b.keys[(size_t)q * b.num_keys], b.num_keys, b.null_policy, reqs.data(), b.num_requests, plan);
  if (st != MGB_SUCCESS) return st;
  // the record layout must be the batch's (same ops at the same offsets)
  const Plan& p0 = b.plan;
  bool same = plan.stride == p0.stride && plan.cols.size() == p0.cols.size() && plan.key_mode == p0.key_mode &&
              plan.rows_off == p0.rows_off && plan.outs.size() == p0.outs.size() &&
              plan.skip_null_keys == p0.skip_null_keys;
  for (size_t c = 0; same && c < plan.cols.size(); ++c)
    same = plan.cols[c].ops == p0.cols[c].ops && plan.cols[c].type == p0.cols[c].type &&
           plan.cols[c].nullable == p0.cols[c].nullable;
  for (size_t k = 0; same && k < plan.keys.size(); ++k) same = plan.keys[k].type_id == p0.keys[k].type_id;
  if (!same)
    return fail(MGB_ERROR_LOGIC,
                "query " + std::to_string(q) + " of the batch has a different schema (types, nulls or aggregations)");
  ap = AggParams{};
  ap.num_rows = b.keys[(size_t)q * b.num_keys].size;
  ap.key_mode = plan.key_mode;
  ap.num_keys = (int32_t)plan.keys.size();
  ap.skip_null_keys = plan.skip_null_keys;
  ap.compare_key_nulls = plan.compare_key_nulls;
  for (size_t k = 0; k < plan.keys.size(); ++k) ap.keys[k] = plan.keys[k];
  ap.stride = plan.stride;
  if (!single_launch_ops(plan, ap)) return fail(MGB_ERROR_UNSUPPORTED, "too many aggregations for a batch");
  // value validity comes from the null masks (flag bits are for partitioned records)
  for (int c = 0; c < ap.num_vals; ++c) ap.vals[c].flag_bit = -1;
  return MGB_SUCCESS;
}

}  // namespace

extern "C" mgb_status mgb_batch_create(const mgb_column* keys, int32_t num_keys, int32_t null_policy,
                                       const mgb_request* requests, int32_t num_requests, int32_t num_queries,
                                       const mgb_batch_options* options, void* stream_v, mgb_batch** out)
{
  if (!out) return fail(MGB_ERROR_LOGIC, "out is NULL");
  *out = nullptr;
  if (num_queries <= 0 || num_keys <= 0 || num_requests < 0 || !keys || (num_requests > 0 && !requests))
    return fail(MGB_ERROR_LOGIC, "empty or malformed batch");
  hipStream_t stream = static_cast<hipStream_t>(stream_v);
  mgb_batch_options opt{};
  opt.force_mode = -1;
  opt.use_graph = 1;
  opt.direct_item_rows = 0;
  opt.estimate_scale = 1.0;
  if (options) opt = *options;
  if (!(opt.estimate_scale > 0.0)) opt.estimate_scale = 1.0;

  // batches allocate stream-ordered from the device's default pool: keep
  // released memory in the pool instead of returning it at every
  // synchronisation, so that planning a batch does not map fresh memory
  // (once per device: changing the attribute trims the pool, which can take
  // tens of milliseconds)
  {
    static std::mutex pool_mu;
    static std::vector<int> pool_done;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
      std::lock_guard<std::mutex> lk(pool_mu);
      if (std::find(pool_done.begin(), pool_done.end(), dev) == pool_done.end()) {
        pool_done.push_back(dev);
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
          uint64_t keep = UINT64_MAX;
          (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
      }
    }
  }
  std::unique_ptr<mgb_batch> b(new mgb_batch());
  b->Q = num_queries;
  b->num_keys = num_keys;
  b->num_requests = num_requests;
  b->null_policy = null_policy;
  b->use_graph = opt.use_graph != 0;
  b->keys.assign(keys, keys + (size_t)num_queries * num_keys);
  b->values.resize((size_t)num_queries * num_requests);
  b->aggs.resize((size_t)num_queries * num_requests);
  for (size_t i = 0; i < (size_t)num_queries * num_requests; ++i) {
    b->values[i] = requests[i].values;
    b->aggs[i].assign(requests[i].aggs, requests[i].aggs + requests[i].num_aggs);
  }
  for (int q = 0; q < num_queries; ++q) {
    const int64_t R = b->keys[(size_t)q * num_keys].size;
    for (int c = 1; c < num_keys; ++c)
      if (b->keys[(size_t)q * num_keys + c].size != R) return fail(MGB_ERROR_LOGIC, "key columns of a query differ in size");
    for (int r = 0; r < num_requests; ++r)
      if (b->values[(size_t)q * num_requests + r].size != R)
        return fail(MGB_ERROR_LOGIC, "Size mismatch between request values and groupby keys.");
    if (R < 0 || R > kMaxRows) return fail(MGB_ERROR_UNSUPPORTED, "more rows than cudf::size_type holds (2^31 - 1)");
    if (R > 0) {
      // the kernels read row 0 of a column in place of an absent row
      for (int c = 0; c < num_keys; ++c)
        if (!b->keys[(size_t)q * num_keys + c].data) return fail(MGB_ERROR_LOGIC, "a key column with rows has no data");
      for (int r = 0; r < num_requests; ++r)
        if (!b->values[(size_t)q * num_requests + r].data)
          return fail(MGB_ERROR_LOGIC, "a value column with rows has no data");
    }
  }
  // the batch plan: query 0's
  {
    std::vector<mgb_request> reqs(num_requests);
    for (int r = 0; r < num_requests; ++r) {
      reqs[r].values = b->values[r];
      reqs[r].aggs = b->aggs[r].data();
      reqs[r].num_aggs = (int32_t)b->aggs[r].size();
    }
    mgb_status st = build_plan(b->keys.data(), num_keys, null_policy, reqs.data(), num_requests, b->plan);
    if (st != MGB_SUCCESS) return st;
  }
  const Plan& plan = b->plan;
  if (plan.compare_key_nulls)
    return fail(MGB_ERROR_UNSUPPORTED, "batched groupby does not group null keys (null_policy INCLUDE with null keys)");
  if (plan.has_pass2 || plan.has_sort)
    return fail(MGB_ERROR_UNSUPPORTED, "batched groupby does not serve VARIANCE, STD, M2, ARGMIN, ARGMAX, MEDIAN, "
                                       "QUANTILE, NUNIQUE, NTH_ELEMENT, COLLECT_LIST or COLLECT_SET (a second "
                                       "pass over the rows): use mgb_groupby_aggregate");
  if ((int)plan.outs.size() > kMaxOuts) return fail(MGB_ERROR_UNSUPPORTED, "too many aggregations for a batch");
  b->num_outs = (int)plan.outs.size();
  const int Q = num_queries;

  std::vector<BatchQuery> hq(Q);
  for (int q = 0; q < Q; ++q) {
    mgb_status st = query_params(*b, q, hq[q].ap);
    if (st != MGB_SUCCESS) return st;
  }

  // ---- distinct-count estimates: given, or sampled on the device (one sync)
  const int lds_words = lds_stride_of(plan.stride) / 8;
  const LdsGeom geom = lds_geometry(plan.stride);
  if (!geom.ok) return fail(MGB_ERROR_UNSUPPORTED, "record too large for the LDS executions of a batch");
  b->qp.resize(Q);
  if (opt.estimates) {
    for (int q = 0; q < Q; ++q) {
      const int64_t R = hq[q].ap.num_rows;
      const int64_t hi = (int64_t)std::ceil((double)opt.estimates[q] * 1.1 * opt.estimate_scale) + 16;
      b->qp[q].rows = R;
      b->qp[q].est = std::max<int64_t>(1, std::min<int64_t>(hi, R));
    }
  } else {
    BatchQuery* d_q = nullptr;
    unsigned long long* d_est = nullptr;
    MGB_HIP(hipMallocAsync((void**)&d_q, sizeof(BatchQuery) * Q, stream));
    MGB_HIP(hipMallocAsync((void**)&d_est, sizeof(unsigned long long) * 4 * Q, stream));
    MGB_HIP(hipMemcpyAsync(d_q, hq.data(), sizeof(BatchQuery) * Q, hipMemcpyHostToDevice, stream));
    batch_sample_kernel<<<(unsigned)Q, kBlock, 0, stream>>>(d_q, d_est);
    MGB_HIP(hipGetLastError());
    std::vector<unsigned long long> est((size_t)4 * Q);
    MGB_HIP(hipMemcpyAsync(est.data(), d_est, sizeof(unsigned long long) * 4 * Q, hipMemcpyDeviceToHost, stream));
    MGB_HIP(hipStreamSynchronize(stream));
    MGB_HIP(hipFreeAsync(d_q, stream));
    MGB_HIP(hipFreeAsync(d_est, stream));
    for (int q = 0; q < Q; ++q) {
      const double d = (double)est[4 * q], f1 = (double)est[4 * q + 1], f2 = (double)est[4 * q + 2];
      const int64_t S = (int64_t)est[4 * q + 3];
      const int64_t R = hq[q].ap.num_rows;
      double g = d;
      if (S < R) g = d + f1 * (f1 - 1.0) / (2.0 * (f2 + 1.0));  // Chao1
      int64_t hi = (int64_t)std::ceil(g * (S < R ? 1.3 : 1.0) * opt.estimate_scale) + (S < R ? 64 : 0);
      b->qp[q].rows = R;
      b->qp[q].est = std::max<int64_t>(1, std::min<int64_t>(hi, R));
    }
  }

  // ---- per-query execution (cost model)
  const CostModel cm;
  const int nops = (int)[&] {
    int n = plan.rows_off >= 0 ? 1 : 0;
    for (auto& pc : plan.cols) n += (int)pc.ops.size();
    return n;
  }();
  // one record less than the plain path's table: the persistent kernel keeps
  // its queue position in static LDS
  const int part_slots = geom.slots - 1, part_cap = (int)((geom.slots - 1) * 0.62);
  int max_log2p = 0;
  // work-item and chunk sizes from the batch's total rows: enough items and
  // chunks to fill every CU a few times over, whether the batch holds one
  // query or thousands
  const double cus = (double)device_cu_count();
  int64_t total_rows = 0;
  for (int q = 0; q < Q; ++q) total_rows += b->qp[q].rows;
  const int64_t fill_rows = std::max<int64_t>(1, total_rows / (int64_t)(4 * cus));
  const int64_t item_rows = opt.direct_item_rows > 0 ? opt.direct_item_rows
                                                     : std::max<int64_t>(16384, std::min<int64_t>(262144, fill_rows));
  // partitions of a partitioned query: enough for its groups, and more when
  // its rows are many, up to one work item of part_item_rows per partition
  // while each keeps 64 groups or more
  const int64_t part_item_rows = std::max<int64_t>(65536, fill_rows);
  for (int q = 0; q < Q; ++q) {
    QueryPlan& P = b->qp[q];
    const int64_t R = P.rows, G = P.est;
    // direct: the smallest LDS table with room for the estimate
    int64_t slots = std::max<int64_t>(256, next_pow2(2 * G));
    const bool direct_ok = slots <= part_slots && G <= part_cap;
    if (slots > part_slots) slots = part_slots;
    const int64_t items = std::max<int64_t>(1, (R + item_rows - 1) / item_rows);
    const double c_direct = direct_ok ? R * cm.lds_row + items * ((cm.item + slots * cm.slot_init) / cus +
                                                                  std::min<int64_t>(G, item_rows) * nops * cm.atomic)
                                      : 1e30;
    const double c_global = R * nops * cm.atomic + G * cm.atomic;
    int log2p = 1;
    while ((int64_t)(part_cap * 0.75) * (1ll << log2p) < G && log2p < 12) ++log2p;
    while (log2p < 12 && (R >> log2p) > part_item_rows && (G >> (log2p + 1)) >= 64) ++log2p;
    const double c_part = R * cm.part_row + (1ll << log2p) * (cm.item + part_slots * cm.slot_init) / cus;
    int mode = c_direct <= c_global && c_direct <= c_part ? BM_DIRECT : c_global <= c_part ? BM_GLOBAL : BM_PARTITIONED;
    if (opt.force_mode == BM_DIRECT && direct_ok) mode = BM_DIRECT;
    if (opt.force_mode == BM_GLOBAL) mode = BM_GLOBAL;
    if (opt.force_mode == BM_PARTITIONED) mode = BM_PARTITIONED;
    if (R == 0) mode = BM_GLOBAL;
    P.mode = mode;
    P.cost = mode == BM_DIRECT ? c_direct : mode == BM_GLOBAL ? c_global : c_part;
    P.slots = (int)slots;
    P.cap = (int)std::max<int64_t>(1, (int64_t)(slots * 0.62));
    P.item_rows = item_rows;
    if (mode == BM_PARTITIONED) {
      P.log2p = log2p;
      max_log2p = std::max(max_log2p, log2p);
      // partitions still much longer than a work item are cut into several
      // items (all of whose groups then go to the query's table)
      P.nsub = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (R >> log2p) / part_item_rows));
      // its table: every group when partitions are cut, else room for two
      // overflowing partitions
      P.capacity = P.nsub > 1 ? next_pow2(2 * G + 1024) : next_pow2(4 * (int64_t)part_cap + 1024);
    } else {
      P.capacity = next_pow2(std::max<int64_t>(64, 2 * G + 16));
    }
  }

  // histogram / scatter chunks of the partitioned queries (multiples of the
  // scatter tile)
  {
    int64_t part_rows = 0;
    for (int q = 0; q < Q; ++q)
      if (b->qp[q].mode == BM_PARTITIONED) part_rows += b->qp[q].rows;
    int64_t chunk = (int64_t)align_up((size_t)std::max<int64_t>(1, part_rows / (int64_t)(4 * cus)), kScatterTile);
    chunk = std::max<int64_t>(8 * kScatterTile, std::min<int64_t>(1 << 19, chunk));
    for (int q = 0; q < Q; ++q) {
      QueryPlan& P = b->qp[q];
      if (P.mode != BM_PARTITIONED) continue;
      P.chunk_rows = chunk;
      P.nchunks = std::max<int64_t>(1, (P.rows + chunk - 1) / chunk);
    }
  }

  // ---- layout: tables, histogram, records, outputs
  const int64_t stride = plan.stride;
  int64_t table_records = 0, hist_entries = 0, recs_total = 0, out_total = 0;
  int rec_bytes = 0, flags_rec_off = 0, flag_bytes = 0;
  int32_t koff[kMaxKeys] = {}, voff[kMaxVals] = {};
  {
    // partitioned record layout: keys, values, flags (8-byte multiple)
    int off = 0;
    for (size_t c = 0; c < plan.keys.size(); ++c) {
      off = (int)align_up((size_t)off, (size_t)plan.keys[c].width);
      koff[c] = off;
      off += plan.keys[c].width;
    }
    const AggParams& a0 = hq[0].ap;
    int nflags = 0;
    for (int c = 0; c < a0.num_vals; ++c) {
      off = (int)align_up((size_t)off, (size_t)a0.vals[c].width);
      voff[c] = off;
      off += a0.vals[c].width;
      if (plan.cols[c].nullable) ++nflags;
    }
    flag_bytes = nflags == 0 ? 0 : nflags <= 8 ? 1 : 2;
    if (flag_bytes) {
      off = (int)align_up((size_t)off, (size_t)flag_bytes);
      flags_rec_off = off;
      off += flag_bytes;
    }
    rec_bytes = (int)align_up((size_t)std::max(off, 8), 8);
  }
  for (int q = 0; q < Q; ++q) {
    QueryPlan& P = b->qp[q];
    table_records += P.capacity + 1;
    if (P.mode == BM_PARTITIONED) {
      P.hist_base = hist_entries;
      hist_entries += ((int64_t)1 << P.log2p) * P.nchunks;
      P.rec_base = recs_total;
      recs_total += P.rows;
      P.out_cap = std::min<int64_t>(P.rows, ((int64_t)1 << P.log2p) * (part_slots + 1) + P.capacity + 1);
    } else {
      P.out_cap = std::min<int64_t>(P.rows, P.capacity + 1);
    }
    P.out_base = out_total;
    out_total += P.out_cap;
  }
  if (recs_total >= (1ll << 32)) return fail(MGB_ERROR_UNSUPPORTED, "partitioned rows of one batch exceed 2^32");
  b->total_out = out_total;
  b->hist_entries = hist_entries;

  // items: largest estimated cost first
  std::vector<BatchItem> items;
  std::vector<double> item_cost;
  std::vector<BatchChunk> chunks;
  for (int q = 0; q < Q; ++q) {
    const QueryPlan& P = b->qp[q];
    if (P.mode == BM_PARTITIONED) {
      for (int64_t c = 0; c < P.nchunks; ++c) chunks.push_back(BatchChunk{q, (int32_t)c});
      const int np = 1 << P.log2p;
      for (int part = 0; part < np; ++part)
        for (int k = 0; k < P.nsub; ++k) {
          items.push_back(BatchItem{q, BI_PART, part, P.nsub, k, 0});
          item_cost.push_back(P.cost / ((double)np * P.nsub));
        }
    } else {
      for (int64_t r = 0; r < P.rows; r += P.item_rows) {
        items.push_back(BatchItem{q, P.mode == BM_DIRECT ? BI_DIRECT : BI_GLOBAL, 0, 0, r,
                                  std::min<int64_t>(P.rows, r + P.item_rows)});
        item_cost.push_back(P.cost * (double)(items.back().end - r) / (double)std::max<int64_t>(1, P.rows));
      }
    }
  }
  {
    std::vector<size_t> order(items.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return item_cost[a] > item_cost[c]; });
    std::vector<BatchItem> sorted(items.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = items[order[i]];
    items.swap(sorted);
  }
  b->num_items = (int64_t)items.size();
  b->num_chunks = (int64_t)chunks.size();
  std::vector<BatchSeg> segs;
  {
    int64_t records = 0;
    for (int q = 0; q < Q; ++q) records += b->qp[q].capacity + 1;
    const int64_t seg = std::max<int64_t>(
        kSegMin, std::min<int64_t>(kSegMax, (int64_t)align_up((size_t)(records / (int64_t)(4 * cus)), kBlock)));
    for (int q = 0; q < Q; ++q) {
      const int64_t n = b->qp[q].capacity + 1;
      for (int64_t f = 0; f < n; f += seg) segs.push_back(BatchSeg{q, 0, f, std::min<int64_t>(seg, n - f)});
    }
  }
  b->num_segs = (int64_t)segs.size();

  // one device allocation
  const int nk = num_keys, no = b->num_outs;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align_up(std::max<size_t>(bytes, 1), 256);
    return o;
  };
  // the uploaded descriptors first: one copy of [0, upload_bytes)
  const size_t o_queries = take(sizeof(BatchQuery) * Q);
  const size_t o_items = take(sizeof(BatchItem) * items.size());
  const size_t o_chunks = take(sizeof(BatchChunk) * chunks.size());
  const size_t o_segs = take(sizeof(BatchSeg) * segs.size());
  const size_t o_tmpl = take(sizeof(uint64_t) * kMaxRecordWords);
  const size_t o_ot = take(sizeof(OutTable));
  const size_t upload_bytes = off;
  const size_t o_counters = take(sizeof(unsigned long long) * (4 * (size_t)Q + 8));
  const size_t o_tables = take((size_t)table_records * stride);
  const size_t o_hist = take(sizeof(uint32_t) * (size_t)(hist_entries + 1));
  const size_t o_offs = take(sizeof(uint32_t) * (size_t)(hist_entries + 1));
  const int64_t nbh = std::max<int64_t>(1, (hist_entries + 1 + kScanTile - 1) / kScanTile);
  const size_t o_bsum = take(sizeof(uint32_t) * (size_t)nbh);
  const size_t o_recs = take((size_t)recs_total * rec_bytes);
  std::vector<size_t> o_key(nk), o_out(no), o_valid(no, SIZE_MAX), o_mask(no, SIZE_MAX);
  const size_t nout = (size_t)std::max<int64_t>(out_total, 1);
  for (int c = 0; c < nk; ++c) o_key[c] = take(nout * plan.keys[c].width);
  for (int o = 0; o < no; ++o) {
    o_out[o] = take(nout * plan.outs[o].spec.out_width);
    if (plan.outs[o].spec.valid) {
      o_valid[o] = take(nout);
      o_mask[o] = take((size_t)bitmask_bytes((int64_t)nout));
    }
  }
  const size_t o_nulls = take(sizeof(unsigned long long) * (no + 1));
  b->alloc_stream = stream;
  b->dev_bytes = off;
  if (opt.alloc) {
    if (!opt.alloc->allocate || !opt.alloc->deallocate) return fail(MGB_ERROR_LOGIC, "allocator without callbacks");
    b->alloc = *opt.alloc;
    b->own_alloc = true;
    b->dev = b->alloc.allocate(b->alloc.ctx, off, stream);
    if (!b->dev) return fail(MGB_ERROR_ALLOC, "allocation of " + std::to_string(off) + " bytes for the batch failed");
  } else {
    MGB_HIP(hipMallocAsync(&b->dev, off, stream));
  }
  uint8_t* base = static_cast<uint8_t*>(b->dev);
  b->counters = reinterpret_cast<unsigned long long*>(base + o_counters);
  b->null_counts = reinterpret_cast<unsigned long long*>(base + o_nulls);

  OutTable ot{};
  for (int c = 0; c < nk; ++c) {
    b->key_out.push_back(base + o_key[c]);
    ot.key_out[c] = base + o_key[c];
    ot.key_valid[c] = nullptr;
  }
  ot.num_outs = no;
  for (int o = 0; o < no; ++o) {
    b->out_data.push_back(base + o_out[o]);
    b->out_valid.push_back(o_valid[o] == SIZE_MAX ? nullptr : base + o_valid[o]);
    b->out_mask.push_back(o_mask[o] == SIZE_MAX ? nullptr : reinterpret_cast<uint32_t*>(base + o_mask[o]));
    ot.outs[o] = plan.outs[o].spec;
    ot.outs[o].out = b->out_data[o];
    ot.outs[o].valid = b->out_valid[o];
  }
  uint64_t tmpl[kMaxRecordWords] = {};
  record_template(plan, plan.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef, tmpl);

  // tile of the batch scatter: the largest whose LDS fits with 2^max_log2p partitions
  const int Pmax = 1 << std::max(max_log2p, 1);
  const size_t fixed_lds = (size_t)Pmax * 8 + 16 * 4 + (size_t)kScatterTile * 2 + 16;
  int tile = 0;
  for (int tt = kScatterTile; tt >= kLdsBlock; tt -= kLdsBlock)
    if (fixed_lds + (size_t)tt * rec_bytes <= lds_limit()) {
      tile = tt;
      break;
    }
  if (hist_entries > 0 && tile == 0) return fail(MGB_ERROR_UNSUPPORTED, "partition records too wide for a batch");
  b->hist_lds = (size_t)Pmax * 4;
  b->scatter_lds = fixed_lds + (size_t)tile * rec_bytes;

  // per-query descriptors
  uint8_t* tables = base + o_tables;
  int64_t trec = 0;
  uint8_t* recs = base + o_recs;
  for (int q = 0; q < Q; ++q) {
    QueryPlan& P = b->qp[q];
    BatchQuery& B = hq[q];
    B.mode = P.mode;
    B.slots = P.slots;
    B.cap = P.cap;
    B.hist_base = P.hist_base;
    B.out_base = P.out_base;
    B.out_count = b->counters + 4 * (size_t)q + 2;
    AggParams& ap = B.ap;
    ap.table = tables + trec * stride;
    ap.capacity = P.capacity;
    ap.counters = b->counters + 4 * (size_t)q;
    ap.max_groups = std::max<int64_t>(1, P.capacity * 3 / 4);
    trec += P.capacity + 1;
    if (P.mode == BM_PARTITIONED) {
      PartParams& pp = B.pp;
      pp = PartParams{};
      pp.src = ap;
      pp.chunk = P.chunk_rows;
      pp.log2p = P.log2p;
      pp.nblocks = (int32_t)P.nchunks;
      pp.hist = reinterpret_cast<uint32_t*>(base + o_hist) + P.hist_base;
      pp.offsets = reinterpret_cast<const uint32_t*>(base + o_offs) + P.hist_base;
      pp.out = recs;  // offsets are record positions of the whole batch
      pp.rec_bytes = rec_bytes;
      pp.rec_magic = ((1ull << 32) + (uint64_t)(rec_bytes / 8) - 1) / (uint64_t)(rec_bytes / 8);
      pp.flag_bytes = flag_bytes;
      pp.flags_off = flags_rec_off;
      pp.tile_rows = tile;
      int fb = 0;
      for (int c = 0; c < ap.num_vals; ++c) {
        pp.val_off[c] = voff[c];
        if (plan.cols[c].nullable) pp.src.vals[c].flag_bit = fb++;
      }
      for (int c = 0; c < ap.num_keys; ++c) pp.key_off[c] = koff[c];
      AggParams& ra = B.rec_ap;
      ra = ap;
      ra.num_rows = recs_total;
      ra.skip_null_keys = 0;
      ra.flags = flag_bytes ? recs + flags_rec_off : nullptr;
      ra.flag_bytes = flag_bytes;
      ra.flag_stride = rec_bytes;
      ra.rec_words = rec_bytes <= 32 ? rec_bytes / 8 : 0;
      ra.rec_base = recs;
      for (int c = 0; c < ra.num_keys; ++c) {
        ra.keys[c].data = recs + koff[c];
        ra.keys[c].stride = rec_bytes;
        ra.keys[c].mask = nullptr;
        ra.keys[c].bit_off = 0;
      }
      for (int c = 0; c < ra.num_vals; ++c) {
        ra.vals[c].data = recs + voff[c];
        ra.vals[c].stride = rec_bytes;
        ra.vals[c].mask = nullptr;
        ra.vals[c].bit_off = 0;
        ra.vals[c].flag_bit = pp.src.vals[c].flag_bit;
        if (c < kProgramCols) ra.prog[c] = pack_program(ra.vals[c], voff[c]);
      }
    }
  }

  BatchParams& bp = b->bp;
  bp.queries = reinterpret_cast<const BatchQuery*>(base + o_queries);
  bp.num_queries = Q;
  bp.words = lds_words;
  bp.items = reinterpret_cast<const BatchItem*>(base + o_items);
  bp.num_items = (int64_t)items.size();
  bp.queue = b->counters + 4 * (size_t)Q;
  bp.chunks = reinterpret_cast<const BatchChunk*>(base + o_chunks);
  bp.num_chunks = (int64_t)chunks.size();
  bp.offsets = reinterpret_cast<const uint32_t*>(base + o_offs);
  bp.tmpl = reinterpret_cast<const uint64_t*>(base + o_tmpl);
  bp.out = reinterpret_cast<const OutTable*>(base + o_ot);
  bp.part_slots = part_slots;
  bp.part_cap = part_cap;
  bp.hist = hist_entries ? reinterpret_cast<uint32_t*>(base + o_hist) : nullptr;
  bp.segs = reinterpret_cast<const BatchSeg*>(base + o_segs);
  bp.num_segs = (int64_t)segs.size();
  bp.hist_entries = hist_entries;
  b->bsum = reinterpret_cast<uint32_t*>(base + o_bsum);
  b->nbh = nbh;

  // stage every descriptor in one pinned block and upload it with one
  // asynchronous copy; the block stays with the batch (the copy may still be
  // in flight when this returns); its tail receives the counters
  {
    const size_t counters_bytes = sizeof(unsigned long long) * (4 * (size_t)Q + 8);
    const size_t need = align_up(upload_bytes, 256) + counters_bytes;
    b->pinned = pinned_get(need, &b->pinned_bytes);
    if (!b->pinned) return fail(MGB_ERROR_ALLOC, "pinned staging allocation failed");
    uint8_t* h = static_cast<uint8_t*>(b->pinned);
    std::memcpy(h + o_queries, hq.data(), sizeof(BatchQuery) * Q);
    if (!items.empty()) std::memcpy(h + o_items, items.data(), sizeof(BatchItem) * items.size());
    if (!chunks.empty()) std::memcpy(h + o_chunks, chunks.data(), sizeof(BatchChunk) * chunks.size());
    if (!segs.empty()) std::memcpy(h + o_segs, segs.data(), sizeof(BatchSeg) * segs.size());
    std::memcpy(h + o_tmpl, tmpl, sizeof(tmpl));
    std::memcpy(h + o_ot, &ot, sizeof(OutTable));
    MGB_HIP(hipMemcpyAsync(base, h, upload_bytes, hipMemcpyHostToDevice, stream));
    b->host_counters = reinterpret_cast<unsigned long long*>(h + align_up(upload_bytes, 256));
  }

  // launch geometry
  b->items_lds = std::max(lds_table_bytes(part_slots, lds_words), (size_t)16);
  // the largest dynamic LDS either kernel may use, set once per process
  static std::once_flag lds_once;
  static hipError_t lds_err = hipSuccess;
  std::call_once(lds_once, [] {
    const int lim = (int)lds_limit();
    // (the items kernel also holds 8 bytes of static LDS)
    lds_err = hipFuncSetAttribute((const void*)batch_items_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  lim - 64);
    if (lds_err == hipSuccess)
      lds_err = hipFuncSetAttribute((const void*)batch_scatter_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
  });
  MGB_HIP(lds_err);
  b->items_grid = (int)std::max<int64_t>(1, std::min<int64_t>(b->num_items, (int64_t)device_cu_count()));

  // statistics
  b->stats.num_queries = Q;
  b->stats.output_rows_reserved = out_total;
  b->stats.device_bytes = (int64_t)off;
  b->stats.num_items = b->num_items;
  for (int q = 0; q < Q; ++q) {
    if (b->qp[q].mode == BM_DIRECT) ++b->stats.direct_queries;
    else if (b->qp[q].mode == BM_GLOBAL) ++b->stats.global_queries;
    else ++b->stats.partitioned_queries;
  }
  b->groups.assign(Q, 0);
  b->fallback.assign(Q, 0);
  b->fb_keys.resize(Q);
  b->fb_results.resize(Q);

  *out = b.release();
  return MGB_SUCCESS;
}

namespace {

// Enqueues every step of one execution on `stream` (nothing here waits for
// the device, so the sequence can be captured in a graph).
mgb_status enqueue_batch(mgb_batch* b, hipStream_t stream)
{
  const BatchParams& bp = b->bp;
  const int Q = b->Q;
  {
    const int64_t n = 4 * (int64_t)Q + 8;
    fill_u64_kernel<<<(unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 1024)), kBlock, 0,
                      stream>>>(b->counters, n, 0ull);
  }
  batch_fill_kernel<<<(unsigned)b->num_segs, kBlock, 0, stream>>>(bp, bp.queries, bp.segs,
                                                                   b->plan.key_mode == KEY_INT ? 0ull : kEmptyRowRef);
  if (b->num_chunks > 0) {
    batch_hist_kernel<<<(unsigned)b->num_chunks, kBlock, b->hist_lds, stream>>>(bp.queries, bp.chunks);
    const int64_t n = b->hist_entries + 1;
    uint32_t* offs = const_cast<uint32_t*>(bp.offsets);
    scan_reduce_u32_kernel<<<(unsigned)b->nbh, 256, 0, stream>>>(bp.hist, n, b->bsum);
    scan_bsums_u32_kernel<<<1, 1024, 0, stream>>>(b->bsum, b->nbh, b->counters + 4 * (size_t)Q + 1);
    scan_apply_u32_kernel<<<(unsigned)b->nbh, 256, 0, stream>>>(bp.hist, n, b->bsum, offs);
    batch_scatter_kernel<<<(unsigned)b->num_chunks, kLdsBlock, b->scatter_lds, stream>>>(bp.queries, bp.chunks);
  }
  if (b->num_items > 0)
    batch_items_kernel<<<(unsigned)b->items_grid, kLdsBlock, b->items_lds, stream>>>(bp, bp.queries, bp.items);
  batch_finalize_kernel<<<(unsigned)b->num_segs, kBlock, 0, stream>>>(bp, bp.queries, bp.segs);
  const int no = b->num_outs;
  const int64_t words = bitmask_bytes(std::max<int64_t>(b->total_out, 1)) / 4;
  for (int o = 0; o < no; ++o) {
    if (!b->out_mask[o]) continue;
    pack_validity_kernel<<<(unsigned)((words + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
      b->out_valid[o], b->total_out, b->out_mask[o], words, b->null_counts + o);
  }
  MGB_HIP(hipGetLastError());
  return MGB_SUCCESS;
}

}  // namespace

extern "C" mgb_status mgb_batch_execute(mgb_batch* b, void* stream_v)
{
  if (!b) return fail(MGB_ERROR_LOGIC, "batch is NULL");
  hipStream_t stream = static_cast<hipStream_t>(stream_v);
  b->release_fallbacks();
  std::fill(b->fallback.begin(), b->fallback.end(), 0);
  if (!b->use_graph) return enqueue_batch(b, stream);
  if (!b->exec) {
    if (!b->capture_stream) MGB_HIP(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
    MGB_HIP(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
    mgb_status st = enqueue_batch(b, b->capture_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->capture_stream, &g);
    if (st != MGB_SUCCESS) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return fail(MGB_ERROR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    MGB_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  {
    ProfScope prof("batch", stream);
    MGB_HIP(hipGraphLaunch(b->exec, stream));
  }
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_batch_finish(mgb_batch* b, void* stream_v)
{
  if (!b) return fail(MGB_ERROR_LOGIC, "batch is NULL");
  hipStream_t stream = static_cast<hipStream_t>(stream_v);
  const int Q = b->Q;
  MGB_HIP(hipMemcpyAsync(b->host_counters, b->counters, sizeof(unsigned long long) * 4 * Q, hipMemcpyDeviceToHost,
                         stream));
  MGB_HIP(hipStreamSynchronize(stream));
  int64_t reruns = 0;
  const mgb_allocator fb{fb_alloc, fb_free, nullptr, nullptr, nullptr};
  const mgb_allocator& out_alloc = b->own_alloc ? b->alloc : fb;
  DeferredTemps temps{out_alloc.temporaries ? out_alloc.temporaries : &out_alloc, {}};
  const mgb_allocator temp_alloc{DeferredTemps::allocate, DeferredTemps::deallocate, &temps, nullptr, nullptr};
  for (int q = 0; q < Q; ++q) {
    const unsigned long long* c = b->host_counters + 4 * (size_t)q;
    b->groups[q] = (int64_t)c[2];
    b->fallback[q] = c[1] != 0 || (int64_t)c[2] > b->qp[q].out_cap;
    if (!b->fallback[q]) continue;
    // the estimate was too low for this query: run it alone on the plain path
    ++reruns;
    std::vector<mgb_request> reqs(b->num_requests);
    for (int r = 0; r < b->num_requests; ++r) {
      reqs[r].values = b->values[(size_t)q * b->num_requests + r];
      reqs[r].aggs = b->aggs[(size_t)q * b->num_requests + r].data();
      reqs[r].num_aggs = (int32_t)b->aggs[(size_t)q * b->num_requests + r].size();
    }
    b->fb_keys[q].assign(b->num_keys, mgb_out_column{});
    b->fb_results[q].assign(b->num_outs, mgb_out_column{});
    mgb_allocator al = out_alloc;
    al.temporaries = &temp_alloc;
    int64_t g = 0;
    mgb_status st = mgb_groupby_aggregate(&b->keys[(size_t)q * b->num_keys], b->num_keys, b->null_policy,
                                          reqs.data(), b->num_requests, b->fb_keys[q].data(), b->fb_results[q].data(),
                                          &g, stream, &al);
    if (st != MGB_SUCCESS) {
      (void)hipStreamSynchronize(stream);
      temps.flush(stream);
      return st;
    }
    b->groups[q] = g;
  }
  // re-run results are read from any stream afterwards; their temporaries go back
  if (reruns) {
    hipError_t e = hipStreamSynchronize(stream);
    temps.flush(stream);
    MGB_HIP(e);
  }
  b->stats.reruns = reruns;
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_batch_result(const mgb_batch* b, int32_t query, mgb_column* out_keys,
                                       mgb_column* out_results, int64_t* num_groups)
{
  if (!b) return fail(MGB_ERROR_LOGIC, "batch is NULL");
  if (query < 0 || query >= b->Q) return fail(MGB_ERROR_LOGIC, "query index out of range");
  const int64_t g = b->groups[query];
  if (num_groups) *num_groups = g;
  const bool fb = b->fallback[query] != 0;
  const int64_t base = fb ? 0 : b->qp[query].out_base;
  for (int c = 0; c < b->num_keys && out_keys; ++c) {
    mgb_column& k = out_keys[c];
    k = mgb_column{};
    k.type_id = b->plan.keys[c].type_id;
    k.size = g;
    k.offset = base;
    k.data = fb ? b->fb_keys[query][c].data : b->key_out[c];
    k.null_mask = fb ? b->fb_keys[query][c].null_mask : nullptr;
    k.null_count = fb ? b->fb_keys[query][c].null_count : 0;
  }
  for (int o = 0; o < b->num_outs && out_results; ++o) {
    mgb_column& r = out_results[o];
    r = mgb_column{};
    r.type_id = b->plan.outs[o].out_type;
    r.size = g;
    r.offset = base;
    r.data = fb ? b->fb_results[query][o].data : b->out_data[o];
    r.null_mask = fb ? b->fb_results[query][o].null_mask : b->out_mask[o];
    r.null_count = fb ? b->fb_results[query][o].null_count : (b->out_mask[o] ? -1 : 0);
  }
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_batch_copy_result(const mgb_batch* b, int32_t query, mgb_out_column* out_keys,
                                            mgb_out_column* out_results, int64_t* num_groups, void* stream_v,
                                            const mgb_allocator* alloc)
{
  if (!b || !alloc) return fail(MGB_ERROR_LOGIC, "NULL argument");
  if (query < 0 || query >= b->Q) return fail(MGB_ERROR_LOGIC, "query index out of range");
  hipStream_t stream = static_cast<hipStream_t>(stream_v);
  std::vector<mgb_column> kv(b->num_keys), rv(b->num_outs);
  int64_t g = 0;
  mgb_status st = mgb_batch_result(b, query, kv.data(), rv.data(), &g);
  if (st != MGB_SUCCESS) return st;
  if (num_groups) *num_groups = g;
  std::vector<std::pair<void*, size_t>> live;
  auto get = [&](size_t bytes) {
    void* p = alloc->allocate(alloc->ctx, bytes, stream_v);
    if (p) live.push_back({p, bytes});
    return p;
  };
  auto release = [&] {
    for (auto& l : live) alloc->deallocate(alloc->ctx, l.first, l.second, stream_v);
  };
  unsigned long long* d_nulls = nullptr;
  const int ncols = b->num_keys + b->num_outs;
  MGB_HIP(hipMallocAsync((void**)&d_nulls, sizeof(unsigned long long) * std::max(ncols, 1), stream));
  MGB_HIP(hipMemsetAsync(d_nulls, 0, sizeof(unsigned long long) * std::max(ncols, 1), stream));
  for (int j = 0; j < ncols; ++j) {
    const mgb_column& v = j < b->num_keys ? kv[j] : rv[j - b->num_keys];
    mgb_out_column& o = j < b->num_keys ? out_keys[j] : out_results[j - b->num_keys];
    o = mgb_out_column{};
    o.type_id = v.type_id;
    o.size = g;
    if (g == 0) continue;
    const int w = type_width(v.type_id);
    o.data_bytes = (size_t)g * w;
    o.data = get(o.data_bytes);
    if (!o.data) {
      release();
      (void)hipFreeAsync(d_nulls, stream);
      return fail(MGB_ERROR_ALLOC, "allocation of a result column failed");
    }
    MGB_HIP(hipMemcpyAsync(o.data, static_cast<const uint8_t*>(v.data) + (size_t)v.offset * w, o.data_bytes,
                           hipMemcpyDeviceToDevice, stream));
    if (v.null_mask) {
      o.mask_bytes = (size_t)bitmask_bytes(g);
      o.null_mask = static_cast<uint32_t*>(get(o.mask_bytes));
      if (!o.null_mask) {
        release();
        (void)hipFreeAsync(d_nulls, stream);
        return fail(MGB_ERROR_ALLOC, "allocation of a result null mask failed");
      }
      const int64_t words = (int64_t)o.mask_bytes / 4;
      copy_bits_kernel<<<(unsigned)((words + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
        v.null_mask, v.offset, g, o.null_mask, words, d_nulls + j);
    }
  }
  MGB_HIP(hipGetLastError());
  std::vector<unsigned long long> nulls(std::max(ncols, 1));
  MGB_HIP(hipMemcpyAsync(nulls.data(), d_nulls, sizeof(unsigned long long) * std::max(ncols, 1), hipMemcpyDeviceToHost,
                         stream));
  MGB_HIP(hipFreeAsync(d_nulls, stream));
  MGB_HIP(hipStreamSynchronize(stream));
  for (int j = 0; j < ncols; ++j) {
    mgb_out_column& o = j < b->num_keys ? out_keys[j] : out_results[j - b->num_keys];
    o.null_count = o.null_mask ? (int64_t)nulls[j] : 0;
  }
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_batch_get_stats(const mgb_batch* b, mgb_batch_stats* out)
{
  if (!b || !out) return fail(MGB_ERROR_LOGIC, "NULL argument");
  *out = b->stats;
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_batch_query_mode(const mgb_batch* b, int32_t query, int32_t* mode, int64_t* estimate)
{
  if (!b) return fail(MGB_ERROR_LOGIC, "batch is NULL");
  if (query < 0 || query >= b->Q) return fail(MGB_ERROR_LOGIC, "query index out of range");
  if (mode) *mode = b->qp[query].mode;
  if (estimate) *estimate = b->qp[query].est;
  return MGB_SUCCESS;
}

extern "C" void mgb_batch_destroy(mgb_batch* b) { delete b; }
