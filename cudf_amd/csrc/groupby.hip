// MI355X hash groupby-aggregate: kernels and the C ABI driver.
//
// The path replaces cpp/src/groupby/hash/groupby.cu:424-620 of libcudf. It
// has three executions, chosen per call from a sampled distinct-count
// estimate (G) and the size of one group's accumulator record:
//
//   * partitioned (G above what one workgroup's LDS holds): the rows are hash
//     partitioned (histogram, scan, scatter) into P = 2^k partitions whose
//     groups fit one workgroup's LDS; each partition is then aggregated by one
//     workgroup in an LDS hash table with LDS atomics and written out as
//     final results. No global atomics touch the per-row path.
//   * direct (small G): each workgroup aggregates a contiguous slice of the
//     input into its own LDS table, then folds its groups into a small global
//     table with device atomics (a handful per group and workgroup).
//   * global (composite keys compared with nulls, or too many aggregations
//     for LDS): one global open-addressing table of records updated with
//     device atomics per row.
//
// Records (global or LDS) are an 8-byte header followed by the group's
// accumulators. The header is claimed with one 64-bit CAS and never
// rewritten: it is the key itself for a single integer-like key, otherwise
// (row hash << 32 | representative row). Nothing else has to be published
// before another row may use the group, so there is no spin on half-built
// slots and no agent-scope release/acquire (an L2 writeback / invalidate on
// gfx950). Float MIN/MAX use order-preserving unsigned codes so they are
// exact and independent of update order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>


#include "groupby_host.hpp"
#include "hot_keys.hpp"
#include "lds_agg.hpp"
#include "tiny_agg.hpp"
#include "partition.hpp"
#include "seg_sort.hpp"
#include "sort_aggs.hpp"

namespace mgb {


// =================================================================== kernels


__global__ __launch_bounds__(kBlock) void fill_table_kernel(const FillParams p)
{
  const int64_t total = p.num_records * p.words;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += step) {
    int64_t r = i / p.words;
    int w = (int)(i - r * p.words);
    uint64_t v = p.tmpl[w];
    if (w == 0 && r == p.num_records - 1) v = p.special_hdr;
    p.table[i] = v;
  }
  for (int z = 0; z < 2; ++z)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.zero_words[z]; i += step) p.zero[z][i] = 0ull;
}

// Per-row insert-or-find and atomic update in the global table.
// New groups are counted per workgroup in LDS and added to the table's
// group counter every 32 rows of a thread and at the end, not at every
// insert (one counter taken by every insert of the launch is one L2
// channel's traffic, which stalls every access that maps there); the
// budget is tested against the counter as last added to. With p.row_rec,
// every row's record index (kNoRecord: none) is written for the sort path.
__global__ __launch_bounds__(kBlock) void aggregate_kernel(const AggParams p)
{
  __shared__ uint32_t inserted;
  if (threadIdx.x == 0) inserted = 0u;
  __syncthreads();
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  uint32_t iter = 0;
  for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < p.num_rows; row += step) {
    if ((++iter & 31u) == 0u) {
      add_inserted(p, atomicExch(&inserted, 0u));
      // an overflowed attempt is repeated with a larger table: stop early
      if (__hip_atomic_load(&p.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    }
    uint8_t* rec = nullptr;
    if (key_row_valid(p, row)) {
      const RowKey rk = row_key(p, row);
      rec = global_find_or_insert(p, rk, row, false, &inserted, true);
      if (rec) apply_row(p, rec, row);
    }
    if (p.row_rec) p.row_rec[row] = rec ? (uint32_t)((rec - p.table) / p.stride) : kNoRecord;
  }
  __syncthreads();
  if (threadIdx.x == 0) add_inserted(p, inserted);
}

// Second pass over the rows (VARIANCE / STD / ARGMIN / ARGMAX fields) on
// the global table of the global execution.
__global__ __launch_bounds__(kBlock) void pass2_kernel(const AggParams p, const Pass2Ops* __restrict__ q)
{
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < p.num_rows; row += step) {
    if (!key_row_valid(p, row)) continue;
    uint8_t* rec = global_find(p, row_key(p, row), row);
    if (rec) pass2_update<false>(q, rec, row, 0u, (uint32_t)row);
  }
}

__device__ __forceinline__ bool record_occupied(const uint8_t* table, int64_t r, int64_t capacity, int32_t stride,
                                                int32_t key_mode)
{
  if (r > capacity) return false;
  uint64_t hdr = *reinterpret_cast<const uint64_t*>(table + r * (int64_t)stride);
  if (r == capacity) return key_mode == KEY_INT && hdr == 1ull;
  return hdr != (key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef);
}

// table_groups (or nullptr): the table's group counter; an empty table is
// not scanned (this kernel, scan_blocks_kernel and finalize_kernel).
__global__ __launch_bounds__(kBlock) void count_occupied_kernel(const uint8_t* table, int64_t capacity,
                                                                int32_t stride, int32_t key_mode,
                                                                int64_t* block_counts,
                                                                const unsigned long long* table_groups)
{
  if (table_groups && *table_groups == 0) {
    if (threadIdx.x == 0) block_counts[blockIdx.x] = 0;
    return;
  }
  int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int c = __syncthreads_count(record_occupied(table, r, capacity, stride, key_mode));
  if (threadIdx.x == 0) block_counts[blockIdx.x] = c;
}

// Exclusive scan of n block counts in place by one workgroup; total -> *total.
// Each pass takes 16 consecutive counts a thread: the threads' sums are
// scanned within each wave by shuffles and across the 16 waves in LDS.
constexpr int kScanBlocksItems = 16;
__global__ __launch_bounds__(1024) void scan_blocks_kernel(int64_t* counts, int64_t n, unsigned long long* total,
                                                          const unsigned long long* table_groups)
{
  __shared__ long long wsum[16];
  __shared__ long long carry;
  if (table_groups && *table_groups == 0) {  // every count is 0
    if (threadIdx.x == 0) *total = 0ull;
    return;
  }
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t base = 0; base < n; base += 1024 * kScanBlocksItems) {
    const int64_t first = base + (int64_t)threadIdx.x * kScanBlocksItems;
    long long v[kScanBlocksItems];
    long long sum = 0;
#pragma unroll
    for (int k = 0; k < kScanBlocksItems; ++k) {
      v[k] = first + k < n ? (long long)counts[first + k] : 0ll;
      sum += v[k];
    }
    long long x = sum;
    for (int d = 1; d < 64; d <<= 1) {
      const long long y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    long long run = carry + x - sum;
    for (int ww = 0; ww < w; ++ww) run += wsum[ww];
#pragma unroll
    for (int k = 0; k < kScanBlocksItems; ++k) {
      if (first + k < n) counts[first + k] = run;
      run += v[k];
    }
    __syncthreads();
    if (threadIdx.x == 1023) carry = run;  // the counts so far
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = (unsigned long long)carry;
}

// Compacts the occupied records of the global table in index order and
// writes keys and results.
__global__ __launch_bounds__(kBlock) void finalize_kernel(const FinParams p)
{
  if (p.table_groups && *p.table_groups == 0) return;
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool occ = record_occupied(p.table, r, p.capacity, p.stride, p.key_mode);
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t ballot = __ballot(occ);
  __shared__ int wave_counts[kBlock / 64];
  if (lane == 0) wave_counts[wave] = __popcll(ballot);
  __syncthreads();
  if (!occ) return;
  int64_t pos = p.block_offsets[blockIdx.x] + __popcll(ballot & ((1ull << lane) - 1ull));
  for (unsigned w = 0; w < wave; ++w) pos += wave_counts[w];
  if (p.pos_base) pos += (int64_t)*p.pos_base;
  const uint8_t* rec = p.table + r * (int64_t)p.stride;
  const uint64_t hdr = *reinterpret_cast<const uint64_t*>(rec);
  finalize_keys(hdr, r == p.capacity, p.key_mode, p.keys, p.num_keys, pos, *p.out);
  finalize_record(rec, pos, *p.out);
  if (p.rec_pos) p.rec_pos[r] = (uint32_t)pos;
}

// Packs validity bytes into a cudf bitmask (words beyond n are zeroed) and
// adds the number of nulls to *null_count.
__global__ __launch_bounds__(kBlock) void pack_validity_kernel(const uint8_t* valid, int64_t n, uint32_t* mask,
                                                               int64_t words_alloc, unsigned long long* null_count)
{
  const int64_t w = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (w >= words_alloc) return;
  uint32_t bits = 0;
  int nulls = 0;
  for (int b = 0; b < 32; ++b) {
    int64_t i = w * 32 + b;
    if (i < n) {
      if (valid[i]) bits |= 1u << b;
      else ++nulls;
    }
  }
  mask[w] = bits;
  if (nulls) atom_add(null_count, (unsigned long long)nulls);
}

// ---- result columns: copy and validity packing in one launch

// One result column: `bytes` of data copied from the temporary outputs and,
// when `mask` is set, `rows` validity bytes packed into `words` bitmask words
// (bits past the rows 0) with the nulls added to *nulls.
struct EmitCol {
  const uint8_t* src;
  uint8_t* dst;
  int64_t bytes;
  const uint8_t* valid;
  uint32_t* mask;
  int64_t words;
  unsigned long long* nulls;
};
constexpr int kEmitCols = 32;
struct EmitParams {
  EmitCol cols[kEmitCols];
  int64_t rows;
};

// blockIdx.y: the column; the blocks along x stride over its units.
__global__ __launch_bounds__(kBlock) void emit_columns_kernel(const EmitParams ep)
{
  __shared__ unsigned int block_nulls;
  const EmitCol& c = ep.cols[blockIdx.y];
  const int64_t first = (int64_t)blockIdx.x * kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
  const global_t<const uint8_t>* src = to_global(c.src);
  global_t<uint8_t>* dst = to_global(c.dst);
  int64_t tail = 0;
  if ((((uintptr_t)c.src | (uintptr_t)c.dst) & 15) == 0) {
    const int64_t n16 = c.bytes >> 4;
    const global_t<const u64x2>* s16 = reinterpret_cast<const global_t<const u64x2>*>(src);
    global_t<u64x2>* d16 = reinterpret_cast<global_t<u64x2>*>(dst);
    for (int64_t i = first; i < n16; i += stride) d16[i] = s16[i];
    tail = n16 << 4;
  }
  for (int64_t i = tail + first; i < c.bytes; i += stride) dst[i] = src[i];
  if (!c.mask) return;
  if (threadIdx.x == 0) block_nulls = 0;
  __syncthreads();
  unsigned int nulls = 0;
  const global_t<const uint8_t>* v = to_global(c.valid);
  for (int64_t w = first; w < c.words; w += stride) {
    uint32_t bits = 0;
    const int64_t r0 = w * 32;
    if (r0 + 32 <= ep.rows) {
      // 32 validity bytes as two 16-byte loads (the array is 256-byte aligned)
      const global_t<const u64x2>* v16 = reinterpret_cast<const global_t<const u64x2>*>(v + r0);
      const u64x2 a = v16[0], b = v16[1];
      const uint64_t q[4] = {a.x, a.y, b.x, b.y};
#pragma unroll
      for (int k = 0; k < 32; ++k)
        if ((q[k >> 3] >> ((k & 7) * 8)) & 0xffu) bits |= 1u << k;
      nulls += 32u - (unsigned)__popc(bits);
    } else {
      for (int k = 0; k < 32 && r0 + k < ep.rows; ++k) {
        if (v[r0 + k]) bits |= 1u << k;
        else ++nulls;
      }
    }
    to_global(c.mask)[w] = bits;
  }
  if (nulls) atomicAdd(&block_nulls, nulls);
  __syncthreads();
  if (threadIdx.x == 0 && block_nulls && c.nulls) atom_add(c.nulls, (unsigned long long)block_nulls);
}

// ---- cardinality estimate: stratified sample into a small table

struct SampleParams {
  AggParams agg;              // key description; table/capacity unused
  unsigned long long* hdr;    // sample table headers
  uint32_t* cnt;              // occurrences per slot; cnt[scap] = the special key
  int64_t scap;
  int64_t samples;
  unsigned long long* stats;  // [0] distinct, [1] f1, [2] f2, [3] sampled rows, [4] keys found, [5..] keys
};

// Occurrence counts are only told apart up to kHotMin (distinct keys, keys
// seen once and twice, hot keys): a count already there is not incremented,
// so that the few keys of a low-cardinality input do not serialise the
// sample on their counters.
__device__ __forceinline__ void sample_count(uint32_t* c)
{
  if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < kHotMin) atom_add(c, 1u);
}

__global__ __launch_bounds__(kBlock) void sample_insert_kernel(const SampleParams s)
{
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const AggParams& p = s.agg;
  const int64_t span = p.num_rows / s.samples;
  int64_t row = j * span + (span > 1 ? (int64_t)(fmix32((uint32_t)j * 0x9e3779b9u + 1u) % (uint32_t)span) : 0);
  const bool valid = j < s.samples && row < p.num_rows && key_row_valid(p, row);
  const int n = __syncthreads_count(valid);  // sampled rows: one add per workgroup
  if (threadIdx.x == 0 && n) atom_add(&s.stats[3], (unsigned long long)n);
  if (!valid) return;
  const RowKey rk = row_key(p, row);
  if (rk.special) {
    sample_count(&s.cnt[s.scap]);
    return;
  }
  const uint64_t empty = p.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef;
  int64_t idx = reduce_range(rk.hash, s.scap);
  for (int64_t probe = 0; probe < s.scap; ++probe) {
    uint64_t cur = s.hdr[idx];
    for (;;) {
      if (cur == empty) {
        uint64_t seen = atom_cas(&s.hdr[idx], empty, rk.key);
        if (seen == empty) {
          sample_count(&s.cnt[idx]);
          return;
        }
        cur = seen;
        continue;
      }
      bool eq = p.key_mode == KEY_INT
                  ? cur == rk.key
                  : ((uint32_t)(cur >> 32) == rk.hash &&
                     rows_equal(p.keys, p.num_keys, p.compare_key_nulls != 0, (int64_t)(cur & 0xffffffffull), row));
      if (eq) {
        sample_count(&s.cnt[idx]);
        return;
      }
      break;
    }
    idx = (idx + 1 == s.scap) ? 0 : idx + 1;
  }
}

// Clears the sample table: empty headers, zero counts and statistics (and
// the hot set's count, when there is one).
__global__ __launch_bounds__(kBlock) void sample_clear_kernel(unsigned long long* hdr, uint32_t* cnt, int64_t scap,
                                                              unsigned long long empty, unsigned long long* stats,
                                                              int nstats, HotSet* hot)
{
  const int64_t step = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i <= scap; i += step) {
    if (i < scap) hdr[i] = empty;
    cnt[i] = 0u;
    if (i < nstats) stats[i] = 0ull;
    if (i == 0 && hot) hot->n = 0u;
  }
}

// The key columns of a compact call whose hot keys are collected: widths
// (k1 = 0: one column) and, for two columns, the columns themselves (the
// sample table then holds row references).
struct HotSpec {
  int32_t k0 = 0, k1 = 0;
  KeyCol c0{}, c1{};
};

// The sample table's summary: distinct keys (occupied slots, the special
// record included: n = scap + 1 counters), keys seen once and twice, into
// stats[0..2]; with `keys`, the distinct integer keys (slot headers), up to
// kTinyKeys of them, into keys[0..] and their number into *count. Each
// workgroup strides over the counters, reduces its counts and keeps its
// first kTinyKeys keys in LDS, then adds to the totals once (the same
// counters written by every workgroup are the slow atomics here).
// With `hot`, keys counted kHotMin times are added to the hot set (their raw
// column values: from the header for one column, from the sampled row for
// two); keys past kHotMax are left out.
constexpr int kSampleSummaryBlocks = 64;
__global__ __launch_bounds__(kBlock) void sample_summary_kernel(const unsigned long long* hdr, const uint32_t* cnt,
                                                               int64_t scap, unsigned long long* stats,
                                                               unsigned long long* count, unsigned long long* keys,
                                                               HotSet* hot, const HotSpec hs)
{
  __shared__ uint32_t red[3][kBlock / 64];
  __shared__ uint32_t nkeys;
  __shared__ unsigned long long kbuf[kTinyKeys];
  if (threadIdx.x == 0) nkeys = 0;
  __syncthreads();
  uint32_t d = 0, f1 = 0, f2 = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i <= scap; i += (int64_t)gridDim.x * kBlock) {
    const uint32_t c = cnt[i];
    d += c > 0;
    f1 += c == 1;
    f2 += c == 2;
    if (keys && c > 0 && i < scap) {
      const uint32_t j = atomicAdd(&nkeys, 1u);
      if (j < (uint32_t)kTinyKeys) kbuf[j] = hdr[i];
    }
    if (hot && c >= kHotMin && i < scap) {
      const uint64_t h = hdr[i];
      HotKey k;
      if (hs.k1 == 0) {
        k.c0 = hs.k0 == 8 ? h : (h & 0xffffffffull);
        k.c1 = 0;
      } else {
        const int64_t row = (int64_t)(h & 0xffffffffull);
        k.c0 = load_at(hs.c0.data, hs.k0, hs.c0.stride, row);
        k.c1 = load_at(hs.c1.data, hs.k1, hs.c1.stride, row);
      }
      const uint32_t j = atomicAdd(&hot->n, 1u);
      if (j < (uint32_t)kHotMax) hot->keys[j] = k;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    d += __shfl_down(d, o, 64);
    f1 += __shfl_down(f1, o, 64);
    f2 += __shfl_down(f2, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = d;
    red[1][threadIdx.x >> 6] = f1;
    red[2][threadIdx.x >> 6] = f2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long v = 0;
    for (int w = 0; w < kBlock / 64; ++w) v += red[threadIdx.x][w];
    if (v) atom_add(&stats[threadIdx.x], v);
  }
  if (keys && threadIdx.x == 0 && nkeys) {
    const unsigned long long base = atomicAdd(count, (unsigned long long)nkeys);
    const uint32_t kept = nkeys < (uint32_t)kTinyKeys ? nkeys : (uint32_t)kTinyKeys;
    for (uint32_t j = 0; j < kept && base + j < (unsigned long long)kTinyKeys; ++j) keys[base + j] = kbuf[j];
  }
}

// ---- cardinality estimate: HyperLogLog over every row's key hash

constexpr int kHllBits = 14;
constexpr int kHllRegs = 1 << kHllBits;

// Per-workgroup registers: register = top 14 bits of the row hash, value =
// leading zeros + 1 of a bijective remix of the hash. The same pass counts
// each chunk's rows per fine bucket (top kFineBits bits of the hash),
// fine[chunk * kFineBuckets + bucket]: partitions of up to 2^kFineBits take
// their histogram from these counts (hist_coarsen_kernel) instead of a
// second pass over the keys. Workgroup b covers rows [b * chunk, +chunk), the
// chunks of the partitioned execution.
constexpr int kFineBits = 13;  // up to 8192 partitions (kLeanMaxLog2P)
constexpr int kFineBuckets = 1 << kFineBits;

// One row's hash into the registers and the fine counts.
__device__ __forceinline__ void hll_add(uint32_t* regs, uint32_t* counts, uint32_t h)
{
  const uint32_t rank = (uint32_t)__clz(fmix32(h ^ 0x9e3779b9u)) + 1u;
  uint32_t* r = &regs[h >> (32 - kHllBits)];
  // registers saturate quickly: most rows only read
  if (rank > __hip_atomic_load(r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) atomicMax(r, rank);
  atomicAdd(&counts[h >> (32 - kFineBits)], 1u);
}

// The rows of a single key column of W bytes (4 or 8) without null
// grouping: kItems rows per thread per step, the next step's keys (and
// null-mask words) loaded before the current ones are hashed.
template <int W>
__device__ __forceinline__ void hll_rows_single(const AggParams& p, uint32_t* regs, uint32_t* counts, int64_t begin,
                                                int64_t end, const HotLds& hl)
{
  constexpr int kItems = 8;
  const KeyCol& k = p.keys[0];
  const bool nulls = p.skip_null_keys && k.mask;
  uint64_t cur[kItems], nxt[kItems];
  uint32_t cm[kItems], nm[kItems];
  auto load = [&](int64_t row0, uint64_t (&kv)[kItems], uint32_t (&mv)[kItems]) {
#pragma unroll
    for (int u = 0; u < kItems; ++u) {
      int64_t r = row0 + (int64_t)u * kLdsBlock;
      r = r < end ? r : end - 1;
      if constexpr (W == 8) kv[u] = to_global(reinterpret_cast<const uint64_t*>(k.data))[r];
      else kv[u] = to_global(reinterpret_cast<const uint32_t*>(k.data))[r];
      mv[u] = nulls ? to_global(k.mask)[(k.bit_off + r) >> 5] : ~0u;
    }
  };
  if (begin < end) load(begin + threadIdx.x, cur, cm);
  for (int64_t row0 = begin + threadIdx.x; row0 < end; row0 += (int64_t)kItems * kLdsBlock) {
    load(row0 + (int64_t)kItems * kLdsBlock, nxt, nm);
    uint32_t h[kItems];
    bool in[kItems];
#pragma unroll
    for (int u = 0; u < kItems; ++u) {
      const int64_t r = row0 + (int64_t)u * kLdsBlock;
      in[u] = r < end && ((cm[u] >> ((k.bit_off + r) & 31)) & 1u);
      h[u] = hash_combine(0u, murmur3_bits(normalize_float_bits(cur[u], k.vclass), W, 0));
    }
    if (hl.n != 0) {
      // hot rows: not partitioned
      const uint64_t zero[kItems] = {};
      int hi[kItems];
      hot_find_n<kItems>(hl, h, cur, zero, in, hi);
#pragma unroll
      for (int u = 0; u < kItems; ++u) in[u] = in[u] && hi[u] < 0;
    }
#pragma unroll
    for (int u = 0; u < kItems; ++u)
      if (in[u]) hll_add(regs, counts, h[u]);
#pragma unroll
    for (int u = 0; u < kItems; ++u) {
      cur[u] = nxt[u];
      cm[u] = nm[u];
    }
  }
}

// The rows of two key columns of W0 and W1 bytes (4 or 8 each) without null
// grouping, as hll_rows_single: both columns' values (and null-mask words)
// of the next step loaded before the current ones are hashed; the hot check
// takes the values already in registers.
template <int W0, int W1>
__device__ __forceinline__ void hll_rows_pair(const AggParams& p, uint32_t* regs, uint32_t* counts, int64_t begin,
                                              int64_t end, const HotLds& hl)
{
  constexpr int kItems = 4;
  const KeyCol& k0 = p.keys[0];
  const KeyCol& k1 = p.keys[1];
  const bool n0 = p.skip_null_keys && k0.mask, n1 = p.skip_null_keys && k1.mask;
  uint64_t a[kItems], b[kItems], na[kItems], nb[kItems];
  uint32_t ma[kItems], mb[kItems], nma[kItems], nmb[kItems];
  auto load = [&](int64_t row0, uint64_t (&va)[kItems], uint64_t (&vb)[kItems], uint32_t (&wa)[kItems],
                  uint32_t (&wb)[kItems]) {
#pragma unroll
    for (int u = 0; u < kItems; ++u) {
      int64_t r = row0 + (int64_t)u * kLdsBlock;
      r = r < end ? r : end - 1;
      if constexpr (W0 == 8) va[u] = to_global(reinterpret_cast<const uint64_t*>(k0.data))[r];
      else va[u] = to_global(reinterpret_cast<const uint32_t*>(k0.data))[r];
      if constexpr (W1 == 8) vb[u] = to_global(reinterpret_cast<const uint64_t*>(k1.data))[r];
      else vb[u] = to_global(reinterpret_cast<const uint32_t*>(k1.data))[r];
      wa[u] = n0 ? to_global(k0.mask)[(k0.bit_off + r) >> 5] : ~0u;
      wb[u] = n1 ? to_global(k1.mask)[(k1.bit_off + r) >> 5] : ~0u;
    }
  };
  if (begin < end) load(begin + threadIdx.x, a, b, ma, mb);
  for (int64_t row0 = begin + threadIdx.x; row0 < end; row0 += (int64_t)kItems * kLdsBlock) {
    load(row0 + (int64_t)kItems * kLdsBlock, na, nb, nma, nmb);
    uint32_t h[kItems];
    bool in[kItems];
#pragma unroll
    for (int u = 0; u < kItems; ++u) {
      const int64_t r = row0 + (int64_t)u * kLdsBlock;
      in[u] = r < end && ((ma[u] >> ((k0.bit_off + r) & 31)) & 1u) && ((mb[u] >> ((k1.bit_off + r) & 31)) & 1u);
      h[u] = hash_combine(hash_combine(0u, murmur3_bits(normalize_float_bits(a[u], k0.vclass), W0, 0)),
                          murmur3_bits(normalize_float_bits(b[u], k1.vclass), W1, 0));
    }
    if (hl.n != 0) {
      int hi[kItems];
      hot_find_n<kItems>(hl, h, a, b, in, hi);
#pragma unroll
      for (int u = 0; u < kItems; ++u) in[u] = in[u] && hi[u] < 0;
    }
#pragma unroll
    for (int u = 0; u < kItems; ++u)
      if (in[u]) hll_add(regs, counts, h[u]);
#pragma unroll
    for (int u = 0; u < kItems; ++u) {
      a[u] = na[u];
      b[u] = nb[u];
      ma[u] = nma[u];
      mb[u] = nmb[u];
    }
  }
}

// Whether a sample (SampleParams::stats) holds every group of its call: every
// sampled key seen at least twice, and few distinct keys. The estimate pass
// is then not needed; it is launched beside the sample and returns at once.
__device__ __forceinline__ bool sample_saturated(const unsigned long long* stats)
{
  return stats && stats[1] == 0 && stats[0] * 8 <= stats[3];
}

// Workgroup w takes chunks w, w + gridDim.x, ... (nw chunks of `chunk`
// rows): each chunk's fine counts are written when it is done, the
// registers once at the end (block_regs[w]). With `hot`, rows of hot keys
// are left out of both (the compact scatter aggregates them apart).
__global__ __launch_bounds__(kLdsBlock) void hll_hist_kernel(const AggParams p, uint8_t* block_regs, uint32_t* fine,
                                                            int64_t chunk, int32_t nw, const unsigned long long* gate,
                                                            const HotSet* hot, const HotSpec hs)
{
  if (sample_saturated(gate)) return;
  extern __shared__ uint32_t regs[];  // kHllRegs registers, kFineBuckets counts, then the hot keys' lookup
  uint32_t* counts = regs + kHllRegs;
  for (int i = threadIdx.x; i < kHllRegs + kFineBuckets; i += kLdsBlock) regs[i] = 0;
  HotLds hl{nullptr, nullptr, 0};
  if (hot) {
    HotKey* hk = reinterpret_cast<HotKey*>(counts + kFineBuckets);
    hl = hot_load(hot, hk, reinterpret_cast<uint32_t*>(hk + kHotMax), hs.k0, hs.k1);
  }
  __syncthreads();
  const bool single = p.num_keys == 1 && !(p.key_mode != KEY_INT && p.compare_key_nulls) && p.keys[0].stride == p.keys[0].width;
  // two 4- or 8-byte columns, nulls not grouped: their own row loop
  const bool pair = p.num_keys == 2 && !p.compare_key_nulls && p.keys[0].stride == p.keys[0].width &&
                    p.keys[1].stride == p.keys[1].width && (p.keys[0].width == 4 || p.keys[0].width == 8) &&
                    (p.keys[1].width == 4 || p.keys[1].width == 8);
  const int pair_shape = pair ? (p.keys[0].width == 8) * 2 + (p.keys[1].width == 8) : -1;
  for (int c = blockIdx.x; c < nw; c += gridDim.x) {
    const int64_t begin = (int64_t)c * chunk;
    const int64_t end = begin + chunk < p.num_rows ? begin + chunk : p.num_rows;
    if (single && p.keys[0].width == 8) {
      hll_rows_single<8>(p, regs, counts, begin, end, hl);
    } else if (single && p.keys[0].width == 4) {
      hll_rows_single<4>(p, regs, counts, begin, end, hl);
    } else if (pair_shape == 0) {
      hll_rows_pair<4, 4>(p, regs, counts, begin, end, hl);
    } else if (pair_shape == 1) {
      hll_rows_pair<4, 8>(p, regs, counts, begin, end, hl);
    } else if (pair_shape == 2) {
      hll_rows_pair<8, 4>(p, regs, counts, begin, end, hl);
    } else if (pair_shape == 3) {
      hll_rows_pair<8, 8>(p, regs, counts, begin, end, hl);
    } else {
      constexpr int kItems = 4;  // rows per thread per step, hashed column by column
      for (int64_t row0 = begin + threadIdx.x; row0 < end; row0 += (int64_t)kItems * kLdsBlock) {
        bool present[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
          present[k] = row0 + k * kLdsBlock < end;
          if (present[k] && p.skip_null_keys) present[k] = key_row_valid(p, row0 + k * kLdsBlock);
        }
        uint32_t h[kItems];
        uint64_t key0[kItems];
        hash_rows(p, row0, kLdsBlock, present, h, key0);
        if (hl.n != 0) {
          // two-column compact keys: hot rows are not partitioned
          uint64_t c0[kItems], c1[kItems];
#pragma unroll
          for (int k = 0; k < kItems; ++k) {
            const int64_t r = present[k] ? row0 + k * kLdsBlock : 0;
            c0[k] = load_at(hs.c0.data, hs.k0, hs.c0.stride, r);
            c1[k] = hs.k1 ? load_at(hs.c1.data, hs.k1, hs.c1.stride, r) : 0ull;
          }
          int hi[kItems];
          hot_find_n<kItems>(hl, h, c0, c1, present, hi);
#pragma unroll
          for (int k = 0; k < kItems; ++k) present[k] = present[k] && hi[k] < 0;
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k)
          if (present[k]) hll_add(regs, counts, h[k]);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kFineBuckets; i += kLdsBlock) {
      fine[(int64_t)c * kFineBuckets + i] = counts[i];
      counts[i] = 0;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < kHllRegs; i += kLdsBlock)
    block_regs[(int64_t)blockIdx.x * kHllRegs + i] = (uint8_t)regs[i];
}

// Histogram of 2^log2p partitions (log2p <= kFineBits) from the fine bucket
// counts: hist[p * nw + c] = sum of the partition's buckets in chunk c. A
// workgroup sums a tile of 64 chunks by 64 partitions (each chunk's buckets
// read contiguously) and writes it transposed through LDS.
__global__ __launch_bounds__(kBlock) void hist_coarsen_kernel(const uint32_t* fine, int32_t nw, int32_t log2p,
                                                             uint32_t* hist)
{
  __shared__ uint32_t tile[64][65];
  const int P = 1 << log2p;
  const int per = kFineBuckets >> log2p;
  const int c0 = blockIdx.x * 64, p0 = blockIdx.y * 64;
  for (int idx = threadIdx.x; idx < 64 * 64; idx += kBlock) {
    const int ci = idx >> 6, pi = idx & 63;
    const int c = c0 + ci, part = p0 + pi;
    uint32_t sum = 0;
    if (c < nw && part < P)
      for (int j = 0; j < per; ++j) sum += fine[(int64_t)c * kFineBuckets + part * per + j];
    tile[pi][ci] = sum;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 64 * 64; idx += kBlock) {
    const int pi = idx >> 6, ci = idx & 63;
    const int c = c0 + ci, part = p0 + pi;
    if (c < nw && part < P) hist[(int64_t)part * nw + c] = tile[pi][ci];
  }
}

// The same histogram when each partition spans at least 64 fine buckets
// (log2p <= kFineBits - 6, so at most kBlock partitions): one workgroup per
// chunk; each wave reads 64 consecutive buckets (one partition's), adds them
// across the wave and one lane adds the sum to the partition's LDS count.
// (The tiled kernel's threads would each sum thousands of buckets in turn.)
__global__ __launch_bounds__(kBlock) void hist_coarsen_few_kernel(const uint32_t* fine, int32_t nw, int32_t log2p,
                                                                 uint32_t* hist)
{
  __shared__ uint32_t s[kBlock];
  const int P = 1 << log2p;
  const int shift = kFineBits - log2p;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < P; i += kBlock) s[i] = 0u;
  __syncthreads();
  const int c = blockIdx.x;
  const uint32_t* f = fine + (int64_t)c * kFineBuckets;
  for (int g = w; g < kFineBuckets / 64; g += kBlock / 64) {
    uint32_t v = f[g * 64 + lane];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0 && v) atomicAdd(&s[(g * 64) >> shift], v);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += kBlock) hist[(int64_t)i * nw + c] = s[i];
}

// Max over workgroups per register, then per block the sum of 2^-M and the
// number of empty registers. A workgroup of 1024 threads covers 128 registers
// (32 words of 4 byte-registers): its 32 groups of 32 threads each take every
// 32nd workgroup's words, and the 32 partial maxima meet in LDS.
constexpr int kHllReduceRegs = 128;
__global__ __launch_bounds__(1024) void hll_reduce_kernel(const uint8_t* block_regs, int nb, double* sums,
                                                         uint32_t* zeros, const unsigned long long* gate)
{
  if (sample_saturated(gate)) return;
  __shared__ uint32_t part[32][32][4];
  __shared__ double s[32];
  __shared__ uint32_t z[32];
  const int wi = threadIdx.x & 31, sl = threadIdx.x >> 5;
  const int64_t word = (int64_t)blockIdx.x * 32 + wi;  // u32 word of the register array
  const global_t<const uint32_t>* w32 = to_global(reinterpret_cast<const uint32_t*>(block_regs));
  uint32_t m[4] = {0u, 0u, 0u, 0u};
  for (int b = sl; b < nb; b += 32) {
    const uint32_t x = w32[(int64_t)b * (kHllRegs / 4) + word];
#pragma unroll
    for (int k = 0; k < 4; ++k) m[k] = max(m[k], (x >> (8 * k)) & 0xffu);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) part[sl][wi][k] = m[k];
  __syncthreads();
  if (sl == 0) {
    double sum = 0.0;
    uint32_t zero = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uint32_t r = 0;
      for (int j = 0; j < 32; ++j) r = max(r, part[j][wi][k]);
      sum += ldexp(1.0, -(int)r);
      zero += r == 0;
    }
    s[wi] = sum;
    z[wi] = zero;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0;
    uint32_t zero = 0;
    for (int j = 0; j < 32; ++j) {
      sum += s[j];
      zero += z[j];
    }
    sums[blockIdx.x] = sum;
    zeros[blockIdx.x] = zero;
  }
}

__global__ __launch_bounds__(kBlock) void fill_u64_kernel(unsigned long long* p, int64_t n, unsigned long long v)
{
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) p[i] = v;
}

// ---- hash partitioning

// Rows per partition of each workgroup's contiguous slice -> hist[p][block].
__global__ __launch_bounds__(kBlock) void part_hist_kernel(const PartParams pp)
{
  extern __shared__ uint32_t counts[];
  hist_chunk(pp, blockIdx.x, counts);
}

// Writes one row as a partition record: key and value columns at their
// record offsets (native width), value validity as bits of a flags word.
__device__ __forceinline__ void write_record(const PartParams& pp, uint8_t* rec, int64_t row)
{
  const AggParams& p = pp.src;
  for (int c = 0; c < p.num_keys; ++c) {
    const KeyCol& k = p.keys[c];
    store_width(rec + pp.key_off[c], 0, k.width, load_at(k.data, k.width, k.stride, row));
  }
  uint32_t flags = 0;
  for (int c = 0; c < p.num_vals; ++c) {
    const ValCol& v = p.vals[c];
    store_width(rec + pp.val_off[c], 0, v.width, load_at(v.data, v.width, v.stride, row));
    if (v.flag_bit >= 0 && bit_valid(v.mask, v.bit_off + row)) flags |= 1u << v.flag_bit;
  }
  if (pp.flag_bytes) store_width(rec + pp.flags_off, 0, pp.flag_bytes, flags);
  if (pp.rowid_off >= 0) store_width(rec + pp.rowid_off, 0, 4, (uint64_t)row);
}

// Copies every row with a valid key to its partition's range of this
// workgroup (offsets[p][block] onward) as a record.
__global__ __launch_bounds__(kBlock) void part_scatter_kernel(const PartParams pp)
{
  extern __shared__ uint32_t cursor[];
  const AggParams& p = pp.src;
  const int P = 1 << pp.log2p;
  for (int i = threadIdx.x; i < P; i += kBlock) cursor[i] = pp.offsets[(int64_t)i * pp.nblocks + blockIdx.x];
  __syncthreads();
  const int64_t begin = (int64_t)blockIdx.x * pp.chunk;
  const int64_t end = begin + pp.chunk < p.num_rows ? begin + pp.chunk : p.num_rows;
  for (int64_t row = begin + threadIdx.x; row < end; row += kBlock) {
    if (!key_row_valid(p, row)) continue;
    const uint32_t part = partition_of(row_key(p, row).hash, pp.log2p);
    const int64_t dst = atomicAdd(&cursor[part], 1u);
    write_record(pp, pp.out + dst * (int64_t)pp.rec_bytes, row);
  }
}

__global__ __launch_bounds__(kLdsBlock) void part_scatter_tiled_kernel(const PartParams pp)
{
  extern __shared__ __align__(16) uint8_t smem[];
  scatter_chunk(pp, blockIdx.x, smem);
}

// ---- exclusive scan of u32 counts (values and total below 2^32)

__global__ __launch_bounds__(256) void scan_reduce_u32_kernel(const uint32_t* in, int64_t n, uint32_t* bsum)
{
  __shared__ uint32_t red[256];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint32_t s = 0;
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += in[base + k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if (threadIdx.x < (unsigned)d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(1024) void scan_bsums_u32_kernel(uint32_t* bsum, int64_t n, unsigned long long* total)
{
  __shared__ uint32_t buf[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < n; base += 1024) {
    int64_t i = base + threadIdx.x;
    uint32_t v = i < n ? bsum[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      uint32_t t = threadIdx.x >= (unsigned)d ? buf[threadIdx.x - d] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n) bsum[i] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = (unsigned long long)carry;
}

__global__ __launch_bounds__(256) void scan_apply_u32_kernel(const uint32_t* in, int64_t n, const uint32_t* bsum,
                                                            uint32_t* out)
{
  __shared__ uint32_t buf[256];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    s += v[k];
  }
  buf[threadIdx.x] = s;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    uint32_t t = threadIdx.x >= (unsigned)d ? buf[threadIdx.x - d] : 0;
    __syncthreads();
    buf[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = bsum[blockIdx.x] + buf[threadIdx.x] - s;
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}


// ---- LDS aggregation of one work item

// Work items of the partitioned execution: partition p (rows
// [offsets[p*nblocks], offsets[(p+1)*nblocks])) is cut into
// ceil(rows / max_item_rows) items. count[p] = its number of items;
// count[P] = 0 so that an exclusive scan of P+1 entries ends with the total.
__global__ __launch_bounds__(kBlock) void count_items_kernel(const uint32_t* offsets, int nblocks, int P,
                                                             const unsigned long long* total_rows,
                                                             int64_t max_item_rows, uint32_t* count)
{
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p > P) return;
  if (p == P) {
    count[P] = 0;
    return;
  }
  const int64_t b = offsets[(int64_t)p * nblocks];
  const int64_t e = p + 1 < P ? (int64_t)offsets[(int64_t)(p + 1) * nblocks] : (int64_t)*total_rows;
  count[p] = (uint32_t)((e - b + max_item_rows - 1) / max_item_rows);
}

// The rows of work item `b` of the partitioned mode: [begin, end) of the
// partitioned copy; `merge` when its partition is split over several items,
// which are the `fold_parts` items of that partition (this one is
// `fold_part`). False when `b` is past the last item.
__device__ __forceinline__ bool part_item(const LdsParams& lp, uint32_t b, int64_t& begin, int64_t& end, bool& merge,
                                          int& fold_part, int& fold_parts)
{
  const int P = 1 << lp.log2p;
  if (b >= lp.item_offsets[P]) return false;  // the grid is an upper bound on the items
  int lo = 0, hi = P - 1;                    // last partition whose first item is <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (lp.item_offsets[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const int part = lo;
  const uint32_t n_items = lp.item_offsets[part + 1] - lp.item_offsets[part];
  const int64_t ps = lp.offsets[(int64_t)part * lp.nblocks];
  const int64_t pe = part + 1 < P ? (int64_t)lp.offsets[(int64_t)(part + 1) * lp.nblocks] : (int64_t)*lp.total_rows;
  begin = ps + (int64_t)(b - lp.item_offsets[part]) * lp.max_item_rows;
  end = begin + lp.max_item_rows < pe ? begin + lp.max_item_rows : pe;
  merge = n_items > 1;
  fold_part = (int)(b - lp.item_offsets[part]);
  fold_parts = (int)n_items;
  return true;
}

// One workgroup aggregates one work item in an LDS table. Rows whose group
// does not fit the table go to the global table. At the end the item either
// writes its groups as final output rows (a partition handled by this item
// alone, without spill), or folds them into the global table (direct mode,
// partitions split over several items, or an item that spilled), whose
// groups are finalised afterwards. kPass2 (partitioned mode): an item that
// writes its groups first runs the second pass over its rows against its LDS
// table; every item records whether its groups went to the global table
// (their second pass is pass2_items_kernel's).
template <bool kPass2>
__global__ __launch_bounds__(kLdsBlock) void lds_agg_kernel(const LdsParams lp)
{
  extern __shared__ __align__(16) uint8_t lds[];
  const AggParams& p = lp.src;

  int64_t begin, end;
  bool merge = true;
  int fold_part = 0, fold_parts = 1;
  if (lp.direct) {
    begin = (int64_t)blockIdx.x * lp.rows_per_item;
    end = begin + lp.rows_per_item < p.num_rows ? begin + lp.rows_per_item : p.num_rows;
  } else if (!part_item(lp, blockIdx.x, begin, end, merge, fold_part, fold_parts)) {
    return;
  }

  const LdsTable t = lds_table_at(lds, lp.slots, lp.words, lp.cap);
  lds_table_init(t, lp.tmpl, lp.words);
  lds_aggregate_rows(p, t, lds, begin, end, lp.log2p, lp.direct != 0);
  const bool fold = merge || *t.overflow;
  if constexpr (kPass2) {
    if (threadIdx.x == 0) lp.p2->item_global[blockIdx.x] = fold ? 1u : 0u;
    if (!fold) lds_pass2_rows(lp.p2->rows, lp.p2, t, begin, end, lp.log2p, false);
  }
  if (fold) lds_fold_to_global(p, t, fold_part, fold_parts);
  else lds_write_groups(p, t, &lp.out_counters[0], 0, *lp.out);
}

// The few-groups direct execution: each workgroup aggregates its slice of
// rows against the candidate groups (tiny_aggregate) and folds its table
// into the global table.
template <int NC, int KW, int TV>
__global__ __launch_bounds__(kLdsBlock) void tiny_agg_kernel(const LdsParams lp, const TinyParams* __restrict__ tp,
                                                             int32_t tiny_off)
{
  extern __shared__ __align__(16) uint8_t lds[];
  const AggParams& p = lp.src;
  const int64_t begin = (int64_t)blockIdx.x * lp.rows_per_item;
  const int64_t end = begin + lp.rows_per_item < p.num_rows ? begin + lp.rows_per_item : p.num_rows;
  const LdsTable t = lds_table_at(lds, lp.slots, lp.words, lp.cap);
  lds_table_init(t, lp.tmpl, lp.words);
  tiny_aggregate<NC, KW, TV>(p, t, lds, (uint32_t)tiny_off, lp.tmpl, lp.words, tp, begin, end);
  lds_fold_to_global(p, t);
}

// Second pass of the partitioned items whose groups went to the global
// table (split partitions, spilled items), once every item has folded.
__global__ __launch_bounds__(kLdsBlock) void pass2_items_kernel(const LdsParams lp)
{
  const AggParams& p = lp.src;
  int64_t begin, end;
  bool merge;
  int fold_part, fold_parts;
  if (!part_item(lp, blockIdx.x, begin, end, merge, fold_part, fold_parts)) return;
  const Pass2Ops* __restrict__ q = lp.p2;
  if (!q->item_global[blockIdx.x]) return;
  for (int64_t row = begin + threadIdx.x; row < end; row += kLdsBlock) {
    const uint32_t flags = p.flags ? load_flags(p.flags, p.flag_bytes, p.flag_stride, row) : 0u;
    const uint32_t src = q->rowid ? *reinterpret_cast<const uint32_t*>(q->rowid + row * (int64_t)q->rowid_stride) : 0u;
    uint8_t* g = global_find(p, row_key(p, row), row);
    if (g) pass2_update<false>(q, g, row, flags, src);
  }
}

// Second pass of the direct mode, once the global table holds every group:
// each workgroup copies the groups into its LDS table, runs the second pass
// over its slice of rows there, and folds the second-pass fields back.
__global__ __launch_bounds__(kLdsBlock) void pass2_direct_kernel(const LdsParams lp)
{
  extern __shared__ __align__(16) uint8_t lds[];
  const AggParams& p = lp.src;
  const int64_t begin = (int64_t)blockIdx.x * lp.rows_per_item;
  const int64_t end = begin + lp.rows_per_item < p.num_rows ? begin + lp.rows_per_item : p.num_rows;
  const LdsTable t = lds_table_at(lds, lp.slots, lp.words, lp.cap);
  lds_table_init(t, lp.tmpl, lp.words);
  lds_load_global(p, t);
  lds_pass2_rows(lp.p2->rows, lp.p2, t, begin, end, 0, true);
  lds_fold_pass2(lp.p2->rows, lp.p2, t);
}

// ---- table hashing entry points

struct HashParams {
  KeyCol cols[kMaxKeys];
  int32_t num_cols;
  int32_t check_nulls;
  uint32_t seed;
  int32_t pad;
  int64_t num_rows;
};

__global__ __launch_bounds__(kBlock) void murmur_table_kernel(const HashParams p, uint32_t* out)
{
  const int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (row >= p.num_rows) return;
  uint32_t h = p.seed;
  for (int c = 0; c < p.num_cols; ++c) {
    const KeyCol& k = p.cols[c];
    uint32_t e;
    if (p.check_nulls && !bit_valid(k.mask, k.bit_off + row)) {
      e = 0xffffffffu;
    } else {
      e = murmur3_bits(normalize_float_bits(load_at(k.data, k.width, k.stride, row), k.vclass), k.width, p.seed);
    }
    h = hash_combine(h, e);
  }
  out[row] = h;
}

__global__ __launch_bounds__(kBlock) void spark_murmur_table_kernel(const HashParams p, int32_t* out)
{
  const int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (row >= p.num_rows) return;
  uint32_t h = p.seed;
  for (int c = 0; c < p.num_cols; ++c) {
    const KeyCol& k = p.cols[c];
    if (!bit_valid(k.mask, k.bit_off + row)) continue;  // a null keeps the running hash
    uint64_t raw = load_at(k.data, k.width, k.stride, row);
    if (k.width < 4) {
      // bool / int8 / uint8 / int16 / uint16 are hashed as 4-byte integers
      uint32_t w = k.vclass == VC_SIGNED ? (uint32_t)(int32_t)widen_int(raw, k.width, k.vclass)
                   : k.vclass == VC_BOOL   ? (uint32_t)(raw != 0)
                                           : (uint32_t)raw;
      h = spark_murmur3_bits(w, 4, h);
    } else {
      if (k.vclass == VC_F32 || k.vclass == VC_F64) {
        // Spark normalises NaN only; -0.0 keeps its sign
        uint64_t nb = normalize_float_bits(raw, k.vclass);
        bool zero = k.vclass == VC_F32 ? ((uint32_t)raw & 0x7fffffffu) == 0 : (raw & 0x7fffffffffffffffull) == 0;
        raw = zero ? raw : nb;
      }
      h = spark_murmur3_bits(raw, k.width, h);
    }
  }
  out[row] = (int32_t)h;
}

// =================================================================== host

thread_local std::string g_last_error;

mgb_status fail(mgb_status s, std::string msg)
{
  g_last_error = std::move(msg);
  return s;
}

int type_width(int t)
{
  switch (t) {
    case MGB_TYPE_INT8: case MGB_TYPE_UINT8: case MGB_TYPE_BOOL8: return 1;
    case MGB_TYPE_INT16: case MGB_TYPE_UINT16: return 2;
    case MGB_TYPE_INT32: case MGB_TYPE_UINT32: case MGB_TYPE_FLOAT32: case MGB_TYPE_TIMESTAMP_DAYS:
    case MGB_TYPE_DURATION_DAYS: return 4;
    case MGB_TYPE_INT64: case MGB_TYPE_UINT64: case MGB_TYPE_FLOAT64: case MGB_TYPE_TIMESTAMP_SECONDS:
    case MGB_TYPE_TIMESTAMP_MILLISECONDS: case MGB_TYPE_TIMESTAMP_MICROSECONDS: case MGB_TYPE_TIMESTAMP_NANOSECONDS:
    case MGB_TYPE_DURATION_SECONDS: case MGB_TYPE_DURATION_MILLISECONDS: case MGB_TYPE_DURATION_MICROSECONDS:
    case MGB_TYPE_DURATION_NANOSECONDS: return 8;
    default: return 0;
  }
}
bool is_float_t(int t) { return t == MGB_TYPE_FLOAT32 || t == MGB_TYPE_FLOAT64; }
bool is_unsigned_t(int t) { return t >= MGB_TYPE_UINT8 && t <= MGB_TYPE_UINT64; }
bool is_signed_int_t(int t) { return t >= MGB_TYPE_INT8 && t <= MGB_TYPE_INT64; }
bool is_integral_t(int t) { return is_signed_int_t(t) || is_unsigned_t(t) || t == MGB_TYPE_BOOL8; }
bool is_timestamp_t(int t) { return t >= MGB_TYPE_TIMESTAMP_DAYS && t <= MGB_TYPE_TIMESTAMP_NANOSECONDS; }
bool is_duration_t(int t) { return t >= MGB_TYPE_DURATION_DAYS && t <= MGB_TYPE_DURATION_NANOSECONDS; }

int value_class(int t)
{
  if (t == MGB_TYPE_FLOAT32) return VC_F32;
  if (t == MGB_TYPE_FLOAT64) return VC_F64;
  if (t == MGB_TYPE_BOOL8) return VC_BOOL;
  if (is_unsigned_t(t)) return VC_UNSIGNED;
  return VC_SIGNED;  // signed ints, timestamps, durations
}

// cudf::detail::target_type (aggregation.hpp:1168-1424) for fixed-width types.
int32_t target_type(int32_t t, int32_t kind)
{
  if (type_width(t) == 0) return -1;
  switch (kind) {
    case MGB_AGG_MIN: case MGB_AGG_MAX: return t;
    case MGB_AGG_COUNT_VALID: case MGB_AGG_COUNT_ALL: return MGB_TYPE_INT32;
    case MGB_AGG_ARGMIN: case MGB_AGG_ARGMAX: return MGB_TYPE_INT32;
    case MGB_AGG_SUM:
      if (is_integral_t(t)) return MGB_TYPE_INT64;
      if (is_float_t(t) || is_duration_t(t)) return t;
      return -1;
    case MGB_AGG_PRODUCT: case MGB_AGG_SUM_OF_SQUARES:
      if (is_integral_t(t)) return MGB_TYPE_INT64;
      if (is_float_t(t)) return t;
      return -1;
    case MGB_AGG_MEAN:
      if (is_duration_t(t)) return t;
      if (is_timestamp_t(t)) return -1;
      return MGB_TYPE_FLOAT64;
    // sort-path aggregations (aggregation.hpp:1290-1330): quantiles of numeric
    // sources are double, NUNIQUE is size_type, NTH_ELEMENT keeps the type
    case MGB_AGG_MEDIAN: case MGB_AGG_QUANTILE:
      if (!is_integral_t(t) && !is_float_t(t)) return -1;
      return MGB_TYPE_FLOAT64;
    case MGB_AGG_NUNIQUE: return MGB_TYPE_INT32;
    case MGB_AGG_NTH_ELEMENT: return t;
    // COLLECT_LIST / COLLECT_SET: a list of the source type (aggregation.hpp:1335-1350)
    case MGB_AGG_COLLECT_LIST: case MGB_AGG_COLLECT_SET: return MGB_TYPE_LIST;
    case MGB_AGG_M2: case MGB_AGG_VARIANCE: case MGB_AGG_STD:
      // numeric sources only (multi_pass_kernels.cuh:59-63)
      if (!is_integral_t(t) && !is_float_t(t)) return -1;
      return MGB_TYPE_FLOAT64;
    default: return -1;
  }
}

std::mutex g_prof_mu;
bool g_prof_enabled = false;
std::vector<ProfRecord> g_prof;




constexpr size_t kStagingBytes = 64 * 1024;  // parameter staging (kUploadSlots slots)
constexpr int kUploadSlots = 8;

// Host memory a stream's calls share: pinned readback slots and the pinned
// source of parameter uploads, taken at a stream's first call and kept.
struct HostStaging {
  std::mutex mu;  // one call at a time per stream
  unsigned long long* pinned = nullptr;
  uint8_t* staging = nullptr;
  ~HostStaging()
  {
    if (pinned) (void)hipHostFree(pinned);
    if (staging) (void)hipHostFree(staging);
  }
};

std::mutex g_ws_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<HostStaging>> g_ws;

HostStaging* staging_for(hipStream_t s)
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto& w = g_ws[{dev, s}];
  if (!w) w.reset(new HostStaging());
  return w.get();
}

// One call's working memory: device temporaries (from the call's allocator
// for temporaries, or a fixed region) and the host staging it uses.
struct Workspace {
  CallMem mem;
  DeviceBuffer main{&mem};    // tables, partitioned columns, histograms, counters, sample table
  DeviceBuffer outs{&mem};    // temporary output columns
  DeviceBuffer params{&mem};  // kernel parameter blocks (output tables, identity templates)
  DeviceBuffer fine{&mem};    // fine bucket counts per chunk of the last key pass (hll_hist_kernel)
  DeviceBuffer sorts{&mem};   // row groups, sort keys and radix-sort space of the sort-path aggregations
  DeviceBuffer lists{&mem};   // COLLECT_LIST / COLLECT_SET results: offsets and elements
  DeviceBuffer hot{&mem};     // the call's hot keys (HotSet), when the compact execution may take it
  DeviceBuffer rowrec{&mem};  // sort path: every row's record in the global table (aggregate_kernel)
  bool fine_valid = false;  // `fine` holds the counts of this call's keys
  bool fine_hot = false;    // ... without the rows of the hot keys in `hot`
  int64_t fine_nw = 0, fine_chunk = 0;
  unsigned long long* pinned = nullptr;  // host readback area (kMaxPinned slots)
  uint8_t* staging = nullptr;            // host source of parameter uploads (kStagingBytes)
  int upload_slot = 0;                   // next staging / params slot of the call
};

#define MGB_RESERVE(buf, n)                                                                              \
  do {                                                                                                   \
    if ((buf).reserve((n), stream) != hipSuccess)                                                        \
      return fail(MGB_ERROR_ALLOC, "allocation of " + std::to_string((size_t)(n)) + " bytes of temporaries failed"); \
  } while (0)

constexpr int kMaxPinned = 1024;  // host readback slots (counters, null counts)
constexpr int kPinnedNulls = 128; // first slot of null counts read back with the group count
constexpr int64_t kNoAllocMaxRows = 1 << 23;  // rows of one mgb_groupby_aggregate_noalloc call

int device_cu_count()
{
  static int cus = [] {
    int dev = 0, n = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    return n;
  }();
  return cus;
}



int sum_op_for(int t) { return is_float_t(t) ? OP_SUM_F64 : OP_SUM_I64; }

int minmax_op_for(int t, bool is_min)
{
  int w = type_width(t);
  if (t == MGB_TYPE_FLOAT32) return is_min ? OP_MIN_F32 : OP_MAX_F32;
  if (t == MGB_TYPE_FLOAT64) return is_min ? OP_MIN_F64 : OP_MAX_F64;
  if (t == MGB_TYPE_UINT64) return is_min ? OP_MIN_U64 : OP_MAX_U64;
  if (w == 8) return is_min ? OP_MIN_I64 : OP_MAX_I64;
  if (is_unsigned_t(t) || t == MGB_TYPE_BOOL8) return is_min ? OP_MIN_U32 : OP_MAX_U32;
  return is_min ? OP_MIN_I32 : OP_MAX_I32;
}

mgb_status build_plan(const mgb_column* keys, int32_t num_keys, int32_t null_policy, const mgb_request* reqs,
                      int32_t nreq, Plan& plan)
{
  if (num_keys > kMaxKeys)
    return fail(MGB_ERROR_UNSUPPORTED, "more than " + std::to_string(kMaxKeys) + " key columns");
  for (int c = 0; c < num_keys; ++c) {
    const mgb_column& k = keys[c];
    int w = type_width(k.type_id);
    if (w == 0) return fail(MGB_ERROR_UNSUPPORTED, "key column type is not a fixed-width numeric type");
    KeyCol kc{};
    kc.data = static_cast<const uint8_t*>(k.data) + k.offset * w;
    kc.mask = (k.null_mask && k.null_count != 0) ? k.null_mask : nullptr;
    kc.bit_off = k.offset;
    kc.type_id = k.type_id;
    kc.width = w;
    kc.stride = w;
    kc.vclass = value_class(k.type_id);
    plan.key_has_nulls |= kc.mask != nullptr;
    plan.keys.push_back(kc);
  }
  plan.skip_null_keys = plan.key_has_nulls && null_policy == MGB_NULL_EXCLUDE;
  plan.compare_key_nulls = plan.key_has_nulls && null_policy == MGB_NULL_INCLUDE;
  bool int_key = num_keys == 1 && !is_float_t(plan.keys[0].type_id) && !plan.compare_key_nulls;
  plan.key_mode = int_key ? KEY_INT : KEY_ROWREF;

  // value columns, deduplicated by view (as cudf's result cache does)
  std::map<ValKey, int> col_index;
  struct Want {
    int col, kind, ddof;
    mgb_aggregation agg;
  };
  std::vector<Want> wants;
  for (int r = 0; r < nreq; ++r) {
    const mgb_column& v = reqs[r].values;
    int w = type_width(v.type_id);
    for (int a = 0; a < reqs[r].num_aggs; ++a) {
      int kind = reqs[r].aggs[a].kind;
      if (target_type(v.type_id, kind) < 0) return fail(MGB_ERROR_LOGIC, "Invalid type/aggregation combination.");
      switch (kind) {
        case MGB_AGG_SUM: case MGB_AGG_PRODUCT: case MGB_AGG_MIN: case MGB_AGG_MAX: case MGB_AGG_COUNT_VALID:
        case MGB_AGG_COUNT_ALL: case MGB_AGG_SUM_OF_SQUARES: case MGB_AGG_MEAN: case MGB_AGG_VARIANCE:
        case MGB_AGG_STD: case MGB_AGG_M2: case MGB_AGG_ARGMIN: case MGB_AGG_ARGMAX: break;
        case MGB_AGG_MEDIAN: case MGB_AGG_QUANTILE: case MGB_AGG_NUNIQUE: case MGB_AGG_NTH_ELEMENT:
        case MGB_AGG_COLLECT_LIST: case MGB_AGG_COLLECT_SET:
          plan.has_sort = true;
          break;
        default:
          return fail(MGB_ERROR_UNSUPPORTED,
                      "aggregation kind " + std::to_string(kind) + " is not served by this hash groupby");
      }
    }
    ValKey vk{v.data, v.null_mask, v.offset, v.size, v.type_id};
    auto it = col_index.find(vk);
    int ci;
    if (it == col_index.end()) {
      ci = (int)plan.cols.size();
      col_index[vk] = ci;
      PlanCol pc{};
      pc.type = v.type_id;
      pc.nullable = v.null_mask != nullptr && v.null_count != 0;
      pc.col.data = static_cast<const uint8_t*>(v.data) + v.offset * w;
      pc.col.mask = pc.nullable ? v.null_mask : nullptr;
      pc.col.bit_off = v.offset;
      pc.col.width = w;
      pc.col.stride = w;
      pc.col.vclass = value_class(v.type_id);
      plan.cols.push_back(pc);
    } else {
      ci = it->second;
    }
    for (int a = 0; a < reqs[r].num_aggs; ++a)
      wants.push_back({ci, reqs[r].aggs[a].kind, reqs[r].aggs[a].ddof, reqs[r].aggs[a]});
  }

  // ops per column (deduplicated by kind); the counts are added below
  bool need_rows = false;
  auto add_op = [](PlanCol& pc, int kind) {
    for (auto& o : pc.ops)
      if (o.first == kind) return;
    pc.ops.push_back({kind, -1});
  };
  auto add_p2 = [&plan](PlanCol& pc, int kind) {
    plan.has_pass2 = true;
    for (auto& o : pc.p2)
      if (o.first == kind) return;
    pc.p2.push_back({kind, -1});
  };
  for (auto& wt : wants) {
    PlanCol& pc = plan.cols[wt.col];
    switch (wt.kind) {
      case MGB_AGG_SUM: case MGB_AGG_MEAN: add_op(pc, sum_op_for(pc.type)); break;
      case MGB_AGG_PRODUCT: add_op(pc, is_float_t(pc.type) ? OP_PROD_F64 : OP_PROD_I64); break;
      case MGB_AGG_SUM_OF_SQUARES: add_op(pc, is_float_t(pc.type) ? OP_SUMSQ_F64 : OP_SUMSQ_I64); break;
      case MGB_AGG_MIN: add_op(pc, minmax_op_for(pc.type, true)); break;
      case MGB_AGG_MAX: add_op(pc, minmax_op_for(pc.type, false)); break;
      case MGB_AGG_COUNT_ALL: need_rows = true; break;
      // second-pass aggregations: the SUM (mean) or MIN / MAX of the first
      // pass, then a field the second pass fills (M2 is shared by every ddof)
      case MGB_AGG_VARIANCE: case MGB_AGG_STD: case MGB_AGG_M2:
        add_op(pc, sum_op_for(pc.type));
        add_p2(pc, OP_M2);
        break;
      case MGB_AGG_ARGMIN:
        add_op(pc, minmax_op_for(pc.type, true));
        add_p2(pc, OP_ARGMIN_IDX);
        break;
      case MGB_AGG_ARGMAX:
        add_op(pc, minmax_op_for(pc.type, false));
        add_p2(pc, OP_ARGMAX_IDX);
        break;
      default: break;
    }
    // COUNT_VALID = rows - nulls: one shared row count per group plus a null
    // count per nullable column (updated only for null values); the
    // sort-path kinds read neither
    if (wt.kind == MGB_AGG_MEDIAN || wt.kind == MGB_AGG_QUANTILE || wt.kind == MGB_AGG_NUNIQUE ||
        wt.kind == MGB_AGG_NTH_ELEMENT || wt.kind == MGB_AGG_COLLECT_LIST || wt.kind == MGB_AGG_COLLECT_SET)
      continue;
    need_rows = true;
    if (wt.kind != MGB_AGG_COUNT_ALL && pc.nullable) add_op(pc, OP_NULLS);
  }

  // record layout: 8-byte header, 8-byte fields, then 4-byte fields
  int off = 8;
  for (auto& pc : plan.cols) {
    for (auto& o : pc.ops)
      if (op_field_bytes(o.first) == 8) { o.second = off; off += 8; }
    for (auto& o : pc.p2)
      if (op_field_bytes(o.first) == 8) { o.second = off; off += 8; }
  }
  if (need_rows) { plan.rows_off = off; off += 4; }
  for (auto& pc : plan.cols) {
    for (auto& o : pc.ops)
      if (op_field_bytes(o.first) == 4) { o.second = off; off += 4; }
    for (auto& o : pc.p2)
      if (op_field_bytes(o.first) == 4) { o.second = off; off += 4; }
  }
  plan.stride = (int)align_up((size_t)off, 8);
  if ((plan.stride + 7) / 8 + 1 > kMaxRecordWords)  // room for the LDS padding word
    return fail(MGB_ERROR_UNSUPPORTED, "too many distinct aggregations for one groupby call");
  for (auto& pc : plan.cols) {
    pc.count_off = plan.rows_off;
    pc.nulls_off = pc.nullable ? pc.op_offset(OP_NULLS) : -1;
  }

  // outputs
  for (auto& wt : wants) {
    PlanCol& pc = plan.cols[wt.col];
    int t = pc.type, w = type_width(t);
    Plan::Out o{};
    o.col = wt.col;
    o.kind = wt.kind;
    o.out_type = target_type(t, wt.kind);
    OutSpec& s = o.spec;
    s.count_off = pc.count_off;
    s.nulls_off = pc.nulls_off;
    s.out_width = type_width(o.out_type);
    s.valid = reinterpret_cast<uint8_t*>(1);  // nullable marker; real pointer set at finalize
    switch (wt.kind) {
      case MGB_AGG_COUNT_VALID:
        s.fin = FIN_COUNT;
        s.valid = nullptr;
        break;
      case MGB_AGG_VARIANCE: case MGB_AGG_STD:
        s.field_off = pc.p2_offset(OP_M2);
        s.fin = wt.kind == MGB_AGG_VARIANCE ? FIN_VAR : FIN_STD;
        s.ddof = wt.ddof;
        break;
      case MGB_AGG_MEDIAN: case MGB_AGG_QUANTILE: case MGB_AGG_NUNIQUE: case MGB_AGG_NTH_ELEMENT:
        s.fin = FIN_NONE;  // the sort path writes it
        o.agg = wt.agg;
        break;
      case MGB_AGG_COLLECT_LIST: case MGB_AGG_COLLECT_SET:
        // a list column: the sort path writes its offsets and elements
        // (TempOuts::lists); the temporary output is unused
        s.fin = FIN_NONE;
        s.out_width = 4;
        s.valid = nullptr;
        o.agg = wt.agg;
        break;
      case MGB_AGG_M2:
        // sum of squared deviations from the MEAN (cpp/src/groupby/sort/group_m2.cu:40-52; cudf 24.04
        // serves it from its sort path only), null where MEAN is
        s.field_off = pc.p2_offset(OP_M2);
        s.fin = FIN_M2;
        break;
      case MGB_AGG_ARGMIN: case MGB_AGG_ARGMAX:
        s.field_off = pc.p2_offset(wt.kind == MGB_AGG_ARGMIN ? OP_ARGMIN_IDX : OP_ARGMAX_IDX);
        s.fin = FIN_ARGIDX;
        break;
      case MGB_AGG_COUNT_ALL:
        s.fin = FIN_COUNT;
        s.count_off = plan.rows_off;
        s.nulls_off = -1;
        s.valid = nullptr;
        break;
      case MGB_AGG_SUM: case MGB_AGG_PRODUCT: case MGB_AGG_SUM_OF_SQUARES: {
        int op = wt.kind == MGB_AGG_SUM ? sum_op_for(t)
                 : wt.kind == MGB_AGG_PRODUCT ? (is_float_t(t) ? OP_PROD_F64 : OP_PROD_I64)
                                              : (is_float_t(t) ? OP_SUMSQ_F64 : OP_SUMSQ_I64);
        s.field_off = pc.op_offset(op);
        s.fin = t == MGB_TYPE_FLOAT32 ? FIN_F64_TO_F32 : t == MGB_TYPE_FLOAT64 ? FIN_F64 : FIN_INT;
        break;
      }
      case MGB_AGG_MIN: case MGB_AGG_MAX: {
        int op = minmax_op_for(t, wt.kind == MGB_AGG_MIN);
        s.field_off = pc.op_offset(op);
        s.fin = t == MGB_TYPE_FLOAT32 ? FIN_F32_CODE : t == MGB_TYPE_FLOAT64 ? FIN_F64_CODE : FIN_INT;
        if (s.fin == FIN_INT && w < 8) s.out_width = w;
        break;
      }
      case MGB_AGG_MEAN:
        s.field_off = pc.op_offset(sum_op_for(t));
        s.fin = is_duration_t(t) ? FIN_MEAN_DURATION
                : t == MGB_TYPE_FLOAT32 ? FIN_MEAN_F32
                : t == MGB_TYPE_FLOAT64 ? FIN_MEAN_F64 : FIN_MEAN_I64;
        break;
      default: break;
    }
    plan.outs.push_back(o);
  }
  return MGB_SUCCESS;
}

// Identity template of one record (header + every op's identity).
void record_template(const Plan& plan, uint64_t header, uint64_t* words)
{
  std::vector<uint8_t> bytes((size_t)plan.stride, 0);
  std::memcpy(bytes.data(), &header, 8);
  for (auto& pc : plan.cols) {
    for (auto& o : pc.ops) {
      uint64_t id = op_identity(o.first);
      std::memcpy(bytes.data() + o.second, &id, (size_t)op_field_bytes(o.first));
    }
    for (auto& o : pc.p2) {
      uint64_t id = op_identity(o.first);
      std::memcpy(bytes.data() + o.second, &id, (size_t)op_field_bytes(o.first));
    }
  }
  std::memcpy(words, bytes.data(), (size_t)plan.stride);
}





// Temporary output columns in the workspace (filled by finalize / LDS
// kernels, then copied into the caller's allocations).
// A COLLECT_LIST / COLLECT_SET result as the sort path leaves it.
struct ListTemp {
  int out = -1;                  // index into plan.outs
  uint32_t* offsets = nullptr;   // groups + 1 entries
  uint8_t* child = nullptr;      // elements, the value column's width each
  uint8_t* child_valid = nullptr;
  bool child_nullable = false;   // the elements may be null
  int64_t total = 0;             // elements (read back)
  int64_t child_nulls = 0;       // null elements (read back)
};

struct TempOuts {
  std::vector<ListTemp> lists;
  std::vector<uint8_t*> key_data, key_valid, out_data, out_valid;
  unsigned long long* null_counts = nullptr;
  int64_t rows = 0;  // capacity in groups
  // null counts read back with the group count (counted as the groups were
  // written), so that the outputs are emitted without a second wait
  bool nulls_known = false;
  std::vector<int64_t> host_nulls;
};

mgb_status layout_temp(Workspace* ws, const Plan& plan, int64_t rows, hipStream_t stream, TempOuts& t,
                       OutTable& ot)
{
  const int nk = (int)plan.keys.size(), no = (int)plan.outs.size();
  if (no > kMaxOuts) return fail(MGB_ERROR_UNSUPPORTED, "more than 64 aggregations in one groupby call");
  const size_t n = (size_t)std::max<int64_t>(rows, 1);
  size_t off = 0;
  std::vector<size_t> koff(nk), kvoff(nk, SIZE_MAX), ooff(no), ovoff(no, SIZE_MAX);
  for (int c = 0; c < nk; ++c) {
    koff[c] = off;
    off += align_up(n * plan.keys[c].width, 256);
    if (plan.compare_key_nulls && plan.keys[c].mask) {
      kvoff[c] = off;
      off += align_up(n, 256);
    }
  }
  for (int o = 0; o < no; ++o) {
    ooff[o] = off;
    off += align_up(n * plan.outs[o].spec.out_width, 256);
    if (plan.outs[o].spec.valid) {
      ovoff[o] = off;
      off += align_up(n, 256);
    }
  }
  size_t nc_off = off;
  off += align_up((size_t)(nk + no) * 8, 256);
  MGB_RESERVE(ws->outs, off);
  uint8_t* b = static_cast<uint8_t*>(ws->outs.ptr);
  t = TempOuts{};
  t.rows = rows;
  ot = OutTable{};
  for (int c = 0; c < nk; ++c) {
    t.key_data.push_back(b + koff[c]);
    t.key_valid.push_back(kvoff[c] == SIZE_MAX ? nullptr : b + kvoff[c]);
    ot.key_out[c] = t.key_data[c];
    ot.key_valid[c] = t.key_valid[c];
  }
  ot.num_outs = no;
  for (int o = 0; o < no; ++o) {
    t.out_data.push_back(b + ooff[o]);
    t.out_valid.push_back(ovoff[o] == SIZE_MAX ? nullptr : b + ovoff[o]);
    ot.outs[o] = plan.outs[o].spec;
    ot.outs[o].out = t.out_data[o];
    ot.outs[o].valid = t.out_valid[o];
  }
  t.null_counts = reinterpret_cast<unsigned long long*>(b + nc_off);
  return MGB_SUCCESS;
}

// Copies an OutTable and the identity template to the device (through the
// pinned staging area, ordered on `stream`).
mgb_status upload_params(Workspace* ws, const OutTable& ot, const uint64_t* tmpl, int words, hipStream_t stream,
                         const OutTable** d_ot, const uint64_t** d_tmpl)
{
  const size_t ot_bytes = align_up(sizeof(OutTable), 256);
  const size_t slot_bytes = kStagingBytes / kUploadSlots;
  static_assert(align_up_c(sizeof(OutTable), 256) + kMaxRecordWords * 8 <= kStagingBytes / kUploadSlots,
                "parameter slot too small");
  if (ws->upload_slot >= kUploadSlots) return fail(MGB_ERROR_UNSUPPORTED, "too many parameter uploads in one call");
  const int slot = ws->upload_slot++;
  MGB_RESERVE(ws->params, kStagingBytes);
  uint8_t* host = ws->staging + slot * slot_bytes;
  uint8_t* dev = static_cast<uint8_t*>(ws->params.ptr) + slot * slot_bytes;
  std::memcpy(host, &ot, sizeof(OutTable));
  std::memcpy(host + ot_bytes, tmpl, (size_t)words * 8);
  MGB_HIP(hipMemcpyAsync(dev, host, ot_bytes + (size_t)words * 8, hipMemcpyHostToDevice, stream));
  *d_ot = reinterpret_cast<const OutTable*>(dev);
  *d_tmpl = reinterpret_cast<const uint64_t*>(dev + ot_bytes);
  return MGB_SUCCESS;
}

// Copies the ops of a second pass into a parameter slot of device memory.
mgb_status upload_pass2(Workspace* ws, const Pass2Ops& q, hipStream_t stream, const Pass2Ops** d)
{
  const size_t slot_bytes = kStagingBytes / kUploadSlots;
  static_assert(sizeof(Pass2Ops) <= kStagingBytes / kUploadSlots, "parameter slot too small");
  if (ws->upload_slot >= kUploadSlots) return fail(MGB_ERROR_UNSUPPORTED, "too many parameter uploads in one call");
  const int slot = ws->upload_slot++;
  MGB_RESERVE(ws->params, kStagingBytes);
  uint8_t* host = ws->staging + slot * slot_bytes;
  uint8_t* dev = static_cast<uint8_t*>(ws->params.ptr) + slot * slot_bytes;
  std::memcpy(host, &q, sizeof(Pass2Ops));
  MGB_HIP(hipMemcpyAsync(dev, host, sizeof(Pass2Ops), hipMemcpyHostToDevice, stream));
  *d = reinterpret_cast<const Pass2Ops*>(dev);
  return MGB_SUCCESS;
}

// Copies the LDS aggregation's parameters into a parameter slot of device
// memory (the compact execution's kernels read them there).
mgb_status upload_lds(Workspace* ws, const LdsParams& lp, hipStream_t stream, const LdsParams** d)
{
  const size_t slot_bytes = kStagingBytes / kUploadSlots;
  static_assert(sizeof(LdsParams) <= kStagingBytes / kUploadSlots, "parameter slot too small");
  if (ws->upload_slot >= kUploadSlots) return fail(MGB_ERROR_UNSUPPORTED, "too many parameter uploads in one call");
  const int slot = ws->upload_slot++;
  MGB_RESERVE(ws->params, kStagingBytes);
  uint8_t* host = ws->staging + slot * slot_bytes;
  uint8_t* dev = static_cast<uint8_t*>(ws->params.ptr) + slot * slot_bytes;
  std::memcpy(host, &lp, sizeof(LdsParams));
  MGB_HIP(hipMemcpyAsync(dev, host, sizeof(LdsParams), hipMemcpyHostToDevice, stream));
  *d = reinterpret_cast<const LdsParams*>(dev);
  return MGB_SUCCESS;
}

// Copies the few-groups parameters into a parameter slot of device memory.
mgb_status upload_tiny(Workspace* ws, const TinyParams& tp, hipStream_t stream, const TinyParams** d)
{
  const size_t slot_bytes = kStagingBytes / kUploadSlots;
  static_assert(sizeof(TinyParams) <= kStagingBytes / kUploadSlots, "parameter slot too small");
  if (ws->upload_slot >= kUploadSlots) return fail(MGB_ERROR_UNSUPPORTED, "too many parameter uploads in one call");
  const int slot = ws->upload_slot++;
  MGB_RESERVE(ws->params, kStagingBytes);
  uint8_t* host = ws->staging + slot * slot_bytes;
  uint8_t* dev = static_cast<uint8_t*>(ws->params.ptr) + slot * slot_bytes;
  std::memcpy(host, &tp, sizeof(TinyParams));
  MGB_HIP(hipMemcpyAsync(dev, host, sizeof(TinyParams), hipMemcpyHostToDevice, stream));
  *d = reinterpret_cast<const TinyParams*>(dev);
  return MGB_SUCCESS;
}

// Where the result columns go: buffers from the caller's allocator, or
// caller-provided buffers of `capacity` rows each (the entry point that
// allocates nothing).
struct OutputSink {
  const mgb_allocator* alloc = nullptr;
  int64_t capacity = -1;  // >= 0: the caller's buffers (data and null_mask of every output)
};

// Copies the temporary outputs into the result columns (data and null_mask
// set, null_mask NULL where the column has no validity) and builds their
// null masks and null counts.
// The list temporaries of output o, or nullptr when it is not a list.
const ListTemp* list_of(const TempOuts& t, int o)
{
  for (const ListTemp& lt : t.lists)
    if (lt.out == o) return &lt;
  return nullptr;
}

mgb_status copy_outputs(Workspace* ws, const Plan& plan, const TempOuts& t, int64_t groups, mgb_out_column* out_keys,
                        mgb_out_column* out_results, hipStream_t stream)
{
  const int nk = (int)plan.keys.size(), no = (int)plan.outs.size();
  if (groups == 0) return MGB_SUCCESS;
  if (!t.nulls_known) MGB_HIP(hipMemsetAsync(t.null_counts, 0, (size_t)(nk + no) * 8, stream));
  const int64_t words = bitmask_bytes(groups) / 4;
  // every column's copy and validity in one launch (per kEmitCols columns)
  for (int j0 = 0; j0 < no + nk; j0 += kEmitCols) {
    EmitParams ep{};
    ep.rows = groups;
    const int n = std::min(kEmitCols, no + nk - j0);
    int64_t units = words;
    for (int i = 0; i < n; ++i) {
      const int j = j0 + i;
      const mgb_out_column& oc = j < no ? out_results[j] : out_keys[j - no];
      const int width = j < no ? plan.outs[j].spec.out_width : plan.keys[j - no].width;
      EmitCol& c = ep.cols[i];
      c.src = j < no ? t.out_data[j] : t.key_data[j - no];
      c.dst = static_cast<uint8_t*>(oc.data);
      c.bytes = groups * width;
      c.valid = j < no ? t.out_valid[j] : t.key_valid[j - no];
      if (const ListTemp* lt = j < no ? list_of(t, j) : nullptr) {
        c.src = reinterpret_cast<const uint8_t*>(lt->offsets);
        c.bytes = (groups + 1) * 4;
        c.valid = nullptr;
      }
      c.mask = oc.null_mask;
      c.words = oc.null_mask ? words : 0;
      c.nulls = t.nulls_known ? nullptr : t.null_counts + j;
      units = std::max(units, (c.bytes + 15) / 16);
    }
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((units + kBlock - 1) / kBlock, 1024));
    emit_columns_kernel<<<dim3(gx, (unsigned)n), kBlock, 0, stream>>>(ep);
  }
  // the lists' elements (rows of their own)
  for (const ListTemp& lt : t.lists) {
    const mgb_out_column& oc = out_results[lt.out];
    if (lt.total == 0) continue;
    EmitParams ep{};
    ep.rows = lt.total;
    EmitCol& c = ep.cols[0];
    c.src = lt.child;
    c.dst = static_cast<uint8_t*>(oc.child_data);
    c.bytes = lt.total * (int64_t)plan.cols[plan.outs[lt.out].col].col.width;
    c.valid = oc.child_null_mask ? lt.child_valid : nullptr;
    c.mask = oc.child_null_mask;
    c.words = oc.child_null_mask ? bitmask_bytes(lt.total) / 4 : 0;
    c.nulls = nullptr;
    const int64_t units = std::max<int64_t>(c.words, (c.bytes + 15) / 16);
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((units + kBlock - 1) / kBlock, 1024));
    emit_columns_kernel<<<dim3(gx, 1u), kBlock, 0, stream>>>(ep);
  }
  MGB_HIP(hipGetLastError());
  const int64_t* counts = nullptr;
  if (t.nulls_known) {
    counts = t.host_nulls.data();
  } else {
    MGB_HIP(hipMemcpyAsync(ws->pinned, t.null_counts, (size_t)(no + nk) * 8, hipMemcpyDeviceToHost, stream));
    MGB_HIP(hipStreamSynchronize(stream));
    counts = reinterpret_cast<const int64_t*>(ws->pinned);
  }
  for (int o = 0; o < no; ++o) out_results[o].null_count = out_results[o].null_mask ? counts[o] : 0;
  for (int c = 0; c < nk; ++c) out_keys[c].null_count = out_keys[c].null_mask ? counts[no + c] : 0;
  for (const ListTemp& lt : t.lists)
    out_results[lt.out].child_null_count = out_results[lt.out].child_null_mask ? lt.child_nulls : 0;
  return MGB_SUCCESS;
}

// The result columns in the caller's buffers: every output's data buffer
// must hold `capacity` rows (data_bytes), and its null_mask buffer
// bitmask_bytes(capacity) bytes (mask_bytes) where the column has validity.
mgb_status emit_into_caller_buffers(Workspace* ws, const Plan& plan, const TempOuts& t, int64_t groups,
                                    mgb_out_column* out_keys, mgb_out_column* out_results, hipStream_t stream,
                                    int64_t capacity)
{
  const int nk = (int)plan.keys.size(), no = (int)plan.outs.size();
  if (!t.lists.empty())
    return fail(MGB_ERROR_UNSUPPORTED, "COLLECT_LIST / COLLECT_SET results need allocated outputs");
  if (groups > capacity)
    return fail(MGB_ERROR_ALLOC, "the result has " + std::to_string(groups) + " groups, more than the " +
                                     std::to_string(capacity) + " rows the output buffers hold");
  auto check_col = [&](mgb_out_column& oc, int type_id, int width, const uint8_t* valid, const char* what) {
    oc.type_id = type_id;
    oc.size = groups;
    oc.null_count = 0;
    if (groups > 0 && (!oc.data || oc.data_bytes < (size_t)(groups * width)))
      return fail(MGB_ERROR_LOGIC, std::string(what) + ": data buffer missing or too small");
    if (valid && groups > 0 && (!oc.null_mask || oc.mask_bytes < (size_t)bitmask_bytes(groups)))
      return fail(MGB_ERROR_LOGIC, std::string(what) + ": null_mask buffer missing or too small");
    if (!valid) oc.null_mask = nullptr;
    return MGB_SUCCESS;
  };
  for (int c = 0; c < nk; ++c) {
    mgb_status st = check_col(out_keys[c], plan.keys[c].type_id, plan.keys[c].width, t.key_valid[c], "key column");
    if (st != MGB_SUCCESS) return st;
  }
  for (int o = 0; o < no; ++o) {
    mgb_status st = check_col(out_results[o], plan.outs[o].out_type, plan.outs[o].spec.out_width, t.out_valid[o],
                              "result column");
    if (st != MGB_SUCCESS) return st;
  }
  return copy_outputs(ws, plan, t, groups, out_keys, out_results, stream);
}

// Gives the caller's columns `groups` rows (allocated, or checked against the
// caller's buffers), copies the temporary outputs into them and builds their
// null masks.
mgb_status emit_outputs(Workspace* ws, const Plan& plan, const TempOuts& t, int64_t groups, mgb_out_column* out_keys,
                        mgb_out_column* out_results, hipStream_t stream, const OutputSink& sink)
{
  if (sink.capacity >= 0)
    return emit_into_caller_buffers(ws, plan, t, groups, out_keys, out_results, stream, sink.capacity);
  const int nk = (int)plan.keys.size(), no = (int)plan.outs.size();
  Allocs al{sink.alloc, stream, {}};
  // the columns' sizes, then every buffer in one request
  std::vector<size_t> sizes;
  std::vector<void**> dests;
  auto fill_col = [&](mgb_out_column& oc, int type_id, int width, const uint8_t* valid) {
    std::memset(&oc, 0, sizeof(oc));
    oc.type_id = type_id;
    oc.size = groups;
    if (groups == 0) return;
    oc.data_bytes = (size_t)groups * width;
    sizes.push_back(oc.data_bytes);
    dests.push_back(&oc.data);
    if (valid) {
      oc.mask_bytes = (size_t)bitmask_bytes(groups);
      sizes.push_back(oc.mask_bytes);
      dests.push_back(reinterpret_cast<void**>(&oc.null_mask));
    }
  };
  for (int c = 0; c < nk; ++c) fill_col(out_keys[c], plan.keys[c].type_id, plan.keys[c].width, t.key_valid[c]);
  for (int o = 0; o < no; ++o) {
    mgb_out_column& oc = out_results[o];
    const ListTemp* lt = list_of(t, o);
    if (!lt) {
      fill_col(oc, plan.outs[o].out_type, plan.outs[o].spec.out_width, t.out_valid[o]);
      if (oc.type_id == MGB_TYPE_LIST) oc.child_type_id = plan.cols[plan.outs[o].col].type;  // no groups
      continue;
    }
    // a list column: groups + 1 offsets, then its elements
    const ValCol& v = plan.cols[plan.outs[o].col].col;
    std::memset(&oc, 0, sizeof(oc));
    oc.type_id = MGB_TYPE_LIST;
    oc.size = groups;
    oc.child_type_id = plan.cols[plan.outs[o].col].type;
    oc.child_size = lt->total;
    if (groups == 0) continue;
    oc.data_bytes = (size_t)(groups + 1) * 4;
    sizes.push_back(oc.data_bytes);
    dests.push_back(&oc.data);
    if (lt->total == 0) continue;
    oc.child_data_bytes = (size_t)lt->total * v.width;
    sizes.push_back(oc.child_data_bytes);
    dests.push_back(&oc.child_data);
    if (lt->child_nullable) {
      oc.child_mask_bytes = (size_t)bitmask_bytes(lt->total);
      sizes.push_back(oc.child_mask_bytes);
      dests.push_back(reinterpret_cast<void**>(&oc.child_null_mask));
    }
  }
  auto clear_ptrs = [&] {
    for (int c = 0; c < nk; ++c) out_keys[c].data = out_keys[c].null_mask = nullptr;
    for (int o = 0; o < no; ++o) {
      out_results[o].data = out_results[o].null_mask = nullptr;
      out_results[o].child_data = out_results[o].child_null_mask = nullptr;
    }
  };
  std::vector<void*> ptrs(sizes.size(), nullptr);
  if (!al.get_many((int)sizes.size(), sizes.data(), ptrs.data())) {
    al.release_all();
    clear_ptrs();
    return fail(MGB_ERROR_ALLOC, "allocation of the output columns failed");
  }
  for (size_t i = 0; i < ptrs.size(); ++i) *dests[i] = ptrs[i];
  mgb_status st = copy_outputs(ws, plan, t, groups, out_keys, out_results, stream);
  if (st != MGB_SUCCESS) {
    al.release_all();
    clear_ptrs();
  }
  return st;
}

// The LDS path's view of a column's ops (ValCol::sum_off ... lds_generic).
void column_program(ValCol& vc, const std::vector<std::pair<int, int>>& ops)
{
  vc.sum_off = vc.min_off = vc.max_off = -1;
  vc.mm_kind = -1;
  vc.lds_generic = 0;
  for (auto& o : ops) {
    switch (o.first) {
      case OP_SUM_I64: case OP_SUM_F64: vc.sum_off = o.second; break;
      case OP_MIN_I32: case OP_MIN_U32: case OP_MIN_I64: case OP_MIN_U64: case OP_MIN_F32: case OP_MIN_F64:
        vc.min_off = o.second;
        vc.mm_kind = o.first;
        break;
      case OP_MAX_I32: case OP_MAX_U32: case OP_MAX_I64: case OP_MAX_U64: case OP_MAX_F32: case OP_MAX_F64:
        vc.max_off = o.second;
        vc.mm_kind = o.first - 1;  // the matching MIN kind
        break;
      case OP_NULLS: break;
      default: vc.lds_generic = 1; break;
    }
  }
}

// Fills AggParams' value columns and ops from the plan; false when they do
// not fit one launch.
bool single_launch_ops(const Plan& plan, AggParams& ap)
{
  ap.num_vals = 0;
  ap.num_ops = 0;
  ap.rows_op = -1;
  if (plan.rows_off >= 0) {
    ap.ops[0] = OpDesc{OP_ROWS, plan.rows_off};
    ap.rows_op = 0;
    ap.num_ops = 1;
  }
  int flag_bit = 0;
  for (const PlanCol& pc : plan.cols) {
    if (ap.num_vals >= kMaxVals || ap.num_ops + (int)pc.ops.size() > kMaxOps) return false;
    ValCol vc = pc.col;
    vc.flag_bit = pc.nullable ? flag_bit++ : -1;
    vc.nulls_off = pc.nulls_off;
    vc.op_begin = ap.num_ops;
    for (auto& o : pc.ops) ap.ops[ap.num_ops++] = OpDesc{o.first, o.second};
    vc.op_end = ap.num_ops;
    column_program(vc, pc.ops);
    ap.vals[ap.num_vals++] = vc;
  }
  ap.program_cols = ap.num_vals <= kProgramCols;
  for (int c = 0; c < ap.num_vals; ++c)
    if (ap.vals[c].lds_generic) ap.program_cols = 0;
  return true;
}

// Second-pass op of field `o` of plan column `pc`, over the value column `v`
// of the rows (its `val`-th).
Pass2Op pass2_op(const Plan& plan, const PlanCol& pc, const std::pair<int, int>& o, const ValCol& v, int val)
{
  Pass2Op op{};
  op.data = v.data;
  op.mask = v.mask;
  op.bit_off = v.bit_off;
  op.width = v.width;
  op.vclass = v.vclass;
  op.stride = v.stride;
  op.flag_bit = v.mask ? -1 : v.flag_bit;
  op.kind = o.first;
  op.val = val;
  op.off = o.second;
  op.ref_kind = o.first == OP_M2 ? sum_op_for(pc.type) : minmax_op_for(pc.type, o.first == OP_ARGMIN_IDX);
  op.ref_off = pc.op_offset(op.ref_kind);
  op.count_off = plan.rows_off;
  op.nulls_off = pc.nulls_off;
  return op;
}

// The second-pass ops of a plan over `rows`, whose value column c is
// plan.cols[c] (as single_launch_ops lays them out); false when they exceed
// one launch.
bool pass2_ops(const Plan& plan, const AggParams& rows, Pass2Ops& q)
{
  q.num_ops = 0;
  for (size_t c = 0; c < plan.cols.size(); ++c) {
    const PlanCol& pc = plan.cols[c];
    if (c < (size_t)kProgramCols) {
      auto f10 = [](int off) { return (uint32_t)(off + 1) & 1023u; };
      q.p2prog[c] = f10(pc.p2_offset(OP_M2)) | f10(pc.p2_offset(OP_ARGMIN_IDX)) << 10 |
                    f10(pc.p2_offset(OP_ARGMAX_IDX)) << 20;
    }
    for (auto& o : pc.p2) {
      if (q.num_ops == kMaxPass2) return false;
      q.ops[q.num_ops++] = pass2_op(plan, pc, o, rows.vals[c], (int)c);
    }
  }
  return true;
}

size_t lds_limit()
{
  static size_t lim = [] {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      v = 64 * 1024;
    return (size_t)std::min(v, 160 * 1024);
  }();
  return lim;
}


// LDS record stride: the record padded to 8 x (odd number) bytes, so that the
// same field of records at random slots spreads over 32 of the 64 LDS banks
// (an 80-byte stride = 20 dwords would reach only 16 of them).
int lds_stride_of(int stride)
{
  int words = (stride + 7) / 8;
  if (words % 2 == 0) ++words;
  return words * 8;
}

// Raises a kernel's dynamic LDS limit to `bytes` when it is below: the
// attribute is set once per kernel, device and size, not at every launch.
hipError_t set_dynamic_lds(const void* fn, size_t bytes)
{
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, size_t> done;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(mu);
  size_t& cur = done[{dev, fn}];
  if (bytes <= cur) return hipSuccess;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) cur = bytes;
  return e;
}

LdsGeom lds_geometry(int stride)
{
  LdsGeom g;
  stride = lds_stride_of(stride);
  const size_t reserve = 16 + (size_t)kLdsBlock * 4;
  const size_t lim = lds_limit();
  if (lim <= reserve + 2 * (size_t)stride) return g;
  int64_t slots = (int64_t)((lim - reserve) / (size_t)stride) - 1;
  g.slots = (int)slots;
  g.cap = (int)(slots * 0.62);
  g.bytes = (size_t)(slots + 1) * stride + reserve;
  g.ok = slots >= 256;
  return g;
}

enum class Path { Auto, Global, Direct, Partitioned };

Path forced_path()
{
  const char* e = std::getenv("MGB_GROUPBY_PATH");
  if (!e) return Path::Auto;
  std::string s(e);
  if (s == "global") return Path::Global;
  if (s == "direct") return Path::Direct;
  if (s == "partitioned") return Path::Partitioned;
  return Path::Auto;
}

// Chunks (workgroups) of the histogram and scatter passes over R rows into
// 2^log2p partitions: up to four per CU, fewer when the partitions are many
// (the histogram holds 2^log2p counts per chunk), at least 4096 rows each.
int64_t partition_chunks(int64_t R, int log2p, int cus)
{
  // 4 chunks per CU up to 2^13 partitions (the same chunks as the estimate's
  // fine histogram), fewer beyond (the histogram is P x chunks words)
  int64_t nw = std::min<int64_t>((int64_t)cus * 4, std::max<int64_t>(64, (1ll << 23) >> log2p));
  return std::max<int64_t>(1, std::min<int64_t>(nw, (R + 4095) / 4096));
}

// The sample's distinct integer keys when it holds every group of the call
// and at most kTinyKeys of them.
struct SampleKeys {
  bool complete = false;
  int n = 0;
  uint64_t keys[kTinyKeys] = {};
};

constexpr int kHllReduceBlocks = kHllRegs / kHllReduceRegs;
constexpr int kPinnedHll = 256;  // first host readback slot of the HyperLogLog sums (then the zero counts)
constexpr int kPinnedHot = kMaxPinned - 8;  // readback slot of the hot keys' count
static_assert(kPinnedHll + kHllReduceBlocks * 3 / 2 <= kMaxPinned, "readback slots past the end");

// Estimate of the number of distinct key rows. A sample of 65536 rows
// (every row of a small input) is hashed into a table first; when it holds
// every group (all rows sampled, or every sampled key seen at least twice
// and few distinct keys) its count is exact. Otherwise a HyperLogLog pass
// over every row's key hash (standard error ~1.04/sqrt(2^14) = 0.8%, linear
// counting for small cardinalities) gives the estimate; that pass also
// counts each chunk's rows per fine bucket (ws->fine). The pass is launched
// right after the sample and reads the sample's counts on the device (it
// returns at once when the sample is complete), so one host wait serves
// both. With `hot` (a call the compact execution may take), the sample's
// hot keys are collected into ws->hot and the pass leaves their rows out
// of the fine counts (ws->fine_hot).
int64_t sample_estimate(Workspace* ws, const AggParams& base, int64_t R, hipStream_t stream, mgb_status& st,
                        SampleKeys* sk = nullptr, const HotSpec* hot = nullptr)
{
  st = MGB_SUCCESS;
  ws->fine_valid = false;
  ws->fine_hot = false;
  // a call of 8M rows or more that collects hot keys samples twice the
  // rows, so that keys of about 1 row in 16K reach kHotMin sightings (the
  // sample table keeps its 2^18 slots: at most half of them taken)
  const int64_t scap = 4 << 16;
  const int64_t kSample = hot && R >= (1ll << 23) ? 1 << 17 : 1 << 16;
  // small inputs are counted exactly; larger ones are sampled
  const int64_t samples = R <= scap / 2 ? R : kSample;
  const bool may_need_hll = samples < R;
  const size_t sample_bytes = align_up((size_t)(scap + 1) * 8, 256) + align_up((size_t)(scap + 1) * 4, 256) +
                              align_up((5 + kTinyKeys) * 8, 256);
  // the HyperLogLog pass: one workgroup per chunk of the partitioned execution
  const int cus = device_cu_count();
  const int64_t nb = partition_chunks(R, kFineBits, cus);
  const int64_t chunk = (R + nb - 1) / nb;
  // workgroups of the pass: each takes several chunks when they are many
  // (fewer workgroups to start and to return at once after a complete sample)
  const int64_t hll_wgs = std::min<int64_t>(nb, 2 * (int64_t)cus);
  const size_t regs_bytes = align_up((size_t)hll_wgs * kHllRegs, 256);
  const size_t hll_bytes = regs_bytes + align_up((size_t)kHllReduceBlocks * 8, 256) +
                           align_up((size_t)kHllReduceBlocks * 4, 256);
  // MGB_HOT_KEYS=off: no hot keys (every row partitioned)
  const char* hot_env = std::getenv("MGB_HOT_KEYS");
  if (!may_need_hll || (hot_env && std::string(hot_env) == "off")) hot = nullptr;
  if (ws->main.reserve(sample_bytes + (may_need_hll ? hll_bytes : 0), stream) != hipSuccess ||
      (may_need_hll && ws->fine.reserve((size_t)kFineBuckets * nb * 4, stream) != hipSuccess) ||
      (hot && ws->hot.reserve(sizeof(HotSet), stream) != hipSuccess)) {
    st = fail(MGB_ERROR_ALLOC, "allocation of the cardinality estimate's temporaries failed");
    return R;
  }
  uint8_t* mem = static_cast<uint8_t*>(ws->main.ptr);
  SampleParams sp{};
  sp.agg = base;
  sp.hdr = reinterpret_cast<unsigned long long*>(mem);
  sp.cnt = reinterpret_cast<uint32_t*>(mem + align_up((size_t)(scap + 1) * 8, 256));
  sp.stats = reinterpret_cast<unsigned long long*>(reinterpret_cast<uint8_t*>(sp.cnt) +
                                                   align_up((size_t)(scap + 1) * 4, 256));
  sp.scap = scap;
  sp.samples = samples;
  const uint64_t empty = base.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef;
  const bool want_keys = sk && base.key_mode == KEY_INT;
  uint8_t* regs = mem + sample_bytes;
  double* sums = reinterpret_cast<double*>(regs + regs_bytes);
  uint32_t* zeros = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(sums) +
                                                align_up((size_t)kHllReduceBlocks * 8, 256));
  hipError_t e = hipSuccess;
  {
    ProfScope prof("estimate", stream);
    HotSet* hot_set = hot ? static_cast<HotSet*>(ws->hot.ptr) : nullptr;
    const HotSpec hs = hot ? *hot : HotSpec{};
    sample_clear_kernel<<<256, kBlock, 0, stream>>>(sp.hdr, sp.cnt, scap, empty, sp.stats, 5, hot_set);
    sample_insert_kernel<<<(unsigned)((samples + kBlock - 1) / kBlock), kBlock, 0, stream>>>(sp);
    sample_summary_kernel<<<kSampleSummaryBlocks, kBlock, 0, stream>>>(sp.hdr, sp.cnt, scap, sp.stats, sp.stats + 4,
                                                                        want_keys ? sp.stats + 5 : nullptr, hot_set, hs);
    if (may_need_hll) {
      const size_t lds = (size_t)(kHllRegs + kFineBuckets) * 4 +
                         (hot ? (size_t)kHotMax * sizeof(HotKey) + (size_t)kHotSlots * 4 : 0);
      e = set_dynamic_lds((const void*)hll_hist_kernel, lds);
      if (e == hipSuccess) {
        hll_hist_kernel<<<(unsigned)hll_wgs, kLdsBlock, lds, stream>>>(base, regs,
                                                                       static_cast<uint32_t*>(ws->fine.ptr), chunk,
                                                                       (int32_t)nb, sp.stats, hot_set, hs);
        hll_reduce_kernel<<<(unsigned)kHllReduceBlocks, 1024, 0, stream>>>(regs, (int)hll_wgs, sums, zeros, sp.stats);
      }
    }
  }
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(ws->pinned, sp.stats, (want_keys ? 5 + kTinyKeys : 4) * sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess && hot)
    e = hipMemcpyAsync(ws->pinned + kPinnedHot, &static_cast<HotSet*>(ws->hot.ptr)->n, 4, hipMemcpyDeviceToHost,
                       stream);
  if (e == hipSuccess && may_need_hll)
    e = hipMemcpyAsync(ws->pinned + kPinnedHll, sums, (size_t)kHllReduceBlocks * 8, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess && may_need_hll)
    e = hipMemcpyAsync(ws->pinned + kPinnedHll + kHllReduceBlocks, zeros, (size_t)kHllReduceBlocks * 4,
                       hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    st = fail(MGB_ERROR_HIP, std::string("cardinality estimate: ") + hipGetErrorString(e));
    return R;
  }
  // hot keys found: the fine counts left their rows out
  ws->fine_hot = hot && (*reinterpret_cast<const uint32_t*>(ws->pinned + kPinnedHot)) > 0u;
  const uint64_t d = ws->pinned[0], f1 = ws->pinned[1], sampled = ws->pinned[3];
  // every row sampled, or a saturated sample (as sample_saturated): all groups found
  const bool complete = (int64_t)sampled >= R || samples == R || (f1 == 0 && d * 8 <= sampled);
  if (want_keys && complete && ws->pinned[4] <= (uint64_t)kTinyKeys) {
    sk->complete = true;
    sk->n = (int)ws->pinned[4];
    for (int i = 0; i < sk->n; ++i) sk->keys[i] = ws->pinned[5 + i];
  }
  if (complete) return (int64_t)d;
  // the HyperLogLog estimate over all rows
  ws->fine_valid = true;
  ws->fine_nw = nb;
  ws->fine_chunk = chunk;
  double sum = 0.0;
  uint64_t z = 0;
  const double* hs = reinterpret_cast<const double*>(ws->pinned + kPinnedHll);
  const uint32_t* hz = reinterpret_cast<const uint32_t*>(ws->pinned + kPinnedHll + kHllReduceBlocks);
  for (int i = 0; i < kHllReduceBlocks; ++i) {
    sum += hs[i];
    z += hz[i];
  }
  const double m = (double)kHllRegs;
  double est = (0.7213 / (1.0 + 1.079 / m)) * m * m / sum;
  if (est <= 2.5 * m && z > 0) est = m * std::log(m / (double)z);
  return std::max<int64_t>(1, std::min<int64_t>(R, (int64_t)std::ceil(est)));
}

// The sort-path aggregations (MEDIAN, QUANTILE, NUNIQUE, NTH_ELEMENT,
// COLLECT_LIST, COLLECT_SET) of a call whose global table is complete and
// finalised: `rec_pos` gives every record's output row. Every row's group is
// looked up once. The outputs then share orderings: one per value column for
// MEDIAN, QUANTILE, NUNIQUE and COLLECT_SET (each group's valid values
// ascending), one per value column for NTH_ELEMENT and COLLECT_LIST EXCLUDE
// and one for every INCLUDE of them (each group's rows in row order).
// cpp/src/groupby/sort/sort_helper.cu orders the rows by key the same way;
// the results are read off the sorted positions.
//
// An ordering is a segmented sort (seg_sort.hpp): each group's keys (32-bit
// value codes for columns of at most 4 bytes, 64-bit codes for 8-byte
// columns, row indices for the row orderings) bucketed into the group's
// segment and sorted there. Results are decoded from the sorted codes, so no
// value is gathered.
mgb_status run_sort_aggs(Workspace* ws, const Plan& plan, const AggParams& rows, const uint32_t* rec_pos,
                         int64_t groups, hipStream_t stream, TempOuts& temp)
{
  const int64_t R = rows.num_rows;
  if (groups == 0 || R == 0) return MGB_SUCCESS;
  if (R > (int64_t)0x7fffffff || groups >= (int64_t)0xffffffff)
    return fail(MGB_ERROR_UNSUPPORTED, "sort-path aggregations of more than 2^31 rows");
  const uint32_t G = (uint32_t)groups;

  // the orderings the outputs need, in order of first use
  struct Ordering { int col; int mode; };
  std::vector<Ordering> orderings;
  std::vector<int> out_ordering(plan.outs.size(), -1);
  for (size_t o = 0; o < plan.outs.size(); ++o) {
    const Plan::Out& out = plan.outs[o];
    if (out.spec.fin != FIN_NONE) continue;
    Ordering want{out.col, SORT_BY_VALUE};
    if (out.kind == MGB_AGG_NTH_ELEMENT || out.kind == MGB_AGG_COLLECT_LIST)
      want = out.agg.null_policy == MGB_NULL_INCLUDE ? Ordering{-1, SORT_ROWS_ALL} : Ordering{out.col, SORT_ROWS_VALID};
    int k = 0;
    while (k < (int)orderings.size() && !(orderings[k].col == want.col && orderings[k].mode == want.mode)) ++k;
    if (k == (int)orderings.size()) orderings.push_back(want);
    out_ordering[o] = k;
  }

  // (R + 1)-entry u32 buffers: rank and group buffers, the distinct-value
  // flags and their exclusive sums
  const size_t r4 = align_up((size_t)(R + 1) * 4, 256), r8 = align_up((size_t)R * 8, 256);
  const size_t g4 = align_up((size_t)G * 4, 256), o4 = align_up(((size_t)G + 1) * 4, 256);
  const int64_t tiles = (R + kRadixTile - 1) / kRadixTile;
  const size_t count_bytes = align_up((size_t)kRadixBins * (size_t)tiles * 4, 256);
  const int64_t scan_n = std::max<int64_t>({R, (int64_t)G, (int64_t)kRadixBins * tiles}) + 1;
  const int64_t scan_blocks = (scan_n + kScanTile - 1) / kScanTile;
  const size_t scan_bytes = align_up((size_t)scan_blocks * 4, 256) + 256;
  const size_t list_bytes_seg = align_up(((size_t)R / kWaveSeg + 2) * 8, 256) + 256;
  // gid | A, B (8 bytes a row) | C, D, E (4 bytes a row) | offset, has_null,
  // group nulls, distinct values, list lengths | scan block sums |
  // long-segment list | radix counts
  const size_t need = 4 * r4 + 2 * r8 + 2 * o4 + 3 * g4 + scan_bytes + list_bytes_seg + count_bytes;
  MGB_RESERVE(ws->sorts, need);
  uint8_t* b = static_cast<uint8_t*>(ws->sorts.ptr);
  size_t off = 0;
  auto take = [&](size_t bytes) { uint8_t* p = b + off; off += bytes; return p; };
  uint32_t* gid = reinterpret_cast<uint32_t*>(take(r4));
  uint8_t* A = take(r8);
  uint8_t* B = take(r8);
  uint32_t* C = reinterpret_cast<uint32_t*>(take(r4));
  uint32_t* D = reinterpret_cast<uint32_t*>(take(r4));
  uint32_t* E = reinterpret_cast<uint32_t*>(take(r4));
  uint32_t* offset = reinterpret_cast<uint32_t*>(take(o4));
  uint32_t* has_null = reinterpret_cast<uint32_t*>(take(g4));
  uint32_t* gnulls = reinterpret_cast<uint32_t*>(take(g4));
  uint32_t* distinct = reinterpret_cast<uint32_t*>(take(g4));
  uint32_t* lens = reinterpret_cast<uint32_t*>(take(o4));
  uint32_t* bsum = reinterpret_cast<uint32_t*>(take(scan_bytes));
  unsigned long long* scan_total = reinterpret_cast<unsigned long long*>(
    reinterpret_cast<uint8_t*>(bsum) + align_up((size_t)scan_blocks * 4, 256));
  unsigned long long* seg_list = reinterpret_cast<unsigned long long*>(take(list_bytes_seg));
  uint32_t* seg_count = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(seg_list) + list_bytes_seg - 256);
  uint32_t* radix_counts = reinterpret_cast<uint32_t*>(take(count_bytes));
  // exclusive sum of in[0, n) into out (in place allowed)
  auto excl_scan = [&](const uint32_t* in, uint32_t* out, int64_t n) -> mgb_status {
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    scan_reduce_u32_kernel<<<(unsigned)nb, 256, 0, stream>>>(in, n, bsum);
    scan_bsums_u32_kernel<<<1, 1024, 0, stream>>>(bsum, nb, scan_total);
    scan_apply_u32_kernel<<<(unsigned)nb, 256, 0, stream>>>(in, n, bsum, out);
    MGB_HIP(hipGetLastError());
    return MGB_SUCCESS;
  };

  // COLLECT_LIST / COLLECT_SET results: offsets, elements and their validity
  // bytes per list output, and two counters (elements, null elements)
  size_t list_bytes = 0;
  for (const Plan::Out& out : plan.outs)
    if (out.kind == MGB_AGG_COLLECT_LIST || out.kind == MGB_AGG_COLLECT_SET)
      list_bytes += o4 + align_up((size_t)R * plan.cols[out.col].col.width, 256) + align_up((size_t)R, 256) + 256;
  temp.lists.clear();
  if (list_bytes) {
    MGB_RESERVE(ws->lists, list_bytes);
    uint8_t* lb = static_cast<uint8_t*>(ws->lists.ptr);
    size_t loff = 0;
    for (size_t o = 0; o < plan.outs.size(); ++o) {
      const Plan::Out& out = plan.outs[o];
      if (out.kind != MGB_AGG_COLLECT_LIST && out.kind != MGB_AGG_COLLECT_SET) continue;
      const ValCol& v = plan.cols[out.col].col;
      ListTemp lt;
      lt.out = (int)o;
      lt.offsets = reinterpret_cast<uint32_t*>(lb + loff);
      loff += o4;
      lt.child = lb + loff;
      loff += align_up((size_t)R * v.width, 256);
      lt.child_valid = lb + loff;
      loff += align_up((size_t)R, 256);
      loff += 256;  // the counters (elements, null elements) sit before the next output
      lt.child_nullable = out.agg.null_policy == MGB_NULL_INCLUDE && v.mask != nullptr;
      temp.lists.push_back(lt);
    }
  }
  auto list_counters = [&](const ListTemp& lt) {
    return reinterpret_cast<unsigned long long*>(lt.child_valid + align_up((size_t)R, 256));
  };
  for (const ListTemp& lt : temp.lists) MGB_HIP(hipMemsetAsync(list_counters(lt), 0, 16, stream));

  const int cus = device_cu_count();
  const unsigned grid_rows = (unsigned)std::max<int64_t>(1, std::min<int64_t>((R + kBlock - 1) / kBlock, (int64_t)cus * 16));
  const unsigned grid_groups = (unsigned)((groups + kBlock - 1) / kBlock);
  ProfScope prof("sort_aggs", stream);
  row_group_kernel<<<grid_rows, kBlock, 0, stream>>>(rows, rec_pos, G, gid);
  MGB_HIP(hipGetLastError());
  const unsigned grid_small = (unsigned)std::max<int64_t>(1, std::min<int64_t>(((int64_t)G + 3) / 4, (int64_t)cus * 32));
  const unsigned grid_long = (unsigned)std::max<int64_t>(1, std::min<int64_t>(R / kWaveSeg + 1, (int64_t)cus * 4));
  const int max_rounds = seg_rounds((uint64_t)R);
  // radix passes over the group keys 0..G (G: rows that take no part)
  int gbits = 1;
  while (gbits < 32 && (1ull << gbits) <= (unsigned long long)G) ++gbits;
  const int passes = (gbits + 7) / 8;
  const int pass_bits = (gbits + passes - 1) / passes;
  for (int k = 0; k < (int)orderings.size(); ++k) {
    const Ordering ord = orderings[k];
    SegParams sp{};
    if (ord.col >= 0) sp.v = plan.cols[ord.col].col;
    sp.num_rows = R;
    sp.groups = G;
    sp.mode = ord.mode;
    sp.narrow = ord.mode == SORT_BY_VALUE && sp.v.width <= 4;
    const bool wide = ord.mode == SORT_BY_VALUE && !sp.narrow;
    bool want_nulls = false;
    for (size_t o = 0; o < plan.outs.size(); ++o)
      if (out_ordering[o] == k && plan.outs[o].kind == MGB_AGG_NUNIQUE &&
          plan.outs[o].agg.null_policy == MGB_NULL_INCLUDE)
        want_nulls = true;
    if (want_nulls) {
      MGB_HIP(hipMemsetAsync(has_null, 0, (size_t)G * 4, stream));
      sp.has_null = has_null;
    }
    sp.gid = gid;
    // group keys C <-> E; payloads: 64-bit codes A <-> B, 32-bit keys D <-> B
    // (the first pass reads the rows themselves and writes gk_b / pl_b)
    uint32_t* gk_a = E;
    uint32_t* gk_b = C;
    void* pl_a = B;
    void* pl_b = wide ? static_cast<void*>(A) : static_cast<void*>(D);
    for (int ps = 0; ps < passes; ++ps) {
      RadixPass rp{};
      rp.n = R;
      rp.tiles = tiles;
      rp.shift = ps * pass_bits;
      rp.bits = std::min(pass_bits, gbits - ps * pass_bits);
      rp.gk_in = gk_a;
      rp.gk_out = gk_b;
      rp.pl_in = pl_a;
      rp.pl_out = pl_b;
      rp.counts = radix_counts;
      if (ps == 0) radix_hist_kernel<true><<<(unsigned)tiles, kRadixBlock, 0, stream>>>(rp, sp);
      else radix_hist_kernel<false><<<(unsigned)tiles, kRadixBlock, 0, stream>>>(rp, sp);
      MGB_HIP(hipGetLastError());
      const mgb_status st = excl_scan(radix_counts, radix_counts, ((int64_t)1 << rp.bits) * tiles);
      if (st != MGB_SUCCESS) return st;
      const unsigned gt = (unsigned)tiles;
      if (wide) {
        if (ps == 0) radix_scatter_kernel<uint64_t, true, 8><<<gt, kRadixBlock, 0, stream>>>(rp, sp);
        else radix_scatter_kernel<uint64_t, false, 0><<<gt, kRadixBlock, 0, stream>>>(rp, sp);
      } else if (ps > 0) {
        radix_scatter_kernel<uint32_t, false, 0><<<gt, kRadixBlock, 0, stream>>>(rp, sp);
      } else {
        // the first pass loads the value at its width (none for the row orderings)
        switch (ord.mode == SORT_BY_VALUE ? sp.v.width : 0) {
          case 1: radix_scatter_kernel<uint32_t, true, 1><<<gt, kRadixBlock, 0, stream>>>(rp, sp); break;
          case 2: radix_scatter_kernel<uint32_t, true, 2><<<gt, kRadixBlock, 0, stream>>>(rp, sp); break;
          case 4: radix_scatter_kernel<uint32_t, true, 4><<<gt, kRadixBlock, 0, stream>>>(rp, sp); break;
          default: radix_scatter_kernel<uint32_t, true, 0><<<gt, kRadixBlock, 0, stream>>>(rp, sp); break;
        }
      }
      MGB_HIP(hipGetLastError());
      std::swap(gk_a, gk_b);
      std::swap(pl_a, pl_b);
    }
    seg_bounds_kernel<<<grid_rows, kBlock, 0, stream>>>(gk_a, R, G, offset);
    MGB_HIP(hipGetLastError());
    if (ord.mode == SORT_BY_VALUE) {
      // each group's codes sorted in place; narrow results packed into A
      sp.off = offset;
      sp.keys = pl_a;
      sp.alt = pl_b;
      sp.key64 = sp.narrow ? reinterpret_cast<uint64_t*>(A) : nullptr;
      sp.gkey = nullptr;  // the sorted group keys are in place already
      sp.list = seg_list;
      sp.list_count = seg_count;
      MGB_HIP(hipMemsetAsync(seg_count, 0, 4, stream));
      if (wide) {
        seg_sort_small_kernel<uint64_t><<<grid_small, 256, 0, stream>>>(sp);
        seg_sort_block_kernel<uint64_t><<<grid_long, kSegBlock, 0, stream>>>(sp);
        for (int r = 0; r < max_rounds; ++r) seg_merge_kernel<uint64_t><<<grid_long, kSegBlock, 0, stream>>>(sp, r);
        if (max_rounds) seg_emit_long_kernel<uint64_t><<<grid_long, kSegBlock, 0, stream>>>(sp);
      } else {
        seg_sort_small_kernel<uint32_t><<<grid_small, 256, 0, stream>>>(sp);
        seg_sort_block_kernel<uint32_t><<<grid_long, kSegBlock, 0, stream>>>(sp);
        for (int r = 0; r < max_rounds; ++r) seg_merge_kernel<uint32_t><<<grid_long, kSegBlock, 0, stream>>>(sp, r);
        if (max_rounds) seg_emit_long_kernel<uint32_t><<<grid_long, kSegBlock, 0, stream>>>(sp);
      }
      MGB_HIP(hipGetLastError());
    }
    SortView view{};
    view.n = R;
    view.groups = G;
    view.narrow = sp.narrow;
    view.offset = offset;
    view.total = offset + G;
    if (sp.narrow) {
      view.key64 = reinterpret_cast<const uint64_t*>(A);
    } else if (wide) {
      view.gkey = gk_a;
      view.code = static_cast<const uint64_t*>(pl_a);
    } else {
      view.gkey = gk_a;
      view.order = static_cast<const uint32_t*>(pl_a);
    }
    // COLLECT_SET's distinct-value flags and their sums (R + 1 entries each,
    // in buffers the view does not use)
    uint32_t* flags = gk_b;
    uint32_t* sums = sp.narrow ? static_cast<uint32_t*>(pl_b) : D;
    bool counted = false;  // NUNIQUE's distinct values (in `distinct`) for this ordering
    for (size_t o = 0; o < plan.outs.size(); ++o) {
      if (out_ordering[o] != k) continue;
      const Plan::Out& out = plan.outs[o];
      SortOutParams p{};
      p.v = plan.cols[out.col].col;
      p.s = view;
      p.kind = out.kind;
      p.interp = out.kind == MGB_AGG_MEDIAN ? 0 : out.agg.interpolation;
      p.n = out.agg.ddof;
      p.nulls = out.agg.null_policy == MGB_NULL_INCLUDE ? 1 : 0;
      p.q = out.kind == MGB_AGG_MEDIAN ? 0.5 : out.agg.quantile;
      p.has_null = has_null;
      p.out = temp.out_data[o];
      p.out_valid = temp.out_valid[o];
      p.out_width = out.spec.out_width;
      const bool is_list = out.kind == MGB_AGG_COLLECT_LIST || out.kind == MGB_AGG_COLLECT_SET;
      if (!p.out_valid && !is_list) return fail(MGB_ERROR_HIP, "sort-path output without validity");
      if (is_list) {
        const ListTemp* lt = nullptr;
        for (const ListTemp& l : temp.lists)
          if (l.out == (int)o) lt = &l;
        if (!lt) return fail(MGB_ERROR_HIP, "list output without its temporaries");
        CollectParams cp{};
        cp.v = p.v;
        cp.s = view;
        cp.list_off = lt->offsets;
        cp.child = lt->child;
        cp.child_valid = lt->child_valid;
        cp.include_nulls = out.agg.null_policy == MGB_NULL_INCLUDE ? 1 : 0;
        cp.nulls_unequal = out.agg.ddof == MGB_NULL_UNEQUAL ? 1 : 0;
        if (out.kind == MGB_AGG_COLLECT_LIST) {
          // the ordering's group bounds are the lists' offsets
          MGB_HIP(hipMemcpyAsync(lt->offsets, offset, (size_t)(G + 1) * 4, hipMemcpyDeviceToDevice, stream));
          collect_gather_kernel<<<grid_rows, kBlock, 0, stream>>>(cp);
          if (lt->child_nullable)
            list_nulls_kernel<<<grid_rows, kBlock, 0, stream>>>(p.v, gid, R, G, list_counters(*lt) + 1);
        } else {
          const int32_t is_float = p.v.vclass == VC_F32 || p.v.vclass == VC_F64;
          distinct_start_kernel<<<grid_rows, kBlock, 0, stream>>>(view, is_float, flags,
                                                                  out.agg.interpolation == MGB_NAN_UNEQUAL ? 1 : 0);
          MGB_HIP(hipGetLastError());
          mgb_status st = excl_scan(flags, sums, R + 1);
          if (st != MGB_SUCCESS) return st;
          MGB_HIP(hipMemsetAsync(gnulls, 0, (size_t)G * 4, stream));
          if (p.v.mask) group_nulls_kernel<<<grid_rows, kBlock, 0, stream>>>(p.v, gid, R, G, gnulls);
          cp.prefix = sums;
          cp.flag = flags;
          cp.nulls = gnulls;
          set_lengths_kernel<<<(unsigned)((G + kBlock) / kBlock), kBlock, 0, stream>>>(cp, lens);
          MGB_HIP(hipGetLastError());
          st = excl_scan(lens, lt->offsets, (int64_t)G + 1);
          if (st != MGB_SUCCESS) return st;
          set_values_kernel<<<grid_rows, kBlock, 0, stream>>>(cp);
          set_nulls_kernel<<<grid_groups, kBlock, 0, stream>>>(cp, list_counters(*lt) + 1);
        }
        MGB_HIP(hipGetLastError());
        continue;
      }
      if (out.kind == MGB_AGG_NUNIQUE) {
        if (!counted) {
          const int32_t is_float = p.v.vclass == VC_F32 || p.v.vclass == VC_F64;
          MGB_HIP(hipMemsetAsync(distinct, 0, (size_t)G * 4, stream));
          nunique_count_kernel<<<grid_rows, kBlock, 0, stream>>>(view, is_float, distinct);
          MGB_HIP(hipGetLastError());
          counted = true;
        }
        p.distinct = distinct;
        nunique_finish_kernel<<<grid_groups, kBlock, 0, stream>>>(p);
      } else {
        sort_select_kernel<<<grid_groups, kBlock, 0, stream>>>(p);
      }
      MGB_HIP(hipGetLastError());
    }
  }
  if (!temp.lists.empty()) {
    // the lists' element and null-element counts: one more wait
    const int nl = (int)temp.lists.size();
    if (2 * nl > kPinnedNulls) return fail(MGB_ERROR_UNSUPPORTED, "too many list aggregations in one call");
    uint32_t* totals = reinterpret_cast<uint32_t*>(ws->pinned);
    for (int l = 0; l < nl; ++l) {
      const ListTemp& lt = temp.lists[l];
      MGB_HIP(hipMemcpyAsync(totals + l, lt.offsets + G, 4, hipMemcpyDeviceToHost, stream));
      MGB_HIP(hipMemcpyAsync(ws->pinned + nl + l, list_counters(lt) + 1, 8, hipMemcpyDeviceToHost, stream));
    }
    MGB_HIP(hipStreamSynchronize(stream));
    for (int l = 0; l < nl; ++l) {
      temp.lists[l].total = totals[l];
      temp.lists[l].child_nulls = (int64_t)ws->pinned[nl + l];
    }
  }
  return MGB_SUCCESS;
}

// Whether a call takes the few-groups direct execution, and its parameters:
// a single integer key, every group among the sample's distinct keys (at
// most kTinyKeys), one or two value columns whose ops are SUM / MIN / MAX
// and counts, no second pass or sort. MGB_FEW_GROUPS=off turns it off,
// =collapse makes every wave collapse each group's lanes first.
bool tiny_plan(const Plan& plan, const AggParams& ap, const SampleKeys& sk, TinyParams& tp)
{
  const char* env = std::getenv("MGB_FEW_GROUPS");
  const std::string mode = env ? env : "";
  if (mode == "off") return false;
  if (!sk.complete || sk.n < 1 || plan.key_mode != KEY_INT || plan.compare_key_nulls || plan.has_pass2 ||
      plan.has_sort || ap.num_vals < 1 || ap.num_vals > kTinyCols)
    return false;
  tp = TinyParams{};
  for (int c = 0; c < ap.num_vals; ++c) {
    ValCol v = ap.vals[c];
    if (v.lds_generic) return false;
    v.flag_bit = -1;  // validity from the column's mask
    tp.prog[c] = pack_program(v, 0);
  }
  tp.collapse = mode == "collapse" ? 1 : 0;
  // candidates (the key equal to the empty header takes the probe)
  int n = 0;
  for (int i = 0; i < sk.n; ++i)
    if (sk.keys[i] != kEmptyIntKey) tp.keys[n++] = sk.keys[i];
  if (n == 0) return false;
  tp.num_keys = n;
  // a lookup multiplier under which the candidates fall into distinct entries
  for (uint32_t trial = 0; trial < 4096; ++trial) {
    const uint32_t mult = (0x9e3779b1u + trial * 0x6a09e668u) | 1u;
    bool used[1 << kTinyLutBits] = {};
    bool ok = true;
    for (int i = 0; i < n && ok; ++i) {
      const uint32_t h = tiny_hash(tp.keys[i], mult);
      ok = !used[h];
      used[h] = true;
    }
    if (!ok) continue;
    tp.mult = mult;
    std::memset(tp.lut, kTinyNone, sizeof(tp.lut));
    for (int i = 0; i < n; ++i) tp.lut[tiny_hash(tp.keys[i], mult)] = (uint8_t)i;
    return true;
  }
  return false;
}

// Global-table executions (global and direct paths): build the table, run
// the second pass, compact it into temporary outputs (null results counted
// as they are written) and read the group count, the overflow flag and the
// null counts back with one wait. A table that overflowed is rebuilt larger.
// The sort-path aggregations need the group count on the host first, so
// their calls wait before compacting.
mgb_status second_pass_table(Workspace* ws, const Plan& plan, const AggParams& base, int64_t R, bool direct,
                             const LdsParams& direct_lp, int64_t direct_items, const LdsGeom& geom,
                             int64_t capacity, hipStream_t stream)
{
  const int cus = device_cu_count();
  if (direct) {
    // the table is complete: every workgroup loads it into LDS for the
    // second pass over its slice
    LdsParams lp = direct_lp;
    Pass2Ops q{};
    if (!pass2_ops(plan, lp.src, q)) return fail(MGB_ERROR_UNSUPPORTED, "too many second-pass aggregations");
    q.rows = lp.src;
    mgb_status st = upload_pass2(ws, q, stream, &lp.p2);
    if (st != MGB_SUCCESS) return st;
    MGB_HIP(set_dynamic_lds((const void*)pass2_direct_kernel, (int)geom.bytes));
    {
      ProfScope prof("pass2", stream);
      pass2_direct_kernel<<<(unsigned)direct_items, kLdsBlock, geom.bytes, stream>>>(lp);
    }
    MGB_HIP(hipGetLastError());
    return MGB_SUCCESS;
  }
  // the table is complete: each row finds its group and updates the
  // second-pass fields, in launches of up to kMaxPass2 ops
  AggParams src = base;
  src.table = static_cast<uint8_t*>(ws->main.ptr);
  src.capacity = capacity;
  src.flags = nullptr;
  std::vector<Pass2Op> all;
  for (size_t c = 0; c < plan.cols.size(); ++c) {
    ValCol v = plan.cols[c].col;
    v.flag_bit = -1;
    for (auto& o : plan.cols[c].p2) all.push_back(pass2_op(plan, plan.cols[c], o, v, (int)c));
  }
  for (size_t first = 0; first < all.size(); first += kMaxPass2) {
    Pass2Ops q{};
    q.num_ops = (int)std::min<size_t>(kMaxPass2, all.size() - first);
    for (int i = 0; i < q.num_ops; ++i) q.ops[i] = all[first + i];
    const Pass2Ops* d_q = nullptr;
    mgb_status st = upload_pass2(ws, q, stream, &d_q);
    if (st != MGB_SUCCESS) return st;
    const int64_t grid = std::min<int64_t>((R + kBlock - 1) / kBlock, (int64_t)cus * 16);
    {
      ProfScope prof("pass2", stream);
      pass2_kernel<<<(unsigned)std::max<int64_t>(grid, 1), kBlock, 0, stream>>>(src, d_q);
    }
    MGB_HIP(hipGetLastError());
  }
  return MGB_SUCCESS;
}

// Compacts the table's groups into the temporary outputs (`temp`, laid out
// for at least the table's groups), counting null results when `nulls`.
mgb_status finalize_table(Workspace* ws, const Plan& plan, int64_t capacity, int64_t nblocks, size_t block_off,
                          uint32_t* rec_pos, hipStream_t stream, const OutTable& ot)
{
  uint64_t tmpl[kMaxRecordWords] = {};
  const OutTable* d_ot = nullptr;
  const uint64_t* d_tmpl = nullptr;
  mgb_status st = upload_params(ws, ot, tmpl, 1, stream, &d_ot, &d_tmpl);
  if (st != MGB_SUCCESS) return st;
  FinParams f{};
  f.table = static_cast<const uint8_t*>(ws->main.ptr);
  f.capacity = capacity;
  f.stride = plan.stride;
  f.key_mode = plan.key_mode;
  f.num_keys = (int)plan.keys.size();
  f.block_offsets = reinterpret_cast<const int64_t*>(static_cast<uint8_t*>(ws->main.ptr) + block_off);
  f.out = d_ot;
  for (size_t c = 0; c < plan.keys.size(); ++c) f.keys[c] = plan.keys[c];
  f.rec_pos = rec_pos;
  {
    ProfScope prof("finalize", stream);
    finalize_kernel<<<(unsigned)nblocks, kBlock, 0, stream>>>(f);
  }
  MGB_HIP(hipGetLastError());
  return MGB_SUCCESS;
}

mgb_status run_table(Workspace* ws, const Plan& plan, const AggParams& base, int64_t R, int64_t G_est, bool direct,
                     const LdsGeom& geom, hipStream_t stream, TempOuts& temp, int64_t& groups,
                     const TinyParams* tiny = nullptr)
{
  bool row_rec_written = false;  // the global build wrote the sort path's row-to-record map
  const int cus = device_cu_count();
  int64_t capacity = std::max<int64_t>(64, 2 * G_est + 1024);
  capacity = std::min<int64_t>(capacity, std::max<int64_t>(2 * R + 64, 64));
  int64_t nblocks = 0;
  size_t block_off = 0;
  size_t rec_pos_off = 0;  // per record: its output row (sort-path aggregations)
  // without sort-path aggregations the groups are compacted before the one
  // wait of the call
  const bool early = !plan.has_sort;
  const int n_nulls = (int)plan.outs.size() + (int)plan.keys.size();
  const int first_slot = ws->upload_slot;
  for (int attempt = 0;; ++attempt) {
    if (capacity > (int64_t)0x7fffffff) return fail(MGB_ERROR_UNSUPPORTED, "groupby table would exceed 2^31 records");
    // the previous attempt's kernels have completed (its counters were
    // read): its parameter slots are free again
    ws->upload_slot = first_slot;
    nblocks = (capacity + 1 + kBlock - 1) / kBlock;
    const size_t table_bytes = align_up((size_t)(capacity + 1) * plan.stride, 256);
    const size_t counters_off = table_bytes;
    block_off = counters_off + 256;
    rec_pos_off = block_off + align_up((size_t)nblocks * 8, 256);
    const size_t need = rec_pos_off + (plan.has_sort ? align_up((size_t)(capacity + 1) * 4, 256) : 0);
    MGB_RESERVE(ws->main, need);
    uint8_t* wsb = static_cast<uint8_t*>(ws->main.ptr);

    FillParams fp{};
    fp.table = reinterpret_cast<uint64_t*>(wsb);
    fp.num_records = capacity + 1;
    fp.words = plan.stride / 8;
    fp.special_hdr = plan.key_mode == KEY_INT ? 0ull : kEmptyRowRef;
    record_template(plan, plan.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef, fp.tmpl);
    const int64_t fill_words = fp.num_records * fp.words;
    const unsigned fill_grid = (unsigned)std::min<int64_t>((fill_words + kBlock - 1) / kBlock, (int64_t)cus * 32);
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(wsb + counters_off);
    fp.zero[0] = counters;  // cleared with the table
    fp.zero_words[0] = 32;
    if (early) {
      // the temporary outputs (and their null counters, cleared with the
      // table) are laid out before the aggregation
      OutTable ot{};
      mgb_status st = layout_temp(ws, plan, std::min<int64_t>(capacity + 1, R), stream, temp, ot);
      if (st != MGB_SUCCESS) return st;
      fp.zero[1] = temp.null_counts;
      fp.zero_words[1] = n_nulls;
    }
    {
      ProfScope prof("fill", stream);
      fill_table_kernel<<<fill_grid, kBlock, 0, stream>>>(fp);
    }

    AggParams ap = base;
    ap.table = wsb;
    ap.capacity = capacity;
    ap.counters = counters;
    ap.max_groups = std::max<int64_t>(1, capacity * 3 / 4);
    LdsParams direct_lp{};  // the direct launch, for its second pass
    int64_t direct_items = 0;
    if (direct) {
      LdsParams lp{};
      lp.src = ap;
      if (!single_launch_ops(plan, lp.src)) return fail(MGB_ERROR_UNSUPPORTED, "plan does not fit the LDS path");
      lp.src.flags = nullptr;
      for (int c = 0; c < lp.src.num_vals; ++c) lp.src.vals[c].flag_bit = -1;
      const int64_t items = std::max<int64_t>(1, std::min<int64_t>((int64_t)cus * 2, (R + 16383) / 16384));
      lp.rows_per_item = (R + items - 1) / items;
      uint64_t tmpl[kMaxRecordWords] = {};
      record_template(plan, plan.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef, tmpl);
      OutTable dummy{};
      const OutTable* d_ot = nullptr;
      const uint64_t* d_tmpl = nullptr;
      mgb_status st = upload_params(ws, dummy, tmpl, lds_stride_of(plan.stride) / 8, stream, &d_ot, &d_tmpl);
      if (st != MGB_SUCCESS) return st;
      lp.tmpl = d_tmpl;
      lp.out = d_ot;
      lp.slots = geom.slots;
      lp.cap = geom.cap;
      lp.words = lds_stride_of(plan.stride) / 8;
      lp.direct = 1;
      lp.log2p = 0;
      if (tiny) {
        // the few-groups execution: a small table (for keys the sample
        // missed) and the striped copies of the candidates' records
        const int words = lds_stride_of(plan.stride) / 8;
        const size_t tiny_bytes = tiny_lds_bytes(words);
        const size_t avail = lds_limit() - tiny_bytes - 64;
        const int slots = (int)std::min<int64_t>(1023, (int64_t)((avail - 16 - (size_t)kLdsBlock * 4) / ((size_t)words * 8)) - 1);
        if (slots < 64) return fail(MGB_ERROR_UNSUPPORTED, "few-groups execution does not fit LDS");
        lp.slots = slots;
        lp.cap = (int)(slots * 0.62);
        const size_t table_bytes = align_up(lds_table_bytes(slots, words), 16);
        const size_t total = table_bytes + tiny_bytes;
        const TinyParams* d_tp = nullptr;
        st = upload_tiny(ws, *tiny, stream, &d_tp);
        if (st != MGB_SUCCESS) return st;
        // compiled for a single value column of a common type and a 4- or
        // 8-byte key; the general instance otherwise
        using TinyFn = void (*)(const LdsParams, const TinyParams*, int32_t);
        TinyFn fn = lp.src.num_vals == 1 ? tiny_agg_kernel<1, 0, TV_ANY> : tiny_agg_kernel<2, 0, TV_ANY>;
        if (lp.src.num_vals == 1) {
          const int kw = lp.src.keys[0].stride == lp.src.keys[0].width ? lp.src.keys[0].width : 0;
          const int tv = lp.src.vals[0].stride == lp.src.vals[0].width
                             ? tiny_val_type(lp.src.vals[0].width, lp.src.vals[0].vclass)
                             : TV_ANY;
          if (kw == 4) {
            fn = tv == TV_I32 ? tiny_agg_kernel<1, 4, TV_I32> : tv == TV_I64 ? tiny_agg_kernel<1, 4, TV_I64>
                 : tv == TV_F32 ? tiny_agg_kernel<1, 4, TV_F32> : tv == TV_F64 ? tiny_agg_kernel<1, 4, TV_F64> : fn;
          } else if (kw == 8) {
            fn = tv == TV_I32 ? tiny_agg_kernel<1, 8, TV_I32> : tv == TV_I64 ? tiny_agg_kernel<1, 8, TV_I64>
                 : tv == TV_F32 ? tiny_agg_kernel<1, 8, TV_F32> : tv == TV_F64 ? tiny_agg_kernel<1, 8, TV_F64> : fn;
          }
        }
        MGB_HIP(set_dynamic_lds((const void*)fn, (int)total));
        {
          ProfScope prof("lds_aggregate", stream);
          hipLaunchKernelGGL(fn, dim3((unsigned)items), dim3(kLdsBlock), total, stream, lp, d_tp, (int32_t)table_bytes);
        }
        MGB_HIP(hipGetLastError());
      } else {
        MGB_HIP(set_dynamic_lds((const void*)lds_agg_kernel<false>, (int)geom.bytes));
        {
          ProfScope prof("lds_aggregate", stream);
          lds_agg_kernel<false><<<(unsigned)items, kLdsBlock, geom.bytes, stream>>>(lp);
        }
        MGB_HIP(hipGetLastError());
      }
      direct_lp = lp;
      direct_items = items;
    } else {
      // launch in batches of value columns / ops
      size_t ci = 0;
      bool first = true;
      do {
        AggParams lp = ap;
        lp.num_vals = 0;
        lp.num_ops = 0;
        lp.rows_op = -1;
        lp.flags = nullptr;
        if (first && plan.rows_off >= 0) {
          lp.ops[0] = OpDesc{OP_ROWS, plan.rows_off};
          lp.rows_op = 0;
          lp.num_ops = 1;
        }
        while (ci < plan.cols.size() && lp.num_vals < kMaxVals &&
               lp.num_ops + (int)plan.cols[ci].ops.size() <= kMaxOps) {
          const PlanCol& pc = plan.cols[ci];
          ValCol vc = pc.col;
          vc.flag_bit = -1;
          vc.nulls_off = pc.nulls_off;
          vc.op_begin = lp.num_ops;
          for (auto& o : pc.ops) lp.ops[lp.num_ops++] = OpDesc{o.first, o.second};
          vc.op_end = lp.num_ops;
          column_program(vc, pc.ops);
          lp.vals[lp.num_vals++] = vc;
          ++ci;
        }
        if (lp.num_vals == 0 && ci < plan.cols.size())
          return fail(MGB_ERROR_UNSUPPORTED, "a value column needs more accumulator ops than one launch holds");
        // the sort path's row-to-record map, from the first launch
        if (first && plan.has_sort) {
          MGB_RESERVE(ws->rowrec, (size_t)R * 4);
          lp.row_rec = static_cast<uint32_t*>(ws->rowrec.ptr);
          row_rec_written = true;
        }
        const int64_t grid = std::min<int64_t>((R + kBlock - 1) / kBlock, (int64_t)cus * 16);
        {
          ProfScope prof("aggregate", stream);
          aggregate_kernel<<<(unsigned)std::max<int64_t>(grid, 1), kBlock, 0, stream>>>(lp);
        }
        MGB_HIP(hipGetLastError());
        first = false;
      } while (ci < plan.cols.size());
    }
    if (plan.has_pass2 && early) {
      // a table that overflowed is missing groups: the second pass skips
      // their rows, and the attempt is discarded below
      mgb_status st = second_pass_table(ws, plan, base, R, direct, direct_lp, direct_items, geom, capacity, stream);
      if (st != MGB_SUCCESS) return st;
    }

    int64_t* block_counts = reinterpret_cast<int64_t*>(wsb + block_off);
    count_occupied_kernel<<<(unsigned)nblocks, kBlock, 0, stream>>>(wsb, capacity, plan.stride, plan.key_mode,
                                                                     block_counts, nullptr);
    scan_blocks_kernel<<<1, 1024, 0, stream>>>(block_counts, nblocks, counters + 2, nullptr);
    MGB_HIP(hipGetLastError());
    if (early) {
      // the temporary outputs for every record the table can hold (laid out
      // above; layout_temp gives the same buffers for the same size)
      OutTable ot{};
      mgb_status st = layout_temp(ws, plan, std::min<int64_t>(capacity + 1, R), stream, temp, ot);
      if (st != MGB_SUCCESS) return st;
      ot.nulls = temp.null_counts;
      st = finalize_table(ws, plan, capacity, nblocks, block_off, nullptr, stream, ot);
      if (st != MGB_SUCCESS) return st;
      MGB_HIP(hipMemcpyAsync(ws->pinned + kPinnedNulls, temp.null_counts, (size_t)n_nulls * 8, hipMemcpyDeviceToHost,
                             stream));
    }
    MGB_HIP(hipMemcpyAsync(ws->pinned, counters, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    MGB_HIP(hipStreamSynchronize(stream));
    if (ws->pinned[1] == 0) {
      groups = (int64_t)ws->pinned[2];
      if (early) {
        temp.nulls_known = true;
        temp.host_nulls.assign(ws->pinned + kPinnedNulls, ws->pinned + kPinnedNulls + n_nulls);
        return MGB_SUCCESS;
      }
      // sort-path aggregations: second pass, compaction and the sorts after the wait
      if (plan.has_pass2 && groups > 0) {
        mgb_status st = second_pass_table(ws, plan, base, R, direct, direct_lp, direct_items, geom, capacity, stream);
        if (st != MGB_SUCCESS) return st;
      }
      break;
    }
    if (attempt >= 6) return fail(MGB_ERROR_HIP, "groupby table kept overflowing");
    capacity = std::min<int64_t>(capacity * 4, std::max<int64_t>(2 * R + 64, 64));
  }

  OutTable ot{};
  mgb_status st = layout_temp(ws, plan, groups, stream, temp, ot);
  if (st != MGB_SUCCESS) return st;
  if (groups == 0) return MGB_SUCCESS;
  uint32_t* rec_pos = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(ws->main.ptr) + rec_pos_off);
  st = finalize_table(ws, plan, capacity, nblocks, block_off, rec_pos, stream, ot);
  if (st != MGB_SUCCESS) return st;
  AggParams rows = base;
  rows.table = static_cast<uint8_t*>(ws->main.ptr);
  rows.capacity = capacity;
  rows.row_rec = row_rec_written ? static_cast<uint32_t*>(ws->rowrec.ptr) : nullptr;
  return run_sort_aggs(ws, plan, rows, rec_pos, groups, stream, temp);
}

// Partitioned execution: hash-partition the rows, then aggregate every
// partition in LDS. Partitions larger than one work item are split over
// several items whose groups meet in a small global table (as do the groups
// of an item that outgrows its LDS table). Returns MGB_SUCCESS with `done`
// false when the estimate asks for more partitions than supported (the
// caller then uses the global path).
// The compact partitioned execution of a call (lean.hip), when it has one: a
// single 4- or 8-byte integer key, at most kLeanCols value columns of 4 or 8
// bytes whose ops the column programs apply, no second pass or sort, at most
// 4096 partitions, null masks starting at a word boundary.
// MGB_LEAN=off turns the execution off.
struct LeanPlan {
  bool on = false;
  bool direct = false;  // the scatter writes rows straight out (partitions far outnumber a tile's rows)
  int k0 = 0, k1 = 0;  // key column widths (k1: a second key column, else 0)
  int kb = 0;          // key bytes of a record
  int tail = 0;        // two-column keys: LDS tail bytes of a record (lean_tail_bytes)
  int w[kLeanCols] = {};
  LeanKeyOffsets ko{};
  LeanOffsets lo{};
  int rec_bytes = 0;
  int tile = 0;        // rows staged per scatter tile
  size_t lds = 0;      // the scatter's LDS bytes
  int flag_bits = 0;   // LeanScatterArgs::flag_bits
};

// Whether a call fits the compact partitioned execution, and its record
// shape (everything but the partition count). MGB_LEAN=off turns it off.
bool lean_shape_of(const Plan& plan, const AggParams& src, LeanPlan& lp)
{
  lp = LeanPlan{};
  const char* e = std::getenv("MGB_LEAN");
  if (e && std::string(e) == "off") return false;
  if (plan.compare_key_nulls || plan.has_pass2 || plan.has_sort || src.num_vals > kLeanCols || !src.program_cols)
    return false;
  auto plain = [](const KeyCol& k) {
    return (k.width == 4 || k.width == 8) && k.stride == k.width && !(k.mask && (k.bit_off & 31));
  };
  if (plan.key_mode == KEY_INT && plan.keys.size() == 1) {
    if (!plain(plan.keys[0])) return false;
    lp.k0 = plan.keys[0].width;
  } else if (plan.key_mode == KEY_ROWREF && plan.keys.size() == 2) {
    // two integer columns: compared as raw bits (no float normalisation)
    for (const KeyCol& k : plan.keys)
      if (!plain(k) || (k.vclass != VC_SIGNED && k.vclass != VC_UNSIGNED)) return false;
    lp.k0 = plan.keys[0].width;
    lp.k1 = plan.keys[1].width;
  } else {
    return false;
  }
  lp.ko = lean_key_offsets(lp.k0, lp.k1);
  lp.kb = lp.ko.bytes;
  lp.tail = lp.k1 ? lean_tail_bytes(lp.kb) : 0;
  lp.flag_bits = 0xfff;
  for (int c = 0; c < src.num_vals; ++c) {
    const ValCol& v = src.vals[c];
    if (v.lds_generic || (v.width != 4 && v.width != 8) || v.stride != v.width || (v.mask && (v.bit_off & 31)) ||
        v.flag_bit >= 8)
      return false;
    lp.w[c] = v.width;
    if (v.mask && v.flag_bit >= 0) lp.flag_bits = (lp.flag_bits & ~(15 << (4 * c))) | v.flag_bit << (4 * c);
  }
  if (!lean_supported(lp.k0, lp.k1, lp.w)) return false;
  lp.lo = lean_offsets(lp.kb, lp.w[0], lp.w[1], lp.w[2]);
  for (int c = 0; c < src.num_vals; ++c)
    if (lp.lo.off[c] < 0) return false;
  lp.rec_bytes = lp.lo.bytes <= 16 ? 16 : 32;
  return true;
}

// The compact scatter's tile for 2^log2p partitions (hot_stride: the LDS
// stride of hot-key records, 0 without hot keys); lp.on when it fits LDS.
void lean_tile(LeanPlan& lp, int log2p, int hot_stride)
{
  lp.on = false;
  lp.direct = false;
  lp.tile = 0;
  if (lp.rec_bytes == 0 || log2p > kLeanMaxLog2P) return;
  const int P = 1 << log2p;
  for (int tt = kScatterTile; tt >= kLdsBlock; tt -= kLdsBlock)
    if (lean_scatter_lds(P, tt, lp.rec_bytes, hot_stride).total <= lds_limit()) {
      lp.tile = tt;
      break;
    }
  lp.lds = lean_scatter_lds(P, lp.tile, lp.rec_bytes, hot_stride).total;
  lp.on = lp.tile > 0;
  // a tile of fewer than two rows per partition makes no runs to copy out
  // together, while every tile scans all partitions: rows go straight out
  // (MGB_LEAN_DIRECT=on / off: always / never)
  const char* e = std::getenv("MGB_LEAN_DIRECT");
  const std::string mode = e ? e : "";
  lp.direct = mode == "on" || (mode != "off" && (!lp.on || 2 * lp.tile <= P));
  if (lp.direct) {
    lp.tile = kScatterItems * kLdsBlock;
    lp.lds = lean_scatter_lds(P, 0, lp.rec_bytes, hot_stride).total;
    lp.on = lp.lds <= lds_limit();
  }
}

// The representative records of two-column hot keys (records rep_base + i):
// their key columns at the record's key offsets, the rest zero. Global-table
// groups of hot keys refer to them.
__global__ __launch_bounds__(kHotMax) void hot_reps_kernel(const HotSet* hot, uint8_t* recs, int64_t rep_base,
                                                          int32_t rec_bytes, int32_t off0, int32_t k0, int32_t off1,
                                                          int32_t k1)
{
  const int i = threadIdx.x;
  const int n = (int)min(hot->n, (uint32_t)kHotMax);
  if (i >= n) return;
  uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
  const uint64_t c0 = hot->keys[i].c0, c1 = hot->keys[i].c1;
  w[off0 / 4] = (uint32_t)c0;
  if (k0 == 8) w[off0 / 4 + 1] = (uint32_t)(c0 >> 32);
  w[off1 / 4] = (uint32_t)c1;
  if (k1 == 8) w[off1 / 4 + 1] = (uint32_t)(c1 >> 32);
  uint32_t* dst = reinterpret_cast<uint32_t*>(recs + (rep_base + i) * (int64_t)rec_bytes);
  for (int j = 0; j < rec_bytes / 4; ++j) dst[j] = w[j];
}

mgb_status run_partitioned(Workspace* ws, const Plan& plan, const AggParams& base, int64_t R, int64_t G_est,
                           const LdsGeom& plan_geom, hipStream_t stream, TempOuts& temp, int64_t& groups, bool& done)
{
  done = false;
  const int cus = device_cu_count();
  AggParams src = base;
  if (!single_launch_ops(plan, src)) return MGB_SUCCESS;
  int nullable = 0;
  for (int c = 0; c < src.num_vals; ++c) nullable += src.vals[c].flag_bit >= 0;
  const int flag_bytes = nullable == 0 ? 0 : nullable <= 8 ? 1 : nullable <= 16 ? 2 : 4;
  // the compact execution's record shape; a two-column key's LDS records
  // carry a tail after the plan's fields (8-byte aligned), so fewer fit
  LeanPlan lean_shape;
  bool lean_ok = lean_shape_of(plan, src, lean_shape);
  const int tail_off = (int)align_up((size_t)plan.stride, 8);
  LdsGeom geom = plan_geom;
  if (lean_ok && lean_shape.tail) {
    const LdsGeom g2 = lds_geometry(tail_off + lean_shape.tail);
    if (g2.ok) geom = g2;
    else lean_ok = false;
  }
  int64_t est = std::max<int64_t>(G_est, 1);
  for (int attempt = 0; attempt < 4; ++attempt) {
    const int slot0 = ws->upload_slot;  // parameter slots of an attempt are reused by the next
    // groups per partition: the table's capacity less four standard deviations
    // of a partition's group count (a partition over it spills the rest)
    auto partitions_for = [&](const LdsGeom& g) {
      const double m_target = std::max(g.cap * 0.5, g.cap - 4.0 * std::sqrt((double)g.cap));
      int l = 1;
      while ((double)(1ll << l) * m_target < (double)est) ++l;
      return l;
    };
    int log2p = partitions_for(geom);
    if (lean_ok && lean_shape.tail && log2p > kLeanMaxLog2P) {
      // more partitions than the compact scatter takes: the general records
      lean_ok = false;
      geom = plan_geom;
      log2p = partitions_for(geom);
    }
    if (log2p > 15) return MGB_SUCCESS;
    // more partitions than the groups need when the rows are many, so that a
    // partition of mean size is one work item (the items of a partition cut
    // in several fold all of its groups into the global table), while every
    // partition keeps 64 groups or more (fewer put many lanes of a wave on
    // one LDS record; such partitions are cut instead, and fold few groups)
    const int64_t item_target = std::max<int64_t>(65536, (R + 2 * (int64_t)cus - 1) / (2 * (int64_t)cus));
    while (log2p < 12 && (R >> log2p) > item_target && (est >> (log2p + 1)) >= 64) ++log2p;
    const int64_t P = 1ll << log2p;
    const int64_t nw = partition_chunks(R, log2p, cus);
    const int64_t chunk = (R + nw - 1) / nw;
    const int64_t nh = P * nw;
    const int64_t nb = (nh + kScanTile - 1) / kScanTile;
    const int64_t nb_items = (P + 1 + kScanTile - 1) / kScanTile;
    // work items: partitions longer than max_item_rows are split, both for
    // skew (four times the mean partition) and for parallelism (few large
    // partitions would leave most CUs idle)
    const int64_t par_rows = (R + 2 * (int64_t)cus - 1) / (2 * (int64_t)cus);
    const int64_t max_item_rows = std::max<int64_t>(65536, std::min<int64_t>(4 * ((R + P - 1) / P), par_rows));
    const int64_t item_grid = P + (R + max_item_rows - 1) / max_item_rows;
    // global table for the groups of split or spilled items: no more than the
    // items could hold, nor than the groups there are
    const int64_t split_bound = (R + max_item_rows - 1) / max_item_rows + P / 16 + 1;
    // (and the hot keys' groups)
    int64_t gcap = 2 * std::min<int64_t>({R, split_bound * (int64_t)(geom.slots + 1), 2 * est}) + 1024 + kHotMax;
    gcap = std::min<int64_t>(gcap, 2 * R + 64 + kHotMax);
    const int64_t gblocks = (gcap + 1 + kBlock - 1) / kBlock;

    // record layout of a partitioned row: 8-byte fields, then 4, 2 and 1-byte
    // fields (the source row index among the 4-byte ones, kept for ARGMIN /
    // ARGMAX), then the validity bits; the size is a multiple of 8
    bool need_rowid = false;
    for (const PlanCol& pc : plan.cols)
      for (auto& o : pc.p2) need_rowid |= o.first != OP_M2;
    std::vector<int> koff(plan.keys.size()), voff(src.num_vals);
    int rec = 0, rowid_off = -1;
    for (int w : {8, 4, 2, 1}) {
      for (size_t c = 0; c < plan.keys.size(); ++c)
        if (plan.keys[c].width == w) { koff[c] = rec; rec += w; }
      for (int c = 0; c < src.num_vals; ++c)
        if (src.vals[c].width == w) { voff[c] = rec; rec += w; }
      if (w == 4 && need_rowid) { rowid_off = rec; rec += 4; }
    }
    int flags_rec_off = (int)align_up((size_t)rec, (size_t)std::max(flag_bytes, 1));
    rec = flags_rec_off + flag_bytes;
    int rec_bytes = (int)align_up((size_t)rec, 8);
    // the compact records, when the call has them. The estimate's fine
    // counts serve as the histogram when they cover these partitions and
    // chunks; when they leave the hot keys' rows out, only the compact
    // scatter (which aggregates those rows apart) may take them.
    const bool fine_ok = ws->fine_valid && log2p <= kFineBits && ws->fine_nw == nw && ws->fine_chunk == chunk;
    const int hot_words = lds_stride_of(plan.stride) / 8;
    LeanPlan lean = lean_shape;
    lean.on = false;
    bool use_hot = false;
    if (lean_ok) {
      if (fine_ok && ws->fine_hot) {
        lean_tile(lean, log2p, hot_words * 8);
        use_hot = lean.on;
      }
      if (!lean.on) lean_tile(lean, log2p, 0);  // without hot keys (then a second pass counts every row)
    }
    const bool coarsen = fine_ok && (!ws->fine_hot || use_hot);
    // LDS record words (two-column compact keys: the plan's fields, then the tail)
    const int lds_words = lds_stride_of(lean.on && lean.tail ? tail_off + lean.tail : plan.stride) / 8;
    if (lean.on) {
      koff[0] = lean.ko.off0;
      if (lean.k1) koff[1] = lean.ko.off1;
      for (int c = 0; c < src.num_vals; ++c) voff[c] = lean.lo.off[c];
      flags_rec_off = lean.kb;
      rec_bytes = lean.rec_bytes;
    }

    // workspace: records, hist, offsets, block sums, item counts/offsets, global table, counters
    size_t off = 0;
    const size_t recs_off = off;  // the partitioned records, then kHotMax representative records of hot keys
    off += align_up((size_t)(R + kHotMax) * rec_bytes, 256);
    const size_t hist_off = off;
    off += align_up((size_t)nh * 4, 256);
    const size_t offs_off = off;
    off += align_up((size_t)nh * 4, 256);
    const size_t bsum_off = off;
    off += align_up((size_t)nb * 4, 256);
    const size_t icnt_off = off;
    off += align_up((size_t)(P + 1) * 4, 256);
    const size_t ioff_off = off;
    off += align_up((size_t)(P + 1) * 4, 256);
    const size_t ibsum_off = off;
    off += align_up((size_t)nb_items * 4, 256);
    const size_t iglob_off = off;  // second pass: per item, its groups went global
    off += plan.has_pass2 ? align_up((size_t)item_grid * 4, 256) : 0;
    const size_t gtab_off = off;
    off += align_up((size_t)(gcap + 1) * plan.stride, 256);
    const size_t gblk_off = off;
    off += align_up((size_t)gblocks * 8, 256);
    const size_t ctr_off = off;
    off += 256;
    MGB_RESERVE(ws->main, off);
    uint8_t* wsb = static_cast<uint8_t*>(ws->main.ptr);
    // ctr: [0] partitioned rows [1] groups written by items [2] unused [3] items
    //      [4] global groups [5] global overflow [6] global occupied
    unsigned long long* ctr = reinterpret_cast<unsigned long long*>(wsb + ctr_off);
    MGB_HIP(hipMemsetAsync(ctr, 0, 256, stream));

    // every group has a row; items write at most P * (slots + 1) groups, the
    // global table holds at most gcap
    const int64_t U = std::min<int64_t>(R, P * (int64_t)(geom.slots + 1) + gcap);
    OutTable ot{};
    mgb_status st = layout_temp(ws, plan, U, stream, temp, ot);
    if (st != MGB_SUCCESS) return st;
    const int n_nulls = (int)plan.outs.size() + (int)plan.keys.size();
    ot.nulls = temp.null_counts;
    MGB_HIP(hipMemsetAsync(temp.null_counts, 0, (size_t)n_nulls * 8, stream));
    uint64_t tmpl[kMaxRecordWords] = {};
    record_template(plan, plan.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef, tmpl);
    if (lean.on && lean.tail) {
      // two-column compact keys: an empty header is kEmptyIntKey (key part
      // A), and the tail reads as not yet written
      tmpl[0] = kEmptyIntKey;
      for (int w = tail_off / 8; w < (tail_off + lean.tail) / 8; ++w) tmpl[w] = ~0ull;
    }
    const OutTable* d_ot = nullptr;
    const uint64_t* d_tmpl = nullptr;
    st = upload_params(ws, ot, tmpl, lds_words, stream, &d_ot, &d_tmpl);
    if (st != MGB_SUCCESS) return st;

    PartParams pp{};
    pp.src = src;
    pp.chunk = chunk;
    pp.log2p = log2p;
    pp.nblocks = (int32_t)nw;
    pp.hist = reinterpret_cast<uint32_t*>(wsb + hist_off);
    pp.offsets = reinterpret_cast<const uint32_t*>(wsb + offs_off);
    pp.out = wsb + recs_off;
    pp.rec_bytes = rec_bytes;
    pp.rec_magic = ((1ull << 32) + (uint64_t)(rec_bytes / 8) - 1) / (uint64_t)(rec_bytes / 8);
    pp.flag_bytes = flag_bytes;
    pp.flags_off = flags_rec_off;
    for (size_t c = 0; c < plan.keys.size(); ++c) pp.key_off[c] = koff[c];
    for (int c = 0; c < src.num_vals; ++c) pp.val_off[c] = voff[c];
    pp.rowid_off = rowid_off;
    const size_t part_lds = (size_t)P * 4;
    MGB_HIP(set_dynamic_lds((const void*)part_hist_kernel, (int)part_lds));
    MGB_HIP(set_dynamic_lds((const void*)part_scatter_kernel, (int)part_lds));
    if (coarsen) {
      // the estimate's pass counted the rows per fine bucket of these chunks
      ProfScope prof("partition_hist", stream);
      if (log2p <= kFineBits - 6) {
        hist_coarsen_few_kernel<<<(unsigned)nw, kBlock, 0, stream>>>(static_cast<const uint32_t*>(ws->fine.ptr),
                                                                    (int32_t)nw, log2p, pp.hist);
      } else {
        const dim3 grid((unsigned)((nw + 63) / 64), (unsigned)((P + 63) / 64));
        hist_coarsen_kernel<<<grid, kBlock, 0, stream>>>(static_cast<const uint32_t*>(ws->fine.ptr), (int32_t)nw, log2p,
                                                        pp.hist);
      }
    } else {
      ProfScope prof("partition_hist", stream);
      part_hist_kernel<<<(unsigned)nw, kBlock, part_lds, stream>>>(pp);
    }
    uint32_t* bsum = reinterpret_cast<uint32_t*>(wsb + bsum_off);
    scan_reduce_u32_kernel<<<(unsigned)nb, 256, 0, stream>>>(pp.hist, nh, bsum);
    scan_bsums_u32_kernel<<<1, 1024, 0, stream>>>(bsum, nb, ctr + 0);
    scan_apply_u32_kernel<<<(unsigned)nb, 256, 0, stream>>>(pp.hist, nh, bsum, reinterpret_cast<uint32_t*>(wsb + offs_off));
    // the largest tile (a multiple of the workgroup size) whose records fit LDS
    const size_t fixed_lds = (size_t)P * 8 + 16 * 4 + (size_t)kScatterTile * 2 + 16;
    int tile = 0;
    int tile_max = kScatterTile;
    if (const char* e = std::getenv("MGB_SCATTER_TILE"))  // tuning override (rows, multiple of 1024)
      tile_max = std::max(kLdsBlock, std::min(kScatterTile, std::atoi(e) / kLdsBlock * kLdsBlock));
    for (int tt = tile_max; tt >= kLdsBlock; tt -= kLdsBlock)
      if (fixed_lds + (size_t)tt * rec_bytes <= lds_limit()) { tile = tt; break; }
    pp.tile_rows = tile;
    const size_t tiled_lds = fixed_lds + (size_t)tile * rec_bytes;

    // work items per partition
    uint32_t* icnt = reinterpret_cast<uint32_t*>(wsb + icnt_off);
    uint32_t* ioff = reinterpret_cast<uint32_t*>(wsb + ioff_off);
    uint32_t* ibsum = reinterpret_cast<uint32_t*>(wsb + ibsum_off);
    count_items_kernel<<<(unsigned)((P + 1 + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
      pp.offsets, (int)nw, (int)P, ctr + 0, max_item_rows, icnt);
    scan_reduce_u32_kernel<<<(unsigned)nb_items, 256, 0, stream>>>(icnt, P + 1, ibsum);
    scan_bsums_u32_kernel<<<1, 1024, 0, stream>>>(ibsum, nb_items, ctr + 3);
    scan_apply_u32_kernel<<<(unsigned)nb_items, 256, 0, stream>>>(icnt, P + 1, ibsum, ioff);

    // the global table
    uint8_t* gtab = wsb + gtab_off;
    FillParams fp{};
    fp.table = reinterpret_cast<uint64_t*>(gtab);
    fp.num_records = gcap + 1;
    fp.words = plan.stride / 8;
    fp.special_hdr = plan.key_mode == KEY_INT ? 0ull : kEmptyRowRef;
    record_template(plan, plan.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef, fp.tmpl);
    {
      const int64_t fill_words = fp.num_records * fp.words;
      const unsigned fill_grid = (unsigned)std::min<int64_t>((fill_words + kBlock - 1) / kBlock, (int64_t)cus * 32);
      ProfScope prof("fill", stream);
      fill_table_kernel<<<fill_grid, kBlock, 0, stream>>>(fp);
    }

    LdsParams lp{};
    lp.src = src;
    lp.src.num_rows = R;
    lp.src.skip_null_keys = 0;
    lp.src.table = gtab;
    lp.src.capacity = gcap;
    lp.src.counters = ctr + 4;
    lp.src.max_groups = std::max<int64_t>(1, gcap * 3 / 4);
    uint8_t* recs = wsb + recs_off;
    lp.src.flags = flag_bytes ? recs + flags_rec_off : nullptr;
    lp.src.flag_bytes = flag_bytes;
    lp.src.flag_stride = rec_bytes;
    lp.src.rec_words = rec_bytes <= 32 ? rec_bytes / 8 : 0;
    lp.src.rec_base = recs;
    for (size_t c = 0; c < plan.keys.size(); ++c) {
      lp.src.keys[c].data = recs + koff[c];
      lp.src.keys[c].stride = rec_bytes;
      lp.src.keys[c].mask = nullptr;
      lp.src.keys[c].bit_off = 0;
    }
    for (int c = 0; c < src.num_vals; ++c) {
      lp.src.vals[c].data = recs + voff[c];
      lp.src.vals[c].stride = rec_bytes;
      lp.src.vals[c].mask = nullptr;
      lp.src.vals[c].bit_off = 0;
      if (c < kProgramCols) lp.src.prog[c] = pack_program(lp.src.vals[c], voff[c]);
    }
    lp.offsets = pp.offsets;
    lp.total_rows = ctr + 0;
    lp.tmpl = d_tmpl;
    lp.out = d_ot;
    lp.out_counters = ctr + 1;
    lp.nblocks = (int32_t)nw;
    lp.log2p = log2p;
    lp.slots = geom.slots;
    lp.cap = geom.cap;
    lp.words = lds_words;
    lp.direct = 0;
    lp.item_offsets = ioff;
    lp.max_item_rows = max_item_rows;
    if (plan.has_pass2) {
      Pass2Ops q{};
      if (!pass2_ops(plan, lp.src, q)) return MGB_SUCCESS;  // the global execution takes it
      q.rows = lp.src;
      q.rowid = need_rowid ? recs + rowid_off : nullptr;
      q.rowid_stride = rec_bytes;
      q.item_global = reinterpret_cast<uint32_t*>(wsb + iglob_off);
      st = upload_pass2(ws, q, stream, &lp.p2);
      if (st != MGB_SUCCESS) return st;
    }
    // the LDS parameters in device memory: the compact kernels' full view
    // (spilled rows, the groups' output, the hot keys' fold)
    const LdsParams* d_lp = nullptr;
    if (lean.on) {
      st = upload_lds(ws, lp, stream, &d_lp);
      if (st != MGB_SUCCESS) return st;
    }
    if (lean.on) {
      LeanScatterArgs sa{};
      const KeyCol& kc = plan.keys[0];
      sa.key = kc.data;
      sa.key_mask = src.skip_null_keys && kc.mask ? kc.mask + (kc.bit_off >> 5) : nullptr;
      if (lean.k1) {
        const KeyCol& kc1 = plan.keys[1];
        sa.key1 = kc1.data;
        sa.key1_mask = src.skip_null_keys && kc1.mask ? kc1.mask + (kc1.bit_off >> 5) : nullptr;
      }
      for (int c = 0; c < src.num_vals; ++c) {
        sa.val[c] = src.vals[c].data;
        sa.mask[c] = src.vals[c].mask ? src.vals[c].mask + (src.vals[c].bit_off >> 5) : nullptr;
      }
      sa.offsets = pp.offsets;
      sa.out = pp.out;
      sa.rows = R;
      sa.chunk = chunk;
      sa.log2p = log2p;
      sa.nblocks = (int32_t)nw;
      sa.tile_rows = lean.tile;
      sa.flag_bits = lean.flag_bits;
      sa.direct = lean.direct ? 1 : 0;
      if (use_hot) {
        sa.hot = static_cast<const HotSet*>(ws->hot.ptr);
        sa.lp = d_lp;
        sa.tmpl = d_tmpl;
        sa.hot_words = hot_words;
        sa.rows_off = lp.src.rows_op >= 0 ? lp.src.ops[lp.src.rows_op].off : -1;
        sa.key_signed = !lean.k1 && plan.keys[0].vclass == VC_SIGNED;
        sa.rowref = lean.k1 != 0;
        sa.rep_base = R;
        for (int c = 0; c < src.num_vals; ++c) {
          const ValCol& v = lp.src.vals[c];
          sa.cols[c] = LeanCol{v.sum_off, v.min_off, v.max_off, v.nulls_off,
                               v.vclass == VC_SIGNED ? LEAN_SIGNED : v.vclass == VC_UNSIGNED ? LEAN_UNSIGNED : LEAN_FLOAT,
                               -1};
        }
        if (lean.k1)
          hot_reps_kernel<<<1, kHotMax, 0, stream>>>(sa.hot, recs, R, rec_bytes, lean.ko.off0, lean.k0, lean.ko.off1,
                                                     lean.k1);
      }
      ProfScope prof("partition_scatter", stream);
      st = lean_scatter(lean.k0, lean.k1, lean.w, sa, (unsigned)nw, lean.lds, stream);
      if (st != MGB_SUCCESS) return st;
    } else if (tile > 0) {
      MGB_HIP(set_dynamic_lds((const void*)part_scatter_tiled_kernel, (int)tiled_lds));
      ProfScope prof("partition_scatter", stream);
      part_scatter_tiled_kernel<<<(unsigned)nw, kLdsBlock, tiled_lds, stream>>>(pp);
    } else {
      ProfScope prof("partition_scatter", stream);
      part_scatter_kernel<<<(unsigned)nw, kBlock, part_lds, stream>>>(pp);
    }
    MGB_HIP(hipGetLastError());

    if (lean.on) {
      LeanAggArgs aa{};
      aa.recs = recs;
      aa.offsets = lp.offsets;
      aa.item_offsets = lp.item_offsets;
      aa.total_rows = lp.total_rows;
      aa.tmpl = lp.tmpl;
      aa.lp = d_lp;
      aa.max_item_rows = lp.max_item_rows;
      aa.nblocks = lp.nblocks;
      aa.log2p = lp.log2p;
      aa.slots = lp.slots;
      aa.cap = lp.cap;
      aa.words = lp.words;
      aa.rows_off = lp.src.rows_op >= 0 ? lp.src.ops[lp.src.rows_op].off : -1;
      aa.key_signed = !lean.k1 && plan.keys[0].vclass == VC_SIGNED;
      aa.tail_off = tail_off;
      for (int c = 0; c < src.num_vals; ++c) {
        const ValCol& v = lp.src.vals[c];
        aa.cols[c] = LeanCol{v.sum_off, v.min_off, v.max_off, v.nulls_off,
                             v.vclass == VC_SIGNED ? LEAN_SIGNED : v.vclass == VC_UNSIGNED ? LEAN_UNSIGNED : LEAN_FLOAT,
                             src.vals[c].mask && v.flag_bit >= 0 ? v.flag_bit : -1};
      }
      ProfScope prof("lds_aggregate", stream);
      st = lean_aggregate(lean.k0, lean.k1, lean.w, aa, d_lp, (unsigned)item_grid, geom.bytes, stream);
      if (st != MGB_SUCCESS) return st;
    } else {
      const void* agg_fn = plan.has_pass2 ? (const void*)lds_agg_kernel<true> : (const void*)lds_agg_kernel<false>;
      MGB_HIP(set_dynamic_lds(agg_fn, (int)geom.bytes));
      ProfScope prof("lds_aggregate", stream);
      if (plan.has_pass2) lds_agg_kernel<true><<<(unsigned)item_grid, kLdsBlock, geom.bytes, stream>>>(lp);
      else lds_agg_kernel<false><<<(unsigned)item_grid, kLdsBlock, geom.bytes, stream>>>(lp);
    }
    MGB_HIP(hipGetLastError());
    if (plan.has_pass2) {
      // the items whose groups went to the global table: their second pass
      // once the table is complete
      ProfScope prof("pass2", stream);
      pass2_items_kernel<<<(unsigned)item_grid, kLdsBlock, 0, stream>>>(lp);
      MGB_HIP(hipGetLastError());
    }

    // finalise the global table's groups after the items' rows
    int64_t* gblk = reinterpret_cast<int64_t*>(wsb + gblk_off);
    count_occupied_kernel<<<(unsigned)gblocks, kBlock, 0, stream>>>(gtab, gcap, plan.stride, plan.key_mode, gblk,
                                                                     ctr + 4);
    scan_blocks_kernel<<<1, 1024, 0, stream>>>(gblk, gblocks, ctr + 6, ctr + 4);
    FinParams f{};
    f.table = gtab;
    f.table_groups = ctr + 4;
    f.capacity = gcap;
    f.stride = plan.stride;
    f.key_mode = plan.key_mode;
    f.num_keys = (int)plan.keys.size();
    f.block_offsets = gblk;
    f.pos_base = ctr + 1;
    f.out = d_ot;
    for (size_t c = 0; c < plan.keys.size(); ++c) f.keys[c] = lp.src.keys[c];
    {
      ProfScope prof("finalize", stream);
      finalize_kernel<<<(unsigned)gblocks, kBlock, 0, stream>>>(f);
    }
    MGB_HIP(hipGetLastError());
    MGB_HIP(hipMemcpyAsync(ws->pinned, ctr, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    MGB_HIP(hipMemcpyAsync(ws->pinned + kPinnedNulls, temp.null_counts, (size_t)n_nulls * 8, hipMemcpyDeviceToHost,
                           stream));
    MGB_HIP(hipStreamSynchronize(stream));
    if (ws->pinned[5] == 0) {
      groups = (int64_t)(ws->pinned[1] + ws->pinned[6]);
      temp.rows = groups;
      temp.nulls_known = true;
      temp.host_nulls.assign(ws->pinned + kPinnedNulls, ws->pinned + kPinnedNulls + n_nulls);
      done = true;
      return MGB_SUCCESS;
    }
    est *= 4;  // the spill table overflowed: repartition finer
    ws->upload_slot = slot0;
  }
  return MGB_SUCCESS;
}


mgb_status run_groupby(const Plan& plan, int64_t R, Workspace* ws, mgb_out_column* out_keys,
                       mgb_out_column* out_results, int64_t* out_num_groups, hipStream_t stream,
                       const OutputSink& sink)
{
  ws->upload_slot = 0;

  AggParams base{};
  base.num_rows = R;
  base.key_mode = plan.key_mode;
  base.num_keys = (int)plan.keys.size();
  base.skip_null_keys = plan.skip_null_keys;
  base.compare_key_nulls = plan.compare_key_nulls;
  base.stride = plan.stride;
  for (size_t c = 0; c < plan.keys.size(); ++c) base.keys[c] = plan.keys[c];

  mgb_status st = MGB_SUCCESS;
  ws->fine_valid = false;
  SampleKeys sk;
  // a call the compact execution may take collects its hot keys
  HotSpec hot_spec;
  const HotSpec* hot = nullptr;
  {
    AggParams src = base;
    LeanPlan shape;
    if (single_launch_ops(plan, src) && lean_shape_of(plan, src, shape)) {
      hot_spec.k0 = shape.k0;
      hot_spec.k1 = shape.k1;
      hot_spec.c0 = plan.keys[0];
      if (shape.k1) hot_spec.c1 = plan.keys[1];
      hot = &hot_spec;
    }
  }
  int64_t G_est = sample_estimate(ws, base, R, stream, st, &sk, hot);
  if (st != MGB_SUCCESS) return st;
  if (const char* e = std::getenv("MGB_ESTIMATE_SCALE")) {
    // test hook: a scaled estimate (below 1 forces the overflow-and-retry
    // paths of the table and partitioned executions)
    const double f = std::atof(e);
    if (f > 0.0) G_est = std::max<int64_t>(1, (int64_t)((double)G_est * f));
  }

  const LdsGeom geom = lds_geometry(plan.stride);
  AggParams probe = base;
  Pass2Ops probe2{};
  const bool lds_ok = geom.ok && single_launch_ops(plan, probe) && !plan.compare_key_nulls &&
                      (int)plan.outs.size() <= kMaxOuts && pass2_ops(plan, probe, probe2);
  Path path = forced_path();
  if (!lds_ok) path = Path::Global;
  // the sort-path aggregations map rows to groups through a global table,
  // which the partitioned execution does not keep; a call of sort-path
  // aggregations only has records without fields, which only the global
  // execution builds
  bool no_fields = plan.rows_off < 0;
  for (const PlanCol& pc : plan.cols) no_fields = no_fields && pc.ops.empty() && pc.p2.empty();
  if (plan.has_sort && (path != Path::Direct || no_fields)) path = Path::Global;
  if (path == Path::Auto) path = (G_est <= geom.cap / 2) ? Path::Direct : Path::Partitioned;

  TempOuts temp;
  int64_t groups = 0;
  bool done = false;
  if (path == Path::Partitioned) {
    st = run_partitioned(ws, plan, base, R, G_est, geom, stream, temp, groups, done);
    if (st != MGB_SUCCESS) return st;
    if (!done) path = Path::Global;
  }
  if (!done) {
    TinyParams tp;
    const bool tiny = path == Path::Direct && !std::getenv("MGB_ESTIMATE_SCALE") && tiny_plan(plan, probe, sk, tp);
    st = run_table(ws, plan, base, R, G_est, path == Path::Direct, geom, stream, temp, groups, tiny ? &tp : nullptr);
    if (st != MGB_SUCCESS) return st;
  }
  *out_num_groups = groups;
  return emit_outputs(ws, plan, temp, groups, out_keys, out_results, stream, sink);
}

// The host staging of `stream`, taken at its first call (pinned memory).
mgb_status lock_staging(hipStream_t stream, HostStaging*& hs, std::unique_lock<std::mutex>& lk)
{
  hs = staging_for(stream);
  lk = std::unique_lock<std::mutex>(hs->mu);
  if (!hs->pinned) {
    MGB_HIP(hipHostMalloc((void**)&hs->pinned, kMaxPinned * sizeof(unsigned long long), hipHostMallocDefault));
    MGB_HIP(hipHostMalloc((void**)&hs->staging, kStagingBytes, hipHostMallocDefault));
  }
  return MGB_SUCCESS;
}

mgb_status make_hash_params(const mgb_column* cols, int32_t n, HashParams& hp, int64_t& rows)
{
  if (n < 1 || n > kMaxKeys) return fail(MGB_ERROR_UNSUPPORTED, "hashing supports 1 to 8 columns");
  rows = cols[0].size;
  hp = HashParams{};
  hp.num_cols = n;
  for (int c = 0; c < n; ++c) {
    int w = type_width(cols[c].type_id);
    if (w == 0) return fail(MGB_ERROR_UNSUPPORTED, "hashing supports fixed-width numeric columns");
    if (cols[c].size != rows) return fail(MGB_ERROR_LOGIC, "columns of a table must have equal sizes");
    KeyCol& k = hp.cols[c];
    k.data = static_cast<const uint8_t*>(cols[c].data) + cols[c].offset * w;
    k.mask = (cols[c].null_mask && cols[c].null_count != 0) ? cols[c].null_mask : nullptr;
    k.bit_off = cols[c].offset;
    k.type_id = cols[c].type_id;
    k.width = w;
    k.stride = w;
    k.vclass = value_class(cols[c].type_id);
    hp.check_nulls |= k.mask != nullptr;
  }
  hp.num_rows = rows;
  return MGB_SUCCESS;
}

}  // namespace mgb

// =================================================================== C ABI

using namespace mgb;

extern "C" int32_t mgb_version(void) { return 100; }

extern "C" const char* mgb_last_error_message(void) { return g_last_error.c_str(); }

extern "C" int32_t mgb_target_type(int32_t value_type_id, int32_t agg_kind) { return target_type(value_type_id, agg_kind); }

namespace mgb {

// Argument checks shared by the groupby entry points (groupby.cu:173-216).
mgb_status check_call(const mgb_column* keys, int32_t num_keys, int32_t null_policy, const mgb_request* requests,
                      int32_t num_requests, int64_t& R)
{
  if (num_keys < 0 || num_requests < 0 || (num_keys > 0 && !keys) || (num_requests > 0 && !requests))
    return fail(MGB_ERROR_LOGIC, "null argument");
  R = num_keys > 0 ? keys[0].size : 0;
  for (int c = 0; c < num_keys; ++c)
    if (keys[c].size != R) return fail(MGB_ERROR_LOGIC, "key columns must have equal sizes");
  for (int r = 0; r < num_requests; ++r)
    if (requests[r].values.size != R)
      return fail(MGB_ERROR_LOGIC, "Size mismatch between request values and groupby keys.");
  // cudf::size_type is int32: rows, groups, per-group counts and ARGMIN /
  // ARGMAX indices are 32-bit throughout (cpp/include/cudf/types.hpp:80)
  if (R < 0 || R > kMaxRows) return fail(MGB_ERROR_UNSUPPORTED, "more rows than cudf::size_type holds (2^31 - 1)");
  // the kernels read row 0 of a column in place of an absent row
  if (R > 0) {
    for (int c = 0; c < num_keys; ++c)
      if (!keys[c].data) return fail(MGB_ERROR_LOGIC, "a key column with rows has no data");
    for (int r = 0; r < num_requests; ++r)
      if (!requests[r].values.data) return fail(MGB_ERROR_LOGIC, "a value column with rows has no data");
  }
  for (int r = 0; r < num_requests; ++r)
    for (int a = 0; a < requests[r].num_aggs; ++a)
      if (target_type(requests[r].values.type_id, requests[r].aggs[a].kind) < 0)
        return fail(MGB_ERROR_LOGIC, "Invalid type/aggregation combination.");
  if (null_policy != MGB_NULL_EXCLUDE && null_policy != MGB_NULL_INCLUDE)
    return fail(MGB_ERROR_LOGIC, "invalid null_policy");
  return MGB_SUCCESS;
}

// Empty input: empty typed columns (groupby.cu:147-170,216). Buffers the
// caller provided stay the caller's (`keep_buffers`).
void empty_result(const mgb_column* keys, int32_t num_keys, const mgb_request* requests, int32_t num_requests,
                  mgb_out_column* out_keys, mgb_out_column* out_results, bool keep_buffers)
{
  auto reset = [&](mgb_out_column& o, int type_id) {
    if (keep_buffers) {
      o.size = 0;
      o.null_count = 0;
      o.null_mask = nullptr;
    } else {
      std::memset(&o, 0, sizeof(mgb_out_column));
    }
    o.type_id = type_id;
  };
  for (int c = 0; c < num_keys; ++c) reset(out_keys[c], keys[c].type_id);
  int j = 0;
  for (int r = 0; r < num_requests; ++r)
    for (int a = 0; a < requests[r].num_aggs; ++a, ++j) {
      const int tt = target_type(requests[r].values.type_id, requests[r].aggs[a].kind);
      reset(out_results[j], tt);
      if (tt == MGB_TYPE_LIST) out_results[j].child_type_id = requests[r].values.type_id;
    }
}

}  // namespace mgb

extern "C" mgb_status mgb_groupby_aggregate(const mgb_column* keys, int32_t num_keys, int32_t null_policy,
                                            const mgb_request* requests, int32_t num_requests,
                                            mgb_out_column* out_keys, mgb_out_column* out_results,
                                            int64_t* out_num_groups, void* stream, const mgb_allocator* alloc)
{
  g_last_error.clear();
  if (!out_num_groups || !alloc || !alloc->allocate || !alloc->deallocate) return fail(MGB_ERROR_LOGIC, "null argument");
  int64_t R = 0;
  mgb_status st = check_call(keys, num_keys, null_policy, requests, num_requests, R);
  if (st != MGB_SUCCESS) return st;
  if (R == 0) {
    *out_num_groups = 0;
    empty_result(keys, num_keys, requests, num_requests, out_keys, out_results, false);
    return MGB_SUCCESS;
  }
  Plan plan;
  st = build_plan(keys, num_keys, null_policy, requests, num_requests, plan);
  if (st != MGB_SUCCESS) return st;
  const mgb_allocator* temps = alloc->temporaries ? alloc->temporaries : alloc;
  if (!temps->allocate || !temps->deallocate) return fail(MGB_ERROR_LOGIC, "allocator of temporaries without callbacks");
  hipStream_t s = static_cast<hipStream_t>(stream);
  HostStaging* hs = nullptr;
  std::unique_lock<std::mutex> lk;
  st = lock_staging(s, hs, lk);
  if (st != MGB_SUCCESS) return st;
  Workspace ws;  // its temporaries go back to `temps` when the call ends, also on failure
  ws.mem.alloc = temps;
  ws.mem.stream = s;
  ws.pinned = hs->pinned;
  ws.staging = hs->staging;
  OutputSink sink;
  sink.alloc = alloc;
  return run_groupby(plan, R, &ws, out_keys, out_results, out_num_groups, s, sink);
}

// ---- the entry point that allocates nothing

namespace mgb {

// Device memory of mgb_groupby_aggregate_noalloc: every temporary of such a
// call is carved from this array (one copy on each device the library runs
// on, held until the process ends). In host code `scratch` only names the
// symbol; its device address comes from hipGetSymbolAddress, per device.
constexpr size_t kScratchBytes = 256ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_scratch_mu;                  // one allocation-free call at a time in the process
std::map<int, uint8_t*> g_scratch_addr;   // device -> address of its `scratch`
unsigned long long g_noalloc_readback[kMaxPinned];  // host readback of those calls (pageable)
uint8_t g_noalloc_staging[kStagingBytes];           // host source of their parameter uploads

// The address of the current device's `scratch`, looked up at the device's
// first call (with that device current) and kept.
mgb_status scratch_address(uint8_t*& out)
{
  int dev = 0;
  MGB_HIP(hipGetDevice(&dev));
  auto it = g_scratch_addr.find(dev);
  if (it != g_scratch_addr.end()) {
    out = it->second;
    return MGB_SUCCESS;
  }
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(MGB_ERROR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_addr[dev] = out = static_cast<uint8_t*>(p);
  return MGB_SUCCESS;
}

}  // namespace mgb

extern "C" mgb_status mgb_groupby_aggregate_noalloc(const mgb_column* keys, int32_t num_keys, int32_t null_policy,
                                                    const mgb_request* requests, int32_t num_requests,
                                                    mgb_out_column* out_keys, mgb_out_column* out_results,
                                                    int64_t max_groups, int64_t* out_num_groups, void* stream)
{
  g_last_error.clear();
  if (!out_num_groups || max_groups < 0 || (num_keys > 0 && !out_keys)) return fail(MGB_ERROR_LOGIC, "null argument");
  int64_t R = 0;
  mgb_status st = check_call(keys, num_keys, null_policy, requests, num_requests, R);
  if (st != MGB_SUCCESS) return st;
  if (R > (int64_t)kNoAllocMaxRows)
    return fail(MGB_ERROR_UNSUPPORTED, "mgb_groupby_aggregate_noalloc takes at most " +
                                           std::to_string(kNoAllocMaxRows) + " rows (its scratch memory is " +
                                           std::to_string(kScratchBytes >> 20) + " MiB)");
  if (R == 0) {
    *out_num_groups = 0;
    empty_result(keys, num_keys, requests, num_requests, out_keys, out_results, true);
    return MGB_SUCCESS;
  }
  Plan plan;
  st = build_plan(keys, num_keys, null_policy, requests, num_requests, plan);
  if (st != MGB_SUCCESS) return st;
  for (const Plan::Out& o : plan.outs)
    if (o.kind == MGB_AGG_COLLECT_LIST || o.kind == MGB_AGG_COLLECT_SET)
      return fail(MGB_ERROR_UNSUPPORTED, "mgb_groupby_aggregate_noalloc: COLLECT_LIST / COLLECT_SET results need "
                                         "allocated outputs");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  uint8_t* region = nullptr;
  st = scratch_address(region);
  if (st != MGB_SUCCESS) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Workspace ws;
  ws.mem.region = region;
  ws.mem.region_bytes = kScratchBytes;
  ws.mem.stream = s;
  ws.pinned = g_noalloc_readback;
  ws.staging = g_noalloc_staging;
  OutputSink sink;
  sink.capacity = max_groups;
  st = run_groupby(plan, R, &ws, out_keys, out_results, out_num_groups, s, sink);
  if (st == MGB_ERROR_ALLOC && ws.mem.region && g_last_error.find("temporaries") != std::string::npos)
    return fail(MGB_ERROR_ALLOC, g_last_error + " (beyond the " + std::to_string(kScratchBytes >> 20) +
                                     " MiB scratch memory of mgb_groupby_aggregate_noalloc)");
  // the scratch is reused by the next call: its work must be complete
  // before another call may start (the copies into the caller's columns)
  if (st == MGB_SUCCESS) MGB_HIP(hipStreamSynchronize(s));
  return st;
}

extern "C" mgb_status mgb_estimate_groups(const mgb_column* keys, int32_t num_keys, int32_t null_policy, void* stream_v,
                                          const mgb_allocator* alloc, int64_t* estimate)
{
  g_last_error.clear();
  if (!estimate) return fail(MGB_ERROR_LOGIC, "estimate is NULL");
  if (!alloc) return fail(MGB_ERROR_LOGIC, "allocator is NULL");
  const mgb_allocator* temps = alloc->temporaries ? alloc->temporaries : alloc;
  if (!temps->allocate || !temps->deallocate) return fail(MGB_ERROR_LOGIC, "allocator without callbacks");
  if (num_keys <= 0 || !keys) return fail(MGB_ERROR_LOGIC, "no key columns");
  const int64_t R = keys[0].size;
  for (int c = 1; c < num_keys; ++c)
    if (keys[c].size != R) return fail(MGB_ERROR_LOGIC, "key columns differ in size");
  if (R < 0 || R > kMaxRows) return fail(MGB_ERROR_UNSUPPORTED, "more rows than cudf::size_type holds (2^31 - 1)");
  Plan plan;
  mgb_status st = build_plan(keys, num_keys, null_policy, nullptr, 0, plan);
  if (st != MGB_SUCCESS) return st;
  if (R == 0) {
    *estimate = 0;
    return MGB_SUCCESS;
  }
  hipStream_t stream = static_cast<hipStream_t>(stream_v);
  HostStaging* hs = nullptr;
  std::unique_lock<std::mutex> lk;
  st = lock_staging(stream, hs, lk);
  if (st != MGB_SUCCESS) return st;
  Workspace call;
  Workspace* ws = &call;
  call.mem.alloc = temps;
  call.mem.stream = stream;
  call.pinned = hs->pinned;
  call.staging = hs->staging;
  AggParams base{};
  base.num_rows = R;
  base.key_mode = plan.key_mode;
  base.num_keys = (int)plan.keys.size();
  base.skip_null_keys = plan.skip_null_keys;
  base.compare_key_nulls = plan.compare_key_nulls;
  base.stride = plan.stride;
  for (size_t c = 0; c < plan.keys.size(); ++c) base.keys[c] = plan.keys[c];
  *estimate = sample_estimate(ws, base, R, stream, st);
  return st;
}

extern "C" mgb_status mgb_murmurhash3_x86_32(const mgb_column* columns, int32_t num_columns, uint32_t seed,
                                             uint32_t* out, void* stream)
{
  g_last_error.clear();
  HashParams hp;
  int64_t rows = 0;
  mgb_status st = make_hash_params(columns, num_columns, hp, rows);
  if (st != MGB_SUCCESS) return st;
  hp.seed = seed;
  if (rows == 0) return MGB_SUCCESS;
  murmur_table_kernel<<<(unsigned)((rows + kBlock - 1) / kBlock), kBlock, 0, static_cast<hipStream_t>(stream)>>>(hp, out);
  MGB_HIP(hipGetLastError());
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_spark_murmurhash3_x86_32(const mgb_column* columns, int32_t num_columns, uint32_t seed,
                                                   int32_t* out, void* stream)
{
  g_last_error.clear();
  HashParams hp;
  int64_t rows = 0;
  mgb_status st = make_hash_params(columns, num_columns, hp, rows);
  if (st != MGB_SUCCESS) return st;
  hp.seed = seed;
  if (rows == 0) return MGB_SUCCESS;
  spark_murmur_table_kernel<<<(unsigned)((rows + kBlock - 1) / kBlock), kBlock, 0,
                              static_cast<hipStream_t>(stream)>>>(hp, out);
  MGB_HIP(hipGetLastError());
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_profiling_enable(int32_t enable)
{
  g_last_error.clear();
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_enabled = enable != 0;
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_profiling_read(const char* kernel, int64_t* launches, double* total_ms)
{
  g_last_error.clear();
  if (!kernel || !launches || !total_ms) return fail(MGB_ERROR_LOGIC, "null argument");
  std::vector<ProfRecord> mine;
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    std::vector<ProfRecord> rest;
    for (auto& r : g_prof) (r.kernel == kernel ? mine : rest).push_back(r);
    g_prof.swap(rest);
  }
  *launches = 0;
  *total_ms = 0.0;
  mgb_status st = MGB_SUCCESS;
  for (auto& r : mine) {
    float ms = 0.f;
    hipError_t e = hipEventSynchronize(r.stop);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, r.start, r.stop);
    if (e != hipSuccess && st == MGB_SUCCESS) st = fail(MGB_ERROR_HIP, std::string("profiling: ") + hipGetErrorString(e));
    *launches += 1;
    *total_ms += ms;
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  return st;
}

extern "C" mgb_status mgb_release_workspace(void* stream)
{
  g_last_error.clear();
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::unique_ptr<HostStaging> w;
  {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_ws.find({dev, static_cast<hipStream_t>(stream)});
    if (it == g_ws.end()) return MGB_SUCCESS;
    w = std::move(it->second);
    g_ws.erase(it);
  }
  // a call still running on the stream holds the staging's lock: wait for it
  // before the staging goes, and for the stream's copies from it
  std::lock_guard<std::mutex> lk(w->mu);
  MGB_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return MGB_SUCCESS;
}
