// The compact partitioned execution: the partitioned execution's two passes
// for calls with an integer key of one or two 4- or 8-byte columns and at
// most kLeanCols value columns of 4 or 8 bytes whose ops are SUM / MIN / MAX
// and counts. Each pass is compiled per key and value widths, so that record
// offsets are constants, and takes a small argument block (LeanScatterArgs,
// LeanAggArgs): the row loops hold few uniform values, all in scalar
// registers, and the rare paths (a spilled row, the groups' output) read the
// full parameters from device memory.
//
// Scatter (lean_scatter_rows): each tile of rows is ranked by partition with
// one LDS atomic per row; every row's record (key, row hash with the
// validity bits, values) is built in registers and written to the tile's LDS
// stage at its position in partition order with 16-byte stores, beside the
// record's destination row; the stage is then copied out 16 bytes per lane,
// so each partition's run leaves as one contiguous stretch. The next tile's
// loads are issued before the copy-out.
//
// Aggregation (lean_aggregate_rows): one work item per workgroup; each
// thread takes a few records at a time (kLeanRows), issuing the loads of its
// next ones before it updates the table with the current ones. The LDS slot
// comes from the hash stored in the record (its bits below the partition's).
// A single-column key is the LDS record's header. A two-column key puts its
// first 8 bytes (part A) in the header and the rest with the group's first
// record index in a tail: a lane claims an empty slot by writing part A,
// then writes the tail; a lane that meets part A before the tail is written
// retries on its next pass of the miss loop (never waits inside it, so a
// claim by a lane of its own wave always completes first). Before the groups
// leave the table, headers become the row-reference headers the general
// routines read (the first record's index).
#pragma once

#include "hot_keys.hpp"
#include "lds_agg.hpp"
#include "partition.hpp"

namespace mgb {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // 16 bytes of a record

// records per thread per step of the aggregation loop (two sets of them are
// held in registers: the current step's and the next step's)
template <int RB16>
constexpr int kLeanRows = RB16 == 1 ? 4 : 3;
constexpr int kLeanProbe = 4;  // slots of a row's probe read before the full probe

// Exclusive scan, in place, of the P counts `cnt` (P a power of two, at most
// 8 * kLdsBlock) into tile starts; cnt[P] = the total, also returned. `tmp`:
// 16 words of LDS. Ends with a workgroup barrier.
__device__ __forceinline__ uint32_t lean_scan_counts(uint32_t* cnt, int P, uint32_t* tmp)
{
  const int per = P > kLdsBlock ? P / kLdsBlock : 1;
  const int i0 = (int)threadIdx.x * per;
  uint32_t sum = 0;
  for (int e = 0; e < per; ++e)
    if (i0 + e < P) sum += cnt[i0 + e];
  uint32_t total;
  uint32_t run = block_exclusive_scan_1024(sum, tmp, &total);
  for (int e = 0; e < per; ++e)
    if (i0 + e < P) {
      const uint32_t c = cnt[i0 + e];
      cnt[i0 + e] = run;
      run += c;
    }
  if (threadIdx.x == 0) cnt[P] = total;
  __syncthreads();
  return total;
}

// Element `r` of a column of W bytes (W: 4 or 8), zero-extended.
template <int W>
__device__ __forceinline__ uint64_t lean_load(const uint8_t* col, int64_t r)
{
  if constexpr (W == 8) return to_global(reinterpret_cast<const uint64_t*>(col))[r];
  else return to_global(reinterpret_cast<const uint32_t*>(col))[r];
}

// Puts the low W bytes of v at byte OFF of the record words w.
template <int W, int OFF>
__device__ __forceinline__ void lean_put(uint32_t (&w)[8], uint64_t v)
{
  if constexpr (W != 0) {
    w[OFF / 4] = (uint32_t)v;
    if constexpr (W == 8) w[OFF / 4 + 1] = (uint32_t)(v >> 32);
  }
}

template <int W, int N>
__device__ __forceinline__ void lean_update(uint8_t* lds, const LeanCol& c, const uint64_t (&x)[N],
                                            const bool (&valid)[N], const uint32_t (&loff)[N], uint32_t dum);

// Scatter of chunk blockIdx.x (rows [chunk * a.chunk, +a.chunk)) into
// records. LDS (lean_scatter_lds): cursor and tile count (then tile start)
// of every partition, 16 scan words, the destination row of every staged
// record, the stage; with hot keys (a.hot), their lookup and records. K1 =
// 0: a single key column of K0 bytes.
//
// A hot key's rows are not staged: they update the chunk's record of the
// key (LDS, in the layout of the plan's records), and when the chunk is
// done the records that rows reached fold into the global table. Their
// headers are the key (one column), or the index of the key's
// representative record (two columns: records past the partitioned ones,
// written beforehand), as the global table's row references expect.
template <int K0, int K1, int W0, int W1, int W2, bool HOT, bool DIRECT>
__device__ __forceinline__ void lean_scatter_rows(const LeanScatterArgs& a, uint8_t* smem)
{
  constexpr LeanKeyOffsets KO = lean_key_offsets(K0, K1);
  constexpr int KB = KO.bytes;
  constexpr LeanOffsets LO = lean_offsets(KB, W0, W1, W2);
  constexpr int RB16 = LO.bytes <= 16 ? 1 : 2;
  constexpr int I = kScatterItems;
  const int P = 1 << a.log2p;
  const int T = a.tile_rows;
  const int items = T / kLdsBlock;
  uint32_t* cursor = reinterpret_cast<uint32_t*>(smem);
  uint32_t* tcount = cursor + P;  // counts of the tile's rows, then their starts (P + 1 words)
  uint32_t* tmp = tcount + P + 1;
  constexpr bool use_hot = HOT;  // a.hot set (the variant compiled for calls with hot keys)
  const LeanScatterLds lay = lean_scatter_lds(P, DIRECT ? 0 : T, RB16 * 16, use_hot ? a.hot_words * 8 : 0);
  uint32_t* dstrow = reinterpret_cast<uint32_t*>(smem + lay.dst);
  u32x4* stage = reinterpret_cast<u32x4*>(smem + lay.stage);
  const int64_t chunk = blockIdx.x;
  for (int i = threadIdx.x; i < P; i += kLdsBlock) {
    cursor[i] = a.offsets[(int64_t)i * a.nblocks + chunk];
    tcount[i] = 0;
  }
  __syncthreads();
  const int64_t begin = chunk * a.chunk;
  const int64_t end = begin + a.chunk < a.rows ? begin + a.chunk : a.rows;
  if (begin >= end) return;  // the whole workgroup
  global_t<u32x4>* out = to_global(reinterpret_cast<u32x4*>(a.out));

  HotLds hl{nullptr, nullptr, 0};
  LdsTable ht{};
  uint32_t* touched = nullptr;  // hot records that rows reached (kHotMax bits)
  uint32_t hdum = 0;            // the thread's dummy word for hot updates
  if constexpr (use_hot) {
    hl = hot_load(a.hot, reinterpret_cast<HotKey*>(smem + lay.hot_keys), reinterpret_cast<uint32_t*>(smem + lay.hot_idx),
                  K0, K1);
    int32_t* ctr = reinterpret_cast<int32_t*>(smem + lay.hot_ctr);  // count, overflow, 2 fold words
    touched = reinterpret_cast<uint32_t*>(ctr + 4);
    hdum = (uint32_t)(lay.hot_ctr + 16 + kHotMax / 8) + (threadIdx.x & 63u) * 8u;  // past the touched bits
    ht.rec = smem + lay.hot_rec;
    ht.slots = hl.n;
    ht.stride = a.hot_words * 8;
    ht.cap = kHotMax;
    ht.count = ctr;
    ht.overflow = ctr + 1;
    const int words = a.hot_words;
    uint64_t* w64 = reinterpret_cast<uint64_t*>(ht.rec);
    for (int i = threadIdx.x; i < (hl.n + 1) * words; i += kLdsBlock) {
      const int r = i / words, w = i - r * words;
      uint64_t v = a.tmpl[w];
      if (w == 0) {
        if (r == hl.n) v = 0ull;  // the table's special record: never occupied
        else if constexpr (K1 != 0) v = (uint64_t)(a.rep_base + r);
        else v = (K0 == 4 && a.key_signed) ? (uint64_t)(int64_t)(int32_t)(uint32_t)hl.keys[r].c0 : hl.keys[r].c0;
      }
      w64[i] = v;
    }
    if (threadIdx.x < 4) ctr[threadIdx.x] = 0;
    if (threadIdx.x < kHotMax / 32) touched[threadIdx.x] = 0u;
    __syncthreads();
  }

  // the thread's rows of one tile, as loaded (rows past the end load the
  // last row and are not staged)
  uint64_t key[I], key1[I], v0[I], v1[I], v2[I];
  // km1: the second key column's mask word, kept apart from km until the
  // row is tested (combining them here would wait for km's load inside the
  // prefetch of the next tile)
  uint32_t km[I], km1[I], m0[I], m1[I], m2[I];
  auto fetch = [&](int64_t tile) {
#pragma unroll
    for (int k = 0; k < I; ++k) {
      if (k >= items) continue;
      int64_t r = tile + threadIdx.x + (int64_t)k * kLdsBlock;
      r = r < end ? r : end - 1;
      key[k] = lean_load<K0>(a.key, r);
      km[k] = a.key_mask ? to_global(a.key_mask)[r >> 5] : ~0u;
      if constexpr (K1 != 0) {
        key1[k] = lean_load<K1>(a.key1, r);
        km1[k] = a.key1_mask ? to_global(a.key1_mask)[r >> 5] : ~0u;
      }
      if constexpr (W0 != 0) {
        v0[k] = lean_load<W0>(a.val[0], r);
        m0[k] = a.mask[0] ? to_global(a.mask[0])[r >> 5] : ~0u;
      }
      if constexpr (W1 != 0) {
        v1[k] = lean_load<W1>(a.val[1], r);
        m1[k] = a.mask[1] ? to_global(a.mask[1])[r >> 5] : ~0u;
      }
      if constexpr (W2 != 0) {
        v2[k] = lean_load<W2>(a.val[2], r);
        m2[k] = a.mask[2] ? to_global(a.mask[2])[r >> 5] : ~0u;
      }
    }
  };
  const int fb0 = a.flag_bits & 15, fb1 = (a.flag_bits >> 4) & 15, fb2 = (a.flag_bits >> 8) & 15;

  fetch(begin);
  for (int64_t tile = begin; tile < end; tile += T) {
    bool present[I];
    uint32_t part[I], rank[I], hf[I], bitk[I], hk[I];
    int hidx[I];
#pragma unroll
    for (int k = 0; k < I; ++k) {
      const int64_t r = tile + threadIdx.x + (int64_t)k * kLdsBlock;
      const uint32_t bit = (uint32_t)r & 31u;
      bitk[k] = bit;
      uint32_t kmk = km[k];
      if constexpr (K1 != 0) kmk &= km1[k];
      present[k] = k < items && r < end && ((kmk >> bit) & 1u);
      // cudf's row hash: hash_combine over the key columns from seed 0
      uint32_t h = hash_combine(0u, murmur3_bits(key[k], K0, 0));
      if constexpr (K1 != 0) h = hash_combine(h, murmur3_bits(key1[k], K1, 0));
      hk[k] = h;
      uint32_t f = 0;
      if constexpr (W0 != 0) f |= fb0 < 8 ? ((m0[k] >> bit) & 1u) << fb0 : 0u;
      if constexpr (W1 != 0) f |= fb1 < 8 ? ((m1[k] >> bit) & 1u) << fb1 : 0u;
      if constexpr (W2 != 0) f |= fb2 < 8 ? ((m2[k] >> bit) & 1u) << fb2 : 0u;
      hf[k] = (h & 0xffffff00u) | f;
      part[k] = partition_of(h, a.log2p);
    }
#pragma unroll
    for (int k = 0; k < I; ++k) hidx[k] = -1;
    if constexpr (use_hot) {
      uint64_t c1[I];
#pragma unroll
      for (int k = 0; k < I; ++k) c1[k] = K1 != 0 ? key1[k] : 0ull;
      hot_find_n<I>(hl, hk, key, c1, present, hidx);
#pragma unroll
      for (int k = 0; k < I; ++k) present[k] = present[k] && hidx[k] < 0;
    }
    if constexpr (DIRECT) {
      // many more partitions than rows a tile holds: each row takes its
      // destination from its partition's cursor and is written at once
#pragma unroll
      for (int k = 0; k < I; ++k) rank[k] = present[k] ? atomicAdd(&cursor[part[k]], 1u) : 0u;
    } else {
#pragma unroll
      for (int k = 0; k < I; ++k) rank[k] = present[k] ? atomicAdd(&tcount[part[k]], 1u) : 0u;
    }
    if constexpr (use_hot) {
      // hot rows into their records
      uint32_t hoff[I];
#pragma unroll
      for (int k = 0; k < I; ++k) {
        hoff[k] = hidx[k] >= 0 ? (uint32_t)(lay.hot_rec + (size_t)hidx[k] * ht.stride) : kRecNone;
        if (hidx[k] >= 0) atomicOr(&touched[hidx[k] >> 5], 1u << (hidx[k] & 31));
      }
      if (a.rows_off >= 0) {
#pragma unroll
        for (int k = 0; k < I; ++k)
          lds_add(reinterpret_cast<uint32_t*>(smem + (hoff[k] < kRecSpill ? hoff[k] + a.rows_off : hdum)), 1u);
      }
      bool valid[I];
      if constexpr (W0 != 0) {
#pragma unroll
        for (int k = 0; k < I; ++k) valid[k] = (m0[k] >> bitk[k]) & 1u;
        lean_update<W0, I>(smem, a.cols[0], v0, valid, hoff, hdum);
      }
      if constexpr (W1 != 0) {
#pragma unroll
        for (int k = 0; k < I; ++k) valid[k] = (m1[k] >> bitk[k]) & 1u;
        lean_update<W1, I>(smem, a.cols[1], v1, valid, hoff, hdum);
      }
      if constexpr (W2 != 0) {
#pragma unroll
        for (int k = 0; k < I; ++k) valid[k] = (m2[k] >> bitk[k]) & 1u;
        lean_update<W2, I>(smem, a.cols[2], v2, valid, hoff, hdum);
      }
    }
    if constexpr (DIRECT) {
#pragma unroll
      for (int k = 0; k < I; ++k) {
        if (!present[k]) continue;
        uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
        lean_put<K0, KO.off0>(w, key[k]);
        if constexpr (K1 != 0) lean_put<K1, KO.off1>(w, key1[k]);
        w[KB / 4] = hf[k];
        if constexpr (W0 != 0) lean_put<W0, LO.off[0]>(w, v0[k]);
        if constexpr (W1 != 0) lean_put<W1, LO.off[1]>(w, v1[k]);
        if constexpr (W2 != 0) lean_put<W2, LO.off[2]>(w, v2[k]);
        out[(uint64_t)rank[k] * RB16] = u32x4{w[0], w[1], w[2], w[3]};
        if constexpr (RB16 == 2) out[(uint64_t)rank[k] * 2 + 1] = u32x4{w[4], w[5], w[6], w[7]};
      }
      if (tile + T < end) fetch(tile + T);
      continue;
    }
    __syncthreads();
    const uint32_t tile_rows = lean_scan_counts(tcount, P, tmp);
    const uint32_t* tstart = tcount;
#pragma unroll
    for (int k = 0; k < I; ++k) {
      if (!present[k]) continue;
      const uint32_t pos = tstart[part[k]] + rank[k];
      dstrow[pos] = cursor[part[k]] + rank[k];
      uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
      lean_put<K0, KO.off0>(w, key[k]);
      if constexpr (K1 != 0) lean_put<K1, KO.off1>(w, key1[k]);
      w[KB / 4] = hf[k];
      if constexpr (W0 != 0) lean_put<W0, LO.off[0]>(w, v0[k]);
      if constexpr (W1 != 0) lean_put<W1, LO.off[1]>(w, v1[k]);
      if constexpr (W2 != 0) lean_put<W2, LO.off[2]>(w, v2[k]);
      stage[pos * RB16] = u32x4{w[0], w[1], w[2], w[3]};
      if constexpr (RB16 == 2) stage[pos * 2 + 1] = u32x4{w[4], w[5], w[6], w[7]};
    }
    __syncthreads();
    // the next tile's loads fly while this one is copied out
    if (tile + T < end) fetch(tile + T);
    const uint32_t units = tile_rows * (uint32_t)RB16;
    for (uint32_t u = threadIdx.x; u < units; u += kLdsBlock) {
      const uint32_t r = RB16 == 2 ? u >> 1 : u;
      out[(uint64_t)dstrow[r] * RB16 + (u & (uint32_t)(RB16 - 1))] = stage[u];
    }
    for (int i = threadIdx.x; i < P; i += kLdsBlock) cursor[i] += tcount[i + 1] - tcount[i];
    __syncthreads();
    for (int i = threadIdx.x; i < P; i += kLdsBlock) tcount[i] = 0;
    __syncthreads();
  }
  if constexpr (use_hot) {
    __syncthreads();  // every hot update done (the direct mode has no barrier per tile)
    // the records that rows reached fold into the global table
    for (int i = threadIdx.x; i < hl.n; i += kLdsBlock)
      if (!((touched[i >> 5] >> (i & 31)) & 1u))
        *reinterpret_cast<uint64_t*>(ht.rec + (int64_t)i * ht.stride) = K1 != 0 ? kEmptyRowRef : kEmptyIntKey;
    if (threadIdx.x == 0) {
      int n = 0;
      for (int j = 0; j < kHotMax / 32; ++j) n += __popc(touched[j]);
      ht.count[0] = n;
    }
    lds_fold_to_global(a.lp->src, ht);
  }
}

// The record of an integer key in the LDS table (lds_find_or_insert for
// KEY_INT headers, without the general parameters); nullptr when full.
__device__ __forceinline__ uint8_t* lean_find_or_insert(const LdsTable& t, uint64_t key, uint32_t slot_hash)
{
  if (key == kEmptyIntKey) {
    uint8_t* rec = lds_record(t, t.slots);
    unsigned long long expect = 0ull;
    if (__hip_atomic_compare_exchange_strong(reinterpret_cast<unsigned long long*>(rec), &expect, 1ull,
                                             __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
      atomicAdd(t.count, 1);
    return rec;
  }
  int32_t idx = (int32_t)reduce_range(slot_hash, t.slots);
  for (int32_t probe = 0; probe < t.slots; ++probe) {
    uint8_t* rec = lds_record(t, idx);
    unsigned long long* hp = reinterpret_cast<unsigned long long*>(rec);
    uint64_t cur = __hip_atomic_load(hp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == kEmptyIntKey) {
      if (__hip_atomic_load(t.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= t.cap) {
        *t.overflow = 1;
        return nullptr;
      }
      unsigned long long expect = kEmptyIntKey;
      if (__hip_atomic_compare_exchange_strong(hp, &expect, (unsigned long long)key, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
        atomicAdd(t.count, 1);
        return rec;
      }
      cur = expect;
    }
    if (cur == key) return rec;
    idx = (idx + 1 == t.slots) ? 0 : idx + 1;
  }
  *t.overflow = 1;
  return nullptr;
}

constexpr uint32_t kLeanRetry = 0xfffffffdu;  // the probe met a slot being claimed: try again

// Byte offset of the record of a two-column key (part A `ka`, part B `kb`:
// none for KB = 8, its low 4 bytes for KB = 12) in the LDS table, inserting it
// with `rowref` (the row's record index) when absent; kRecSpill when the
// table is full or part A equals the empty header (such rows always take the
// global table); kLeanRetry when a slot holding part A has no tail yet.
template <int KB>
__device__ __forceinline__ uint32_t lean2_find_or_insert(const LdsTable& t, uint8_t* lds, uint64_t ka, uint64_t kb,
                                                         uint32_t rowref, uint32_t slot_hash, uint32_t tail_off)
{
  if (ka == kEmptyIntKey) return kRecSpill;
  int32_t idx = (int32_t)reduce_range(slot_hash, t.slots);
  for (int32_t probe = 0; probe < t.slots; ++probe) {
    uint8_t* rec = lds_record(t, idx);
    unsigned long long* hp = reinterpret_cast<unsigned long long*>(rec);
    unsigned long long* tail = reinterpret_cast<unsigned long long*>(rec + tail_off);
    uint64_t cur = __hip_atomic_load(hp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == kEmptyIntKey) {
      if (__hip_atomic_load(t.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= t.cap) {
        *t.overflow = 1;
        return kRecSpill;
      }
      unsigned long long expect = kEmptyIntKey;
      if (__hip_atomic_compare_exchange_strong(hp, &expect, (unsigned long long)ka, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
        if constexpr (KB <= 12) {
          // one word: record index in the high half, part B (if any) in the low
          __hip_atomic_store(tail, (unsigned long long)rowref << 32 | (uint32_t)kb, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
          __hip_atomic_store(tail, (unsigned long long)kb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_store(tail + 1, (unsigned long long)rowref << 32, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        atomicAdd(t.count, 1);
        return (uint32_t)(rec - lds);
      }
      cur = expect;
    }
    if (cur == ka) {
      if constexpr (KB == 8) {
        return (uint32_t)(rec - lds);
      } else if constexpr (KB == 12) {
        const uint64_t w = __hip_atomic_load(tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if ((uint32_t)(w >> 32) == kLeanNotReady) return kLeanRetry;
        if ((uint32_t)w == (uint32_t)kb) return (uint32_t)(rec - lds);
      } else {
        const uint64_t w1 = __hip_atomic_load(tail + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if ((uint32_t)(w1 >> 32) == kLeanNotReady) return kLeanRetry;
        if (__hip_atomic_load(tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == kb)
          return (uint32_t)(rec - lds);
      }
    }
    idx = (idx + 1 == t.slots) ? 0 : idx + 1;
  }
  *t.overflow = 1;
  return kRecSpill;
}

// Two-column keys: every record's header becomes the row-reference header of
// the general routines (the group's first record index; kEmptyRowRef where
// empty), for the groups' output or fold. Ends with a workgroup barrier.
template <int KB>
__device__ __forceinline__ void lean2_headers_to_rowref(const LdsTable& t, uint32_t tail_off)
{
  for (int i = threadIdx.x; i <= t.slots; i += kLdsBlock) {
    uint8_t* rec = lds_record(t, i);
    uint64_t* hp = reinterpret_cast<uint64_t*>(rec);
    const uint64_t* tail = reinterpret_cast<const uint64_t*>(rec + tail_off);
    const bool used = i < t.slots && *hp != kEmptyIntKey;
    *hp = used ? (tail[KB <= 12 ? 0 : 1] >> 32) : kEmptyRowRef;
  }
  __syncthreads();
}

// A row whose group did not fit the LDS table: the global table of the full
// parameters (device memory). Out of line: rare, and its parameter loads stay
// out of the row loop.
__device__ __noinline__ void lean_spill_row(const LdsParams* lp, int64_t row)
{
  const AggParams& p = lp->src;
  spill_row(p, row_key(p, row), row);
}

// 32-bit word D of a record held as RB16 16-byte units.
template <int D, int RB16>
__device__ __forceinline__ uint32_t lean_word(const u32x4 (&r)[RB16])
{
  return r[D / 4][D % 4];
}

// SUM / MIN / MAX and the null count of one value column of W bytes over N
// rows: values x (zero-extended), their validity, the byte offsets of their
// LDS records (kRecSpill / kRecNone: none). The column's fields and class
// are uniform: every branch on them is scalar. Rows without a record or a
// valid value update the thread's dummy word `dum` instead.
template <int W, int N>
__device__ __forceinline__ void lean_update(uint8_t* lds, const LeanCol& c, const uint64_t (&x)[N],
                                            const bool (&valid)[N], const uint32_t (&loff)[N], uint32_t dum)
{
  bool ok[N];
#pragma unroll
  for (int u = 0; u < N; ++u) ok[u] = loff[u] < kRecSpill && valid[u];
  if (c.nulls_off >= 0) {
#pragma unroll
    for (int u = 0; u < N; ++u)
      lds_add(reinterpret_cast<uint32_t*>(lds + (loff[u] < kRecSpill && !ok[u] ? loff[u] + c.nulls_off : dum)), 1u);
  }
  uint32_t a[N];
  if (c.cls == LEAN_FLOAT) {
    if (c.sum_off >= 0) {
#pragma unroll
      for (int u = 0; u < N; ++u) {
        a[u] = ok[u] ? loff[u] + c.sum_off : dum;
        const double d = W == 8 ? __longlong_as_double((long long)x[u]) : (double)__uint_as_float((uint32_t)x[u]);
        lds_add(reinterpret_cast<double*>(lds + a[u]), d);
      }
    }
    if constexpr (W == 8) {
      uint64_t lo_code[N], hi_code[N];
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint64_t b = x[u];
        const bool nan = (b & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (b & 0x000fffffffffffffull);
        const uint64_t code = f64_code(b);
        lo_code[u] = nan ? kF64MinNaN : code;
        hi_code[u] = nan ? kF64MaxNaN : code;
      }
      if (c.min_off >= 0) {
#pragma unroll
        for (int u = 0; u < N; ++u)
          lds_min(reinterpret_cast<unsigned long long*>(lds + (ok[u] ? loff[u] + c.min_off : dum)),
                  (unsigned long long)lo_code[u]);
      }
      if (c.max_off >= 0) {
#pragma unroll
        for (int u = 0; u < N; ++u)
          lds_max(reinterpret_cast<unsigned long long*>(lds + (ok[u] ? loff[u] + c.max_off : dum)),
                  (unsigned long long)hi_code[u]);
      }
    } else {
      uint32_t lo_code[N], hi_code[N];
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint32_t b = (uint32_t)x[u];
        const bool nan = (b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu);
        const uint32_t code = f32_code(b);
        lo_code[u] = nan ? kF32MinNaN : code;
        hi_code[u] = nan ? kF32MaxNaN : code;
      }
      if (c.min_off >= 0) {
#pragma unroll
        for (int u = 0; u < N; ++u)
          lds_min(reinterpret_cast<uint32_t*>(lds + (ok[u] ? loff[u] + c.min_off : dum)), lo_code[u]);
      }
      if (c.max_off >= 0) {
#pragma unroll
        for (int u = 0; u < N; ++u)
          lds_max(reinterpret_cast<uint32_t*>(lds + (ok[u] ? loff[u] + c.max_off : dum)), hi_code[u]);
      }
    }
    return;
  }
  const bool sgn = c.cls == LEAN_SIGNED;
  if (c.sum_off >= 0) {
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const uint64_t v = W == 8 ? x[u] : sgn ? (uint64_t)(int64_t)(int32_t)x[u] : x[u];
      lds_add(reinterpret_cast<unsigned long long*>(lds + (ok[u] ? loff[u] + c.sum_off : dum)),
              (unsigned long long)v);
    }
  }
  for (int which = 0; which < 2; ++which) {
    const int off = which == 0 ? c.min_off : c.max_off;
    if (off < 0) continue;
#pragma unroll
    for (int u = 0; u < N; ++u) a[u] = ok[u] ? loff[u] + off : dum;
    if constexpr (W == 8) {
      if (sgn) {
#pragma unroll
        for (int u = 0; u < N; ++u) {
          if (which == 0) lds_min(reinterpret_cast<long long*>(lds + a[u]), (long long)x[u]);
          else lds_max(reinterpret_cast<long long*>(lds + a[u]), (long long)x[u]);
        }
      } else {
#pragma unroll
        for (int u = 0; u < N; ++u) {
          if (which == 0) lds_min(reinterpret_cast<unsigned long long*>(lds + a[u]), (unsigned long long)x[u]);
          else lds_max(reinterpret_cast<unsigned long long*>(lds + a[u]), (unsigned long long)x[u]);
        }
      }
    } else {
      if (sgn) {
#pragma unroll
        for (int u = 0; u < N; ++u) {
          if (which == 0) lds_min(reinterpret_cast<int32_t*>(lds + a[u]), (int32_t)x[u]);
          else lds_max(reinterpret_cast<int32_t*>(lds + a[u]), (int32_t)x[u]);
        }
      } else {
#pragma unroll
        for (int u = 0; u < N; ++u) {
          if (which == 0) lds_min(reinterpret_cast<uint32_t*>(lds + a[u]), (uint32_t)x[u]);
          else lds_max(reinterpret_cast<uint32_t*>(lds + a[u]), (uint32_t)x[u]);
        }
      }
    }
  }
}

// lean_update of the value column at byte OFF of N records (RB16 16-byte
// units each), its validity from bit c.flag_bit of the records' hash words.
template <int W, int OFF, int N, int RB16>
__device__ __forceinline__ void lean_column(uint8_t* lds, const LeanCol& c, const u32x4 (&rec)[N][RB16],
                                            const uint32_t (&loff)[N], const uint32_t (&hf)[N], uint32_t dum)
{
  uint64_t x[N];
  bool valid[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    const uint32_t lo = lean_word<OFF / 4, RB16>(rec[u]);
    if constexpr (W == 8) x[u] = (uint64_t)lo | (uint64_t)lean_word<OFF / 4 + 1, RB16>(rec[u]) << 32;
    else x[u] = lo;
    valid[u] = c.flag_bit < 0 || ((hf[u] >> c.flag_bit) & 1u);
  }
  lean_update<W, N>(lds, c, x, valid, loff, dum);
}

// Aggregates the records [begin, end) into the LDS table. Rows whose group
// does not fit go to the global table. Ends with a workgroup barrier.
template <int K0, int K1, int W0, int W1, int W2>
__device__ __forceinline__ void lean_aggregate_rows(const LeanAggArgs& a, const LdsTable& t, uint8_t* lds,
                                                    int64_t begin, int64_t end)
{
  constexpr int KB = lean_key_offsets(K0, K1).bytes;
  constexpr int KW = K1 == 0 ? K0 : 8;  // bytes of the key's header part
  constexpr LeanOffsets LO = lean_offsets(KB, W0, W1, W2);
  constexpr int RB16 = LO.bytes <= 16 ? 1 : 2;
  constexpr int R = kLeanRows<RB16>;
  const global_t<const u32x4>* recs = to_global(reinterpret_cast<const u32x4*>(a.recs));
  const uint32_t dum = (uint32_t)(reinterpret_cast<uint8_t*>(t.count + 4) - lds) + (threadIdx.x & 63u) * 8u;
  const uint32_t stride = (uint32_t)t.stride;
  u32x4 cur[R][RB16], nxt[R][RB16];
  auto load = [&](int64_t b, u32x4 (&d)[R][RB16]) {
#pragma unroll
    for (int u = 0; u < R; ++u) {
      int64_t r = b + (int64_t)u * kLdsBlock;
      r = r < end ? r : end - 1;
#pragma unroll
      for (int j = 0; j < RB16; ++j) d[u][j] = recs[r * RB16 + j];
    }
  };
  if (begin < end) load(begin + threadIdx.x, cur);
  LeanCol cols[kLeanCols];
#pragma unroll
  for (int c = 0; c < kLeanCols; ++c) cols[c] = a.cols[c];
  int rows_off = a.rows_off, log2p = a.log2p;
  const uint32_t tail_off = (uint32_t)a.tail_off;
  for (int64_t base = begin + threadIdx.x; base < end; base += (int64_t)R * kLdsBlock) {
    // the uniform values are redefined (in scalar registers) every step, so
    // that the compiler tests them where they are used instead of keeping a
    // lane mask of every test across the loop
#pragma unroll
    for (int c = 0; c < kLeanCols; ++c)
      asm volatile("" : "+s"(cols[c].sum_off), "+s"(cols[c].min_off), "+s"(cols[c].max_off), "+s"(cols[c].nulls_off),
                   "+s"(cols[c].cls), "+s"(cols[c].flag_bit));
    asm volatile("" : "+s"(rows_off), "+s"(log2p));
    load(base + (int64_t)R * kLdsBlock, nxt);
    uint64_t kk[R], kb[R];
    uint32_t hf[R], sh[R], loff[R], i0[R];
    uint64_t hv[R][kLeanProbe];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      if constexpr (KW == 8) kk[u] = (uint64_t)cur[u][0][0] | (uint64_t)cur[u][0][1] << 32;
      else kk[u] = a.key_signed ? (uint64_t)(int64_t)(int32_t)cur[u][0][0] : (uint64_t)cur[u][0][0];
      kb[u] = 0;
      if constexpr (KB == 12) kb[u] = cur[u][0][2];
      if constexpr (KB == 16) kb[u] = (uint64_t)cur[u][0][2] | (uint64_t)cur[u][0][3] << 32;
      hf[u] = lean_word<KB / 4, RB16>(cur[u]);
      sh[u] = (hf[u] & 0xffffff00u) << log2p;
      i0[u] = (uint32_t)reduce_range(sh[u], t.slots);
      uint32_t i = i0[u];
#pragma unroll
      for (int j = 0; j < kLeanProbe; ++j) {
        hv[u][j] = __hip_atomic_load(reinterpret_cast<unsigned long long*>(t.rec + i * stride), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
        i = i + 1 == (uint32_t)t.slots ? 0u : i + 1;
      }
    }
    uint32_t miss = 0;
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const bool present = base + (int64_t)u * kLdsBlock < end;
      // the key's record among the first kLeanProbe slots of its probe
      // sequence, before any empty one (the full probe otherwise)
      uint32_t found = kRecSpill;
      bool open = true;
      uint32_t i = i0[u];
#pragma unroll
      for (int j = 0; j < kLeanProbe; ++j) {
        if (open && found == kRecSpill && hv[u][j] == kk[u]) found = i * stride;
        open = open && hv[u][j] != kEmptyIntKey;
        i = i + 1 == (uint32_t)t.slots ? 0u : i + 1;
      }
      loff[u] = !present ? kRecNone : kk[u] == kEmptyIntKey ? kRecSpill : found;
    }
    if constexpr (K1 != 0 && KB > 8) {
      // part B of the slot whose header is part A: read once the tail is
      // written (a slot still being claimed, or another part B: the full probe)
#pragma unroll
      for (int u = 0; u < R; ++u) {
        if (loff[u] >= kRecSpill) continue;
        const unsigned long long* tl = reinterpret_cast<const unsigned long long*>(t.rec + loff[u] + tail_off);
        if constexpr (KB == 12) {
          const uint64_t w = __hip_atomic_load(tl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if ((uint32_t)(w >> 32) == kLeanNotReady || (uint32_t)w != (uint32_t)kb[u]) loff[u] = kRecSpill;
        } else {
          const uint64_t w1 = __hip_atomic_load(tl + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
          if ((uint32_t)(w1 >> 32) == kLeanNotReady ||
              __hip_atomic_load(tl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != kb[u])
            loff[u] = kRecSpill;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < R; ++u)
      if (loff[u] == kRecSpill) miss |= 1u << u;
#pragma unroll 1
    while (miss) {
      const int u = __builtin_ctz(miss);
      uint64_t key = 0, part_b = 0;
      uint32_t slot_hash = 0;
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (q == u) {
          key = kk[q];
          part_b = kb[q];
          slot_hash = sh[q];
        }
      uint32_t v;
      if constexpr (K1 == 0) {
        uint8_t* r = lean_find_or_insert(t, key, slot_hash);
        v = r ? (uint32_t)(r - lds) : kRecSpill;
      } else {
        v = lean2_find_or_insert<KB>(t, lds, key, part_b, (uint32_t)(base + (int64_t)u * kLdsBlock), slot_hash,
                                     tail_off);
        if (v == kLeanRetry) continue;  // the slot's claim completes before the next pass
      }
      miss &= miss - 1;
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (q == u) loff[q] = v;
    }

    if (rows_off >= 0) {
#pragma unroll
      for (int u = 0; u < R; ++u)
        lds_add(reinterpret_cast<uint32_t*>(lds + (loff[u] < kRecSpill ? loff[u] + rows_off : dum)), 1u);
    }
    if constexpr (W0 != 0) lean_column<W0, LO.off[0], R, RB16>(lds, cols[0], cur, loff, hf, dum);
    if constexpr (W1 != 0) lean_column<W1, LO.off[1], R, RB16>(lds, cols[1], cur, loff, hf, dum);
    if constexpr (W2 != 0) lean_column<W2, LO.off[2], R, RB16>(lds, cols[2], cur, loff, hf, dum);

    uint32_t spill = 0;
#pragma unroll
    for (int u = 0; u < R; ++u)
      if (loff[u] == kRecSpill) spill |= 1u << u;
#pragma unroll 1
    while (spill) {
      const int u = __builtin_ctz(spill);
      spill &= spill - 1;
      lean_spill_row(a.lp, base + (int64_t)u * kLdsBlock);
    }
#pragma unroll
    for (int u = 0; u < R; ++u)
#pragma unroll
      for (int j = 0; j < RB16; ++j) cur[u][j] = nxt[u][j];
  }
  __syncthreads();
}

// The rows of work item `b` (part_item with the compact arguments).
__device__ __forceinline__ bool lean_item(const LeanAggArgs& a, uint32_t b, int64_t& begin, int64_t& end, bool& merge,
                                          int& fold_part, int& fold_parts)
{
  const int P = 1 << a.log2p;
  if (b >= a.item_offsets[P]) return false;
  int lo = 0, hi = P - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.item_offsets[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const int part = lo;
  const uint32_t n_items = a.item_offsets[part + 1] - a.item_offsets[part];
  const int64_t ps = a.offsets[(int64_t)part * a.nblocks];
  const int64_t pe = part + 1 < P ? (int64_t)a.offsets[(int64_t)(part + 1) * a.nblocks] : (int64_t)*a.total_rows;
  begin = ps + (int64_t)(b - a.item_offsets[part]) * a.max_item_rows;
  end = begin + a.max_item_rows < pe ? begin + a.max_item_rows : pe;
  merge = n_items > 1;
  fold_part = (int)(b - a.item_offsets[part]);
  fold_parts = (int)n_items;
  return true;
}

}  // namespace mgb
