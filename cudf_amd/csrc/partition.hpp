// Device routines of the hash partitioning passes (histogram and tiled
// scatter of one chunk of rows), shared by the plain and the batched
// executions.
#pragma once

#include "groupby_host.hpp"

namespace mgb {

__device__ __forceinline__ uint32_t partition_of(uint32_t hash, int log2p) { return log2p ? hash >> (32 - log2p) : 0u; }

constexpr int kScatterPreVals = 3;  // value columns loaded with the keys (see scatter_chunk)

// Rows per partition of chunk `chunk` (rows [chunk * pp.chunk, +pp.chunk))
// -> hist[p * nblocks + chunk]. `counts`: 2^log2p words of LDS. Workgroups
// of kBlock threads.
__device__ __forceinline__ void hist_chunk(const PartParams& pp, int64_t chunk, uint32_t* counts)
{
  const AggParams& p = pp.src;
  const int P = 1 << pp.log2p;
  for (int i = threadIdx.x; i < P; i += kBlock) counts[i] = 0;
  __syncthreads();
  const int64_t begin = chunk * pp.chunk;
  const int64_t end = begin + pp.chunk < p.num_rows ? begin + pp.chunk : p.num_rows;
  constexpr int kItems = 4;  // rows per thread per step, hashed column by column
  for (int64_t row0 = begin + threadIdx.x; row0 < end; row0 += (int64_t)kItems * kBlock) {
    bool present[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      present[k] = row0 + k * kBlock < end;
      if (present[k] && p.skip_null_keys) present[k] = key_row_valid(p, row0 + k * kBlock);
    }
    uint32_t h[kItems];
    uint64_t key0[kItems];
    hash_rows(p, row0, kBlock, present, h, key0);
#pragma unroll
    for (int k = 0; k < kItems; ++k)
      if (present[k]) atomicAdd(&counts[partition_of(h[k], pp.log2p)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += kBlock) pp.hist[(int64_t)i * pp.nblocks + chunk] = counts[i];
  __syncthreads();
}

// Exclusive scan of one value per thread across a 1024-thread workgroup
// (16 wave64); returns the thread's prefix, and the total through *total.
// `tmp` holds 16 words of LDS.
__device__ __forceinline__ uint32_t block_exclusive_scan_1024(uint32_t v, uint32_t* tmp, uint32_t* total)
{
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= (unsigned)d) x += y;
  }
  if (lane == 63) tmp[wave] = x;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < 16 ? tmp[lane] : 0u;
    for (int d = 1; d < 16; d <<= 1) {
      uint32_t y = __shfl_up(w, d, 64);
      if (lane >= (unsigned)d) w += y;
    }
    if (lane < 16) tmp[lane] = w;  // inclusive wave totals
  }
  __syncthreads();
  const uint32_t before = wave ? tmp[wave - 1] : 0u;
  *total = tmp[15];
  __syncthreads();
  return before + x - v;
}

constexpr int kScatterItems = 4;                        // rows per thread per tile (at most)
constexpr int kScatterTile = kLdsBlock * kScatterItems;  // 4096 rows

// Tiled scatter of chunk `chunk` (rows [chunk * pp.chunk, +pp.chunk)): each
// tile of rows is counting-sorted by partition in LDS (one LDS atomic per row
// gives its rank inside its partition's run) and staged as records;
// consecutive threads then copy consecutive 8-byte words of the sorted
// records, so each partition's run leaves as one contiguous stretch. A
// thread's rows of a tile are handled column by column. The loads of the next
// tile (the first two key columns, the first kScatterPreVals value columns and
// their null-mask words) are issued before the current tile's copy-out, so that
// reading the input overlaps writing the records. LDS: per-partition cursor
// and tile count, partition id per staged row, the staged records.
__device__ __forceinline__ void scatter_chunk(const PartParams& pp, int64_t chunk, uint8_t* smem)
{
  const AggParams& p = pp.src;
  const int P = 1 << pp.log2p;
  const int tile_rows_max = pp.tile_rows;             // multiple of kLdsBlock, <= kScatterTile
  const int rec_words = pp.rec_bytes / 8;
  uint32_t* cursor = reinterpret_cast<uint32_t*>(smem);
  uint32_t* tcount = cursor + P;
  uint32_t* scan_tmp = tcount + P;  // 16 words
  uint16_t* pid = reinterpret_cast<uint16_t*>(scan_tmp + 16);
  // staged records start 16-byte aligned after the partition ids (offset
  // computed from smem itself so the compiler keeps the LDS address space)
  const int stage_off = ((P * 8 + 16 * 4 + kScatterTile * 2) + 15) & ~15;
  uint8_t* stage_bytes = smem + stage_off;
  const uint64_t* stage = reinterpret_cast<const uint64_t*>(smem + stage_off);
  uint64_t* out = reinterpret_cast<uint64_t*>(pp.out);
  for (int i = threadIdx.x; i < P; i += kLdsBlock) {
    cursor[i] = pp.offsets[(int64_t)i * pp.nblocks + chunk];
    tcount[i] = 0;
  }
  __syncthreads();
  const int64_t begin = chunk * pp.chunk;
  const int64_t end = begin + pp.chunk < p.num_rows ? begin + pp.chunk : p.num_rows;
  const int items = tile_rows_max / kLdsBlock;
  const int64_t step = kLdsBlock;
  // up to two key columns without null grouping: the hash comes from the
  // prefetched keys alone
  bool pre_keys = p.num_keys <= 2;
  for (int c = 0; c < p.num_keys && c < 2; ++c) pre_keys = pre_keys && !(p.compare_key_nulls && p.keys[c].mask);

  // the thread's rows of one tile, as loaded
  bool present[kScatterItems];
  uint64_t kpre[2][kScatterItems];
  uint64_t vpre[kScatterPreVals][kScatterItems];
  uint32_t mword[kScatterPreVals][kScatterItems];
  auto fetch = [&](int64_t tile) {
    const int64_t row0 = tile + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kScatterItems; ++k) present[k] = k < items && row0 + k * step < end;
#pragma unroll
    for (int c = 0; c < 2; ++c)
      if (c < p.num_keys)
        load_rows(p.keys[c].data, p.keys[c].width, p.keys[c].stride, VC_UNSIGNED, row0, step, present, kpre[c]);
#pragma unroll
    for (int c = 0; c < kScatterPreVals; ++c) {
      if (c >= p.num_vals) continue;
      const ValCol& v = p.vals[c];
      load_rows(v.data, v.width, v.stride, VC_UNSIGNED, row0, step, present, vpre[c]);
#pragma unroll
      for (int k = 0; k < kScatterItems; ++k)
        mword[c][k] = (v.mask && present[k]) ? to_global(v.mask)[(v.bit_off + row0 + k * step) >> 5] : ~0u;
    }
  };

  if (begin < end) fetch(begin);
  for (int64_t tile = begin; tile < end; tile += tile_rows_max) {
    const int64_t row0 = tile + threadIdx.x;
    if (p.skip_null_keys) {
#pragma unroll
      for (int k = 0; k < kScatterItems; ++k)
        if (present[k]) present[k] = key_row_valid(p, row0 + k * step);
    }
    uint32_t h[kScatterItems];
    if (pre_keys) {
#pragma unroll
      for (int k = 0; k < kScatterItems; ++k) h[k] = 0u;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        if (c >= p.num_keys) continue;
        const KeyCol& kc = p.keys[c];
#pragma unroll
        for (int k = 0; k < kScatterItems; ++k)
          h[k] = hash_combine(h[k], murmur3_bits(normalize_float_bits(kpre[c][k], kc.vclass), kc.width, 0));
      }
    } else {
      uint64_t ignored[kScatterItems];
      hash_rows(p, row0, step, present, h, ignored);
    }
    uint32_t flags[kScatterItems] = {};
    if (pp.flag_bytes) {
#pragma unroll
      for (int c = 0; c < kScatterPreVals; ++c) {
        if (c >= p.num_vals || p.vals[c].flag_bit < 0) continue;
        const ValCol& v = p.vals[c];
#pragma unroll
        for (int k = 0; k < kScatterItems; ++k)
          if (present[k] && ((mword[c][k] >> ((v.bit_off + row0 + k * step) & 31)) & 1u))
            flags[k] |= 1u << v.flag_bit;
      }
#pragma unroll 1
      for (int c = kScatterPreVals; c < p.num_vals; ++c) {
        const ValCol& v = p.vals[c];
        if (v.flag_bit < 0) continue;
#pragma unroll
        for (int k = 0; k < kScatterItems; ++k)
          if (present[k] && bit_valid(v.mask, v.bit_off + row0 + k * step)) flags[k] |= 1u << v.flag_bit;
      }
    }
    uint32_t part[kScatterItems], rank[kScatterItems];
#pragma unroll
    for (int k = 0; k < kScatterItems; ++k) {
      part[k] = partition_of(h[k], pp.log2p);
      rank[k] = present[k] ? atomicAdd(&tcount[part[k]], 1u) : 0u;
    }
    __syncthreads();
    // tile-local start of each partition's run
    uint32_t tile_rows = 0;
    for (int base = 0; base < P; base += kLdsBlock) {
      const int i = base + threadIdx.x;
      const uint32_t c = i < P ? tcount[i] : 0u;
      uint32_t chunk_total;
      const uint32_t ex = block_exclusive_scan_1024(c, scan_tmp, &chunk_total);
      if (i < P) tcount[i] = tile_rows + ex;  // now the run start
      tile_rows += chunk_total;
      __syncthreads();
    }
    // stage rows in partition order, column by column
    uint32_t pos[kScatterItems];
#pragma unroll
    for (int k = 0; k < kScatterItems; ++k) {
      pos[k] = present[k] ? tcount[part[k]] + rank[k] : 0u;
      if (present[k]) pid[pos[k]] = (uint16_t)part[k];
    }
#pragma unroll 1
    for (int c = 0; c < p.num_keys; ++c) {
      const KeyCol& kc = p.keys[c];
      uint64_t x[kScatterItems];
      switch (c) {
        case 0: copy_rows(kpre[0], x); break;
        case 1: copy_rows(kpre[1], x); break;
        default: load_rows(kc.data, kc.width, kc.stride, VC_UNSIGNED, row0, step, present, x); break;
      }
      store_rows(stage_bytes, pos, pp.rec_bytes, pp.key_off[c], kc.width, x, present);
    }
#pragma unroll 1
    for (int c = 0; c < p.num_vals; ++c) {
      const ValCol& v = p.vals[c];
      uint64_t x[kScatterItems];
      switch (c) {
        case 0: copy_rows(vpre[0], x); break;
        case 1: copy_rows(vpre[1], x); break;
        case 2: copy_rows(vpre[2], x); break;
        default: load_rows(v.data, v.width, v.stride, VC_UNSIGNED, row0, step, present, x); break;
      }
      store_rows(stage_bytes, pos, pp.rec_bytes, pp.val_off[c], v.width, x, present);
    }
    if (pp.flag_bytes) {
      uint64_t f[kScatterItems];
#pragma unroll
      for (int k = 0; k < kScatterItems; ++k) f[k] = flags[k];
      store_rows(stage_bytes, pos, pp.rec_bytes, pp.flags_off, pp.flag_bytes, f, present);
    }
    if (pp.rowid_off >= 0) {
      uint64_t rid[kScatterItems];
#pragma unroll
      for (int k = 0; k < kScatterItems; ++k) rid[k] = (uint64_t)(row0 + k * step);
      store_rows(stage_bytes, pos, pp.rec_bytes, pp.rowid_off, 4, rid, present);
    }
    __syncthreads();
    // the next tile's loads fly while this one is copied out
    if (tile + tile_rows_max < end) fetch(tile + tile_rows_max);
    // copy out word by word: the words of one partition's run are contiguous
    const uint32_t words = tile_rows * (uint32_t)rec_words;
    for (uint32_t w = threadIdx.x; w < words; w += kLdsBlock) {
      const uint32_t r = (uint32_t)(((uint64_t)w * pp.rec_magic) >> 32);  // w / rec_words
      const uint32_t q = pid[r];
      const int64_t dst = (int64_t)cursor[q] + (r - tcount[q]);
      out[dst * rec_words + (w - r * (uint32_t)rec_words)] = stage[w];
    }
    __syncthreads();
    // advance cursors by this tile's runs and clear the counts
    for (int i = threadIdx.x; i < P; i += kLdsBlock) {
      const uint32_t next = (i + 1 < P) ? tcount[i + 1] : tile_rows;
      cursor[i] += next - tcount[i];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P; i += kLdsBlock) tcount[i] = 0;
    __syncthreads();
  }
  __syncthreads();
}

}  // namespace mgb
