// Device routines of LDS aggregation work items (one workgroup of kLdsBlock
// threads per item): set-up of the item's LDS table, aggregation of a range
// of rows into it, and the two ways an item ends: folding its groups into a
// global table, or writing them as output rows.
#pragma once

#include "groupby_host.hpp"

namespace mgb {

// A row whose group did not fit the LDS table: device atomics on the global
// table. A call of its own, so that the LDS path's record pointer and this
// global one never meet in one variable (which would make both flat).
__device__ __forceinline__ void spill_row(const AggParams& p, const RowKey& rk, int64_t row)
{
  uint8_t* g = global_find_or_insert(p, rk, row);
  if (g) apply_row(p, g, row);
}

// The LDS table of a work item: `slots` records of `words` 8-byte words, one
// record for the 8-byte key equal to the empty header, then counters
// (count, overflow, a 64-bit output base) and kLdsBlock ints of scan space.
__device__ __forceinline__ LdsTable lds_table_at(uint8_t* lds, int slots, int words, int cap)
{
  LdsTable t;
  t.rec = lds;
  t.slots = slots;
  t.stride = words * 8;  // LDS record stride (the record padded to 8 x odd bytes)
  t.cap = cap;
  int32_t* ctr = reinterpret_cast<int32_t*>(lds + (int64_t)(slots + 1) * t.stride);
  t.count = ctr;
  t.overflow = ctr + 1;
  return t;
}

// LDS bytes of a table of `slots` records of `words` words.
__host__ __device__ constexpr size_t lds_table_bytes(int slots, int words)
{
  return (size_t)(slots + 1) * words * 8 + 16 + (size_t)kLdsBlock * 4;
}

// Writes the identity record into every slot (the special record starts
// unoccupied) and clears the counters.
__device__ __forceinline__ void lds_table_init(const LdsTable& t, const uint64_t* tmpl, int words)
{
  uint64_t* w64 = reinterpret_cast<uint64_t*>(t.rec);
  const int total_words = (t.slots + 1) * words;
  for (int i = threadIdx.x; i < total_words; i += kLdsBlock) {
    const int wi = i % words;
    uint64_t v = tmpl[wi];
    if (wi == 0 && i / words == t.slots) v = 0ull;
    w64[i] = v;
  }
  if (threadIdx.x < 4) t.count[threadIdx.x] = 0;
  __syncthreads();
}

// The row loop of lds_aggregate_rows; kRec: rows are whole records
// (p.rec_words), loaded with one or two wide loads each.
template <bool kRec>
__device__ __forceinline__ void lds_aggregate_rows_impl(const AggParams& p, const LdsTable& t, uint8_t* lds,
                                                        int64_t begin, int64_t end, int log2p, bool check_keys)
{
  // Rows are taken kRows per thread at a time (rows base + u * kLdsBlock):
  // the loads of the key, the flags and the first kPreCols value columns (or
  // the whole records) of all of them are issued together, then the rows are probed (the home slot
  // headers of all rows read together, the full probe only on a miss), then
  // each value column updates the rows' records (lds_column_rows). Rows whose
  // group does not fit go to the global table afterwards.
  constexpr int kRows = 3;
  constexpr int kPreCols = kRec ? 0 : 3;  // the switch below picks pre[0..2]
  const int64_t step = kLdsBlock;
  const KeyCol& k0 = p.keys[0];
  // the thread's dummy word for updates of rows that have none: in the scan
  // space after the table, unused until the groups are written out
  const uint32_t dum = (uint32_t)(reinterpret_cast<uint8_t*>(t.count + 4) - lds) + (threadIdx.x & 63u) * 8u;
  // Records: the uniform values of the row loop are held in vector
  // registers and made scalar at each use (v_readfirstlane). Held in scalar
  // registers, the compiler reloads them from memory when it runs short of
  // those, and each such load waits for every LDS update in flight.
  uint32_t pin_a = 0, pin_b = 0, pin_c = 0;
  uint64_t pin_base = 0;
  uint64_t pin_prog[kRec ? kProgramCols : 1] = {};
  if constexpr (kRec) {
    const int has_flags = p.flags != nullptr;
    pin_a = (uint32_t)(k0.width & 15) | (uint32_t)(k0.vclass & 7) << 4 |
            (uint32_t)((k0.data - p.rec_base) & 63) << 7 |
            (uint32_t)(has_flags ? (p.flags - p.rec_base) & 63 : 0) << 13 | (uint32_t)(p.flag_bytes & 7) << 19 |
            (uint32_t)has_flags << 22 | (uint32_t)(p.key_mode == KEY_INT) << 23 | (uint32_t)(p.rec_words & 7) << 24;
    pin_b = (uint32_t)t.slots | (uint32_t)t.stride << 16;
    pin_c = (uint32_t)((p.rows_op >= 0 ? p.ops[p.rows_op].off : -1) + 1) | (uint32_t)(p.num_vals & 31) << 10 |
            (uint32_t)p.program_cols << 15;
    pin_base = (uint64_t)p.rec_base;
    asm volatile("" : "+v"(pin_a), "+v"(pin_b), "+v"(pin_c), "+v"(pin_base));
#pragma unroll
    for (int c = 0; c < kProgramCols; ++c) {
      pin_prog[c] = p.prog[c];
      asm volatile("" : "+v"(pin_prog[c]));
    }
  }
  for (int64_t base = begin + threadIdx.x; base < end; base += (int64_t)kRows * kLdsBlock) {
    bool present[kRows];
#pragma unroll
    for (int u = 0; u < kRows; ++u) present[u] = base + u * step < end;
    // this iteration's uniform values
    int kw, kvc, slots, stride, rows_off, nvals;
    bool int_key, prog_cols;
    if constexpr (kRec) {
      const uint32_t a = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_a);
      const uint32_t b = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_b);
      const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_c);
      kw = (int)(a & 15);
      kvc = (int)((a >> 4) & 7);
      int_key = (a >> 23) & 1;
      slots = (int)(b & 0xffff);
      stride = (int)(b >> 16);
      rows_off = (int)(c & 1023) - 1;
      nvals = (int)((c >> 10) & 31);
      prog_cols = (c >> 15) & 1;
    } else {
      kw = k0.width;
      kvc = k0.vclass;
      int_key = p.key_mode == KEY_INT;
      slots = t.slots;
      stride = t.stride;
      rows_off = p.rows_op >= 0 ? p.ops[p.rows_op].off : -1;
      nvals = p.num_vals;
      prog_cols = p.program_cols;
    }
    uint64_t key[kRows];
    uint32_t flags[kRows];
    // records: every field of the rows comes from registers (rw)
    RecWords rw[kRec ? kRows : 1];
    uint64_t pre[kRec ? 1 : kPreCols][kRows];
    if constexpr (kRec) {
      const uint32_t a = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_a);
      const uint64_t rb = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pin_base) |
                          (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pin_base >> 32)) << 32;
      load_recs(reinterpret_cast<const uint8_t*>(rb), (int)((a >> 24) & 7), base, step, present, base, rw);
      if (int_key) extract_rows(rw, (int)((a >> 7) & 63), kw, VC_UNSIGNED, key);
      const bool has_flags = (a >> 22) & 1;
      uint64_t f[kRows];
      if (has_flags) extract_rows(rw, (int)((a >> 13) & 63), (int)((a >> 19) & 7), VC_UNSIGNED, f);
#pragma unroll
      for (int u = 0; u < kRows; ++u) flags[u] = has_flags ? (uint32_t)f[u] : 0u;
    } else {
      if (int_key) load_rows(k0.data, k0.width, k0.stride, VC_UNSIGNED, base, step, present, key);
      {
        uint64_t f[kRows];
        if (p.flags) load_rows(p.flags, p.flag_bytes, p.flag_stride, VC_UNSIGNED, base, step, present, f);
#pragma unroll
        for (int u = 0; u < kRows; ++u) flags[u] = p.flags ? (uint32_t)f[u] : 0u;
      }
#pragma unroll
      for (int c = 0; c < kPreCols; ++c)
        if (c < p.num_vals)
          load_rows(p.vals[c].data, p.vals[c].width, p.vals[c].stride, p.vals[c].vclass, base, step, present,
                    pre[c]);
    }
    if (check_keys && p.skip_null_keys) {
#pragma unroll
      for (int u = 0; u < kRows; ++u)
        if (present[u]) present[u] = key_row_valid(p, base + u * step);
    }

    // probe: the home slot and the next one of every row are read together;
    // rows found in neither (new groups, longer probe runs, row-reference
    // keys) take the full probe one row at a time
    uint32_t loff[kRows];
    uint32_t miss = 0;
    // int keys: the key and hash of each row, for the full probe of a miss
    uint64_t kk[kRows];
    uint32_t hsh[kRows];
    if (int_key) {
      uint64_t h0[kRows], h1[kRows];
      uint32_t i0[kRows], i1[kRows];
#pragma unroll
      for (int u = 0; u < kRows; ++u) {
        kk[u] = (uint64_t)widen_int(key[u], kw, kvc);
        // the slot hash: the row hash's bits below the ones that chose the
        // partition. (A multiplicative hash of the key is cheaper, but keys
        // in an arithmetic progression modulo 2^32 cluster under it: batches
        // of such keys ran 6.6 times slower.)
        hsh[u] = hash_combine(0u, murmur3_bits(normalize_float_bits(key[u], kvc), kw, 0)) << log2p;
        i0[u] = (uint32_t)reduce_range(hsh[u], slots);
        i1[u] = i0[u] + 1 == (uint32_t)slots ? 0u : i0[u] + 1;
        h0[u] = h1[u] = kEmptyIntKey;
        if (present[u]) {
          h0[u] = __hip_atomic_load(reinterpret_cast<unsigned long long*>(t.rec + i0[u] * (uint32_t)stride),
                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          h1[u] = __hip_atomic_load(reinterpret_cast<unsigned long long*>(t.rec + i1[u] * (uint32_t)stride),
                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
#pragma unroll
      for (int u = 0; u < kRows; ++u) {
        loff[u] = kRecNone;
        if (!present[u]) continue;
        if (kk[u] != kEmptyIntKey && h0[u] == kk[u]) loff[u] = i0[u] * (uint32_t)stride;
        else if (kk[u] != kEmptyIntKey && h0[u] != kEmptyIntKey && h1[u] == kk[u]) loff[u] = i1[u] * (uint32_t)stride;
        else miss |= 1u << u;
      }
    } else {
#pragma unroll
      for (int u = 0; u < kRows; ++u) {
        loff[u] = kRecNone;
        if (present[u]) miss |= 1u << u;
      }
    }
#pragma unroll 1
    while (miss) {
      const int u = __builtin_ctz(miss);
      miss &= miss - 1;
      const int64_t row = base + u * step;
      RowKey rk;
      uint32_t slot_hash;
      if (int_key) {
        // from the probe's registers (no second read of the key); rk.hash
        // is not read for integer keys
#pragma unroll
        for (int q = 0; q < kRows; ++q)
          if (q == u) {
            rk.key = kk[q];
            slot_hash = hsh[q];
          }
        rk.hash = slot_hash;
        rk.special = rk.key == kEmptyIntKey;
      } else {
        rk = row_key(p, row);
        slot_hash = rk.hash << log2p;
      }
      uint8_t* r = lds_find_or_insert(t, p, rk, slot_hash, row);
      const uint32_t v = r ? (uint32_t)(r - lds) : kRecSpill;
#pragma unroll
      for (int q = 0; q < kRows; ++q)
        if (q == u) loff[q] = v;
    }

    // updates
    if (rows_off >= 0) {
#pragma unroll
      for (int u = 0; u < kRows; ++u)
        lds_add(reinterpret_cast<uint32_t*>(lds + (loff[u] < kRecSpill ? loff[u] + rows_off : dum)), 1u);
    }
    // records whose columns all take the column program: the column loop
    // unrolled, so that no column's descriptor is loaded inside the row loop
    // (a scalar load there would wait for every LDS update issued before it)
    bool unrolled = false;
    if constexpr (kRec) {
      if (prog_cols) {
        unrolled = true;
#pragma unroll
        for (int c = 0; c < kProgramCols; ++c) {
          if (c >= nvals) break;
          const uint64_t pg = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pin_prog[c]) |
                              (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pin_prog[c] >> 32))
                                  << 32;
          const Program v = unpack_program(pg);
          uint64_t x[kRows];
          extract_rows(rw, v.rec_off, v.width, v.vclass, x);
          bool ok[kRows];
#pragma unroll
          for (int u = 0; u < kRows; ++u)
            ok[u] = loff[u] < kRecSpill && (v.flag_bit < 0 || ((flags[u] >> v.flag_bit) & 1u));
          if (v.nulls_off >= 0) {
#pragma unroll
            for (int u = 0; u < kRows; ++u)
              lds_add(reinterpret_cast<uint32_t*>(lds + (loff[u] < kRecSpill && !ok[u] ? loff[u] + v.nulls_off : dum)),
                      1u);
          }
          lds_column_program(lds, v, loff, x, ok, dum);
        }
      }
    }
    // one call site for all columns keeps the kernel's code small; the
    // preloaded values are picked by a uniform switch
#pragma unroll 1
    for (int c = 0; c < (unrolled ? 0 : p.num_vals); ++c) {
      uint64_t x[kRows];
      if constexpr (kRec) {
        extract_rows(rw, (int)(p.vals[c].data - p.rec_base), p.vals[c].width, p.vals[c].vclass, x);
      } else {
        switch (c) {
          case 0: copy_rows(pre[0], x); break;
          case 1: copy_rows(pre[1], x); break;
          case 2: copy_rows(pre[2], x); break;
          default:
            load_rows(p.vals[c].data, p.vals[c].width, p.vals[c].stride, p.vals[c].vclass, base, step, present, x);
            break;
        }
      }
      lds_column_rows(p, lds, p.vals[c], loff, x, flags, base, step, dum);
    }

    // groups that did not fit: the global table (a separate call keeps the
    // two pointers' address spaces apart)
    uint32_t spill = 0;
#pragma unroll
    for (int u = 0; u < kRows; ++u)
      if (loff[u] == kRecSpill) spill |= 1u << u;
#pragma unroll 1
    while (spill) {
      const int u = __builtin_ctz(spill);
      spill &= spill - 1;
      spill_row(p, row_key(p, base + u * step), base + u * step);
    }
  }
  __syncthreads();
}

// Aggregates rows [begin, end) of `p` into the LDS table. `log2p` is the
// number of hash bits that chose the rows' partition (the slot comes from the
// bits below them); rows with a null key are dropped when `check_keys`.
// Ends with a workgroup barrier.
__device__ __forceinline__ void lds_aggregate_rows(const AggParams& p, const LdsTable& t, uint8_t* lds, int64_t begin,
                                                   int64_t end, int log2p, bool check_keys)
{
  if (p.rec_words > 0) lds_aggregate_rows_impl<true>(p, t, lds, begin, end, log2p, check_keys);
  else lds_aggregate_rows_impl<false>(p, t, lds, begin, end, log2p, check_keys);
}

// Folds every group of the LDS table into p's global table (device atomics).
// `part` of `nparts`: items that cover parts of one partition hold mostly the
// same groups in the same slots and tend to finish together; each starts its
// pass at a different slot, so that their atomics on one record do not meet.
//
// The group budget is checked once for the whole table (its groups at most
// all new): within it, inserts are counted in LDS and added to the table's
// group counter once at the end. Tables folded together may then overshoot
// the budget by at most one table each, which the global table's slack past
// its budget holds; past the budget, every insert checks it as usual.
__device__ __forceinline__ void lds_fold_to_global(const AggParams& p, const LdsTable& t, int part = 0, int nparts = 1)
{
  const int n = t.slots + 1;
  const int first = nparts > 1 ? (int)(((int64_t)n * part) / nparts) : 0;
  uint32_t* inserted = reinterpret_cast<uint32_t*>(t.count + 2);  // free until groups are written out
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long used = __hip_atomic_load(&p.counters[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long over = __hip_atomic_load(&p.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    inserted[0] = 0u;
    inserted[1] = over == 0ull && used + (unsigned long long)*t.count <= (unsigned long long)p.max_groups ? 1u : 0u;
  }
  __syncthreads();
  uint32_t* local = inserted[1] ? inserted : nullptr;
  for (int j = threadIdx.x; j < n; j += kLdsBlock) {
    const int i = j + first < n ? j + first : j + first - n;
    if (!lds_occupied(t, i, p.key_mode)) continue;
    const uint8_t* rec = lds_record(t, i);
    const uint64_t hdr = *reinterpret_cast<const uint64_t*>(rec);
    RowKey rk;
    int64_t rep = 0;
    if (p.key_mode == KEY_INT) {
      const KeyCol& k = p.keys[0];
      rk.special = i == t.slots;
      rk.key = rk.special ? kEmptyIntKey : hdr;
      const uint64_t raw = k.width == 8 ? rk.key : (rk.key & ((1ull << (8 * k.width)) - 1ull));
      rk.hash = hash_combine(0u, murmur3_bits(normalize_float_bits(raw, k.vclass), k.width, 0));
    } else {
      rep = (int64_t)(hdr & 0xffffffffull);
      rk = row_key(p, rep);
    }
    uint8_t* g = global_find_or_insert(p, rk, rep, local == nullptr, local);
    if (g) merge_record(p, g, rec);
  }
  if (local) {
    __syncthreads();
    if (threadIdx.x == 0 && inserted[0]) atom_add(&p.counters[0], (unsigned long long)inserted[0]);
  }
}

// Writes every group of the LDS table as an output row: rows pos_base + n
// onward, n taken from *out_counter (one atomic per workgroup).
__device__ __forceinline__ void lds_write_groups(const AggParams& p, const LdsTable& t,
                                                 unsigned long long* out_counter, int64_t pos_base, const OutTable& ot)
{
  int32_t* scan = t.count + 4;  // kLdsBlock ints
  unsigned long long* base = reinterpret_cast<unsigned long long*>(t.count + 2);
  int mine = 0;
  for (int i = threadIdx.x; i <= t.slots; i += kLdsBlock) mine += lds_occupied(t, i, p.key_mode) ? 1 : 0;
  scan[threadIdx.x] = mine;
  __syncthreads();
  for (int d = 1; d < kLdsBlock; d <<= 1) {
    int v = threadIdx.x >= (unsigned)d ? scan[threadIdx.x - d] : 0;
    __syncthreads();
    scan[threadIdx.x] += v;
    __syncthreads();
  }
  if (threadIdx.x == kLdsBlock - 1) *base = atomicAdd(out_counter, (unsigned long long)scan[kLdsBlock - 1]);
  __syncthreads();
  int64_t pos = pos_base + (int64_t)*base + scan[threadIdx.x] - mine;
  for (int i = threadIdx.x; i <= t.slots; i += kLdsBlock) {
    if (!lds_occupied(t, i, p.key_mode)) continue;
    const uint8_t* rec = lds_record(t, i);
    finalize_keys(*reinterpret_cast<const uint64_t*>(rec), i == t.slots, p.key_mode, p.keys, p.num_keys, pos, ot);
    finalize_record(rec, pos, ot);
    ++pos;
  }
}

// ------------------------------------------------------------- second pass

// The second pass of a row whose group is not in the LDS table: its record
// in the global table (a call of its own, as spill_row).
__device__ __forceinline__ void pass2_spill(const AggParams& p, const Pass2Ops* __restrict__ q, const RowKey& rk,
                                            int64_t row, uint32_t flags, uint32_t src_row)
{
  uint8_t* g = global_find(p, rk, row);
  if (g) pass2_update<false>(q, g, row, flags, src_row);
}

// lds_pass2_rows over whole records (p.rec_words > 0) with an integer key
// and value columns that all take a column program (p.program_cols): the
// loop of lds_aggregate_rows_impl<true>. The uniform values are held in
// vector registers and made scalar at each use, so that no scalar load (which
// waits for every LDS update in flight) falls inside the loop; each column's
// first-pass program (p.prog) gives its value, validity bit and SUM / MIN /
// MAX fields, q->p2prog its second-pass fields. Every group is in the table
// (a whole partition without spill). Ends with a workgroup barrier.
__device__ __forceinline__ void lds_pass2_recs(const AggParams& p, const Pass2Ops* __restrict__ q, const LdsTable& t,
                                               int64_t begin, int64_t end, int log2p)
{
  constexpr int kRows = 3;
  const int64_t step = kLdsBlock;
  const KeyCol& k0 = p.keys[0];
  const int has_flags = p.flags != nullptr;
  const int has_rowid = q->rowid != nullptr;
  uint32_t pin_a = (uint32_t)(k0.width & 15) | (uint32_t)(k0.vclass & 7) << 4 |
                   (uint32_t)((k0.data - p.rec_base) & 63) << 7 |
                   (uint32_t)(has_flags ? (p.flags - p.rec_base) & 63 : 0) << 13 | (uint32_t)(p.flag_bytes & 7) << 19 |
                   (uint32_t)has_flags << 22 | (uint32_t)has_rowid << 23 | (uint32_t)(p.rec_words & 7) << 24;
  uint32_t pin_b = (uint32_t)t.slots | (uint32_t)t.stride << 16;
  uint32_t pin_c = (uint32_t)((p.rows_op >= 0 ? p.ops[p.rows_op].off : -1) + 1) | (uint32_t)(p.num_vals & 31) << 10 |
                   (uint32_t)(has_rowid ? (q->rowid - p.rec_base) & 63 : 0) << 15 | (uint32_t)(log2p & 31) << 21;
  uint64_t pin_base = (uint64_t)p.rec_base;
  asm volatile("" : "+v"(pin_a), "+v"(pin_b), "+v"(pin_c), "+v"(pin_base));
  uint64_t pin_prog[kProgramCols];
  uint32_t pin_p2[kProgramCols];
#pragma unroll
  for (int c = 0; c < kProgramCols; ++c) {
    pin_prog[c] = p.prog[c];
    pin_p2[c] = q->p2prog[c];
    asm volatile("" : "+v"(pin_prog[c]), "+v"(pin_p2[c]));
  }
  for (int64_t base = begin + threadIdx.x; base < end; base += (int64_t)kRows * kLdsBlock) {
    bool present[kRows];
#pragma unroll
    for (int u = 0; u < kRows; ++u) present[u] = base + u * step < end;
    const uint32_t a = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_a);
    const uint32_t b = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_b);
    const uint32_t cc = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_c);
    const int kw = (int)(a & 15), kvc = (int)((a >> 4) & 7);
    const int slots = (int)(b & 0xffff), stride = (int)(b >> 16);
    const int rows_off = (int)(cc & 1023) - 1, nvals = (int)((cc >> 10) & 31), shift = (int)((cc >> 21) & 31);
    const uint64_t rb = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pin_base) |
                        (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pin_base >> 32)) << 32;
    RecWords rw[kRows];
    load_recs(reinterpret_cast<const uint8_t*>(rb), (int)((a >> 24) & 7), base, step, present, base, rw);
    uint64_t key[kRows], f[kRows], rid[kRows];
    extract_rows(rw, (int)((a >> 7) & 63), kw, VC_UNSIGNED, key);
    if ((a >> 22) & 1) extract_rows(rw, (int)((a >> 13) & 63), (int)((a >> 19) & 7), VC_UNSIGNED, f);
    else {
#pragma unroll
      for (int u = 0; u < kRows; ++u) f[u] = 0ull;
    }
    if ((a >> 23) & 1) extract_rows(rw, (int)((cc >> 15) & 63), 4, VC_UNSIGNED, rid);
    else {
#pragma unroll
      for (int u = 0; u < kRows; ++u) rid[u] = (uint64_t)(base + u * step);
    }
    // the row's record: the slot the first pass gave its key
    uint32_t loff[kRows];
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      loff[u] = kRecNone;
      if (!present[u]) continue;
      const uint64_t kk = (uint64_t)widen_int(key[u], kw, kvc);
      if (kk == kEmptyIntKey) {
        loff[u] = (uint32_t)slots * (uint32_t)stride;
        continue;
      }
      const uint32_t hsh = hash_combine(0u, murmur3_bits(normalize_float_bits(key[u], kvc), kw, 0)) << shift;
      uint32_t i = (uint32_t)reduce_range(hsh, slots);
      for (int probe = 0; probe < slots; ++probe) {
        const uint64_t h = *reinterpret_cast<const uint64_t*>(t.rec + i * (uint32_t)stride);
        if (h == kk) {
          loff[u] = i * (uint32_t)stride;
          break;
        }
        if (h == kEmptyIntKey) break;
        i = i + 1 == (uint32_t)slots ? 0u : i + 1;
      }
    }
#pragma unroll
    for (int c = 0; c < kProgramCols; ++c) {
      if (c >= nvals) break;
      const uint32_t p2 = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_p2[c]);
      if (p2 == 0) continue;
      const uint64_t pg = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pin_prog[c]) |
                          (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pin_prog[c] >> 32)) << 32;
      const Program v = unpack_program(pg);
      uint64_t x[kRows];
      extract_rows(rw, v.rec_off, v.width, v.vclass, x);
      bool ok[kRows];
#pragma unroll
      for (int u = 0; u < kRows; ++u)
        ok[u] = loff[u] != kRecNone && (v.flag_bit < 0 || ((f[u] >> v.flag_bit) & 1u));
      const int m2_off = (int)(p2 & 1023u) - 1;
      const int amin_off = (int)((p2 >> 10) & 1023u) - 1;
      const int amax_off = (int)((p2 >> 20) & 1023u) - 1;
      const bool fl = v.vclass == VC_F32 || v.vclass == VC_F64;
      if (m2_off >= 0) {
#pragma unroll
        for (int u = 0; u < kRows; ++u) {
          if (!ok[u]) continue;
          uint8_t* rec = t.rec + loff[u];
          uint32_t cnt = *reinterpret_cast<const uint32_t*>(rec + rows_off);
          if (v.nulls_off >= 0) cnt -= *reinterpret_cast<const uint32_t*>(rec + v.nulls_off);
          double sum;
          if (fl) {
            sum = *reinterpret_cast<const double*>(rec + v.sum_off);
            if (v.vclass == VC_F32) sum = (double)(float)sum;
          } else {
            sum = (double)*reinterpret_cast<const int64_t*>(rec + v.sum_off);
          }
          const double d = widen_f64(x[u], v.width, v.vclass) - sum / (double)cnt;
          lds_add(reinterpret_cast<double*>(rec + m2_off), d * d);
        }
      }
#pragma unroll
      for (int which = 0; which < 2; ++which) {
        const int arg_off = which ? amax_off : amin_off;
        if (arg_off < 0) continue;
        const int kind = v.mm_kind + which;
        const int ref_off = which ? v.max_off : v.min_off;
        const bool wide = op_field_bytes(kind) == 8;
#pragma unroll
        for (int u = 0; u < kRows; ++u) {
          if (!ok[u]) continue;
          uint8_t* rec = t.rec + loff[u];
          const uint64_t code = minmax_code(kind, x[u], v.width, v.vclass);
          const uint64_t best = wide ? *reinterpret_cast<const uint64_t*>(rec + ref_off)
                                     : *reinterpret_cast<const uint32_t*>(rec + ref_off);
          if (code == best) lds_min(reinterpret_cast<uint32_t*>(rec + arg_off), (uint32_t)rid[u]);
        }
      }
    }
  }
  __syncthreads();
}

// The second pass over rows [begin, end) of `p` against the LDS table that
// their first pass completed. Rows are taken kRows per thread at a time:
// each finds its group (the same slot as lds_aggregate_rows: the row hash's
// bits below the partition's), then the ops are applied op by op to all of them,
// each op's descriptor read once and each column's values loaded once. A
// row whose group is not in the table (direct mode with more groups than the
// table holds) updates the global table instead. Ends with a workgroup
// barrier.
__device__ __forceinline__ void lds_pass2_rows(const AggParams& p, const Pass2Ops* __restrict__ q, const LdsTable& t,
                                               int64_t begin, int64_t end, int log2p, bool check_keys)
{
  if (p.rec_words > 0 && p.program_cols && p.key_mode == KEY_INT && !check_keys) {
    lds_pass2_recs(p, q, t, begin, end, log2p);
    return;
  }
  constexpr int kRows = 4;
  const int nops = q->num_ops;
  const uint8_t* rowid = q->rowid;
  const int rowid_stride = q->rowid_stride;
  for (int64_t base = begin + threadIdx.x; base < end; base += (int64_t)kRows * kLdsBlock) {
    uint32_t loff[kRows], flags[kRows], src[kRows];
    bool present[kRows];
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      const int64_t row = base + (int64_t)u * kLdsBlock;
      present[u] = row < end && (!check_keys || !p.skip_null_keys || key_row_valid(p, row));
      loff[u] = kRecNone;
      flags[u] = 0u;
      src[u] = (uint32_t)row;
      if (!present[u]) continue;
      flags[u] = p.flags ? load_flags(p.flags, p.flag_bytes, p.flag_stride, row) : 0u;
      if (rowid) src[u] = *reinterpret_cast<const uint32_t*>(rowid + row * (int64_t)rowid_stride);
      RowKey rk;
      uint32_t slot_hash;
      if (p.key_mode == KEY_INT) {
        const KeyCol& k = p.keys[0];
        const uint64_t raw = load_at(k.data, k.width, k.stride, row);
        rk.key = (uint64_t)widen_int(raw, k.width, k.vclass);
        rk.special = rk.key == kEmptyIntKey;
        slot_hash = hash_combine(0u, murmur3_bits(normalize_float_bits(raw, k.vclass), k.width, 0)) << log2p;
        rk.hash = slot_hash;
      } else {
        rk = row_key(p, row);
        slot_hash = rk.hash << log2p;
      }
      uint8_t* rec = lds_find(t, p, rk, slot_hash, row);
      if (rec) {
        loff[u] = (uint32_t)(rec - t.rec);
      } else {
        if (p.key_mode == KEY_INT) {
          // the global table's probe takes the row hash
          const KeyCol& k = p.keys[0];
          rk.hash = hash_combine(0u, murmur3_bits(normalize_float_bits(load_at(k.data, k.width, k.stride, row),
                                                                       k.vclass), k.width, 0));
        }
        pass2_spill(p, q, rk, row, flags[u], src[u]);
      }
    }
    uint64_t x[kRows];
    int last = -1;
    for (int i = 0; i < nops; ++i) {
      const Pass2Op& o = q->ops[i];
      if (o.val != last) {
        last = o.val;
#pragma unroll
        for (int u = 0; u < kRows; ++u)
          x[u] = loff[u] != kRecNone ? load_at(o.data, o.width, o.stride, base + (int64_t)u * kLdsBlock) : 0ull;
      }
#pragma unroll
      for (int u = 0; u < kRows; ++u) {
        if (loff[u] == kRecNone || !pass2_valid(o, base + (int64_t)u * kLdsBlock, flags[u])) continue;
        pass2_apply<true>(o, t.rec + loff[u], x[u], src[u]);
      }
    }
  }
  __syncthreads();
}

// Direct mode: copies every group of the completed global table `p.table`
// into the LDS table (all of them, when they fit; the rest stay global),
// first-pass fields included. Ends with a workgroup barrier.
__device__ __forceinline__ void lds_load_global(const AggParams& p, const LdsTable& t)
{
  const int words = p.stride / 8;
  for (int64_t i = threadIdx.x; i <= p.capacity; i += kLdsBlock) {
    const uint64_t* g = reinterpret_cast<const uint64_t*>(p.table + i * (int64_t)p.stride);
    const uint64_t hdr = g[0];
    RowKey rk;
    int64_t rep = 0;
    if (i == p.capacity) {
      if (!(p.key_mode == KEY_INT && hdr == 1ull)) continue;
      rk.key = kEmptyIntKey;
      rk.special = true;
      rk.hash = 0;
    } else {
      if (hdr == (p.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef)) continue;
      rk.special = false;
      rk.key = hdr;
      if (p.key_mode == KEY_INT) {
        const KeyCol& k = p.keys[0];
        const uint64_t raw = k.width == 8 ? hdr : (hdr & ((1ull << (8 * k.width)) - 1ull));
        rk.hash = hash_combine(0u, murmur3_bits(normalize_float_bits(raw, k.vclass), k.width, 0));
      } else {
        rep = (int64_t)(hdr & 0xffffffffull);
        rk.hash = (uint32_t)(hdr >> 32);
      }
    }
    uint8_t* rec = lds_find_or_insert(t, p, rk, rk.hash, rep);
    if (!rec) continue;
    uint64_t* w = reinterpret_cast<uint64_t*>(rec);
    for (int j = 1; j < words; ++j) w[j] = g[j];
  }
  __syncthreads();
}

// Folds the second-pass fields of every group of the LDS table into its
// record of the global table (direct mode).
__device__ __forceinline__ void lds_fold_pass2(const AggParams& p, const Pass2Ops* __restrict__ q, const LdsTable& t)
{
  for (int i = threadIdx.x; i <= t.slots; i += kLdsBlock) {
    if (!lds_occupied(t, i, p.key_mode)) continue;
    const uint8_t* rec = lds_record(t, i);
    const uint64_t hdr = *reinterpret_cast<const uint64_t*>(rec);
    RowKey rk;
    int64_t rep = 0;
    if (p.key_mode == KEY_INT) {
      const KeyCol& k = p.keys[0];
      rk.special = i == t.slots;
      rk.key = rk.special ? kEmptyIntKey : hdr;
      const uint64_t raw = k.width == 8 ? rk.key : (rk.key & ((1ull << (8 * k.width)) - 1ull));
      rk.hash = hash_combine(0u, murmur3_bits(normalize_float_bits(raw, k.vclass), k.width, 0));
    } else {
      rep = (int64_t)(hdr & 0xffffffffull);
      rk.special = false;
      rk.key = hdr;
      rk.hash = (uint32_t)(hdr >> 32);
    }
    uint8_t* g = global_find(p, rk, rep);
    if (!g) continue;
    for (int o = 0; o < q->num_ops; ++o) {
      const Pass2Op& op = q->ops[o];
      if (op.kind == OP_M2) {
        const double m2 = *reinterpret_cast<const double*>(rec + op.off);
        if (m2 != 0.0) atom_add(reinterpret_cast<double*>(g + op.off), m2);
      } else {
        const uint32_t idx = *reinterpret_cast<const uint32_t*>(rec + op.off);
        if (idx != 0xffffffffu) atom_min(reinterpret_cast<uint32_t*>(g + op.off), idx);
      }
    }
  }
}

}  // namespace mgb
