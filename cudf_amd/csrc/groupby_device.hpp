// Device routines shared by the groupby kernels: row access, row hashing and
// equality, the global record table, the LDS record table, and finalisation
// of a record into output columns.
#pragma once

#include "groupby_kernels.hpp"

namespace mgb {

// ------------------------------------------------------------- row access

__device__ __forceinline__ uint32_t load_flags(const uint8_t* flags, int flag_bytes, int flag_stride, int64_t row)
{
  return (uint32_t)load_at(flags, flag_bytes, flag_stride, row);
}

__device__ __forceinline__ bool value_valid(const AggParams& p, const ValCol& v, int64_t row, uint32_t flags)
{
  if (v.mask) return bit_valid(v.mask, v.bit_off + row);
  if (v.flag_bit >= 0) return (flags >> v.flag_bit) & 1u;
  return true;
}

__device__ __forceinline__ bool key_row_valid(const AggParams& p, int64_t row)
{
  if (!p.skip_null_keys) return true;
  for (int c = 0; c < p.num_keys; ++c)
    if (!bit_valid(p.keys[c].mask, p.keys[c].bit_off + row)) return false;
  return true;
}

__device__ __forceinline__ bool rows_equal(const KeyCol* keys, int num_keys, bool compare_nulls, int64_t a, int64_t b)
{
  for (int c = 0; c < num_keys; ++c) {
    const KeyCol& k = keys[c];
    if (compare_nulls) {
      bool va = bit_valid(k.mask, k.bit_off + a), vb = bit_valid(k.mask, k.bit_off + b);
      if (va != vb) return false;
      if (!va) continue;
    }
    uint64_t x = load_at(k.data, k.width, k.stride, a), y = load_at(k.data, k.width, k.stride, b);
    if (k.vclass == VC_F32) {
      float fx = __uint_as_float((uint32_t)x), fy = __uint_as_float((uint32_t)y);
      if (!((isnan(fx) && isnan(fy)) || fx == fy)) return false;
    } else if (k.vclass == VC_F64) {
      double fx = __longlong_as_double((long long)x), fy = __longlong_as_double((long long)y);
      if (!((isnan(fx) && isnan(fy)) || fx == fy)) return false;
    } else if (k.vclass == VC_BOOL) {
      if ((x != 0) != (y != 0)) return false;
    } else if (x != y) {
      return false;
    }
  }
  return true;
}

// cudf's row hash: hash_combine fold of per-column MurmurHash3_x86_32 from
// seed 0, nulls hashing to 0xFFFFFFFF (row_operators.cuh:1799-1883).
__device__ __forceinline__ uint32_t row_hash(const KeyCol* keys, int num_keys, int64_t row, bool check_nulls)
{
  uint32_t h = 0;
  for (int c = 0; c < num_keys; ++c) {
    const KeyCol& k = keys[c];
    uint32_t e;
    if (check_nulls && !bit_valid(k.mask, k.bit_off + row)) {
      e = 0xffffffffu;
    } else {
      e = murmur3_bits(normalize_float_bits(load_at(k.data, k.width, k.stride, row), k.vclass), k.width, 0);
    }
    h = hash_combine(h, e);
  }
  return h;
}

// The group identity of a row: the 64-bit header value stored in its record
// and its row hash. For KEY_INT the header is the sign-extended key.
struct RowKey {
  uint64_t key;
  uint32_t hash;
  bool special;   // the 8-byte key equal to the empty header (kEmptyIntKey)
};

__device__ __forceinline__ RowKey row_key(const AggParams& p, int64_t row)
{
  RowKey r;
  if (p.key_mode == KEY_INT) {
    const KeyCol& k = p.keys[0];
    uint64_t raw = load_at(k.data, k.width, k.stride, row);
    r.key = (uint64_t)widen_int(raw, k.width, k.vclass);
    r.hash = hash_combine(0u, murmur3_bits(normalize_float_bits(raw, k.vclass), k.width, 0));
    r.special = r.key == kEmptyIntKey;
  } else {
    r.hash = row_hash(p.keys, p.num_keys, row, p.compare_key_nulls != 0);
    r.key = ((uint64_t)r.hash << 32) | (uint64_t)row;
    r.special = false;
  }
  return r;
}

__device__ __forceinline__ int64_t reduce_range(uint32_t h, int64_t n) { return (int64_t)(((uint64_t)h * (uint64_t)n) >> 32); }

// ------------------------------------------------------------- global table

__device__ __forceinline__ void flag_overflow(unsigned long long* counters) { atom_max(&counters[1], 1ull); }

// n inserts counted apart (in LDS) added to the table's group counter; the
// overflow flag set if they took it to the budget.
__device__ __forceinline__ void add_inserted(const AggParams& p, uint32_t n)
{
  if (n == 0) return;
  const unsigned long long total = atomicAdd(&p.counters[0], (unsigned long long)n) + n;
  if (total >= (unsigned long long)p.max_groups) flag_overflow(p.counters);
}

// Record of the group of (rk, row) in the global table, inserting it if
// needed; nullptr when the table is over its group budget (overflow flag set).
// `entry_check`: test the overflow flag first (a caller that probes every row
// tests it itself, every few rows: one flag read by every row of a launch
// queues on a single L2 channel). `local_count`: the caller checked the
// budget for all of its inserts beforehand; inserts are counted there (LDS)
// instead of in the table's group counter, which the caller adds them to
// once (a counter and flag taken by every insert of a launch are one L2
// channel's traffic, and stall every access that maps there).
// `budget_read`: with local_count, still test the budget (the counter as last
// flushed) before each insert.
__device__ __forceinline__ uint8_t* global_find_or_insert(const AggParams& p, const RowKey& rk, int64_t row,
                                                         bool entry_check = true, uint32_t* local_count = nullptr,
                                                         bool budget_read = false)
{
  if (rk.special) {
    uint8_t* rec = p.table + p.capacity * (int64_t)p.stride;
    unsigned long long* hp = reinterpret_cast<unsigned long long*>(rec);
    if (*hp == 0ull && atom_cas(hp, 0ull, 1ull) == 0ull) {
      if (local_count) atomicAdd(local_count, 1u);
      else atom_add(&p.counters[0], 1ull);
    }
    return rec;
  }
  // once the table has overflowed the call is repeated with a larger one:
  // stop work on this attempt at once
  if (entry_check && __hip_atomic_load(&p.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return nullptr;
  const uint64_t empty = p.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef;
  int64_t idx = reduce_range(rk.hash, p.capacity);
  for (int64_t probe = 0; probe < p.capacity; ++probe) {
    if ((probe & 63) == 63 &&
        __hip_atomic_load(&p.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return nullptr;
    uint8_t* rec = p.table + idx * (int64_t)p.stride;
    unsigned long long* hp = reinterpret_cast<unsigned long long*>(rec);
    uint64_t cur = *hp;  // a stale read can only be EMPTY; the CAS below resolves it
    for (;;) {
      if (cur == empty) {
        if ((!local_count || budget_read) &&
            __hip_atomic_load(&p.counters[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
              (unsigned long long)p.max_groups) {
          flag_overflow(p.counters);
          return nullptr;
        }
        uint64_t seen = atom_cas(hp, empty, rk.key);
        if (seen == empty) {
          if (local_count) atomicAdd(local_count, 1u);
          else atom_add(&p.counters[0], 1ull);
          return rec;
        }
        cur = seen;
        continue;  // re-examine the value that won the slot
      }
      if (p.key_mode == KEY_INT) {
        if (cur == rk.key) return rec;
      } else if ((uint32_t)(cur >> 32) == rk.hash &&
                 rows_equal(p.keys, p.num_keys, p.compare_key_nulls != 0, (int64_t)(cur & 0xffffffffull), row)) {
        return rec;
      }
      break;
    }
    idx = (idx + 1 == p.capacity) ? 0 : idx + 1;
  }
  flag_overflow(p.counters);
  return nullptr;
}

// Record of the row's group in a completed global table (every group of the
// rows inserted, so the probe meets the row's key before any empty slot);
// nullptr if the key is absent.
__device__ __forceinline__ uint8_t* global_find(const AggParams& p, const RowKey& rk, int64_t row)
{
  if (rk.special) return p.table + p.capacity * (int64_t)p.stride;
  const uint64_t empty = p.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef;
  int64_t idx = reduce_range(rk.hash, p.capacity);
  for (int64_t probe = 0; probe < p.capacity; ++probe) {
    uint8_t* rec = p.table + idx * (int64_t)p.stride;
    const uint64_t cur = *reinterpret_cast<const uint64_t*>(rec);
    if (cur == empty) return nullptr;
    if (p.key_mode == KEY_INT) {
      if (cur == rk.key) return rec;
    } else if ((uint32_t)(cur >> 32) == rk.hash &&
               rows_equal(p.keys, p.num_keys, p.compare_key_nulls != 0, (int64_t)(cur & 0xffffffffull), row)) {
      return rec;
    }
    idx = (idx + 1 == p.capacity) ? 0 : idx + 1;
  }
  return nullptr;
}

// All ops of one row on a record (global or LDS).
__device__ __forceinline__ void apply_row(const AggParams& p, uint8_t* rec, int64_t row)
{
  uint32_t flags = p.flags ? load_flags(p.flags, p.flag_bytes, p.flag_stride, row) : 0u;
  if (p.rows_op >= 0) apply_op(rec, OP_ROWS, p.ops[p.rows_op].off, 0, 4, VC_UNSIGNED);
  for (int c = 0; c < p.num_vals; ++c) {
    const ValCol& v = p.vals[c];
    if (v.op_begin == v.op_end && v.nulls_off < 0) continue;  // a column of sort-path aggregations only
    if (!value_valid(p, v, row, flags)) {
      if (v.nulls_off >= 0) atom_add(reinterpret_cast<uint32_t*>(rec + v.nulls_off), 1u);
      continue;
    }
    uint64_t raw = load_at(v.data, v.width, v.stride, row);
    for (int o = v.op_begin; o < v.op_end; ++o) apply_op(rec, p.ops[o].kind, p.ops[o].off, raw, v.width, v.vclass);
  }
}

// Folds one partial record (`src`) into `dst` with device atomics.
__device__ __forceinline__ void merge_record(const AggParams& p, uint8_t* dst, const uint8_t* src)
{
  for (int o = 0; o < p.num_ops; ++o) {
    const int kind = p.ops[o].kind, off = p.ops[o].off;
    const uint8_t* s = src + off;
    uint8_t* d = dst + off;
    switch (kind) {
      case OP_ROWS: case OP_COUNT: case OP_NULLS: {
        uint32_t v = *reinterpret_cast<const uint32_t*>(s);
        if (v) atom_add(reinterpret_cast<uint32_t*>(d), v);
        break;
      }
      case OP_SUM_I64: case OP_SUMSQ_I64:
        atom_add(reinterpret_cast<unsigned long long*>(d), *reinterpret_cast<const unsigned long long*>(s));
        break;
      case OP_SUM_F64: case OP_SUMSQ_F64:
        atom_add(reinterpret_cast<double*>(d), *reinterpret_cast<const double*>(s));
        break;
      case OP_PROD_I64: case OP_PROD_F64: {
        unsigned long long* dp = reinterpret_cast<unsigned long long*>(d);
        unsigned long long x = *reinterpret_cast<const unsigned long long*>(s);
        if (x == op_identity(kind)) break;
        // a float32 column's partial products are float32 values held in a
        // double: they multiply at float32 precision, as each row's update
        // does (apply_op, lds_op_rows)
        bool f32 = false;
        for (int c = 0; c < p.num_vals; ++c)
          f32 |= p.vals[c].vclass == VC_F32 && o >= p.vals[c].op_begin && o < p.vals[c].op_end;
        unsigned long long old = __hip_atomic_load(dp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (;;) {
          unsigned long long nxt;
          if (kind == OP_PROD_I64) {
            nxt = old * x;
          } else if (f32) {
            nxt = (unsigned long long)__double_as_longlong(
                (double)((float)__longlong_as_double((long long)old) * (float)__longlong_as_double((long long)x)));
          } else {
            nxt = (unsigned long long)__double_as_longlong(__longlong_as_double((long long)old) *
                                                          __longlong_as_double((long long)x));
          }
          unsigned long long seen = atom_cas(dp, old, nxt);
          if (seen == old) break;
          old = seen;
        }
        break;
      }
      case OP_MIN_I32: atom_min(reinterpret_cast<int32_t*>(d), *reinterpret_cast<const int32_t*>(s)); break;
      case OP_MAX_I32: atom_max(reinterpret_cast<int32_t*>(d), *reinterpret_cast<const int32_t*>(s)); break;
      case OP_MIN_U32: case OP_MIN_F32:
        atom_min(reinterpret_cast<uint32_t*>(d), *reinterpret_cast<const uint32_t*>(s));
        break;
      case OP_MAX_U32: case OP_MAX_F32:
        atom_max(reinterpret_cast<uint32_t*>(d), *reinterpret_cast<const uint32_t*>(s));
        break;
      case OP_MIN_I64: atom_min(reinterpret_cast<long long*>(d), *reinterpret_cast<const long long*>(s)); break;
      case OP_MAX_I64: atom_max(reinterpret_cast<long long*>(d), *reinterpret_cast<const long long*>(s)); break;
      case OP_MIN_U64: case OP_MIN_F64:
        atom_min(reinterpret_cast<unsigned long long*>(d), *reinterpret_cast<const unsigned long long*>(s));
        break;
      case OP_MAX_U64: case OP_MAX_F64:
        atom_max(reinterpret_cast<unsigned long long*>(d), *reinterpret_cast<const unsigned long long*>(s));
        break;
      default: break;
    }
  }
}

// ------------------------------------------------------------- LDS table

// LDS layout of a work item: `slots` records, then one record for the 8-byte
// key equal to the empty header, then two counters.
struct LdsTable {
  uint8_t* rec;
  int32_t slots;
  int32_t stride;
  int32_t cap;
  int32_t* count;       // groups inserted
  int32_t* overflow;    // set when a new group did not fit
};

__device__ __forceinline__ uint8_t* lds_record(const LdsTable& t, int64_t i) { return t.rec + i * (int64_t)t.stride; }

// Record of the row's group in the LDS table, inserting it if needed; nullptr
// when the table is full (overflow set). `slot_hash` must be independent of
// the bits that chose the partition.
__device__ __forceinline__ uint8_t* lds_find_or_insert(const LdsTable& t, const AggParams& p, const RowKey& rk,
                                                       uint32_t slot_hash, int64_t row)
{
  if (rk.special) {
    uint8_t* rec = lds_record(t, t.slots);
    unsigned long long* hp = reinterpret_cast<unsigned long long*>(rec);
    unsigned long long expect = 0ull;
    if (__hip_atomic_compare_exchange_strong(hp, &expect, 1ull, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP))
      atomicAdd(t.count, 1);
    return rec;
  }
  const uint64_t empty = p.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef;
  int32_t idx = (int32_t)reduce_range(slot_hash, t.slots);
  for (int32_t probe = 0; probe < t.slots; ++probe) {
    uint8_t* rec = lds_record(t, idx);
    unsigned long long* hp = reinterpret_cast<unsigned long long*>(rec);
    uint64_t cur = __hip_atomic_load(hp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (;;) {
      if (cur == empty) {
        if (__hip_atomic_load(t.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= t.cap) {
          *t.overflow = 1;
          return nullptr;
        }
        unsigned long long expect = empty;
        if (__hip_atomic_compare_exchange_strong(hp, &expect, (unsigned long long)rk.key, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
          atomicAdd(t.count, 1);
          return rec;
        }
        cur = expect;
        continue;
      }
      if (p.key_mode == KEY_INT) {
        if (cur == rk.key) return rec;
      } else if ((uint32_t)(cur >> 32) == rk.hash &&
                 rows_equal(p.keys, p.num_keys, p.compare_key_nulls != 0, (int64_t)(cur & 0xffffffffull), row)) {
        return rec;
      }
      break;
    }
    idx = (idx + 1 == t.slots) ? 0 : idx + 1;
  }
  *t.overflow = 1;
  return nullptr;
}

// Record of the row's group in an LDS table that holds it (a probe without
// insert); nullptr when the key is not in the table.
__device__ __forceinline__ uint8_t* lds_find(const LdsTable& t, const AggParams& p, const RowKey& rk,
                                             uint32_t slot_hash, int64_t row)
{
  if (rk.special) {
    uint8_t* rec = lds_record(t, t.slots);
    return *reinterpret_cast<const uint64_t*>(rec) == 1ull ? rec : nullptr;
  }
  const uint64_t empty = p.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef;
  int32_t idx = (int32_t)reduce_range(slot_hash, t.slots);
  for (int32_t probe = 0; probe < t.slots; ++probe) {
    uint8_t* rec = lds_record(t, idx);
    const uint64_t cur = *reinterpret_cast<const uint64_t*>(rec);
    if (cur == empty) return nullptr;
    if (p.key_mode == KEY_INT) {
      if (cur == rk.key) return rec;
    } else if ((uint32_t)(cur >> 32) == rk.hash &&
               rows_equal(p.keys, p.num_keys, p.compare_key_nulls != 0, (int64_t)(cur & 0xffffffffull), row)) {
      return rec;
    }
    idx = idx + 1 == t.slots ? 0 : idx + 1;
  }
  return nullptr;
}

__device__ __forceinline__ bool lds_occupied(const LdsTable& t, int32_t i, int key_mode)
{
  uint64_t hdr = *reinterpret_cast<const uint64_t*>(lds_record(t, i));
  if (i == t.slots) return key_mode == KEY_INT && hdr == 1ull;
  return hdr != (key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef);
}

// ------------------------------------------------------------- batched LDS updates
// A thread takes N rows at a time (row0 + u * step) and updates their LDS
// records column by column: the dispatch on a column's width, class and ops
// is uniform across the wavefront and is paid once per column for all N rows
// instead of once per row and field.

constexpr uint32_t kRecNone = 0xffffffffu;   // row not aggregated (past the end, or a dropped null key)
constexpr uint32_t kRecSpill = 0xfffffffeu;  // the row's group did not fit the LDS table

template <typename T>
__device__ __forceinline__ void lds_add(T* p, T v)
{
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void lds_min(T* p, T v)
{
  __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void lds_max(T* p, T v)
{
  __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Whether a row's value of a second-pass op's column is valid (`flags`: the
// row's validity bits, partitioned records).
__device__ __forceinline__ bool pass2_valid(const Pass2Op& o, int64_t row, uint32_t flags)
{
  if (o.mask) return bit_valid(o.mask, o.bit_off + row);
  if (o.flag_bit >= 0) return (flags >> o.flag_bit) & 1u;
  return true;
}

// One second-pass op for a valid value (`raw`) of a row on its group's
// record, whose first-pass fields are final (kLds: an LDS record,
// workgroup-scope atomics; otherwise a global one); `src_row` is the row's
// index in the input.
//  OP_M2: adds (x - mean)^2 with mean = SUM / COUNT_VALID, the SUM of a
//    float32 column taken at float32 precision as its SUM result is
//    (reference: cpp/src/groupby/hash/multi_pass_kernels.cuh:76-90).
//  OP_ARGMIN_IDX / OP_ARGMAX_IDX: the smallest row index whose value is the
//    group's MIN / MAX (the reference's CAS loop in
//    cpp/include/cudf/detail/aggregation/aggregation.cuh:436-490 returns one
//    of the tied rows; this is the first of them).
template <bool kLds>
__device__ __forceinline__ void pass2_apply(const Pass2Op& o, uint8_t* rec, uint64_t raw, uint32_t src_row)
{
  if (o.kind == OP_M2) {
    uint32_t cnt = *reinterpret_cast<const uint32_t*>(rec + o.count_off);
    if (o.nulls_off >= 0) cnt -= *reinterpret_cast<const uint32_t*>(rec + o.nulls_off);
    double sum;
    if (o.ref_kind == OP_SUM_I64) {
      sum = (double)*reinterpret_cast<const int64_t*>(rec + o.ref_off);
    } else {
      sum = *reinterpret_cast<const double*>(rec + o.ref_off);
      if (o.vclass == VC_F32) sum = (double)(float)sum;
    }
    const double d = widen_f64(raw, o.width, o.vclass) - sum / (double)cnt;
    if constexpr (kLds) lds_add(reinterpret_cast<double*>(rec + o.off), d * d);
    else atom_add(reinterpret_cast<double*>(rec + o.off), d * d);
  } else {
    const uint64_t code = minmax_code(o.ref_kind, raw, o.width, o.vclass);
    const uint64_t best = op_field_bytes(o.ref_kind) == 8 ? *reinterpret_cast<const uint64_t*>(rec + o.ref_off)
                                                          : *reinterpret_cast<const uint32_t*>(rec + o.ref_off);
    if (code == best) {
      if constexpr (kLds) lds_min(reinterpret_cast<uint32_t*>(rec + o.off), src_row);
      else atom_min(reinterpret_cast<uint32_t*>(rec + o.off), src_row);
    }
  }
}

// Every second-pass op of one row on its group's record.
template <bool kLds>
__device__ __forceinline__ void pass2_update(const Pass2Ops* __restrict__ q, uint8_t* rec, int64_t row,
                                             uint32_t flags, uint32_t src_row)
{
  const int n = q->num_ops;
  for (int i = 0; i < n; ++i) {
    const Pass2Op& o = q->ops[i];
    if (!pass2_valid(o, row, flags)) continue;
    pass2_apply<kLds>(o, rec, load_at(o.data, o.width, o.stride, row), src_row);
  }
}

// Elements row0 + u * step (u < N, where present[u]) of a column, in the
// canonical form the ops take: integers widened to 64 bits (sign- or
// zero-extended, bool -> 0/1), floats as their raw bits. Absent rows read 0.
// Every lane issues its loads: an absent row loads row 0 (in bounds for any
// column with rows) and its value is replaced by 0 afterwards, so no load
// sits under a branch (a narrow load under a branch was waited for before
// the branch closed, serialising a thread's loads).
template <int N>
__device__ __forceinline__ void load_rows(const uint8_t* data, int width, int stride, int vclass, int64_t row0,
                                          int64_t step, const bool (&present)[N], uint64_t (&x)[N])
{
  const bool sgn = vclass == VC_SIGNED;
  const global_t<const uint8_t>* g = to_global(data);
  const global_t<const uint8_t>* a[N];
#pragma unroll
  for (int u = 0; u < N; ++u) a[u] = present[u] ? g + (row0 + u * step) * stride : g;
  switch (width) {
    case 1:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint8_t v = *a[u];
        const uint8_t b = present[u] ? v : (uint8_t)0;
        x[u] = sgn ? (uint64_t)(int64_t)(int8_t)b : vclass == VC_BOOL ? (uint64_t)(b != 0) : (uint64_t)b;
      }
      break;
    case 2:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint16_t v = *reinterpret_cast<const global_t<uint16_t>*>(a[u]);
        const uint16_t b = present[u] ? v : (uint16_t)0;
        x[u] = sgn ? (uint64_t)(int64_t)(int16_t)b : (uint64_t)b;
      }
      break;
    case 4:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint32_t v = *reinterpret_cast<const global_t<uint32_t>*>(a[u]);
        const uint32_t b = present[u] ? v : 0u;
        x[u] = sgn ? (uint64_t)(int64_t)(int32_t)b : (uint64_t)b;
      }
      break;
    default:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint64_t v = *reinterpret_cast<const global_t<uint64_t>*>(a[u]);
        x[u] = present[u] ? v : 0ull;
      }
      break;
  }
}

// Elements row0 + u * step of a column as raw bits, zero-extended (an
// absent row loads row 0: its bits are arbitrary). Nothing is computed from
// the loaded values, so the loads stay in flight until the bits are used;
// widen_raw gives their canonical form.
template <int N>
__device__ __forceinline__ void load_rows_raw(const uint8_t* data, int width, int stride, int64_t row0, int64_t step,
                                              const bool (&present)[N], uint64_t (&x)[N])
{
  const global_t<const uint8_t>* g = to_global(data);
  const global_t<const uint8_t>* a[N];
#pragma unroll
  for (int u = 0; u < N; ++u) a[u] = present[u] ? g + (row0 + u * step) * stride : g;
  switch (width) {
    case 1:
#pragma unroll
      for (int u = 0; u < N; ++u) x[u] = (uint64_t)*a[u];
      break;
    case 2:
#pragma unroll
      for (int u = 0; u < N; ++u) x[u] = (uint64_t)*reinterpret_cast<const global_t<uint16_t>*>(a[u]);
      break;
    case 4:
#pragma unroll
      for (int u = 0; u < N; ++u) x[u] = (uint64_t)*reinterpret_cast<const global_t<uint32_t>*>(a[u]);
      break;
    default:
#pragma unroll
      for (int u = 0; u < N; ++u) x[u] = *reinterpret_cast<const global_t<uint64_t>*>(a[u]);
      break;
  }
}

// The canonical form (see load_rows) of the low `width` bytes of raw.
__device__ __forceinline__ uint64_t widen_raw(uint64_t raw, int width, int vclass)
{
  const bool sgn = vclass == VC_SIGNED;
  switch (width) {
    case 1: {
      const uint8_t b = (uint8_t)raw;
      return sgn ? (uint64_t)(int64_t)(int8_t)b : vclass == VC_BOOL ? (uint64_t)(b != 0) : (uint64_t)b;
    }
    case 2: {
      const uint16_t b = (uint16_t)raw;
      return sgn ? (uint64_t)(int64_t)(int16_t)b : (uint64_t)b;
    }
    case 4: {
      const uint32_t b = (uint32_t)raw;
      return sgn ? (uint64_t)(int64_t)(int32_t)b : (uint64_t)b;
    }
    default:
      return raw;
  }
}

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

// One record of up to four 8-byte words, held in registers. A vector value
// and not an array: its words are picked by a uniform offset
// (extract_rows), which on an array becomes an indexed copy in scratch
// memory; on a vector it is a register select.
typedef uint64_t RecWords __attribute__((ext_vector_type(4)));

// Whole records row0 + u * step (u < N) of `words` (<= 4) 8-byte words each
// into rw[u]; words past the record read 0. Rows past the range (not
// present) load the last present row instead of branching around the load:
// `last` is a row known to be in range.
template <int N>
__device__ __forceinline__ void load_recs(const uint8_t* base, int words, int64_t row0, int64_t step,
                                          const bool (&present)[N], int64_t last, RecWords (&rw)[N])
{
  if (words == 4) {
    // 32-byte records of a 16-byte aligned array: two 16-byte loads each
    const global_t<const u64x2>* g = to_global(reinterpret_cast<const u64x2*>(base));
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const int64_t r = present[u] ? row0 + u * step : last;
      const u64x2 a = g[r * 2], b = g[r * 2 + 1];
      rw[u] = __builtin_shufflevector(a, b, 0, 1, 2, 3);
    }
  } else {
    const global_t<const uint64_t>* g = to_global(reinterpret_cast<const uint64_t*>(base));
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const global_t<const uint64_t>* r = g + (present[u] ? row0 + u * step : last) * words;
      const uint64_t w0 = r[0];
      const uint64_t w1 = words > 1 ? r[1] : 0ull;
      const uint64_t w2 = words > 2 ? r[2] : 0ull;
      rw[u] = RecWords{w0, w1, w2, 0ull};
    }
  }
}

// The field at byte `off` (naturally aligned, so inside one 8-byte word) of
// records rw[u], in canonical form. One dispatch on the width for all rows.
template <int N>
__device__ __forceinline__ void extract_rows(const RecWords (&rw)[N], int off, int width, int vclass, uint64_t (&x)[N])
{
  const int wi = (off >> 3) & 3, sh = (off & 7) * 8;
  uint64_t w[N];
#pragma unroll
  for (int u = 0; u < N; ++u) w[u] = rw[u][wi] >> sh;
  const bool sgn = vclass == VC_SIGNED;
  switch (width) {
    case 1:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint8_t b = (uint8_t)w[u];
        x[u] = sgn ? (uint64_t)(int64_t)(int8_t)b : vclass == VC_BOOL ? (uint64_t)(b != 0) : (uint64_t)b;
      }
      break;
    case 2:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint16_t b = (uint16_t)w[u];
        x[u] = sgn ? (uint64_t)(int64_t)(int16_t)b : (uint64_t)b;
      }
      break;
    case 4:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint32_t b = (uint32_t)w[u];
        x[u] = sgn ? (uint64_t)(int64_t)(int32_t)b : (uint64_t)b;
      }
      break;
    default:
#pragma unroll
      for (int u = 0; u < N; ++u) x[u] = w[u];
      break;
  }
}

template <int N>
__device__ __forceinline__ void copy_rows(const uint64_t (&from)[N], uint64_t (&to)[N])
{
#pragma unroll
  for (int u = 0; u < N; ++u) to[u] = from[u];
}

// Writes the low `width` bytes of x[u] at base + pos[u] * rec_bytes + off for
// the present rows.
template <int N>
__device__ __forceinline__ void store_rows(uint8_t* base, const uint32_t (&pos)[N], int rec_bytes, int off, int width,
                                           const uint64_t (&x)[N], const bool (&present)[N])
{
  switch (width) {
    case 1:
#pragma unroll
      for (int u = 0; u < N; ++u)
        if (present[u]) base[pos[u] * rec_bytes + off] = (uint8_t)x[u];
      break;
    case 2:
#pragma unroll
      for (int u = 0; u < N; ++u)
        if (present[u]) *reinterpret_cast<uint16_t*>(base + pos[u] * rec_bytes + off) = (uint16_t)x[u];
      break;
    case 4:
#pragma unroll
      for (int u = 0; u < N; ++u)
        if (present[u]) *reinterpret_cast<uint32_t*>(base + pos[u] * rec_bytes + off) = (uint32_t)x[u];
      break;
    default:
#pragma unroll
      for (int u = 0; u < N; ++u)
        if (present[u]) *reinterpret_cast<uint64_t*>(base + pos[u] * rec_bytes + off) = x[u];
      break;
  }
}

// Row hashes (row_hash / row_key) of N rows, one key column at a time; the
// raw bits of the first key column are kept in key0.
template <int N>
__device__ __forceinline__ void hash_rows(const AggParams& p, int64_t row0, int64_t step, const bool (&present)[N],
                                          uint32_t (&h)[N], uint64_t (&key0)[N])
{
  const bool check_nulls = p.key_mode != KEY_INT && p.compare_key_nulls != 0;
#pragma unroll
  for (int u = 0; u < N; ++u) h[u] = 0u;
#pragma unroll 1
  for (int c = 0; c < p.num_keys; ++c) {
    const KeyCol& k = p.keys[c];
    uint64_t raw[N];
    load_rows(k.data, k.width, k.stride, VC_UNSIGNED, row0, step, present, raw);
#pragma unroll
    for (int u = 0; u < N; ++u) {
      uint32_t e;
      if (check_nulls && present[u] && !bit_valid(k.mask, k.bit_off + row0 + u * step)) e = 0xffffffffu;
      else e = murmur3_bits(normalize_float_bits(raw[u], k.vclass), k.width, 0);
      h[u] = hash_combine(h[u], e);
    }
    if (c == 0) {
#pragma unroll
      for (int u = 0; u < N; ++u) key0[u] = raw[u];
    }
  }
}

__device__ __forceinline__ double canon_f64(uint64_t x, int vclass)
{
  return vclass == VC_F32 ? (double)__uint_as_float((uint32_t)x) : __longlong_as_double((long long)x);
}

// One accumulator op on the LDS records of N rows (ok[u]: the row has a
// record and a valid value). x[u] is the canonical value (see load_rows).
// Rows that are not ok update the thread's dummy word `dum` (LDS byte
// offset, 8-byte aligned, never read) instead of branching around the update.
template <int N>
__device__ __forceinline__ void lds_op_rows(uint8_t* lds, const uint32_t (&loff)[N], const uint64_t (&x)[N],
                                            const bool (&ok)[N], int kind, int off, int vclass, uint32_t dum)
{
  switch (kind) {
    case OP_ROWS: case OP_COUNT:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_add(reinterpret_cast<uint32_t*>(lds + (ok[u] ? loff[u] + off : dum)), 1u);
      break;
    case OP_SUM_I64:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_add(reinterpret_cast<unsigned long long*>(lds + (ok[u] ? loff[u] + off : dum)), (unsigned long long)x[u]);
      break;
    case OP_SUM_F64:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_add(reinterpret_cast<double*>(lds + (ok[u] ? loff[u] + off : dum)), canon_f64(x[u], vclass));
      break;
    case OP_SUMSQ_I64:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_add(reinterpret_cast<unsigned long long*>(lds + (ok[u] ? loff[u] + off : dum)), (unsigned long long)(x[u] * x[u]));
      break;
    case OP_SUMSQ_F64:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        double sq;
        if (vclass == VC_F32) {
          const float f = __uint_as_float((uint32_t)x[u]);
          sq = (double)(f * f);
        } else {
          const double d = __longlong_as_double((long long)x[u]);
          sq = d * d;
        }
        lds_add(reinterpret_cast<double*>(lds + (ok[u] ? loff[u] + off : dum)), sq);
      }
      break;
    case OP_PROD_I64: case OP_PROD_F64:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        if (!ok[u]) continue;
        unsigned long long* q = reinterpret_cast<unsigned long long*>(lds + loff[u] + off);
        unsigned long long old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        for (;;) {
          unsigned long long nxt;
          if (kind == OP_PROD_I64) {
            nxt = old * (unsigned long long)x[u];
          } else if (vclass == VC_F32) {
            nxt = (unsigned long long)__double_as_longlong(
                (double)((float)__longlong_as_double((long long)old) * __uint_as_float((uint32_t)x[u])));
          } else {
            nxt = (unsigned long long)__double_as_longlong(__longlong_as_double((long long)old) *
                                                          __longlong_as_double((long long)x[u]));
          }
          unsigned long long expect = old;
          if (__hip_atomic_compare_exchange_strong(q, &expect, nxt, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP))
            break;
          old = expect;
        }
      }
      break;
    case OP_MIN_I32:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_min(reinterpret_cast<int32_t*>(lds + (ok[u] ? loff[u] + off : dum)), (int32_t)x[u]);
      break;
    case OP_MAX_I32:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_max(reinterpret_cast<int32_t*>(lds + (ok[u] ? loff[u] + off : dum)), (int32_t)x[u]);
      break;
    case OP_MIN_U32:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_min(reinterpret_cast<uint32_t*>(lds + (ok[u] ? loff[u] + off : dum)), (uint32_t)x[u]);
      break;
    case OP_MAX_U32:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_max(reinterpret_cast<uint32_t*>(lds + (ok[u] ? loff[u] + off : dum)), (uint32_t)x[u]);
      break;
    case OP_MIN_I64:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_min(reinterpret_cast<long long*>(lds + (ok[u] ? loff[u] + off : dum)), (long long)x[u]);
      break;
    case OP_MAX_I64:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_max(reinterpret_cast<long long*>(lds + (ok[u] ? loff[u] + off : dum)), (long long)x[u]);
      break;
    case OP_MIN_U64:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_min(reinterpret_cast<unsigned long long*>(lds + (ok[u] ? loff[u] + off : dum)), (unsigned long long)x[u]);
      break;
    case OP_MAX_U64:
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_max(reinterpret_cast<unsigned long long*>(lds + (ok[u] ? loff[u] + off : dum)), (unsigned long long)x[u]);
      break;
    case OP_MIN_F32:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint32_t b = (uint32_t)x[u];
        const bool nan = (b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu);
        lds_min(reinterpret_cast<uint32_t*>(lds + (ok[u] ? loff[u] + off : dum)), nan ? kF32MinNaN : f32_code(b));
      }
      break;
    case OP_MAX_F32:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint32_t b = (uint32_t)x[u];
        const bool nan = (b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu);
        lds_max(reinterpret_cast<uint32_t*>(lds + (ok[u] ? loff[u] + off : dum)), nan ? kF32MaxNaN : f32_code(b));
      }
      break;
    case OP_MIN_F64:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const unsigned long long b = x[u];
        const bool nan = (b & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (b & 0x000fffffffffffffull);
        lds_min(reinterpret_cast<unsigned long long*>(lds + (ok[u] ? loff[u] + off : dum)), nan ? kF64MinNaN : f64_code(b));
      }
      break;
    case OP_MAX_F64:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const unsigned long long b = x[u];
        const bool nan = (b & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (b & 0x000fffffffffffffull);
        lds_max(reinterpret_cast<unsigned long long*>(lds + (ok[u] ? loff[u] + off : dum)), nan ? kF64MaxNaN : f64_code(b));
      }
      break;
    default: break;  // OP_NULLS: updated for null values by the caller
  }
}

// One value column's program for records, packed into 64 bits: LDS field
// offsets + 1 (0: none) of SUM (bits 0-9), MIN (10-19), MAX (20-29) and the
// null count (30-39), the MIN op kind (40-44), flag bit + 1 (45-49), value
// class (50-52), width (53-56) and the value's byte offset in the record
// (57-62).
__host__ __device__ inline uint64_t pack_program(const ValCol& v, int rec_off)
{
  auto f10 = [](int off) { return (uint64_t)(off + 1) & 1023u; };
  return f10(v.sum_off) | f10(v.min_off) << 10 | f10(v.max_off) << 20 | f10(v.nulls_off) << 30 |
         (uint64_t)(v.mm_kind & 31) << 40 | (uint64_t)((v.flag_bit + 1) & 31) << 45 | (uint64_t)(v.vclass & 7) << 50 |
         (uint64_t)(v.width & 15) << 53 | (uint64_t)(rec_off & 63) << 57;
}

// A packed program's fields.
struct Program {
  int sum_off, min_off, max_off, nulls_off, mm_kind, flag_bit, vclass, width, rec_off;
};
__device__ __forceinline__ Program unpack_program(uint64_t pg)
{
  Program g;
  g.sum_off = (int)(pg & 1023u) - 1;
  g.min_off = (int)((pg >> 10) & 1023u) - 1;
  g.max_off = (int)((pg >> 20) & 1023u) - 1;
  g.nulls_off = (int)((pg >> 30) & 1023u) - 1;
  const int mm = (int)((pg >> 40) & 31u);
  g.mm_kind = mm == 31 ? -1 : mm;
  g.flag_bit = (int)((pg >> 45) & 31u) - 1;
  g.vclass = (int)((pg >> 50) & 7u);
  g.width = (int)((pg >> 53) & 15u);
  g.rec_off = (int)((pg >> 57) & 63u);
  return g;
}

// SUM, MIN and MAX of one column (ValCol::sum_off ...) on the LDS records of
// N rows: one dispatch on the column's class for all three, the min/max key
// (order-preserving code for floats) computed once for both.
template <int N, typename V>
__device__ __forceinline__ void lds_column_program(uint8_t* lds, const V& v, const uint32_t (&loff)[N],
                                                   const uint64_t (&x)[N], const bool (&ok)[N], uint32_t dum)
{
  if (v.sum_off >= 0) {
    if (v.vclass == VC_F32 || v.vclass == VC_F64) {
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_add(reinterpret_cast<double*>(lds + (ok[u] ? loff[u] + v.sum_off : dum)), canon_f64(x[u], v.vclass));
    } else {
#pragma unroll
      for (int u = 0; u < N; ++u)
        lds_add(reinterpret_cast<unsigned long long*>(lds + (ok[u] ? loff[u] + v.sum_off : dum)),
                (unsigned long long)x[u]);
    }
  }
  if (v.mm_kind < 0) return;
  // the min/max key of every row, then the MIN and the MAX updates
  uint64_t k[N];
  switch (v.mm_kind) {
    case OP_MIN_F32:
#pragma unroll
      for (int u = 0; u < N; ++u) {
        const uint32_t b = (uint32_t)x[u];
        k[u] = f32_code(b);
      }
      break;
    case OP_MIN_F64:
#pragma unroll
      for (int u = 0; u < N; ++u) k[u] = f64_code(x[u]);
      break;
    default:
#pragma unroll
      for (int u = 0; u < N; ++u) k[u] = x[u];
      break;
  }
  for (int which = 0; which < 2; ++which) {
    const int off = which == 0 ? v.min_off : v.max_off;
    if (off < 0) continue;
    uint32_t a[N];
#pragma unroll
    for (int u = 0; u < N; ++u) a[u] = ok[u] ? loff[u] + off : dum;
    switch (v.mm_kind * 2 + which) {
      case OP_MIN_I32 * 2:
#pragma unroll
        for (int u = 0; u < N; ++u) lds_min(reinterpret_cast<int32_t*>(lds + a[u]), (int32_t)k[u]);
        break;
      case OP_MIN_I32 * 2 + 1:
#pragma unroll
        for (int u = 0; u < N; ++u) lds_max(reinterpret_cast<int32_t*>(lds + a[u]), (int32_t)k[u]);
        break;
      case OP_MIN_U32 * 2:
#pragma unroll
        for (int u = 0; u < N; ++u) lds_min(reinterpret_cast<uint32_t*>(lds + a[u]), (uint32_t)k[u]);
        break;
      case OP_MIN_U32 * 2 + 1:
#pragma unroll
        for (int u = 0; u < N; ++u) lds_max(reinterpret_cast<uint32_t*>(lds + a[u]), (uint32_t)k[u]);
        break;
      case OP_MIN_I64 * 2:
#pragma unroll
        for (int u = 0; u < N; ++u) lds_min(reinterpret_cast<long long*>(lds + a[u]), (long long)k[u]);
        break;
      case OP_MIN_I64 * 2 + 1:
#pragma unroll
        for (int u = 0; u < N; ++u) lds_max(reinterpret_cast<long long*>(lds + a[u]), (long long)k[u]);
        break;
      case OP_MIN_U64 * 2:
#pragma unroll
        for (int u = 0; u < N; ++u) lds_min(reinterpret_cast<unsigned long long*>(lds + a[u]), (unsigned long long)k[u]);
        break;
      case OP_MIN_U64 * 2 + 1:
#pragma unroll
        for (int u = 0; u < N; ++u) lds_max(reinterpret_cast<unsigned long long*>(lds + a[u]), (unsigned long long)k[u]);
        break;
      case OP_MIN_F32 * 2:
#pragma unroll
        for (int u = 0; u < N; ++u) {
          const uint32_t b = (uint32_t)x[u];
          const bool nan = (b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu);
          lds_min(reinterpret_cast<uint32_t*>(lds + a[u]), nan ? kF32MinNaN : (uint32_t)k[u]);
        }
        break;
      case OP_MIN_F32 * 2 + 1:
#pragma unroll
        for (int u = 0; u < N; ++u) {
          const uint32_t b = (uint32_t)x[u];
          const bool nan = (b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu);
          lds_max(reinterpret_cast<uint32_t*>(lds + a[u]), nan ? kF32MaxNaN : (uint32_t)k[u]);
        }
        break;
      case OP_MIN_F64 * 2:
#pragma unroll
        for (int u = 0; u < N; ++u) {
          const unsigned long long b = x[u];
          const bool nan = (b & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (b & 0x000fffffffffffffull);
          lds_min(reinterpret_cast<unsigned long long*>(lds + a[u]), nan ? kF64MinNaN : (unsigned long long)k[u]);
        }
        break;
      default:  // OP_MIN_F64 * 2 + 1
#pragma unroll
        for (int u = 0; u < N; ++u) {
          const unsigned long long b = x[u];
          const bool nan = (b & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (b & 0x000fffffffffffffull);
          lds_max(reinterpret_cast<unsigned long long*>(lds + a[u]), nan ? kF64MaxNaN : (unsigned long long)k[u]);
        }
        break;
    }
  }
}

// All ops of one value column on the LDS records of N rows. Validity comes
// from the column's null mask (input rows) or from bit `flag_bit` of the
// row's flags word (partitioned records).
template <int N>
__device__ __forceinline__ void lds_column_rows(const AggParams& p, uint8_t* lds, const ValCol& v,
                                                const uint32_t (&loff)[N], const uint64_t (&x)[N],
                                                const uint32_t (&flags)[N], int64_t row0, int64_t step, uint32_t dum)
{
  bool ok[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    ok[u] = loff[u] < kRecSpill;
    if (!ok[u]) continue;
    if (v.mask) ok[u] = bit_valid(v.mask, v.bit_off + row0 + u * step);
    else if (v.flag_bit >= 0) ok[u] = (flags[u] >> v.flag_bit) & 1u;
  }
  if (v.nulls_off >= 0) {
#pragma unroll
    for (int u = 0; u < N; ++u)
      lds_add(reinterpret_cast<uint32_t*>(lds + (loff[u] < kRecSpill && !ok[u] ? loff[u] + v.nulls_off : dum)), 1u);
  }
  if (!v.lds_generic) {
    lds_column_program(lds, v, loff, x, ok, dum);
    return;
  }
  // the column's op descriptors are read together, before any update: a
  // scalar load inside the op loop would wait for every LDS update issued
  // before it
  OpDesc cop[kMaxColOps];
  const OpDesc* ops = p.ops + v.op_begin;
#pragma unroll
  for (int o = 0; o < kMaxColOps; ++o) cop[o] = ops[o];
  const int nops = v.op_end - v.op_begin;
#pragma unroll
  for (int o = 0; o < kMaxColOps; ++o)
    if (o < nops) lds_op_rows(lds, loff, x, ok, cop[o].kind, cop[o].off, v.vclass, dum);
}

// ------------------------------------------------------------- finalisation

__device__ __forceinline__ void store_width(void* out, int64_t pos, int width, uint64_t bits)
{
  global_t<uint8_t>* g = to_global(static_cast<uint8_t*>(out));
  switch (width) {
    case 1: g[pos] = (uint8_t)bits; break;
    case 2: reinterpret_cast<global_t<uint16_t>*>(g)[pos] = (uint16_t)bits; break;
    case 4: reinterpret_cast<global_t<uint32_t>*>(g)[pos] = (uint32_t)bits; break;
    default: reinterpret_cast<global_t<uint64_t>*>(g)[pos] = bits; break;
  }
}

// Adds the active lanes' `flag`s to *ctr (one atomic per wave).
__device__ __forceinline__ void count_wave(unsigned long long* ctr, bool flag)
{
  const uint64_t m = __ballot(flag);
  if (m && (threadIdx.x & 63u) == (unsigned)__builtin_ctzll(__ballot(1))) atom_add(ctr, (unsigned long long)__popcll(m));
}

// Writes the requested results of one group's record at output row `pos`.
__device__ __forceinline__ void finalize_record(const uint8_t* rec, int64_t pos, const OutTable& ot)
{
  for (int o = 0; o < ot.num_outs; ++o) {
    const OutSpec& s = ot.outs[o];
    if (s.fin == FIN_NONE) continue;
    uint32_t cnt = *reinterpret_cast<const uint32_t*>(rec + s.count_off);
    if (s.nulls_off >= 0) cnt -= *reinterpret_cast<const uint32_t*>(rec + s.nulls_off);
    const uint8_t* f = rec + s.field_off;
    uint64_t bits = 0;
    bool ok = cnt != 0;
    switch (s.fin) {
      case FIN_COUNT: bits = cnt; break;
      case FIN_INT:
        bits = (s.out_width == 8) ? *reinterpret_cast<const uint64_t*>(f) : *reinterpret_cast<const uint32_t*>(f);
        break;
      case FIN_F64: bits = *reinterpret_cast<const uint64_t*>(f); break;
      case FIN_F64_TO_F32: bits = __float_as_uint((float)*reinterpret_cast<const double*>(f)); break;
      case FIN_F32_CODE: {
        uint32_t c = *reinterpret_cast<const uint32_t*>(f);
        bits = (c == kF32MinNaN || c == kF32MaxNaN) ? 0x7fc00000u : f32_decode(c);
        break;
      }
      case FIN_F64_CODE: {
        uint64_t c = *reinterpret_cast<const uint64_t*>(f);
        bits = (c == kF64MinNaN || c == kF64MaxNaN) ? 0x7ff8000000000000ull : f64_decode(c);
        break;
      }
      case FIN_MEAN_I64: {
        double m = cnt ? (double)*reinterpret_cast<const int64_t*>(f) / (double)cnt : 0.0;
        bits = (uint64_t)__double_as_longlong(m);
        break;
      }
      case FIN_MEAN_F64: {
        double m = cnt ? *reinterpret_cast<const double*>(f) / (double)cnt : 0.0;
        bits = (uint64_t)__double_as_longlong(m);
        break;
      }
      case FIN_MEAN_F32: {
        double m = cnt ? (double)(float)*reinterpret_cast<const double*>(f) / (double)cnt : 0.0;
        bits = (uint64_t)__double_as_longlong(m);
        break;
      }
      case FIN_MEAN_DURATION: {
        int64_t m = cnt ? *reinterpret_cast<const int64_t*>(f) / (int64_t)cnt : 0;
        bits = (uint64_t)m;
        break;
      }
      case FIN_VAR: case FIN_STD: {
        // valid when the group has more valid values than ddof
        // (multi_pass_kernels.cuh:80; STD is the square root, groupby.cu:339-347)
        ok = cnt != 0 && (int64_t)cnt > (int64_t)s.ddof;
        double r = ok ? *reinterpret_cast<const double*>(f) / (double)((int64_t)cnt - s.ddof) : 0.0;
        if (s.fin == FIN_STD) r = sqrt(r);
        bits = (uint64_t)__double_as_longlong(r);
        break;
      }
      case FIN_M2: bits = *reinterpret_cast<const uint64_t*>(f); break;
      case FIN_ARGIDX:
        // a group without valid values keeps the sentinel -1 (groupby.cu:231-240)
        bits = cnt ? *reinterpret_cast<const uint32_t*>(f) : 0xffffffffu;
        break;
      default: break;
    }
    store_width(s.out, pos, s.out_width, bits);
    if (s.valid) {
      to_global(s.valid)[pos] = ok;
      if (ot.nulls) count_wave(ot.nulls + o, !ok);
    }
  }
}

// Writes the key columns of a group whose header is `hdr` (`special`: the
// 8-byte key equal to the empty header) at output row `pos`.
__device__ __forceinline__ void finalize_keys(uint64_t hdr, bool special, int key_mode, const KeyCol* keys,
                                              int num_keys, int64_t pos, const OutTable& ot)
{
  if (key_mode == KEY_INT) {
    store_width(ot.key_out[0], pos, keys[0].width, special ? kEmptyIntKey : hdr);
    return;
  }
  const int64_t rep = (int64_t)(hdr & 0xffffffffull);
  for (int c = 0; c < num_keys; ++c) {
    const KeyCol& k = keys[c];
    store_width(ot.key_out[c], pos, k.width, load_at(k.data, k.width, k.stride, rep));
    if (ot.key_valid[c]) {
      const bool valid = bit_valid(k.mask, k.bit_off + rep);
      to_global(ot.key_valid[c])[pos] = valid ? 1 : 0;
      if (ot.nulls) count_wave(ot.nulls + ot.num_outs + c, !valid);
    }
  }
}

}  // namespace mgb
