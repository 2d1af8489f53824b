// MI355X hash groupby-aggregate: kernels and the C ABI driver.
//
// Design (replacing cpp/src/groupby/hash/groupby.cu:424-620 of libcudf):
//   * One open-addressing table of fixed-stride records in HBM. A record is an
//     8-byte header followed by the group's accumulators (array of structs),
//     so a row touches one record instead of one slot per flattened
//     aggregation column.
//   * The header is claimed with a single 64-bit CAS and is never rewritten.
//     It holds the key itself for a single integer-like key column, or
//     (row hash << 32 | representative row) otherwise; equality then gathers
//     the representative row's key elements from the (read-only) input. No
//     other field has to be published before a second row may use the group,
//     so there is no spin on a half-built slot and no agent-scope
//     release/acquire (which on gfx950 means an L2 writeback / invalidate).
//   * Accumulators start at their identities (one fill of the record array)
//     and are updated with device atomics; float MIN/MAX use order-preserving
//     unsigned codes so they are exact and order independent.
//   * A finalize pass compacts occupied records in index order (block counts,
//     a one-workgroup scan, then a ballot-based intra-block prefix) and writes
//     SoA result columns directly, including MEAN and the validity bytes that
//     a pack kernel turns into cudf bitmasks.
//   * Table capacity comes from a sampled distinct-count estimate (Chao1 over
//     a stratified sample); an overflow flag makes the driver rebuild with a
//     4x larger table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "groupby_kernels.hpp"
#include "mgb_groupby.h"

namespace mgb {

// =================================================================== kernels

__device__ __forceinline__ bool rows_equal(const AggParams& p, int64_t a, int64_t b)
{
  for (int c = 0; c < p.num_keys; ++c) {
    const KeyCol& k = p.keys[c];
    if (p.compare_key_nulls) {
      bool va = bit_valid(k.mask, k.bit_off + a), vb = bit_valid(k.mask, k.bit_off + b);
      if (va != vb) return false;
      if (!va) continue;
    }
    uint64_t x = load_raw(k.data, k.width, a), y = load_raw(k.data, k.width, b);
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

__device__ __forceinline__ uint32_t row_hash(const KeyCol* keys, int num_keys, int64_t row, bool check_nulls)
{
  uint32_t h = 0;  // DEFAULT_HASH_SEED
  for (int c = 0; c < num_keys; ++c) {
    const KeyCol& k = keys[c];
    uint32_t e;
    if (check_nulls && !bit_valid(k.mask, k.bit_off + row)) {
      e = 0xffffffffu;
    } else {
      uint64_t raw = normalize_float_bits(load_raw(k.data, k.width, row), k.vclass);
      e = murmur3_bits(raw, k.width, 0);
    }
    h = hash_combine(h, e);
  }
  return h;
}

__device__ __forceinline__ int64_t slot_of(uint32_t h, int64_t capacity)
{
  return (int64_t)(((uint64_t)h * (uint64_t)capacity) >> 32);
}

__device__ __forceinline__ void flag_overflow(unsigned long long* counters)
{
  atom_max(&counters[1], 1ull);
}

// Returns the record of `row`'s group, inserting the group if needed, or
// nullptr when the table is over its group budget (overflow flag set).
__device__ __forceinline__ uint8_t* find_or_insert(const AggParams& p, int64_t row)
{
  uint64_t key;
  uint32_t h;
  if (p.key_mode == KEY_INT) {
    const KeyCol& k = p.keys[0];
    uint64_t raw = load_raw(k.data, k.width, row);
    key = (uint64_t)widen_int(raw, k.width, k.vclass);
    if (key == kEmptyIntKey) {
      uint8_t* rec = p.table + p.capacity * (int64_t)p.stride;
      unsigned long long* hp = reinterpret_cast<unsigned long long*>(rec);
      if (*hp == 0ull && atom_cas(hp, 0ull, 1ull) == 0ull) atom_add(&p.counters[0], 1ull);
      return rec;
    }
    h = hash_combine(0u, murmur3_bits(normalize_float_bits(raw, k.vclass), k.width, 0));
  } else {
    h = row_hash(p.keys, p.num_keys, row, p.compare_key_nulls != 0);
    key = ((uint64_t)h << 32) | (uint64_t)row;
  }
  const uint64_t empty = p.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef;
  int64_t idx = slot_of(h, p.capacity);
  for (int64_t probe = 0; probe < p.capacity; ++probe) {
    uint8_t* rec = p.table + idx * (int64_t)p.stride;
    unsigned long long* hp = reinterpret_cast<unsigned long long*>(rec);
    uint64_t cur = *hp;  // a stale read can only be EMPTY; the CAS below resolves it
    for (;;) {
      if (cur == empty) {
        if (__hip_atomic_load(&p.counters[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
            (unsigned long long)p.max_groups) {
          flag_overflow(p.counters);
          return nullptr;
        }
        uint64_t seen = atom_cas(hp, empty, key);
        if (seen == empty) {
          atom_add(&p.counters[0], 1ull);
          return rec;
        }
        cur = seen;
        continue;  // re-examine the value that won the slot
      }
      if (p.key_mode == KEY_INT) {
        if (cur == key) return rec;
      } else if ((uint32_t)(cur >> 32) == h && rows_equal(p, (int64_t)(cur & 0xffffffffull), row)) {
        return rec;
      }
      break;
    }
    idx = (idx + 1 == p.capacity) ? 0 : idx + 1;
  }
  flag_overflow(p.counters);
  return nullptr;
}

__global__ __launch_bounds__(kBlock) void aggregate_kernel(const AggParams p)
{
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < p.num_rows; row += step) {
    if (p.skip_null_keys) {
      bool ok = true;
      for (int c = 0; c < p.num_keys; ++c) ok = ok && bit_valid(p.keys[c].mask, p.keys[c].bit_off + row);
      if (!ok) continue;
    }
    uint8_t* rec = find_or_insert(p, row);
    if (rec == nullptr) continue;
    if (p.rows_op >= 0) apply_op(rec, OP_ROWS, p.ops[p.rows_op].off, 0, 4, VC_UNSIGNED);
    for (int c = 0; c < p.num_vals; ++c) {
      const ValCol& v = p.vals[c];
      if (!bit_valid(v.mask, v.bit_off + row)) continue;
      uint64_t raw = load_raw(v.data, v.width, row);
      for (int o = v.op_begin; o < v.op_end; ++o) apply_op(rec, p.ops[o].kind, p.ops[o].off, raw, v.width, v.vclass);
    }
  }
}

// Writes `tmpl` (header + identities) into every record; record `capacity`
// gets `special_hdr` as header.
constexpr int kMaxRecordWords = 64;
struct FillParams {
  uint64_t* table;
  int64_t num_records;
  int32_t words;
  int32_t pad;
  uint64_t special_hdr;
  uint64_t tmpl[kMaxRecordWords];
};

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
}

__device__ __forceinline__ bool record_occupied(const uint8_t* table, int64_t r, int64_t capacity, int32_t stride,
                                                int32_t key_mode)
{
  if (r > capacity) return false;
  uint64_t hdr = *reinterpret_cast<const uint64_t*>(table + r * (int64_t)stride);
  if (r == capacity) return key_mode == KEY_INT && hdr == 1ull;
  return hdr != (key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef);
}

__global__ __launch_bounds__(kBlock) void count_occupied_kernel(const uint8_t* table, int64_t capacity,
                                                                int32_t stride, int32_t key_mode,
                                                                int64_t* block_counts)
{
  int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int c = __syncthreads_count(record_occupied(table, r, capacity, stride, key_mode));
  if (threadIdx.x == 0) block_counts[blockIdx.x] = c;
}

// Exclusive scan of n block counts in place by one workgroup; total -> *total.
__global__ __launch_bounds__(1024) void scan_blocks_kernel(int64_t* counts, int64_t n, unsigned long long* total)
{
  __shared__ int64_t buf[1024];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < n; base += 1024) {
    int64_t i = base + threadIdx.x;
    int64_t v = i < n ? counts[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      int64_t t = threadIdx.x >= (unsigned)d ? buf[threadIdx.x - d] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n) counts[i] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = (unsigned long long)carry;
}

__device__ __forceinline__ void store_width(void* out, int64_t pos, int width, uint64_t bits)
{
  switch (width) {
    case 1: reinterpret_cast<uint8_t*>(out)[pos] = (uint8_t)bits; break;
    case 2: reinterpret_cast<uint16_t*>(out)[pos] = (uint16_t)bits; break;
    case 4: reinterpret_cast<uint32_t*>(out)[pos] = (uint32_t)bits; break;
    default: reinterpret_cast<uint64_t*>(out)[pos] = bits; break;
  }
}

__global__ __launch_bounds__(kBlock) void finalize_kernel(const FinParams p)
{
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

  const uint8_t* rec = p.table + r * (int64_t)p.stride;
  const uint64_t hdr = *reinterpret_cast<const uint64_t*>(rec);
  if (p.key_mode == KEY_INT) {
    uint64_t k = (r == p.capacity) ? kEmptyIntKey : hdr;
    store_width(p.key_out[0], pos, p.keys[0].width, k);
  } else {
    const int64_t rep = (int64_t)(hdr & 0xffffffffull);
    for (int c = 0; c < p.num_keys; ++c) {
      const KeyCol& k = p.keys[c];
      store_width(p.key_out[c], pos, k.width, load_raw(k.data, k.width, rep));
      if (p.key_valid[c]) p.key_valid[c][pos] = bit_valid(k.mask, k.bit_off + rep) ? 1 : 0;
    }
  }
  for (int o = 0; o < p.num_outs; ++o) {
    const OutSpec& s = p.outs[o];
    const uint32_t cnt = *reinterpret_cast<const uint32_t*>(rec + s.count_off);
    const uint8_t* f = rec + s.field_off;
    uint64_t bits = 0;
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
      default: break;
    }
    store_width(s.out, pos, s.out_width, bits);
    if (s.valid) s.valid[pos] = cnt != 0;
  }
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

// ---- cardinality estimate: stratified sample into a small table

struct SampleParams {
  AggParams agg;              // key description; table/capacity unused
  unsigned long long* hdr;    // sample table headers
  uint32_t* cnt;              // occurrences per slot; cnt[scap] = the special key
  int64_t scap;
  int64_t samples;
  unsigned long long* stats;  // [0] distinct, [1] f1, [2] f2, [3] sampled rows
};

__device__ __forceinline__ bool sample_rows_equal(const AggParams& p, int64_t a, int64_t b) { return rows_equal(p, a, b); }

__global__ __launch_bounds__(kBlock) void sample_insert_kernel(const SampleParams s)
{
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= s.samples) return;
  const AggParams& p = s.agg;
  const int64_t span = p.num_rows / s.samples;
  int64_t row = j * span + (span > 1 ? (int64_t)(fmix32((uint32_t)j * 0x9e3779b9u + 1u) % (uint32_t)span) : 0);
  if (row >= p.num_rows) return;
  if (p.skip_null_keys) {
    for (int c = 0; c < p.num_keys; ++c)
      if (!bit_valid(p.keys[c].mask, p.keys[c].bit_off + row)) return;
  }
  atom_add(&s.stats[3], 1ull);
  uint64_t key;
  uint32_t h;
  if (p.key_mode == KEY_INT) {
    const KeyCol& k = p.keys[0];
    uint64_t raw = load_raw(k.data, k.width, row);
    key = (uint64_t)widen_int(raw, k.width, k.vclass);
    if (key == kEmptyIntKey) {
      atom_add(&s.cnt[s.scap], 1u);
      return;
    }
    h = hash_combine(0u, murmur3_bits(normalize_float_bits(raw, k.vclass), k.width, 0));
  } else {
    h = row_hash(p.keys, p.num_keys, row, p.compare_key_nulls != 0);
    key = ((uint64_t)h << 32) | (uint64_t)row;
  }
  const uint64_t empty = p.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef;
  int64_t idx = slot_of(h, s.scap);
  for (int64_t probe = 0; probe < s.scap; ++probe) {
    uint64_t cur = s.hdr[idx];
    for (;;) {
      if (cur == empty) {
        uint64_t seen = atom_cas(&s.hdr[idx], empty, key);
        if (seen == empty) {
          atom_add(&s.cnt[idx], 1u);
          return;
        }
        cur = seen;
        continue;
      }
      bool eq = p.key_mode == KEY_INT
                  ? cur == key
                  : ((uint32_t)(cur >> 32) == h && sample_rows_equal(p, (int64_t)(cur & 0xffffffffull), row));
      if (eq) {
        atom_add(&s.cnt[idx], 1u);
        return;
      }
      break;
    }
    idx = (idx + 1 == s.scap) ? 0 : idx + 1;
  }
}

__global__ __launch_bounds__(kBlock) void sample_stats_kernel(const uint32_t* cnt, int64_t n, unsigned long long* stats)
{
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  uint32_t c = i < n ? cnt[i] : 0u;
  int d = __syncthreads_count(c > 0);
  int f1 = __syncthreads_count(c == 1);
  int f2 = __syncthreads_count(c == 2);
  if (threadIdx.x == 0) {
    if (d) atom_add(&stats[0], (unsigned long long)d);
    if (f1) atom_add(&stats[1], (unsigned long long)f1);
    if (f2) atom_add(&stats[2], (unsigned long long)f2);
  }
}

__global__ __launch_bounds__(kBlock) void fill_u64_kernel(unsigned long long* p, int64_t n, unsigned long long v)
{
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) p[i] = v;
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
      e = murmur3_bits(normalize_float_bits(load_raw(k.data, k.width, row), k.vclass), k.width, p.seed);
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
    uint64_t raw = load_raw(k.data, k.width, row);
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

namespace {

thread_local std::string g_last_error;

mgb_status fail(mgb_status s, std::string msg)
{
  g_last_error = std::move(msg);
  return s;
}

#define MGB_HIP(call)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return fail(MGB_ERROR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

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

}  // namespace

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
    case MGB_AGG_M2: case MGB_AGG_VARIANCE: case MGB_AGG_STD:
      if (is_timestamp_t(t)) return -1;
      return MGB_TYPE_FLOAT64;
    default: return -1;
  }
}

namespace {

// ------------------------------------------------------------ profiling

struct ProfRecord {
  std::string kernel;
  hipEvent_t start, stop;
};
std::mutex g_prof_mu;
bool g_prof_enabled = false;
std::vector<ProfRecord> g_prof;

// Brackets one kernel launch with events when profiling is on.
struct ProfScope {
  hipEvent_t start = nullptr, stop = nullptr;
  hipStream_t stream;
  const char* name;
  ProfScope(const char* kernel, hipStream_t s) : stream{s}, name{kernel}
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_enabled) return;
    if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) {
      start = stop = nullptr;
      return;
    }
    (void)hipEventRecord(start, stream);
  }
  ~ProfScope()
  {
    if (!start) return;
    (void)hipEventRecord(stop, stream);
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.push_back(ProfRecord{name, start, stop});
  }
};

// ------------------------------------------------------------ workspace

struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t need, hipStream_t stream)
  {
    if (need <= bytes) return hipSuccess;
    if (ptr) {
      hipError_t e = hipStreamSynchronize(stream);
      if (e != hipSuccess) return e;
      (void)hipFree(ptr);
      ptr = nullptr;
      bytes = 0;
    }
    size_t want = std::max(need, bytes + bytes / 4);
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
      ptr = nullptr;
      return e;
    }
    bytes = want;
    return hipSuccess;
  }
};

struct Workspace {
  std::mutex mu;
  DeviceBuffer main;   // table, counters, block counts, sample table
  DeviceBuffer outs;   // validity bytes of the outputs
  unsigned long long* pinned = nullptr;  // host readback area
  ~Workspace()
  {
    if (main.ptr) (void)hipFree(main.ptr);
    if (outs.ptr) (void)hipFree(outs.ptr);
    if (pinned) (void)hipHostFree(pinned);
  }
};

std::mutex g_ws_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<Workspace>> g_ws;

Workspace* workspace_for(hipStream_t s)
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto& w = g_ws[{dev, s}];
  if (!w) w.reset(new Workspace());
  return w.get();
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

constexpr int kMaxPinned = 1024;  // host readback slots (counters, null counts)

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

int64_t bitmask_bytes(int64_t n) { return (int64_t)align_up((size_t)((n + 31) / 32) * 4, 64); }

// ------------------------------------------------------------ plan

struct ValKey {
  const void* data;
  const uint32_t* mask;
  int64_t offset, size;
  int32_t type;
  bool operator<(const ValKey& o) const
  {
    return std::tie(data, mask, offset, size, type) < std::tie(o.data, o.mask, o.offset, o.size, o.type);
  }
};

struct PlanCol {
  ValCol col;               // op_begin/op_end filled when batching
  int type;
  bool nullable;
  std::vector<std::pair<int, int>> ops;  // (op kind, byte offset)
  int count_off = -1;       // OP_COUNT field, or the rows field when not nullable
  int op_offset(int kind) const
  {
    for (auto& o : ops)
      if (o.first == kind) return o.second;
    return -1;
  }
};

struct Plan {
  std::vector<KeyCol> keys;
  int key_mode = KEY_INT;
  bool skip_null_keys = false;
  bool compare_key_nulls = false;
  bool key_has_nulls = false;
  std::vector<PlanCol> cols;
  int rows_off = -1;
  int stride = 8;
  struct Out {
    int col;        // index into cols
    int kind;       // aggregation kind
    int out_type;
    OutSpec spec;
  };
  std::vector<Out> outs;
};

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
        case MGB_AGG_COUNT_ALL: case MGB_AGG_SUM_OF_SQUARES: case MGB_AGG_MEAN: break;
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
      pc.col.vclass = value_class(v.type_id);
      plan.cols.push_back(pc);
    } else {
      ci = it->second;
    }
    for (int a = 0; a < reqs[r].num_aggs; ++a) wants.push_back({ci, reqs[r].aggs[a].kind, reqs[r].aggs[a].ddof});
  }

  // ops per column (deduplicated by kind); the counts are added below
  bool need_rows = false;
  auto add_op = [](PlanCol& pc, int kind) {
    for (auto& o : pc.ops)
      if (o.first == kind) return;
    pc.ops.push_back({kind, -1});
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
      default: break;
    }
    if (wt.kind != MGB_AGG_COUNT_ALL) {
      if (pc.nullable) add_op(pc, OP_COUNT);
      else need_rows = true;
    }
  }

  // record layout: 8-byte header, 8-byte fields, then 4-byte fields
  int off = 8;
  for (auto& pc : plan.cols)
    for (auto& o : pc.ops)
      if (op_field_bytes(o.first) == 8) { o.second = off; off += 8; }
  if (need_rows) { plan.rows_off = off; off += 4; }
  for (auto& pc : plan.cols)
    for (auto& o : pc.ops)
      if (op_field_bytes(o.first) == 4) { o.second = off; off += 4; }
  plan.stride = (int)align_up((size_t)off, 8);
  if (plan.stride > kMaxRecordWords * 8)
    return fail(MGB_ERROR_UNSUPPORTED, "too many distinct aggregations for one groupby call");
  for (auto& pc : plan.cols) pc.count_off = pc.nullable ? pc.op_offset(OP_COUNT) : plan.rows_off;

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
    s.out_width = type_width(o.out_type);
    s.valid = reinterpret_cast<uint8_t*>(1);  // nullable marker; real pointer set at finalize
    switch (wt.kind) {
      case MGB_AGG_COUNT_VALID:
        s.fin = FIN_COUNT;
        s.valid = nullptr;
        break;
      case MGB_AGG_COUNT_ALL:
        s.fin = FIN_COUNT;
        s.count_off = plan.rows_off;
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
  for (auto& pc : plan.cols)
    for (auto& o : pc.ops) {
      uint64_t id = op_identity(o.first);
      std::memcpy(bytes.data() + o.second, &id, (size_t)op_field_bytes(o.first));
    }
  std::memcpy(words, bytes.data(), (size_t)plan.stride);
}

// Chao1 distinct-count estimate from a sample's frequency counts.
int64_t estimate_groups(uint64_t d, uint64_t f1, uint64_t f2, uint64_t sampled, int64_t rows)
{
  if (sampled == 0) return 1;
  if ((int64_t)sampled >= rows) return (int64_t)d;
  double est = (double)d + (double)f1 * (double)(f1 > 0 ? f1 - 1 : 0) / (2.0 * (double)(f2 + 1));
  est = std::min(est, (double)rows);
  return std::max<int64_t>((int64_t)d, (int64_t)std::ceil(est));
}

struct Allocs {
  const mgb_allocator* alloc;
  hipStream_t stream;
  std::vector<std::pair<void*, size_t>> live;
  void* get(size_t bytes)
  {
    void* p = alloc->allocate(alloc->ctx, bytes, (void*)stream);
    if (p) live.push_back({p, bytes});
    return p;
  }
  void release_all()
  {
    for (auto& l : live) alloc->deallocate(alloc->ctx, l.first, l.second, (void*)stream);
    live.clear();
  }
};

}  // namespace

// Runs the full groupby for a validated plan. Exposed inside the library so
// the batched executor can reuse it.
mgb_status run_groupby(const Plan& plan, int64_t num_rows, mgb_out_column* out_keys, mgb_out_column* out_results,
                       int64_t* out_num_groups, hipStream_t stream, const mgb_allocator* alloc);

mgb_status run_groupby(const Plan& plan, int64_t R, mgb_out_column* out_keys, mgb_out_column* out_results,
                       int64_t* out_num_groups, hipStream_t stream, const mgb_allocator* alloc)
{
  Workspace* ws = workspace_for(stream);
  std::lock_guard<std::mutex> lk(ws->mu);
  if (!ws->pinned)
    MGB_HIP(hipHostMalloc((void**)&ws->pinned, kMaxPinned * sizeof(unsigned long long), hipHostMallocDefault));
  const int cus = device_cu_count();

  AggParams base{};
  base.num_rows = R;
  base.key_mode = plan.key_mode;
  base.num_keys = (int)plan.keys.size();
  base.skip_null_keys = plan.skip_null_keys;
  base.compare_key_nulls = plan.compare_key_nulls;
  base.stride = plan.stride;
  for (size_t c = 0; c < plan.keys.size(); ++c) base.keys[c] = plan.keys[c];

  // ---- capacity from a sampled estimate
  const int64_t kSample = 1 << 16;
  int64_t capacity;
  if (R <= 4 * kSample) {
    capacity = std::max<int64_t>(64, 2 * R);
  } else {
    int64_t scap = 4 * kSample;
    size_t need = align_up((size_t)(scap + 1) * 8, 256) + align_up((size_t)(scap + 1) * 4, 256) + 256;
    MGB_HIP(ws->main.reserve(need, stream));
    SampleParams sp{};
    sp.agg = base;
    sp.hdr = static_cast<unsigned long long*>(ws->main.ptr);
    sp.cnt = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(ws->main.ptr) + align_up((size_t)(scap + 1) * 8, 256));
    sp.stats = reinterpret_cast<unsigned long long*>(reinterpret_cast<uint8_t*>(sp.cnt) +
                                                     align_up((size_t)(scap + 1) * 4, 256));
    sp.scap = scap;
    sp.samples = kSample;
    uint64_t empty = plan.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef;
    fill_u64_kernel<<<256, kBlock, 0, stream>>>(sp.hdr, scap, empty);
    MGB_HIP(hipMemsetAsync(sp.cnt, 0, (size_t)(scap + 1) * 4, stream));
    MGB_HIP(hipMemsetAsync(sp.stats, 0, 4 * sizeof(unsigned long long), stream));
    sample_insert_kernel<<<(unsigned)((kSample + kBlock - 1) / kBlock), kBlock, 0, stream>>>(sp);
    sample_stats_kernel<<<(unsigned)((scap + 1 + kBlock - 1) / kBlock), kBlock, 0, stream>>>(sp.cnt, scap + 1, sp.stats);
    MGB_HIP(hipGetLastError());
    MGB_HIP(hipMemcpyAsync(ws->pinned, sp.stats, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    MGB_HIP(hipStreamSynchronize(stream));
    int64_t est = estimate_groups(ws->pinned[0], ws->pinned[1], ws->pinned[2], ws->pinned[3], R);
    capacity = std::max<int64_t>(64, 2 * est + 1024);
  }

  // ---- build the table (retry with a larger one on overflow)
  int64_t groups = 0;
  int64_t nblocks = 0;
  size_t table_bytes = 0, block_off = 0, counters_off = 0;
  for (int attempt = 0;; ++attempt) {
    if (capacity > (int64_t)0x7fffffff) return fail(MGB_ERROR_UNSUPPORTED, "groupby table would exceed 2^31 records");
    nblocks = (capacity + 1 + kBlock - 1) / kBlock;
    table_bytes = align_up((size_t)(capacity + 1) * plan.stride, 256);
    counters_off = table_bytes;
    block_off = counters_off + 256;
    size_t need = block_off + align_up((size_t)nblocks * 8, 256);
    MGB_HIP(ws->main.reserve(need, stream));
    uint8_t* wsb = static_cast<uint8_t*>(ws->main.ptr);

    FillParams fp{};
    fp.table = reinterpret_cast<uint64_t*>(wsb);
    fp.num_records = capacity + 1;
    fp.words = plan.stride / 8;
    fp.special_hdr = plan.key_mode == KEY_INT ? 0ull : kEmptyRowRef;
    record_template(plan, plan.key_mode == KEY_INT ? kEmptyIntKey : kEmptyRowRef, fp.tmpl);
    int64_t fill_words = fp.num_records * fp.words;
    unsigned fill_grid = (unsigned)std::min<int64_t>((fill_words + kBlock - 1) / kBlock, (int64_t)cus * 32);
    {
      ProfScope prof("fill", stream);
      fill_table_kernel<<<fill_grid, kBlock, 0, stream>>>(fp);
    }
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(wsb + counters_off);
    MGB_HIP(hipMemsetAsync(counters, 0, 256, stream));

    AggParams ap = base;
    ap.table = wsb;
    ap.capacity = capacity;
    ap.counters = counters;
    ap.max_groups = std::max<int64_t>(1, capacity * 3 / 4);
    // launch in batches of value columns / ops
    size_t ci = 0;
    bool first = true;
    do {
      AggParams lp = ap;
      lp.num_vals = 0;
      lp.num_ops = 0;
      lp.rows_op = -1;
      if (first && plan.rows_off >= 0) {
        lp.ops[0] = OpDesc{OP_ROWS, plan.rows_off};
        lp.rows_op = 0;
        lp.num_ops = 1;
      }
      while (ci < plan.cols.size() && lp.num_vals < kMaxVals &&
             lp.num_ops + (int)plan.cols[ci].ops.size() <= kMaxOps) {
        const PlanCol& pc = plan.cols[ci];
        ValCol vc = pc.col;
        vc.op_begin = lp.num_ops;
        for (auto& o : pc.ops) lp.ops[lp.num_ops++] = OpDesc{o.first, o.second};
        vc.op_end = lp.num_ops;
        lp.vals[lp.num_vals++] = vc;
        ++ci;
      }
      if (lp.num_vals == 0 && ci < plan.cols.size())
        return fail(MGB_ERROR_UNSUPPORTED, "a value column needs more accumulator ops than one launch holds");
      int64_t grid = std::min<int64_t>((R + kBlock - 1) / kBlock, (int64_t)cus * 16);
      {
        ProfScope prof("aggregate", stream);
        aggregate_kernel<<<(unsigned)std::max<int64_t>(grid, 1), kBlock, 0, stream>>>(lp);
      }
      MGB_HIP(hipGetLastError());
      first = false;
    } while (ci < plan.cols.size());

    int64_t* block_counts = reinterpret_cast<int64_t*>(wsb + block_off);
    count_occupied_kernel<<<(unsigned)nblocks, kBlock, 0, stream>>>(wsb, capacity, plan.stride, plan.key_mode,
                                                                     block_counts);
    scan_blocks_kernel<<<1, 1024, 0, stream>>>(block_counts, nblocks, counters + 2);
    MGB_HIP(hipGetLastError());
    MGB_HIP(hipMemcpyAsync(ws->pinned, counters, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    MGB_HIP(hipStreamSynchronize(stream));
    if (ws->pinned[1] == 0) {
      groups = (int64_t)ws->pinned[2];
      break;
    }
    if (attempt >= 6) return fail(MGB_ERROR_HIP, "groupby table kept overflowing");
    capacity = std::min<int64_t>(capacity * 4, std::max<int64_t>(2 * R + 64, 64));
  }
  *out_num_groups = groups;

  // ---- outputs
  Allocs al{alloc, stream, {}};
  const int nk = (int)plan.keys.size();
  const int no = (int)plan.outs.size();
  if (no + nk > kMaxPinned) return fail(MGB_ERROR_UNSUPPORTED, "too many output columns in one groupby call");
  size_t vbytes = align_up((size_t)std::max<int64_t>(groups, 1), 256);
  int n_valid = no + nk;
  size_t nc_bytes = align_up((size_t)n_valid * 8, 256);
  MGB_HIP(ws->outs.reserve(vbytes * (size_t)n_valid + nc_bytes, stream));
  uint8_t* vb = static_cast<uint8_t*>(ws->outs.ptr);
  unsigned long long* null_counts = reinterpret_cast<unsigned long long*>(vb + vbytes * (size_t)n_valid);
  MGB_HIP(hipMemsetAsync(null_counts, 0, nc_bytes, stream));

  FinParams fin{};
  fin.table = static_cast<const uint8_t*>(ws->main.ptr);
  fin.capacity = capacity;
  fin.stride = plan.stride;
  fin.key_mode = plan.key_mode;
  fin.num_keys = nk;
  fin.block_offsets = reinterpret_cast<const int64_t*>(static_cast<uint8_t*>(ws->main.ptr) + block_off);
  for (int c = 0; c < nk; ++c) {
    const KeyCol& k = plan.keys[c];
    fin.keys[c] = k;
    mgb_out_column& oc = out_keys[c];
    std::memset(&oc, 0, sizeof(oc));
    oc.type_id = k.type_id;
    oc.size = groups;
    if (groups > 0) {
      oc.data_bytes = (size_t)groups * k.width;
      oc.data = al.get(oc.data_bytes);
      if (!oc.data) { al.release_all(); return fail(MGB_ERROR_ALLOC, "allocation of a key column failed"); }
      if (plan.compare_key_nulls && k.mask) {
        oc.mask_bytes = (size_t)bitmask_bytes(groups);
        oc.null_mask = static_cast<uint32_t*>(al.get(oc.mask_bytes));
        if (!oc.null_mask) { al.release_all(); return fail(MGB_ERROR_ALLOC, "allocation of a key mask failed"); }
        fin.key_valid[c] = vb + vbytes * (size_t)(no + c);
      }
    }
    fin.key_out[c] = oc.data;
  }
  for (int o = 0; o < no; ++o) {
    const Plan::Out& po = plan.outs[o];
    mgb_out_column& oc = out_results[o];
    std::memset(&oc, 0, sizeof(oc));
    oc.type_id = po.out_type;
    oc.size = groups;
    bool nullable = po.spec.valid != nullptr;
    if (groups > 0) {
      oc.data_bytes = (size_t)groups * type_width(po.out_type);
      oc.data = al.get(oc.data_bytes);
      if (!oc.data) { al.release_all(); return fail(MGB_ERROR_ALLOC, "allocation of a result column failed"); }
      if (nullable) {
        oc.mask_bytes = (size_t)bitmask_bytes(groups);
        oc.null_mask = static_cast<uint32_t*>(al.get(oc.mask_bytes));
        if (!oc.null_mask) { al.release_all(); return fail(MGB_ERROR_ALLOC, "allocation of a result mask failed"); }
      }
    }
  }
  if (groups > 0) {
    for (int base_o = 0; base_o < no || base_o == 0; base_o += kMaxOuts) {
      FinParams f = fin;
      f.num_outs = std::min(kMaxOuts, no - base_o);
      for (int o = 0; o < f.num_outs; ++o) {
        const Plan::Out& po = plan.outs[base_o + o];
        f.outs[o] = po.spec;
        f.outs[o].out = out_results[base_o + o].data;
        f.outs[o].valid = po.spec.valid ? vb + vbytes * (size_t)(base_o + o) : nullptr;
      }
      if (base_o > 0) {
        // keys were written by the first launch; later launches only write results
        for (int c = 0; c < nk; ++c) f.key_valid[c] = nullptr;
      }
      {
        ProfScope prof("finalize", stream);
        finalize_kernel<<<(unsigned)nblocks, kBlock, 0, stream>>>(f);
      }
      MGB_HIP(hipGetLastError());
      if (no <= kMaxOuts) break;
    }
    for (int j = 0; j < no + nk; ++j) {
      uint32_t* mask = j < no ? out_results[j].null_mask : out_keys[j - no].null_mask;
      if (!mask) continue;
      int64_t words = bitmask_bytes(groups) / 4;
      pack_validity_kernel<<<(unsigned)((words + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
        vb + vbytes * (size_t)j, groups, mask, words, null_counts + j);
    }
    MGB_HIP(hipGetLastError());
    MGB_HIP(hipMemcpyAsync(ws->pinned, null_counts, (size_t)(no + nk) * 8, hipMemcpyDeviceToHost, stream));
    MGB_HIP(hipStreamSynchronize(stream));
    for (int o = 0; o < no; ++o) out_results[o].null_count = (int64_t)ws->pinned[o];
    for (int c = 0; c < nk; ++c) out_keys[c].null_count = (int64_t)ws->pinned[no + c];
  }
  al.live.clear();  // ownership passes to the caller
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

extern "C" mgb_status mgb_groupby_aggregate(const mgb_column* keys, int32_t num_keys, int32_t null_policy,
                                            const mgb_request* requests, int32_t num_requests,
                                            mgb_out_column* out_keys, mgb_out_column* out_results,
                                            int64_t* out_num_groups, void* stream, const mgb_allocator* alloc)
{
  g_last_error.clear();
  if (num_keys < 0 || num_requests < 0 || (num_keys > 0 && !keys) || (num_requests > 0 && !requests) ||
      !out_num_groups || !alloc || !alloc->allocate || !alloc->deallocate)
    return fail(MGB_ERROR_LOGIC, "null argument");
  const int64_t R = num_keys > 0 ? keys[0].size : 0;
  for (int c = 0; c < num_keys; ++c)
    if (keys[c].size != R) return fail(MGB_ERROR_LOGIC, "key columns must have equal sizes");
  for (int r = 0; r < num_requests; ++r)
    if (requests[r].values.size != R)
      return fail(MGB_ERROR_LOGIC, "Size mismatch between request values and groupby keys.");
  for (int r = 0; r < num_requests; ++r)
    for (int a = 0; a < requests[r].num_aggs; ++a)
      if (target_type(requests[r].values.type_id, requests[r].aggs[a].kind) < 0)
        return fail(MGB_ERROR_LOGIC, "Invalid type/aggregation combination.");
  if (null_policy != MGB_NULL_EXCLUDE && null_policy != MGB_NULL_INCLUDE)
    return fail(MGB_ERROR_LOGIC, "invalid null_policy");

  Plan plan;
  if (R == 0) {
    // empty input: empty typed columns (groupby.cu:147-170,216)
    *out_num_groups = 0;
    for (int c = 0; c < num_keys; ++c) {
      std::memset(&out_keys[c], 0, sizeof(mgb_out_column));
      out_keys[c].type_id = keys[c].type_id;
    }
    int j = 0;
    for (int r = 0; r < num_requests; ++r)
      for (int a = 0; a < requests[r].num_aggs; ++a, ++j) {
        std::memset(&out_results[j], 0, sizeof(mgb_out_column));
        out_results[j].type_id = target_type(requests[r].values.type_id, requests[r].aggs[a].kind);
      }
    return MGB_SUCCESS;
  }
  mgb_status st = build_plan(keys, num_keys, null_policy, requests, num_requests, plan);
  if (st != MGB_SUCCESS) return st;
  return run_groupby(plan, R, out_keys, out_results, out_num_groups, static_cast<hipStream_t>(stream), alloc);
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
  std::unique_ptr<Workspace> w;
  {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_ws.find({dev, static_cast<hipStream_t>(stream)});
    if (it == g_ws.end()) return MGB_SUCCESS;
    w = std::move(it->second);
    g_ws.erase(it);
  }
  MGB_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return MGB_SUCCESS;
}
