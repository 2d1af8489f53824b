// Device routines of the aggregations cudf serves from its sort-based groupby
// (MEDIAN, QUANTILE, NUNIQUE, NTH_ELEMENT, COLLECT_LIST, COLLECT_SET;
// cpp/src/groupby/sort/). The hash table assigns every row its group (the
// group's output row); segmented sorts (seg_sort.hpp) then order each group's
// values, or its rows, and small kernels read the results off the sorted
// positions.
#pragma once

#include "groupby_host.hpp"

namespace mgb {

// The output row of every row's group (`groups` for rows whose key is dropped).
// With p.row_rec (the build wrote each row's record index), the group is read
// through it instead of probing the table again.
__global__ __launch_bounds__(kBlock) void row_group_kernel(const AggParams p, const uint32_t* rec_pos, uint32_t groups,
                                                           uint32_t* gid)
{
  if (p.row_rec) {
    // 8 rows a thread at a time: their record loads, then their output-row
    // loads, each batch in flight together
    constexpr int U = 8;
    const int64_t step = (int64_t)gridDim.x * blockDim.x * U;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x * U + threadIdx.x; base < p.num_rows; base += step) {
      uint32_t r[U], g[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t row = base + (int64_t)u * blockDim.x;
        r[u] = row < p.num_rows ? p.row_rec[row] : kNoRecord;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) g[u] = r[u] == kNoRecord ? groups : rec_pos[r[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t row = base + (int64_t)u * blockDim.x;
        if (row < p.num_rows) gid[row] = g[u];
      }
    }
    return;
  }
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < p.num_rows; row += step) {
    uint32_t g = groups;
    if (key_row_valid(p, row)) {
      const uint8_t* rec = global_find(p, row_key(p, row), row);
      if (rec) g = rec_pos[(rec - p.table) / p.stride];
    }
    gid[row] = g;
  }
}

// Values in the order of cudf's sort (ascending; NaN after every number) as
// unsigned codes whose order is the value order. Columns of at most 4 bytes
// take 32-bit codes (the group goes above them in one 64-bit sort key),
// 8-byte columns 64-bit codes. -0.0 and +0.0 get adjacent codes (equal
// values; their relative order does not change a result's value).
__device__ __forceinline__ uint32_t sort_code32(uint64_t raw, int width, int vclass)
{
  if (vclass == VC_F32) {
    const float x = __uint_as_float((uint32_t)raw);
    if (x != x) return ~0u;
    const uint32_t b = (uint32_t)raw;
    return (b >> 31) ? ~b : (b | 0x80000000u);
  }
  if (vclass == VC_SIGNED) return (uint32_t)(int32_t)widen_int(raw, width, vclass) ^ 0x80000000u;
  return (uint32_t)widen_int(raw, width, vclass);
}

__device__ __forceinline__ uint64_t sort_code64(uint64_t raw, int vclass)
{
  if (vclass == VC_F64) {
    const double x = __longlong_as_double((long long)raw);
    if (x != x) return ~0ull;
    return (raw >> 63) ? ~raw : (raw | (1ull << 63));
  }
  if (vclass == VC_SIGNED) return raw ^ (1ull << 63);
  return raw;
}

// The value a code stands for, as double (MEDIAN and QUANTILE results).
__device__ __forceinline__ double code_value(uint64_t c, bool narrow, int vclass)
{
  if (narrow) {
    const uint32_t c32 = (uint32_t)c;
    if (vclass == VC_F32) return (double)__uint_as_float((c32 >> 31) ? (c32 & 0x7fffffffu) : ~c32);
    if (vclass == VC_SIGNED) return (double)(int32_t)(c32 ^ 0x80000000u);
    return (double)c32;
  }
  if (vclass == VC_F64) return __longlong_as_double((long long)((c >> 63) ? (c & ~(1ull << 63)) : ~c));
  if (vclass == VC_SIGNED) return (double)(int64_t)(c ^ (1ull << 63));
  return (double)c;
}

// The code of -0.0 (+0.0's is the next one up).
__device__ __forceinline__ uint64_t neg_zero_code(bool narrow) { return narrow ? 0x7fffffffull : ~(1ull << 63); }

// What to sort the rows of one value column by.
enum SortMode : int32_t {
  SORT_BY_VALUE = 0,   // valid values, by (group, value): MEDIAN, QUANTILE, NUNIQUE
  SORT_ROWS_VALID = 1, // valid values, by group in row order: NTH_ELEMENT EXCLUDE
  SORT_ROWS_ALL = 2,   // every row of a group, by group in row order: NTH_ELEMENT INCLUDE
};

// One ordering's sorted rows.
struct SortView {
  const uint64_t* key64;   // narrow: sorted group << 32 | code
  const uint32_t* gkey;    // otherwise: sorted groups
  const uint64_t* code;    // 8-byte values: sorted codes
  const uint32_t* order;   // row order: the sorted rows' indices
  const uint32_t* offset;  // groups + 1 entries: first sorted position of every group
  const uint32_t* total;   // sorted positions in use (offset[groups]; device memory)
  int64_t n;               // positions the buffers hold (at least *total)
  uint32_t groups;
  int32_t narrow;
};

__device__ __forceinline__ uint32_t view_group(const SortView& s, int64_t i)
{
  return s.narrow ? (uint32_t)(s.key64[i] >> 32) : s.gkey[i];
}

__device__ __forceinline__ uint64_t view_code(const SortView& s, int64_t i)
{
  return s.narrow ? (s.key64[i] & 0xffffffffull) : s.code[i];
}

struct SortOutParams {
  ValCol v;               // the value column (NTH_ELEMENT reads its rows)
  SortView s;
  int32_t kind;           // MGB_AGG_MEDIAN / QUANTILE / NUNIQUE / NTH_ELEMENT
  int32_t interp;         // QUANTILE: 0 LINEAR, 1 LOWER, 2 HIGHER, 3 MIDPOINT, 4 NEAREST
  int32_t n;              // NTH_ELEMENT
  int32_t nulls;          // NUNIQUE: 1 = INCLUDE
  double q;               // QUANTILE
  const uint32_t* distinct; // NUNIQUE: distinct values per group
  const uint32_t* has_null;
  void* out;              // per group: result
  uint8_t* out_valid;     // per group: validity byte
  int32_t out_width;
};

// MEDIAN / QUANTILE / NTH_ELEMENT of every group (one thread per group).
__global__ __launch_bounds__(kBlock) void sort_select_kernel(const SortOutParams p)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= p.s.groups) return;
  const int64_t off = p.s.offset[g], n = (int64_t)p.s.offset[g + 1] - off;
  if (p.kind == MGB_AGG_NTH_ELEMENT) {
    // group_nth_element.cu:60-120: the n-th row (from the end when n < 0)
    const int64_t i = p.n < 0 ? n + p.n : p.n;
    bool ok = i >= 0 && i < n;
    uint64_t bits = 0;
    if (ok) {
      const uint32_t row = p.s.order[off + i];
      ok = bit_valid(p.v.mask, p.v.bit_off + row);
      if (ok) bits = load_at(p.v.data, p.v.width, p.v.stride, row);
    }
    store_width(p.out, g, p.out_width, bits);
    p.out_valid[g] = ok;
    return;
  }
  // MEDIAN (0.5, LINEAR) and QUANTILE: quantiles_util.hpp:90-105 (index),
  // :161-181 (selection), :40-80 (interpolation); null without values
  if (n == 0) {
    reinterpret_cast<double*>(p.out)[g] = 0.0;
    p.out_valid[g] = 0;
    return;
  }
  const bool narrow = p.s.narrow != 0;
  const int vc = p.v.vclass;
  const double q = p.q < 0.0 ? 0.0 : p.q > 1.0 ? 1.0 : p.q;
  const double val = q * (double)(n - 1);
  const int64_t lo = (int64_t)floor(val), hi = (int64_t)ceil(val), ne = (int64_t)nearbyint(val);
  const double frac = val - (double)lo;
  double r;
  switch (p.interp) {
    case 1: r = code_value(view_code(p.s, off + lo), narrow, vc); break;
    case 2: r = code_value(view_code(p.s, off + hi), narrow, vc); break;
    case 4: r = code_value(view_code(p.s, off + ne), narrow, vc); break;
    case 3:
      if (!narrow && vc == VC_SIGNED) {
        // exact for int64: halves and the remainders' half
        const int64_t a = (int64_t)(view_code(p.s, off + lo) ^ (1ull << 63));
        const int64_t b = (int64_t)(view_code(p.s, off + hi) ^ (1ull << 63));
        const int64_t half = a / 2 + b / 2, rest = a % 2 + b % 2;
        r = (double)half + (double)rest * 0.5;
      } else {
        r = code_value(view_code(p.s, off + lo), narrow, vc) / 2 + code_value(view_code(p.s, off + hi), narrow, vc) / 2;
      }
      break;
    default: {
      const double a = code_value(view_code(p.s, off + lo), narrow, vc);
      const double b = code_value(view_code(p.s, off + hi), narrow, vc);
      r = (1.0 - frac) * a + frac * b;
      break;
    }
  }
  reinterpret_cast<double*>(p.out)[g] = r;
  p.out_valid[g] = 1;
}

// Whether sorted position i (0 < i < *total) starts its group or differs
// from the previous value (group_nunique.cu:60-70; NaN == NaN, -0.0 ==
// +0.0, as cudf's row equality). nan_distinct (COLLECT_SET under
// nan_equality UNEQUAL): every NaN starts an element of its own.
__device__ __forceinline__ bool distinct_at(const SortView& s, int64_t i, int32_t is_float, int32_t nan_distinct)
{
  const uint64_t nz = neg_zero_code(s.narrow != 0);
  const uint64_t nan_code = s.narrow ? 0xffffffffull : ~0ull;
  if (nan_distinct && is_float && view_code(s, i) == nan_code) return true;
  if (s.narrow) {
    const uint64_t a = s.key64[i - 1], b = s.key64[i];
    const bool zeros = is_float && (a >> 32) == (b >> 32) && (a & 0xffffffffull) == nz && (b & 0xffffffffull) == nz + 1;
    return a != b && !zeros;
  }
  const uint64_t a = s.code[i - 1], b = s.code[i];
  const bool same = a == b || (is_float && a == nz && b == nz + 1);
  return s.gkey[i - 1] != s.gkey[i] || !same;
}

// 1 at every sorted position distinct_at marks (and at position 0), 0 at the
// positions from *total to n (so that the exclusive sum of n + 1 flags holds
// every group's count).
__global__ __launch_bounds__(kBlock) void distinct_start_kernel(const SortView s, int32_t is_float, uint32_t* flag,
                                                               int32_t nan_distinct = 0)
{
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  const int64_t used = *s.total;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= s.n; i += step)
    flag[i] = i < used && (i == 0 || distinct_at(s, i, is_float, nan_distinct)) ? 1u : 0u;
}

// NUNIQUE's distinct values of every group, added into cnt[group] (zero on
// entry): each wave takes 64 consecutive sorted positions, and each run of
// one group's positions there adds its count of distinct_at positions once.
__global__ __launch_bounds__(kBlock) void nunique_count_kernel(const SortView s, int32_t is_float, uint32_t* cnt)
{
  const int64_t used = *s.total;
  const int lane = threadIdx.x & 63;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < used; i0 += step) {
    const int64_t i = i0 + threadIdx.x;
    const bool ok = i < used;
    uint32_t g = 0xffffffffu;
    bool f = false;
    if (ok) {
      g = view_group(s, i);
      f = i == 0 || distinct_at(s, i, is_float, 0);
    }
    const uint32_t gp = __shfl_up(g, 1, 64);
    const bool start = ok && (lane == 0 || gp != g);
    const uint64_t starts = __ballot(start);
    const uint64_t flags = __ballot(f);
    if (start) {
      const uint64_t after = starts & ~((2ull << lane) - 1ull);  // lane 63: 2 << 63 wraps to 0, none after
      const uint64_t below_next = after ? (after & (~after + 1ull)) - 1ull : ~0ull;
      const uint64_t run = below_next & ~((1ull << lane) - 1ull);
      const uint32_t c = (uint32_t)__popcll(flags & run);
      if (c) atomicAdd(&cnt[g], c);
    }
  }
}

// NUNIQUE per group from its distinct values; INCLUDE counts a null value once.
__global__ __launch_bounds__(kBlock) void nunique_finish_kernel(const SortOutParams p)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= p.s.groups) return;
  uint32_t c = p.distinct[g];
  if (p.nulls && p.has_null[g]) c += 1u;
  reinterpret_cast<uint32_t*>(p.out)[g] = c;
  p.out_valid[g] = 1;
}

// ---- COLLECT_LIST and COLLECT_SET (cpp/src/groupby/sort/group_collect.cu,
// cpp/src/groupby/sort/aggregate.cpp collect_set: the lists' elements then
// lists::distinct)

// Null values per group (gid: every row's group, `groups` when dropped).
__global__ __launch_bounds__(kBlock) void group_nulls_kernel(const ValCol v, const uint32_t* gid, int64_t n,
                                                            uint32_t groups, uint32_t* nulls)
{
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < n; row += step) {
    const uint32_t g = gid[row];
    if (g < groups && !bit_valid(v.mask, v.bit_off + row)) atomicAdd(&nulls[g], 1u);
  }
}

// The raw bits of a value from its sort code (sort_code32 / sort_code64
// inverted; a NaN's code gives a quiet NaN).
__device__ __forceinline__ uint64_t code_bits(uint64_t c, bool narrow, int vclass)
{
  if (narrow) {
    const uint32_t c32 = (uint32_t)c;
    if (vclass == VC_F32) return (c32 >> 31) ? (c32 & 0x7fffffffu) : ~c32;
    if (vclass == VC_SIGNED) return (uint64_t)(uint32_t)(c32 ^ 0x80000000u);
    return c32;
  }
  if (vclass == VC_F64) return (c >> 63) ? (c & ~(1ull << 63)) : ~c;
  if (vclass == VC_SIGNED) return c ^ (1ull << 63);
  return c;
}

struct CollectParams {
  ValCol v;
  SortView s;
  const uint32_t* prefix;   // COLLECT_SET: exclusive sum of the distinct-value starts
  const uint32_t* flag;     // COLLECT_SET: the starts
  const uint32_t* nulls;    // COLLECT_SET: null values per group
  uint32_t* list_off;       // groups + 1 entries: first element of every list
  uint8_t* child;           // elements (v.width bytes each)
  uint8_t* child_valid;     // validity byte per element
  int32_t include_nulls;    // null_policy INCLUDE
  int32_t nulls_unequal;    // COLLECT_SET: null_equality UNEQUAL (each null an element)
};

// COLLECT_LIST: element i is the value of sorted row i (the group's rows in
// row order, or its valid rows): a gather up to the last group's end.
__global__ __launch_bounds__(kBlock) void collect_gather_kernel(const CollectParams p)
{
  const int64_t total = p.s.offset[p.s.groups];
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += step) {
    const uint32_t row = p.s.order[i];
    const bool ok = bit_valid(p.v.mask, p.v.bit_off + row);
    store_width(p.child, i, p.v.width, ok ? load_at(p.v.data, p.v.width, p.v.stride, row) : 0ull);
    p.child_valid[i] = ok;
  }
}

// The group's distinct values and the elements its nulls take.
__device__ __forceinline__ uint32_t set_distinct(const CollectParams& p, int64_t g)
{
  return p.prefix[p.s.offset[g + 1]] - p.prefix[p.s.offset[g]];
}
__device__ __forceinline__ uint32_t set_null_elems(const CollectParams& p, int64_t g)
{
  const uint32_t n = p.nulls[g];
  return !p.include_nulls ? 0u : p.nulls_unequal ? n : (n ? 1u : 0u);
}

// COLLECT_SET: each list's length into lens[g] (lens[groups] = 0); their
// exclusive sum is the lists' offsets.
__global__ __launch_bounds__(kBlock) void set_lengths_kernel(const CollectParams p, uint32_t* lens)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g == p.s.groups) lens[g] = 0;
  if (g >= p.s.groups) return;
  lens[g] = set_distinct(p, g) + set_null_elems(p, g);
}

// COLLECT_SET: every distinct value (the start of a run of equal sorted
// values) at its place in its group's list.
__global__ __launch_bounds__(kBlock) void set_values_kernel(const CollectParams p)
{
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  const bool narrow = p.s.narrow != 0;
  const int64_t used = *p.s.total;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < used; i += step) {
    if (!p.flag[i]) continue;
    const uint32_t g = view_group(p.s, i);
    const uint32_t k = p.prefix[i] - p.prefix[p.s.offset[g]];
    const int64_t e = (int64_t)p.list_off[g] + k;
    store_width(p.child, e, p.v.width, code_bits(view_code(p.s, i), narrow, p.v.vclass));
    p.child_valid[e] = 1;
  }
}

// COLLECT_SET: the null elements at the end of every list (one thread per group).
// The null elements are counted into *child_nulls.
__global__ __launch_bounds__(kBlock) void set_nulls_kernel(const CollectParams p, unsigned long long* child_nulls)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t k = 0;
  if (g < p.s.groups) {
    const int64_t first = (int64_t)p.list_off[g] + set_distinct(p, g);
    k = set_null_elems(p, g);
    for (uint32_t j = 0; j < k; ++j) {
      store_width(p.child, first + j, p.v.width, 0ull);
      p.child_valid[first + j] = 0;
    }
  }
  for (int o = 32; o > 0; o >>= 1) k += __shfl_down(k, o, 64);
  if ((threadIdx.x & 63) == 0 && k) atomicAdd(child_nulls, (unsigned long long)k);
}

// COLLECT_LIST under INCLUDE: the null elements (null values of rows whose
// key is kept) counted into *child_nulls.
__global__ __launch_bounds__(kBlock) void list_nulls_kernel(const ValCol v, const uint32_t* gid, int64_t n,
                                                           uint32_t groups, unsigned long long* child_nulls)
{
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  uint32_t k = 0;
  for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < n; row += step)
    k += gid[row] < groups && !bit_valid(v.mask, v.bit_off + row);
  for (int o = 32; o > 0; o >>= 1) k += __shfl_down(k, o, 64);
  if ((threadIdx.x & 63) == 0 && k) atomicAdd(child_nulls, (unsigned long long)k);
}

}  // namespace mgb
