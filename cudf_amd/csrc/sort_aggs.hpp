// Device routines of the aggregations cudf serves from its sort-based groupby
// (MEDIAN, QUANTILE, NUNIQUE, NTH_ELEMENT; cpp/src/groupby/sort/). The hash
// table assigns every row its group (the group's output row); two stable
// radix sorts (hipCUB) then order each value column's rows by group and, for
// the value-ordered aggregations, by value within the group; small kernels
// read the results off the sorted rows.
#pragma once

#include "groupby_host.hpp"

namespace mgb {

// The output row of every row's group (`groups` for rows whose key is dropped).
__global__ __launch_bounds__(kBlock) void row_group_kernel(const AggParams p, const uint32_t* rec_pos, uint32_t groups,
                                                           uint32_t* gid)
{
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

// A value in the order of cudf's sort (ascending; NaN after every number):
// an unsigned code whose order is the value order. -0.0 and +0.0 get adjacent
// codes (equal values; their relative order does not change a result's value).
__device__ __forceinline__ uint64_t sort_code(uint64_t raw, int width, int vclass)
{
  if (vclass == VC_F32 || vclass == VC_F64) {
    const double x = widen_f64(raw, width, vclass);
    if (x != x) return ~0ull;
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | (1ull << 63));
  }
  if (vclass == VC_SIGNED) return (uint64_t)widen_int(raw, width, vclass) ^ (1ull << 63);
  return (uint64_t)widen_int(raw, width, vclass);
}

struct SortAggParams {
  ValCol v;               // the value column (input column, its own validity mask)
  int64_t num_rows;
  uint32_t groups;
  int32_t kind;           // MGB_AGG_MEDIAN / QUANTILE / NUNIQUE / NTH_ELEMENT
  int32_t nulls;          // NUNIQUE / NTH_ELEMENT: 1 = INCLUDE
  int32_t interp;         // QUANTILE: 0 LINEAR, 1 LOWER, 2 HIGHER, 3 MIDPOINT, 4 NEAREST
  int32_t n;              // NTH_ELEMENT
  int32_t pad;
  double q;               // QUANTILE
  const uint32_t* gid;    // group of every row (row_group_kernel)
  uint32_t* gkey;         // per row: the row's group if it takes part, else `groups`
  uint64_t* code;         // per row: sort_code of its value (value-ordered kinds)
  uint32_t* idx;          // per row: its index (the sorts' payload)
  uint32_t* count;        // per group: rows taking part
  uint32_t* has_null;     // per group: 1 when a row of the group has a null value (NUNIQUE INCLUDE)
  const uint32_t* order;  // the rows sorted by (group, value) or by group (row order kept)
  const uint32_t* sorted_gkey;  // gkey of order[i]
  const uint32_t* offset; // per group: first position of its rows in `order`
  void* out;              // per group: result
  uint8_t* out_valid;     // per group: validity byte
  int32_t out_width;
};

__device__ __forceinline__ bool sort_value_valid(const ValCol& v, int64_t row)
{
  return bit_valid(v.mask, v.bit_off + row);
}

// Which rows take part, their sort codes, and the per-group counts.
__global__ __launch_bounds__(kBlock) void sort_prepare_kernel(const SortAggParams s)
{
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  const bool by_value = s.kind != MGB_AGG_NTH_ELEMENT;
  for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < s.num_rows; row += step) {
    const uint32_t g = s.gid[row];
    const bool valid = sort_value_valid(s.v, row);
    // value-ordered kinds take the valid values; NTH_ELEMENT every row of
    // the group (INCLUDE) or its valid values (EXCLUDE)
    const bool takes_part = g < s.groups && (valid || (!by_value && s.nulls));
    s.gkey[row] = takes_part ? g : s.groups;
    s.idx[row] = (uint32_t)row;
    if (by_value) s.code[row] = valid ? sort_code(load_at(s.v.data, s.v.width, s.v.stride, row), s.v.width, s.v.vclass) : 0ull;
    if (takes_part) atomicAdd(&s.count[g], 1u);
    if (g < s.groups && !valid && s.has_null) s.has_null[g] = 1u;
  }
}

// gkey of the rows in `order` (between the value sort and the group sort).
__global__ __launch_bounds__(kBlock) void gather_u32_kernel(const uint32_t* src, const uint32_t* order, int64_t n,
                                                            uint32_t* dst)
{
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) dst[i] = src[order[i]];
}

__device__ __forceinline__ double sort_value_f64(const ValCol& v, uint32_t row)
{
  return widen_f64(load_at(v.data, v.width, v.stride, row), v.width, v.vclass);
}

// MEDIAN / QUANTILE / NTH_ELEMENT of every group (one thread per group).
__global__ __launch_bounds__(kBlock) void sort_select_kernel(const SortAggParams s)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= s.groups) return;
  const int64_t n = s.count[g], off = s.offset[g];
  if (s.kind == MGB_AGG_NTH_ELEMENT) {
    // group_nth_element.cu:60-120: the n-th row (from the end when n < 0)
    const int64_t i = s.n < 0 ? n + s.n : s.n;
    bool ok = i >= 0 && i < n;
    uint64_t bits = 0;
    if (ok) {
      const uint32_t row = s.order[off + i];
      ok = sort_value_valid(s.v, row);
      if (ok) bits = load_at(s.v.data, s.v.width, s.v.stride, row);
    }
    store_width(s.out, g, s.out_width, bits);
    s.out_valid[g] = ok;
    return;
  }
  // MEDIAN (0.5, LINEAR) and QUANTILE: quantiles_util.hpp:90-105 (index),
  // :161-181 (selection), :40-80 (interpolation); null without values
  if (n == 0) {
    reinterpret_cast<double*>(s.out)[g] = 0.0;
    s.out_valid[g] = 0;
    return;
  }
  const double q = s.q < 0.0 ? 0.0 : s.q > 1.0 ? 1.0 : s.q;
  const double val = q * (double)(n - 1);
  const int64_t lo = (int64_t)floor(val), hi = (int64_t)ceil(val), ne = (int64_t)nearbyint(val);
  const double frac = val - (double)lo;
  double r;
  switch (s.interp) {
    case 1: r = sort_value_f64(s.v, s.order[off + lo]); break;
    case 2: r = sort_value_f64(s.v, s.order[off + hi]); break;
    case 4: r = sort_value_f64(s.v, s.order[off + ne]); break;
    case 3:
      if (s.v.vclass == VC_SIGNED && s.v.width == 8) {
        const int64_t a = (int64_t)load_at(s.v.data, 8, s.v.stride, s.order[off + lo]);
        const int64_t b = (int64_t)load_at(s.v.data, 8, s.v.stride, s.order[off + hi]);
        const int64_t half = a / 2 + b / 2, rest = a % 2 + b % 2;
        r = (double)half + (double)rest * 0.5;
      } else {
        r = sort_value_f64(s.v, s.order[off + lo]) / 2 + sort_value_f64(s.v, s.order[off + hi]) / 2;
      }
      break;
    default: {
      const double a = sort_value_f64(s.v, s.order[off + lo]), b = sort_value_f64(s.v, s.order[off + hi]);
      r = (1.0 - frac) * a + frac * b;
      break;
    }
  }
  reinterpret_cast<double*>(s.out)[g] = r;
  s.out_valid[g] = 1;
}

// NUNIQUE: every sorted position that starts its group or differs from the
// previous value counts one (group_nunique.cu:60-70; NaN == NaN, -0.0 ==
// +0.0, as cudf's row equality); `total` positions take part.
__global__ __launch_bounds__(kBlock) void nunique_count_kernel(const SortAggParams s, int64_t total)
{
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += step) {
    const uint32_t g = s.sorted_gkey[i];
    if (g >= s.groups) continue;  // rows that take no part sort last
    bool first = i == (int64_t)s.offset[g];
    if (!first) {
      const uint64_t a = load_at(s.v.data, s.v.width, s.v.stride, s.order[i - 1]);
      const uint64_t b = load_at(s.v.data, s.v.width, s.v.stride, s.order[i]);
      if (s.v.vclass == VC_F32 || s.v.vclass == VC_F64) {
        const double x = widen_f64(a, s.v.width, s.v.vclass), y = widen_f64(b, s.v.width, s.v.vclass);
        first = !((x == y) || (x != x && y != y));
      } else {
        first = widen_int(a, s.v.width, s.v.vclass) != widen_int(b, s.v.width, s.v.vclass);
      }
    }
    if (first) atomicAdd(reinterpret_cast<uint32_t*>(s.out) + g, 1u);
  }
}

// NUNIQUE INCLUDE: a group with a null value counts one more; every group's
// count is valid.
__global__ __launch_bounds__(kBlock) void nunique_finish_kernel(const SortAggParams s)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= s.groups) return;
  if (s.nulls && s.has_null[g]) reinterpret_cast<uint32_t*>(s.out)[g] += 1u;
  s.out_valid[g] = 1;
}

}  // namespace mgb
