// Kernels of the compact partitioned execution (lean_part.hpp) and the table
// of one key shape's kernels (LeanFns, MGB_LEAN_SHAPES): each lean_k*.hip
// instantiates one key shape, so that they compile in parallel.
#pragma once

#include "groupby_host.hpp"
#include "lean_part.hpp"

namespace mgb {

template <int K0, int K1, int W0, int W1, int W2, bool HOT, bool DIRECT>
__global__ __launch_bounds__(kLdsBlock) void lean_scatter_kernel(const LeanScatterArgs a)
{
  extern __shared__ __align__(16) uint8_t smem[];
  lean_scatter_rows<K0, K1, W0, W1, W2, HOT, DIRECT>(a, smem);
}

// One work item: its records into the workgroup's LDS table, then its groups
// written as output rows (a partition held whole) or folded into the global
// table (a split partition, or an item that spilled), as lds_agg_kernel does.
template <int K0, int K1, int W0, int W1, int W2>
__global__ __launch_bounds__(kLdsBlock) void lean_agg_kernel(const LeanAggArgs a, const LdsParams* __restrict__ lp)
{
  extern __shared__ __align__(16) uint8_t lds[];
  int64_t begin, end;
  bool merge = true;
  int fold_part = 0, fold_parts = 1;
  if (!lean_item(a, blockIdx.x, begin, end, merge, fold_part, fold_parts)) return;
  const LdsTable t = lds_table_at(lds, a.slots, a.words, a.cap);
  lds_table_init(t, a.tmpl, a.words);
  lean_aggregate_rows<K0, K1, W0, W1, W2>(a, t, lds, begin, end);
  if constexpr (K1 != 0) lean2_headers_to_rowref<lean_key_offsets(K0, K1).bytes>(t, (uint32_t)a.tail_off);
  if (merge || *t.overflow) lds_fold_to_global(lp->src, t, fold_part, fold_parts);
  else lds_write_groups(lp->src, t, &lp->out_counters[0], 0, *lp->out);
}

// The kernels of one key and value-column shape: the scatter (staged or
// direct, without or with hot keys) and the aggregation.
struct LeanFns {
  const void* scatter[2][2];  // [hot keys][direct]
  const void* agg;
};

// The kernels of a shape, or none when its record exceeds 32 bytes.
template <int K0, int K1, int W0, int W1, int W2>
constexpr LeanFns lean_fns()
{
  constexpr LeanOffsets lo = lean_offsets(lean_key_offsets(K0, K1).bytes, W0, W1, W2);
  if constexpr (lo.bytes > 32 || (W0 != 0 && lo.off[0] < 0) || (W1 != 0 && lo.off[1] < 0) ||
                (W2 != 0 && lo.off[2] < 0)) {
    return LeanFns{{{nullptr, nullptr}, {nullptr, nullptr}}, nullptr};
  } else {
    return LeanFns{{{(const void*)lean_scatter_kernel<K0, K1, W0, W1, W2, false, false>,
                     (const void*)lean_scatter_kernel<K0, K1, W0, W1, W2, false, true>},
                    {(const void*)lean_scatter_kernel<K0, K1, W0, W1, W2, true, false>,
                     (const void*)lean_scatter_kernel<K0, K1, W0, W1, W2, true, true>}},
                   (const void*)lean_agg_kernel<K0, K1, W0, W1, W2>};
  }
}

// The 15 value-column shapes of a key shape, in lean_shape order.
#define MGB_LEAN_SHAPES(K0, K1)                                                                                 \
  lean_fns<K0, K1, 0, 0, 0>(), lean_fns<K0, K1, 4, 0, 0>(), lean_fns<K0, K1, 8, 0, 0>(),                        \
    lean_fns<K0, K1, 4, 4, 0>(), lean_fns<K0, K1, 8, 4, 0>(), lean_fns<K0, K1, 4, 8, 0>(),                      \
    lean_fns<K0, K1, 8, 8, 0>(), lean_fns<K0, K1, 4, 4, 4>(), lean_fns<K0, K1, 8, 4, 4>(),                      \
    lean_fns<K0, K1, 4, 8, 4>(), lean_fns<K0, K1, 8, 8, 4>(), lean_fns<K0, K1, 4, 4, 8>(),                      \
    lean_fns<K0, K1, 8, 4, 8>(), lean_fns<K0, K1, 4, 8, 8>(), lean_fns<K0, K1, 8, 8, 8>()

// The tables, one per key shape (lean_key_shape order), each in its own file.
extern const LeanFns kLeanFns0[15], kLeanFns1[15], kLeanFns2[15], kLeanFns3[15], kLeanFns4[15], kLeanFns5[15];

}  // namespace mgb
