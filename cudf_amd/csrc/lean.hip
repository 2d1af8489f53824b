// Kernels of the compact partitioned execution (lean_part.hpp), compiled
// for every key shape (one column of 4 or 8 bytes, or two such columns) and
// every value-column shape of up to kLeanCols columns of 4 or 8 bytes whose
// record fits 32 bytes, and their launchers.
#include <hip/hip_runtime.h>

#include <string>

#include "groupby_host.hpp"
#include "lean_part.hpp"

namespace mgb {

template <int K0, int K1, int W0, int W1, int W2>
__global__ __launch_bounds__(kLdsBlock) void lean_scatter_kernel(const LeanScatterArgs a)
{
  extern __shared__ __align__(16) uint8_t smem[];
  lean_scatter_rows<K0, K1, W0, W1, W2>(a, smem);
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

namespace {

struct LeanFns {
  const void* scatter;
  const void* agg;
};

// The kernels of a shape, or none when its record exceeds 32 bytes.
template <int K0, int K1, int W0, int W1, int W2>
constexpr LeanFns lean_fns()
{
  if constexpr (lean_offsets(lean_key_offsets(K0, K1).bytes, W0, W1, W2).bytes > 32 ||
                (W0 != 0 && lean_offsets(lean_key_offsets(K0, K1).bytes, W0, W1, W2).off[0] < 0) ||
                (W1 != 0 && lean_offsets(lean_key_offsets(K0, K1).bytes, W0, W1, W2).off[1] < 0) ||
                (W2 != 0 && lean_offsets(lean_key_offsets(K0, K1).bytes, W0, W1, W2).off[2] < 0)) {
    return LeanFns{nullptr, nullptr};
  } else {
    return LeanFns{(const void*)lean_scatter_kernel<K0, K1, W0, W1, W2>,
                   (const void*)lean_agg_kernel<K0, K1, W0, W1, W2>};
  }
}

// Index of a value-column shape: widths of the present columns (4 or 8),
// absent columns 0 and last.
int lean_shape(const int (&w)[kLeanCols])
{
  int nc = 0;
  while (nc < kLeanCols && w[nc] != 0) ++nc;
  for (int c = nc; c < kLeanCols; ++c)
    if (w[c] != 0) return -1;
  int bits = 0;
  for (int c = 0; c < nc; ++c) {
    if (w[c] != 4 && w[c] != 8) return -1;
    bits |= (w[c] == 8) << c;
  }
  return (1 << nc) - 1 + bits;  // nc = 0: 0; 1: 1-2; 2: 3-6; 3: 7-14
}

// Index of a key shape: one column of 4 or 8 bytes (0, 1), two columns
// (4,4), (4,8), (8,4), (8,8) (2-5).
int lean_key_shape(int k0, int k1)
{
  if (k0 != 4 && k0 != 8) return -1;
  if (k1 == 0) return k0 == 8;
  if (k1 != 4 && k1 != 8) return -1;
  return 2 + (k0 == 8) * 2 + (k1 == 8);
}

#define MGB_LEAN_SHAPES(K0, K1)                                                                                 \
  lean_fns<K0, K1, 0, 0, 0>(), lean_fns<K0, K1, 4, 0, 0>(), lean_fns<K0, K1, 8, 0, 0>(),                        \
    lean_fns<K0, K1, 4, 4, 0>(), lean_fns<K0, K1, 8, 4, 0>(), lean_fns<K0, K1, 4, 8, 0>(),                      \
    lean_fns<K0, K1, 8, 8, 0>(), lean_fns<K0, K1, 4, 4, 4>(), lean_fns<K0, K1, 8, 4, 4>(),                      \
    lean_fns<K0, K1, 4, 8, 4>(), lean_fns<K0, K1, 8, 8, 4>(), lean_fns<K0, K1, 4, 4, 8>(),                      \
    lean_fns<K0, K1, 8, 4, 8>(), lean_fns<K0, K1, 4, 8, 8>(), lean_fns<K0, K1, 8, 8, 8>()

const LeanFns kLeanFns[6][15] = {{MGB_LEAN_SHAPES(4, 0)}, {MGB_LEAN_SHAPES(8, 0)}, {MGB_LEAN_SHAPES(4, 4)},
                                 {MGB_LEAN_SHAPES(4, 8)}, {MGB_LEAN_SHAPES(8, 4)}, {MGB_LEAN_SHAPES(8, 8)}};

const LeanFns* lean_find(int k0, int k1, const int (&w)[kLeanCols])
{
  const int s = lean_shape(w), ks = lean_key_shape(k0, k1);
  if (s < 0 || ks < 0 || !kLeanFns[ks][s].scatter) return nullptr;
  return &kLeanFns[ks][s];
}

}  // namespace

bool lean_supported(int k0, int k1, const int (&w)[kLeanCols]) { return lean_find(k0, k1, w) != nullptr; }

mgb_status lean_scatter(int k0, int k1, const int (&w)[kLeanCols], const LeanScatterArgs& a, unsigned grid,
                        size_t lds, hipStream_t stream)
{
  const LeanFns* f = lean_find(k0, k1, w);
  if (!f) return fail(MGB_ERROR_UNSUPPORTED, "compact partitioned execution: unsupported shape");
  MGB_HIP(set_dynamic_lds(f->scatter, lds));
  LeanScatterArgs args = a;
  void* kargs[] = {&args};
  MGB_HIP(hipLaunchKernel(f->scatter, dim3(grid), dim3(kLdsBlock), kargs, lds, stream));
  return MGB_SUCCESS;
}

mgb_status lean_aggregate(int k0, int k1, const int (&w)[kLeanCols], const LeanAggArgs& a, const LdsParams* lp,
                          unsigned grid, size_t lds, hipStream_t stream)
{
  const LeanFns* f = lean_find(k0, k1, w);
  if (!f) return fail(MGB_ERROR_UNSUPPORTED, "compact partitioned execution: unsupported shape");
  MGB_HIP(set_dynamic_lds(f->agg, lds));
  LeanAggArgs args = a;
  const LdsParams* p = lp;
  void* kargs[] = {&args, &p};
  MGB_HIP(hipLaunchKernel(f->agg, dim3(grid), dim3(kLdsBlock), kargs, lds, stream));
  return MGB_SUCCESS;
}

}  // namespace mgb
