// Launchers of the compact partitioned execution's kernels (lean_kernels.hpp;
// instantiated per key shape in lean_k0.hip .. lean_k5.hip): every key shape
// (one column of 4 or 8 bytes, or two such columns) and every value-column
// shape of up to kLeanCols columns of 4 or 8 bytes whose record fits 32 bytes.
#include <hip/hip_runtime.h>

#include <string>

#include "lean_kernels.hpp"

namespace mgb {

namespace {

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

const LeanFns* const kLeanFns[6] = {kLeanFns0, kLeanFns1, kLeanFns2, kLeanFns3, kLeanFns4, kLeanFns5};

const LeanFns* lean_find(int k0, int k1, const int (&w)[kLeanCols])
{
  const int s = lean_shape(w), ks = lean_key_shape(k0, k1);
  if (s < 0 || ks < 0 || !kLeanFns[ks][s].agg) return nullptr;
  return &kLeanFns[ks][s];
}

}  // namespace

bool lean_supported(int k0, int k1, const int (&w)[kLeanCols]) { return lean_find(k0, k1, w) != nullptr; }

mgb_status lean_scatter(int k0, int k1, const int (&w)[kLeanCols], const LeanScatterArgs& a, unsigned grid,
                        size_t lds, hipStream_t stream)
{
  const LeanFns* f = lean_find(k0, k1, w);
  if (!f) return fail(MGB_ERROR_UNSUPPORTED, "compact partitioned execution: unsupported shape");
  const void* fn = f->scatter[a.hot != nullptr][a.direct != 0];
  MGB_HIP(set_dynamic_lds(fn, lds));
  LeanScatterArgs args = a;
  void* kargs[] = {&args};
  MGB_HIP(hipLaunchKernel(fn, dim3(grid), dim3(kLdsBlock), kargs, lds, stream));
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
