// Hot keys of the compact partitioned execution (HotSet): their lookup in
// LDS, shared by the estimate pass (which leaves their rows out of the fine
// histogram) and the compact scatter (which aggregates their rows apart).
#pragma once

#include "groupby_kernels.hpp"
#include "mgb_device.hpp"

namespace mgb {

constexpr uint32_t kHotEmpty = 0xffffffffu;

// cudf's row hash of a key given as its raw column values (k1 = 0: one
// column): the hash the compact execution partitions by.
__device__ __forceinline__ uint32_t hot_hash(uint64_t c0, uint64_t c1, int k0, int k1)
{
  uint32_t h = hash_combine(0u, murmur3_bits(c0, k0, 0));
  if (k1) h = hash_combine(h, murmur3_bits(c1, k1, 0));
  return h;
}

// The hot keys in LDS: the keys, and kHotSlots lookup slots (hot index, or
// kHotEmpty) probed linearly from the row hash's low bits.
struct HotLds {
  const HotKey* keys;
  const uint32_t* idx;
  int n;
};

// Copies the hot set into LDS (`keys`: kHotMax entries, `idx`: kHotSlots
// words) and builds the lookup. Ends with a workgroup barrier.
__device__ __forceinline__ HotLds hot_load(const HotSet* hs, HotKey* keys, uint32_t* idx, int k0, int k1)
{
  const int n = (int)min(to_global(&hs->n)[0], (uint32_t)kHotMax);
  for (int i = threadIdx.x; i < kHotSlots; i += blockDim.x) idx[i] = kHotEmpty;
  for (int i = threadIdx.x; i < n; i += blockDim.x) keys[i] = hs->keys[i];
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    uint32_t s = hot_hash(keys[i].c0, keys[i].c1, k0, k1) & (kHotSlots - 1);
    while (atomicCAS(&idx[s], kHotEmpty, (uint32_t)i) != kHotEmpty) s = (s + 1) & (kHotSlots - 1);
  }
  __syncthreads();
  return HotLds{keys, idx, n};
}

// The hot index of the key with row hash h and raw column values (c0, c1),
// or -1. At most kHotMax of the kHotSlots slots are used, so the probe meets
// an empty slot.
__device__ __forceinline__ int hot_find(const HotLds& hl, uint32_t h, uint64_t c0, uint64_t c1)
{
  uint32_t s = h & (kHotSlots - 1);
  for (int probe = 0; probe < kHotSlots; ++probe) {
    const uint32_t e = hl.idx[s];
    if (e == kHotEmpty) return -1;
    if (hl.keys[e].c0 == c0 && hl.keys[e].c1 == c1) return (int)e;
    s = (s + 1) & (kHotSlots - 1);
  }
  return -1;
}

// hot_find of N rows at once: their home slots read together, then the
// keys those slots name, so that the LDS reads of all N rows overlap; a row
// whose home slot names another key takes the full probe (rare: at most
// kHotMax keys in kHotSlots slots).
template <int N>
__device__ __forceinline__ void hot_find_n(const HotLds& hl, const uint32_t (&h)[N], const uint64_t (&c0)[N],
                                           const uint64_t (&c1)[N], const bool (&look)[N], int (&out)[N])
{
  uint32_t e[N];
#pragma unroll
  for (int u = 0; u < N; ++u) e[u] = look[u] ? hl.idx[h[u] & (kHotSlots - 1)] : kHotEmpty;
  uint64_t k0[N], k1[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    const HotKey& k = hl.keys[e[u] == kHotEmpty ? 0 : e[u]];
    k0[u] = k.c0;
    k1[u] = k.c1;
  }
#pragma unroll
  for (int u = 0; u < N; ++u) {
    if (e[u] == kHotEmpty) out[u] = -1;
    else if (k0[u] == c0[u] && k1[u] == c1[u]) out[u] = (int)e[u];
    else out[u] = hot_find(hl, h[u], c0[u], c1[u]);
  }
}

}  // namespace mgb
