// The orderings of the sort-based aggregations (MEDIAN, QUANTILE, NUNIQUE,
// NTH_ELEMENT, COLLECT_LIST, COLLECT_SET), built from the groups the hash
// table already assigned. cpp/src/groupby/sort/sort_helper.cu sorts every
// row by key and then by value; here every row already carries its group (a
// dense index below 2^32), so:
//   1. seg_load / seg_key: every row's group key (`groups` for a row that
//      takes no part, which sorts last) and payload (value code, or the row
//      index for the row orderings), computed as the first radix pass reads
//      the row;
//   2. a stable LSD radix sort of (group key, payload) over only the bits
//      the group count needs (radix_hist_kernel, a scan, radix_scatter_kernel
//      per pass of at most 8 bits): the rows come out group-contiguous and in
//      row order within each group, which is already the row orderings;
//   3. seg_bounds_kernel: every group's first position;
//   4. for the value orderings, every group's segment sorted by code: up to
//      kWaveSeg keys by one wave in registers (bitonic over lane shuffles),
//      up to kSegChunk by one workgroup in LDS (bitonic), longer ones in
//      kSegChunk chunks then merged pairwise (merge path) in rounds until
//      each segment is one run.
// The sorted keys and every position's group are written in the layout the
// result kernels read (SortView).
#pragma once

#include <algorithm>

#include "sort_aggs.hpp"

namespace mgb {

constexpr int kRadixBlock = 256;            // four waves
constexpr int kRadixItems = 16;             // keys a lane
constexpr int kRadixWaveKeys = 64 * kRadixItems;
constexpr int kRadixTile = kRadixBlock * kRadixItems;
constexpr int kRadixBins = 256;             // digits of at most 8 bits
constexpr int kWaveSeg = 256;               // longest segment one wave sorts (4 keys a lane)
constexpr int kSegChunk = 4096;             // longest segment one workgroup sorts; the merge tile
constexpr int kSegBlock = 512;

struct SegParams {
  ValCol v;
  int64_t num_rows;
  uint32_t groups;
  int32_t mode;               // SortMode
  int32_t narrow;             // SORT_BY_VALUE with 32-bit codes
  int32_t pad;
  const uint32_t* gid;        // group of every row (`groups`: dropped)
  uint32_t* off;              // groups + 1: first position of every group (off[groups] = positions in use)
  uint32_t* has_null;         // optional, per group: 1 when a row of the group has a null value
  void* keys;                 // per position: 32-bit (code or row) or 64-bit code keys
  void* alt;                  // merge rounds' second buffer, as large
  uint64_t* key64;            // narrow: sorted group << 32 | code (else null)
  uint32_t* gkey;             // otherwise, when set: group of every sorted position
  unsigned long long* list;   // group << 32 | chunk of every segment longer than kWaveSeg
  uint32_t* list_count;
};

// Merge rounds a segment of m keys takes after its chunk sorts.
__host__ __device__ __forceinline__ int seg_rounds(uint64_t m)
{
  const uint64_t chunks = (m + kSegChunk - 1) / kSegChunk;
  int r = 0;
  while ((1ull << r) < chunks) ++r;
  return r;
}

__device__ __forceinline__ bool seg_takes_part(const SegParams& s, int64_t row, uint32_t g)
{
  if (g >= s.groups) return false;
  return s.mode == SORT_ROWS_ALL || bit_valid(s.v.mask, s.v.bit_off + row);
}

// Step 1: the group key of a row (`groups` for a row that takes no part,
// which sorts last) and its payload (never read for such a row). The first
// radix pass reads the rows through this; later passes read the previous
// pass's output.
// The first pass's loads of one row, without branches, so that all of a
// thread's loads are in flight together: its group, the word of its
// validity bit (a word of the group array when no mask applies: any in-range
// word, never used), and VW bytes of its value (VW = 0: none).
struct SegRaw {
  uint32_t g;
  uint32_t mword;
  uint64_t raw;
};

__device__ __forceinline__ bool seg_masked(const SegParams& s) { return s.v.mask != nullptr && s.mode != SORT_ROWS_ALL; }

template <int VW>
__device__ __forceinline__ SegRaw seg_load(const SegParams& s, int64_t row)
{
  SegRaw r;
  r.g = s.gid[row];
  const bool masked = seg_masked(s);
  const uint32_t* mp = masked ? s.v.mask : s.gid;
  const int64_t bit = masked ? s.v.bit_off + row : row;
  r.mword = to_global(mp)[bit >> 5];
  const global_t<const uint8_t>* q = to_global(s.v.data) + (VW ? row * (int64_t)s.v.stride : 0);
  if constexpr (VW == 1) r.raw = *q;
  else if constexpr (VW == 2) r.raw = *reinterpret_cast<const global_t<uint16_t>*>(q);
  else if constexpr (VW == 4) r.raw = *reinterpret_cast<const global_t<uint32_t>*>(q);
  else if constexpr (VW == 8) r.raw = *reinterpret_cast<const global_t<uint64_t>*>(q);
  else r.raw = 0;
  return r;
}

// Whether the row takes part: its key's group is kept and (value orderings,
// the valid-row ordering) its value is valid.
__device__ __forceinline__ bool seg_valid(const SegParams& s, int64_t row, const SegRaw& r)
{
  const bool masked = seg_masked(s);
  const int64_t bit = masked ? s.v.bit_off + row : row;
  return !masked || ((r.mword >> (bit & 31)) & 1u);
}

// The group key (`groups` for a row that takes no part) and the payload.
template <typename K, int VW>
__device__ __forceinline__ uint32_t seg_key(const SegParams& s, int64_t row, const SegRaw& r, K& pl)
{
  if (s.mode != SORT_BY_VALUE) pl = (K)row;
  else if constexpr (sizeof(K) == 4) pl = sort_code32(r.raw, VW, s.v.vclass);
  else pl = sort_code64(r.raw, s.v.vclass);
  return r.g < s.groups && seg_valid(s, row, r) ? r.g : s.groups;
}

// One pass of the radix sort: digit (gk >> shift) & (2^bits - 1); tiles of
// kRadixTile keys, one per workgroup.
struct RadixPass {
  int64_t n;
  int64_t tiles;
  int32_t shift;
  int32_t bits;
  const uint32_t* gk_in;
  uint32_t* gk_out;
  const void* pl_in;
  void* pl_out;
  uint32_t* counts;  // [2^bits][tiles]: the tiles' digit counts, then (scanned) their first positions
};

// Step 2a: every tile's digit counts, digit-major (so that their exclusive
// sum is, per digit and tile, the tile's first position for the digit).
// FIRST: the keys come from the rows (seg_load, seg_key), and groups with a null
// value are flagged in s.has_null when it is set.
template <bool FIRST>
__global__ __launch_bounds__(kRadixBlock) void radix_hist_kernel(const RadixPass p, const SegParams s)
{
  __shared__ uint32_t h[kRadixBlock / 64][kRadixBins];  // one histogram a wave
  const uint32_t bins = 1u << p.bits, mask = bins - 1u;
  const int w = threadIdx.x >> 6;
  for (uint32_t b = threadIdx.x; b < (uint32_t)(kRadixBlock / 64 * kRadixBins); b += kRadixBlock) (&h[0][0])[b] = 0u;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
  uint32_t key[kRadixItems];
  if constexpr (FIRST) {
    SegRaw raw[kRadixItems];
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
      const int64_t i = base + k * kRadixBlock + threadIdx.x;
      raw[k] = seg_load<0>(s, i < p.n ? i : p.n - 1);
    }
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
      const int64_t i = base + k * kRadixBlock + threadIdx.x;
      uint32_t dummy;
      key[k] = seg_key<uint32_t, 0>(s, i, raw[k], dummy);
      if (i < p.n && s.has_null && raw[k].g < s.groups && !seg_valid(s, i, raw[k])) s.has_null[raw[k].g] = 1u;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
      const int64_t i = base + k * kRadixBlock + threadIdx.x;
      key[k] = i < p.n ? p.gk_in[i] : 0u;
    }
  }
#pragma unroll
  for (int k = 0; k < kRadixItems; ++k)
    if (base + k * kRadixBlock + threadIdx.x < p.n) atomicAdd(&h[w][(key[k] >> p.shift) & mask], 1u);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < bins; b += kRadixBlock) {
    uint32_t c = 0;
    for (int ww = 0; ww < kRadixBlock / 64; ++ww) c += h[ww][b];
    p.counts[(int64_t)b * p.tiles + blockIdx.x] = c;
  }
}

// Step 2b: the tile's keys to their places, stably. Wave w ranks keys
// [w, w + 1) * kRadixWaveKeys of the tile 64 at a time: lanes of one digit
// find each other with one ballot per digit bit, and a running count per
// wave and digit in LDS orders the rounds. The waves' counts then give each
// key its place in the tile's digit-sorted order, where it is staged in LDS
// (the keys, then the payloads in the same space); consecutive threads then
// write consecutive places of each digit's run.
template <typename K, bool FIRST, int VW>
__global__ __launch_bounds__(kRadixBlock) void radix_scatter_kernel(const RadixPass p, const SegParams s)
{
  __shared__ uint32_t wcnt[kRadixBlock / 64][kRadixBins];
  __shared__ uint32_t tstart[kRadixBins];
  __shared__ uint32_t gbase[kRadixBins];
  __shared__ uint32_t wsum[kRadixBlock / 64];
  __shared__ uint8_t dig[kRadixTile];
  __shared__ K stage[kRadixTile];  // the keys (as 32-bit words), then the payloads
  const uint32_t bins = 1u << p.bits, mask = bins - 1u;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint32_t b = threadIdx.x; b < (uint32_t)(kRadixBlock / 64 * kRadixBins); b += kRadixBlock) (&wcnt[0][0])[b] = 0u;
  // the tile's first place per digit, read while the keys load (bins <= kRadixBlock)
  const uint32_t my_base = threadIdx.x < bins ? p.counts[(int64_t)threadIdx.x * p.tiles + blockIdx.x] : 0u;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
  const int64_t tile_n = std::min<int64_t>(kRadixTile, p.n - base);
  const uint32_t* gin = p.gk_in;
  const K* pin = static_cast<const K*>(p.pl_in);
  uint32_t g[kRadixItems];
  K v[kRadixItems];
  uint32_t rk[kRadixItems];
  const uint64_t lt = (1ull << lane) - 1ull;
  // every load issued before the ranking rounds (which wait on LDS)
  if constexpr (FIRST) {
    SegRaw raw[kRadixItems];
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
      const int64_t li = (int64_t)w * kRadixWaveKeys + k * 64 + lane;
      raw[k] = seg_load<VW>(s, base + (li < tile_n ? li : tile_n - 1));
    }
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
      const int64_t li = (int64_t)w * kRadixWaveKeys + k * 64 + lane;
      g[k] = seg_key<K, VW>(s, base + li, raw[k], v[k]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
      const int64_t li = (int64_t)w * kRadixWaveKeys + k * 64 + lane;
      const bool ok = li < tile_n;
      g[k] = ok ? gin[base + li] : 0u;
      v[k] = ok ? pin[base + li] : (K)0;
    }
  }
#pragma unroll
  for (int k = 0; k < kRadixItems; ++k) {
    const int64_t li = (int64_t)w * kRadixWaveKeys + k * 64 + lane;
    const bool ok = li < tile_n;
    const uint32_t d = (g[k] >> p.shift) & mask;
    uint64_t peers = __ballot(ok);
    for (int b = 0; b < p.bits; ++b) {
      const bool on = (d >> b) & 1u;
      const uint64_t m = __ballot(on);
      peers &= on ? m : ~m;
    }
    // every lane of the digit reads the wave's count; the lowest adds the round's keys
    const uint32_t before = ok ? wcnt[w][d] : 0u;
    __builtin_amdgcn_wave_barrier();
    if (ok && (peers & lt) == 0ull) wcnt[w][d] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    rk[k] = before + (uint32_t)__popcll(peers & lt);
  }
  __syncthreads();
  // per digit: the waves' exclusive sums, the tile's total and its global place
  uint32_t tot = 0;
  for (uint32_t b = threadIdx.x; b < bins; b += kRadixBlock) {
    uint32_t run = 0;
    for (int ww = 0; ww < kRadixBlock / 64; ++ww) {
      const uint32_t c = wcnt[ww][b];
      wcnt[ww][b] = run;
      run += c;
    }
    tot = run;  // bins <= kRadixBlock: one digit a thread
    gbase[b] = my_base;
  }
  // exclusive sum of the digits' totals over the workgroup (thread b holds digit b)
  uint32_t x = tot;
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t wb = 0;
  for (int ww = 0; ww < w; ++ww) wb += wsum[ww];
  if (threadIdx.x < bins) tstart[threadIdx.x] = wb + x - tot;
  __syncthreads();
  // places in the tile's digit-sorted order; keys staged and written first,
  // then the payloads through the same LDS
  uint32_t* stage32 = reinterpret_cast<uint32_t*>(stage);
#pragma unroll
  for (int k = 0; k < kRadixItems; ++k) {
    const int64_t li = (int64_t)w * kRadixWaveKeys + k * 64 + lane;
    if (li < tile_n) {
      const uint32_t d = (g[k] >> p.shift) & mask;
      rk[k] += tstart[d] + wcnt[w][d];
      stage32[rk[k]] = g[k];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < tile_n; i += kRadixBlock) {
    const uint32_t key = stage32[i];
    const uint32_t d = (key >> p.shift) & mask;
    dig[i] = (uint8_t)d;
    p.gk_out[gbase[d] + ((uint32_t)i - tstart[d])] = key;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kRadixItems; ++k) {
    const int64_t li = (int64_t)w * kRadixWaveKeys + k * 64 + lane;
    if (li < tile_n) stage[rk[k]] = v[k];
  }
  __syncthreads();
  K* pout = static_cast<K*>(p.pl_out);
  for (int i = threadIdx.x; i < tile_n; i += kRadixBlock) {
    const uint32_t d = dig[i];
    pout[gbase[d] + ((uint32_t)i - tstart[d])] = stage[i];
  }
}

// Step 3: off[g] = first sorted position of group g, off[groups] = the
// positions in use (every position i from 1 to n fills the groups between its
// predecessor's group and its own).
__global__ __launch_bounds__(kBlock) void seg_bounds_kernel(const uint32_t* gk, int64_t n, uint32_t groups, uint32_t* off)
{
  // 4 positions a thread at a time, their loads in flight together
  constexpr int U = 4;
  const int64_t step = (int64_t)gridDim.x * blockDim.x * U;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x * U + threadIdx.x; base <= n; base += step) {
    int64_t prev[U], cur[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = base + (int64_t)u * blockDim.x;
      prev[u] = i == 0 || i > n ? -1 : (int64_t)gk[i - 1];
      cur[u] = i >= n ? (int64_t)groups : (int64_t)gk[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = base + (int64_t)u * blockDim.x;
      if (i > n) continue;
      for (int64_t g = prev[u] + 1; g <= cur[u]; ++g) off[g] = (uint32_t)i;
    }
  }
}

// A sorted key at its final position: the narrow layout packs the group
// above the code; the others keep the key in place and the group beside it.
template <typename K>
__device__ __forceinline__ void seg_emit(const SegParams& s, K* keys, uint32_t g, int64_t pos, K k, bool in_place)
{
  if (s.key64) {
    s.key64[pos] = ((uint64_t)g << 32) | (uint64_t)(uint32_t)k;
  } else {
    if (!in_place) keys[pos] = k;
    if (s.gkey) s.gkey[pos] = g;
  }
}

// The value of lane (lane ^ m). Partners within a row of 16 lanes move by
// DPP (no LDS crossbar): m = 1 and 2 by quad permutes, 8 by a row rotate of
// 8, 4 by a half-row mirror (lane i of 8 from 7 - i) then a quad mirror
// (lane i of 4 from 3 - i); 16 and 32 by ds_bpermute. m is a constant once
// the networks are unrolled.
__device__ __forceinline__ uint32_t shfl_xor_k(uint32_t v, int m)
{
  switch (m) {
    case 1: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);
    case 2: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);
    case 4: return (uint32_t)__builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, true), 0x1B, 0xF,
                                                      0xF, true);
    case 8: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, true);
    default: return (uint32_t)__shfl_xor((int)v, m, 64);
  }
}
__device__ __forceinline__ uint64_t shfl_xor_k(uint64_t v, int m)
{
  const uint32_t lo = shfl_xor_k((uint32_t)v, m);
  const uint32_t hi = shfl_xor_k((uint32_t)(v >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}

// Bitonic sort of N independent sets of the wave's 64*E keys (key q*64 +
// lane of set n in v[n][q] of `lane`; the sets' networks interleave, which
// hides the shuffles' latency): partners 64 or more apart sit in the same
// lane, nearer ones one shuffle away.
template <typename K, int E, int N>
__device__ __forceinline__ void wave_bitonic(K (&v)[N][E], int lane)
{
  constexpr int P = 64 * E;
#pragma unroll
  for (int k = 2; k <= P; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
          const int qp = q ^ (j >> 6);
          if (q < qp) {
            const bool up = ((((unsigned)q << 6) | (unsigned)lane) & (unsigned)k) == 0;
#pragma unroll
            for (int n = 0; n < N; ++n) {
              const K a = v[n][q], b = v[n][qp];
              const bool sw = (a > b) == up;
              v[n][q] = sw ? b : a;
              v[n][qp] = sw ? a : b;
            }
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < E; ++q) {
          const bool up = ((((unsigned)q << 6) | (unsigned)lane) & (unsigned)k) == 0;
          const bool keep_min = ((lane & j) == 0) == up;
#pragma unroll
          for (int n = 0; n < N; ++n) {
            const K o = shfl_xor_k(v[n][q], j);
            v[n][q] = keep_min ? (o < v[n][q] ? o : v[n][q]) : (o > v[n][q] ? o : v[n][q]);
          }
        }
      }
    }
  }
}

// Two segments of at most 64*E keys each (m = 0: none) sorted by one wave
// (padding keys are the largest key: the first m sorted keys are the
// segment's).
template <typename K, int E>
__device__ __forceinline__ void seg_wave2(const SegParams& s, K* keys, const uint32_t (&g)[2], const uint32_t (&lo)[2],
                                          const uint32_t (&m)[2], int lane)
{
  K v[2][E];
#pragma unroll
  for (int n = 0; n < 2; ++n)
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const uint32_t i = (uint32_t)(q * 64 + lane);
      v[n][q] = i < m[n] ? keys[lo[n] + i] : (K)~(K)0;
    }
  wave_bitonic<K, E, 2>(v, lane);
#pragma unroll
  for (int n = 0; n < 2; ++n)
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const uint32_t i = (uint32_t)(q * 64 + lane);
      if (i < m[n]) seg_emit(s, keys, g[n], (int64_t)lo[n] + i, v[n][q], false);
    }
}

// Step 4a: a wave per group; segments longer than kWaveSeg are listed, one
// entry per kSegChunk chunk, for the workgroup sorts.
template <typename K>
__global__ __launch_bounds__(256) void seg_sort_small_kernel(const SegParams s)
{
  K* keys = static_cast<K*>(s.keys);
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  const int64_t nw = (int64_t)gridDim.x * 4;
  // two groups an iteration: wid + 2*i*nw and the one nw after it
  for (int64_t g0 = wid; g0 < (int64_t)s.groups; g0 += 2 * nw) {
    uint32_t g[2], lo[2], m[2];
    uint32_t mmax = 0;
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const int64_t gg = g0 + n * nw;
      g[n] = (uint32_t)gg;
      lo[n] = 0;
      m[n] = 0;
      if (gg >= (int64_t)s.groups) continue;
      lo[n] = s.off[gg];
      m[n] = s.off[gg + 1] - lo[n];
      if (m[n] > (uint32_t)kWaveSeg) {
        const uint32_t chunks = (m[n] + kSegChunk - 1) / kSegChunk;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(s.list_count, chunks);
        base = (uint32_t)__shfl((int)base, 0, 64);
        for (uint32_t c = lane; c < chunks; c += 64) s.list[base + c] = ((unsigned long long)gg << 32) | c;
        m[n] = 0;  // sorted by the workgroup kernels
      }
      mmax = m[n] > mmax ? m[n] : mmax;
    }
    if (mmax == 0) continue;
    if (mmax <= 64) seg_wave2<K, 1>(s, keys, g, lo, m, lane);
    else if (mmax <= 128) seg_wave2<K, 2>(s, keys, g, lo, m, lane);
    else seg_wave2<K, 4>(s, keys, g, lo, m, lane);
  }
}

// Ascending bitonic sort of t[0, P) (P a power of two) by the workgroup.
template <typename K>
__device__ __forceinline__ void block_bitonic(K* t, uint32_t P)
{
  for (uint32_t k = 2; k <= P; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t c = threadIdx.x; c < P / 2; c += blockDim.x) {
        const uint32_t i = 2 * c - (c & (j - 1));  // c with a 0 bit inserted at j
        const uint32_t l = i + j;
        const bool up = (i & k) == 0;
        const K a = t[i], b = t[l];
        if ((a > b) == up) {
          t[i] = b;
          t[l] = a;
        }
      }
      __syncthreads();
    }
  }
}

// Step 4b: one workgroup per listed chunk. A segment of one chunk is final
// here; the chunks of a longer one go to the buffer its first merge round
// reads (keys after an even number of rounds, alt after an odd one, so that
// the last round writes keys).
template <typename K>
__global__ __launch_bounds__(kSegBlock) void seg_sort_block_kernel(const SegParams s)
{
  __shared__ K t[kSegChunk];
  K* keys = static_cast<K*>(s.keys);
  const uint32_t count = *s.list_count;
  for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
    const unsigned long long ent = s.list[e];
    const uint32_t g = (uint32_t)(ent >> 32), c = (uint32_t)ent;
    const uint32_t lo = s.off[g], m = s.off[g + 1] - lo;
    const int nr = seg_rounds(m);
    const uint32_t b = lo + c * (uint32_t)kSegChunk;
    const uint32_t n = std::min<uint32_t>((uint32_t)kSegChunk, m - c * (uint32_t)kSegChunk);
    uint32_t P = 1;
    while (P < n) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) t[i] = i < n ? keys[b + i] : (K)~(K)0;
    __syncthreads();
    block_bitonic(t, P);
    K* dst = (nr & 1) ? static_cast<K*>(s.alt) : keys;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      if (nr == 0) seg_emit(s, keys, g, (int64_t)b + i, t[i], false);
      else dst[b + i] = t[i];
    }
    __syncthreads();
  }
}

// Number of a's keys among the first d keys of the stable merge of sorted
// a[0, la) and b[0, lb) (a's key first on ties).
template <typename K>
__device__ __forceinline__ int64_t merge_path(const K* a, int64_t la, const K* b, int64_t lb, int64_t d)
{
  int64_t lo = d > lb ? d - lb : 0, hi = d < la ? d : la;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] <= b[d - mid - 1]) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Step 4c, round r: runs of kSegChunk << r keys merged in pairs. Listed
// chunk c of a segment produces the merged pair's keys [c, c + 1) *
// kSegChunk: its two input stretches (found by merge path) staged in LDS,
// merged there by every thread from its own merge-path split, and written
// out in order.
template <typename K>
__global__ __launch_bounds__(kSegBlock) void seg_merge_kernel(const SegParams s, int r)
{
  __shared__ K in[kSegChunk];
  __shared__ K out[kSegChunk];
  __shared__ int64_t split[2];
  K* keys = static_cast<K*>(s.keys);
  K* alt = static_cast<K*>(s.alt);
  const uint32_t count = *s.list_count;
  for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
    const unsigned long long ent = s.list[e];
    const uint32_t g = (uint32_t)(ent >> 32), c = (uint32_t)ent;
    const int64_t lo = s.off[g], m = (int64_t)s.off[g + 1] - lo;
    const int nr = seg_rounds((uint64_t)m);
    if (r >= nr) continue;  // the same for the whole workgroup
    const K* src = ((nr - r) & 1) ? alt : keys;
    K* dst = ((nr - r - 1) & 1) ? alt : keys;
    const int64_t L = (int64_t)kSegChunk << r;
    const int64_t pb = (int64_t)(c >> (r + 1)) * 2 * L;  // the pair's first key
    const int64_t la = std::min<int64_t>(L, m - pb);
    const int64_t lb = m - pb > L ? std::min<int64_t>(L, m - pb - L) : 0;
    const int64_t d0 = (int64_t)c * kSegChunk - pb;
    const int64_t d1 = std::min<int64_t>(d0 + kSegChunk, la + lb);
    const K* a = src + lo + pb;
    const K* b = a + la;
    if (threadIdx.x == 0) split[0] = merge_path(a, la, b, lb, d0);
    if (threadIdx.x == 64) split[1] = merge_path(a, la, b, lb, d1);
    __syncthreads();
    const int64_t i0 = split[0], i1 = split[1];
    const int na = (int)(i1 - i0), nb = (int)((d1 - i1) - (d0 - i0));
    const int64_t j0 = d0 - i0;
    for (int i = threadIdx.x; i < na; i += blockDim.x) in[i] = a[i0 + i];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) in[na + i] = b[j0 + i];
    __syncthreads();
    const int n = na + nb;
    const int per = (n + (int)blockDim.x - 1) / (int)blockDim.x;
    const int dl = std::min(n, (int)threadIdx.x * per), de = std::min(n, dl + per);
    if (dl < de) {
      int i = (int)merge_path(in, na, in + na, nb, dl), j = dl - i;
      for (int d = dl; d < de; ++d) {
        const bool take_a = i < na && (j >= nb || in[i] <= in[na + j]);
        out[d] = take_a ? in[i++] : in[na + j++];
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[lo + pb + d0 + i] = out[i];
    __syncthreads();
  }
}

// Step 4d: the merged segments' groups (and narrow keys) at their positions.
template <typename K>
__global__ __launch_bounds__(kSegBlock) void seg_emit_long_kernel(const SegParams s)
{
  K* keys = static_cast<K*>(s.keys);
  const uint32_t count = *s.list_count;
  for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
    const unsigned long long ent = s.list[e];
    const uint32_t g = (uint32_t)(ent >> 32), c = (uint32_t)ent;
    const uint32_t lo = s.off[g], m = s.off[g + 1] - lo;
    if (seg_rounds(m) == 0) continue;
    const uint32_t b = lo + c * (uint32_t)kSegChunk;
    const uint32_t n = std::min<uint32_t>((uint32_t)kSegChunk, m - c * (uint32_t)kSegChunk);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) seg_emit(s, keys, g, (int64_t)b + i, keys[b + i], true);
  }
}

}  // namespace mgb
