// The few-groups execution: a direct execution for calls whose sample holds
// every group and at most kTinyKeys of them (single integer key, value
// columns whose ops are SUM / MIN / MAX / counts). Each row finds its group
// among the candidate keys with one lookup (a multiplicative hash into a
// 256-entry table, then one key compare) instead of hashing and probing, and
// updates one of kTinyCopies copies of the group's record: lane l updates
// copy l % kTinyCopies. Copy c of group g sits at record g * kTinyCopies + c
// of a striped region whose stride is an odd multiple of 8 bytes, so the
// lanes of a wave reach distinct LDS banks and never the same address,
// whatever the keys (with one record per group, 64 lanes over 16 groups
// queue on every address). The updates are the column programs of the LDS
// path (lds_column_program). Optionally a wave first collapses each group's
// lanes (ballot of the lanes whose group is the leader's, then a butterfly
// reduction per field) and only the leader updates a copy. At the end the
// copies are merged into the group's record of the workgroup's LDS table,
// which is folded into the global table as the direct execution does. Rows
// whose key is not a candidate (a key the sample missed) take the table's
// probe.
#pragma once

#include "lds_agg.hpp"

namespace mgb {

constexpr int kTinyKeys = 16;     // candidate groups
constexpr int kTinyCopies = 32;   // copies of every group's record
constexpr int kTinyLutBits = 8;
constexpr int kTinyCols = 2;      // value columns
constexpr uint8_t kTinyNone = 0xff;

struct TinyParams {
  uint64_t keys[kTinyKeys];        // candidate keys as record headers (sign-extended)
  uint64_t prog[kTinyCols];        // each value column's program (pack_program, rec_off unused)
  uint32_t mult;                   // lookup hash multiplier
  int32_t num_keys;
  int32_t collapse;                // 1: each wave collapses a group's lanes before one update
  int32_t pad;
  uint8_t lut[1 << kTinyLutBits];  // hash -> candidate index, kTinyNone when none
};

__host__ __device__ __forceinline__ uint32_t tiny_hash(uint64_t key, uint32_t mult)
{
  return ((uint32_t)key ^ (uint32_t)(key >> 32)) * mult >> (32 - kTinyLutBits);
}

// LDS bytes after the table: striped records, lookup table, candidate keys,
// their records' offsets, one dummy word per lane.
__host__ __device__ constexpr size_t tiny_lds_bytes(int words)
{
  return (size_t)kTinyKeys * kTinyCopies * words * 8 + (1 << kTinyLutBits) + kTinyKeys * 8 + kTinyKeys * 4 + 64 * 8;
}

// A row whose key is not a candidate: the LDS table's probe, the global
// table when the LDS table is full.
__device__ __forceinline__ void tiny_miss_row(const AggParams& p, const LdsTable& t, int64_t row)
{
  const RowKey rk = row_key(p, row);
  uint8_t* r = lds_find_or_insert(t, p, rk, rk.hash, row);
  if (r) apply_row(p, r, row);
  else spill_row(p, rk, row);
}

// Value column types the row loop is compiled for (TV_ANY: read from the
// column's description at run time).
enum TinyValType : int { TV_ANY = 0, TV_I32, TV_I64, TV_F32, TV_F64 };

__host__ __device__ constexpr int tiny_val_type(int width, int vclass)
{
  return vclass == VC_SIGNED && width == 4 ? TV_I32
         : vclass == VC_SIGNED && width == 8 ? TV_I64
         : vclass == VC_F32 ? TV_F32
         : vclass == VC_F64 ? TV_F64
                            : TV_ANY;
}

// SUM / MIN / MAX of N rows of a column of known type on their record
// copies (the column program's updates with the dispatch resolved at
// compile time).
template <int TV, int N>
__device__ __forceinline__ void tiny_column_fixed(uint8_t* lds, const Program& g, const uint32_t (&lo)[N],
                                                  const uint64_t (&x)[N], const bool (&ok)[N], uint32_t dum)
{
  if (g.sum_off >= 0) {
#pragma unroll
    for (int u = 0; u < N; ++u) {
      uint8_t* a = lds + (ok[u] ? lo[u] + g.sum_off : dum);
      if constexpr (TV == TV_F64 || TV == TV_F32)
        lds_add(reinterpret_cast<double*>(a), canon_f64(x[u], TV == TV_F32 ? VC_F32 : VC_F64));
      else
        lds_add(reinterpret_cast<unsigned long long*>(a), (unsigned long long)x[u]);
    }
  }
  for (int which = 0; which < 2; ++which) {
    const int off = which == 0 ? g.min_off : g.max_off;
    if (off < 0) continue;
#pragma unroll
    for (int u = 0; u < N; ++u) {
      uint8_t* a = lds + (ok[u] ? lo[u] + off : dum);
      if constexpr (TV == TV_F64) {
        const unsigned long long b = x[u];
        const bool nan = (b & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (b & 0x000fffffffffffffull);
        if (which == 0) lds_min(reinterpret_cast<unsigned long long*>(a), nan ? kF64MinNaN : f64_code(b));
        else lds_max(reinterpret_cast<unsigned long long*>(a), nan ? kF64MaxNaN : f64_code(b));
      } else if constexpr (TV == TV_F32) {
        const uint32_t b = (uint32_t)x[u];
        const bool nan = (b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu);
        if (which == 0) lds_min(reinterpret_cast<uint32_t*>(a), nan ? kF32MinNaN : f32_code(b));
        else lds_max(reinterpret_cast<uint32_t*>(a), nan ? kF32MaxNaN : f32_code(b));
      } else if constexpr (TV == TV_I32) {
        if (which == 0) lds_min(reinterpret_cast<int32_t*>(a), (int32_t)x[u]);
        else lds_max(reinterpret_cast<int32_t*>(a), (int32_t)x[u]);
      } else {
        if (which == 0) lds_min(reinterpret_cast<long long*>(a), (long long)x[u]);
        else lds_max(reinterpret_cast<long long*>(a), (long long)x[u]);
      }
    }
  }
}

// Wave-wide combination of one field (every lane ends with the result).
template <typename T, typename F>
__device__ __forceinline__ T tiny_wave_reduce(T v, F op)
{
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) v = op(v, __shfl_xor(v, d, 64));
  return v;
}

// In-wave collapse of N rows' updates of one column: for each group present
// in the wave, the leader lane takes the combined SUM / MIN / MAX / null
// count of the group's lanes; the other lanes' rows are dropped from the
// updates (ok[u] false). Min/max values are the program's keys (order codes).
template <int N>
__device__ __forceinline__ void tiny_collapse(const Program& g, const uint32_t (&cidx)[N], bool (&hit)[N],
                                              bool (&ok)[N], uint64_t (&x)[N], uint32_t (&nulls)[N], bool rows_only)
{
  const unsigned lane = threadIdx.x & 63u;
#pragma unroll
  for (int u = 0; u < N; ++u) {
    uint64_t todo = __ballot(hit[u]);
    bool lead = false;
    uint64_t nx = x[u];
    uint32_t nn = nulls[u];
    bool nok = ok[u];
    while (todo) {
      const int leader = __builtin_ctzll(todo);
      const uint32_t grp = (uint32_t)__builtin_amdgcn_readlane((int)cidx[u], leader);
      const bool mine = hit[u] && cidx[u] == grp;
      const uint64_t m = __ballot(mine);
      todo &= ~m;
      if (rows_only) {
        if (lane == (unsigned)leader) {
          lead = true;
          nn = (uint32_t)__popcll(m);  // the group's rows in the wave
        }
        continue;
      }
      const bool v = mine && ok[u];
      const uint32_t vn = (uint32_t)__popcll(__ballot(mine && !ok[u]));
      const bool any = __ballot(v) != 0;
      uint64_t sum = 0, mn = 0, mx = 0;
      if (g.sum_off >= 0) {
        if (g.vclass == VC_F32 || g.vclass == VC_F64) {
          const double d = tiny_wave_reduce(v ? canon_f64(x[u], g.vclass) : 0.0, [](double a, double b) { return a + b; });
          sum = (uint64_t)__double_as_longlong(d);
        } else {
          sum = tiny_wave_reduce(v ? x[u] : 0ull, [](uint64_t a, uint64_t b) { return a + b; });
        }
      }
      if (lane == (unsigned)leader) {
        lead = true;
        nn = vn;
        nok = any;
        nx = sum;
      }
      (void)mn;
      (void)mx;
    }
    if (!lead) {
      hit[u] = false;
      ok[u] = false;
    } else {
      x[u] = nx;
      nulls[u] = nn;
      ok[u] = nok;
    }
  }
}

// Rows [begin, end) of a direct work item with NC value columns (KW: the
// key's width and TV: the first value column's type when compiled for them,
// 0 otherwise). `tp`
// (device memory): the candidates and programs. LDS: the item's table, then
// at `tiny` the striped records and the lookup data.
template <int NC, int KW, int TV>
__device__ __forceinline__ void tiny_aggregate(const AggParams& p, const LdsTable& t, uint8_t* lds, uint32_t tiny_off,
                                               const uint64_t* tmpl, int words, const TinyParams* __restrict__ tp,
                                               int64_t begin, int64_t end)
{
  const int K = tp->num_keys;
  const uint32_t stride = (uint32_t)words * 8u;
  uint8_t* tiny = lds + tiny_off;
  uint64_t* striped = reinterpret_cast<uint64_t*>(tiny);
  uint8_t* lut = tiny + (size_t)kTinyKeys * kTinyCopies * stride;
  uint64_t* ckeys = reinterpret_cast<uint64_t*>(lut + (1 << kTinyLutBits));
  uint32_t* coff = reinterpret_cast<uint32_t*>(ckeys + kTinyKeys);
  const uint32_t dum_base = (uint32_t)(reinterpret_cast<uint8_t*>(coff + kTinyKeys) - lds);
  for (int i = threadIdx.x; i < kTinyKeys * kTinyCopies * words; i += kLdsBlock) striped[i] = tmpl[i % words];
  for (int i = threadIdx.x; i < (1 << kTinyLutBits); i += kLdsBlock) lut[i] = tp->lut[i];
  if (threadIdx.x < kTinyKeys) ckeys[threadIdx.x] = threadIdx.x < (unsigned)K ? tp->keys[threadIdx.x] : kEmptyIntKey;
  __syncthreads();
  // the candidates' records in the table (one thread each)
  const KeyCol& k0 = p.keys[0];
  if (threadIdx.x < (unsigned)K) {
    RowKey rk;
    rk.key = ckeys[threadIdx.x];
    const uint64_t raw = k0.width == 8 ? rk.key : (rk.key & ((1ull << (8 * k0.width)) - 1ull));
    rk.hash = hash_combine(0u, murmur3_bits(normalize_float_bits(raw, k0.vclass), k0.width, 0));
    rk.special = false;
    uint8_t* r = lds_find_or_insert(t, p, rk, rk.hash, 0);
    coff[threadIdx.x] = r ? (uint32_t)(r - lds) : kRecNone;
  }
  // the loop's uniform values, held in vector registers and made scalar at
  // each use (a scalar load inside the loop waits for the LDS updates in flight)
  uint64_t pin_kdata = (uint64_t)k0.data;
  uint32_t pin_k = (uint32_t)(k0.width & 15) | (uint32_t)(k0.vclass & 7) << 4 | (uint32_t)(k0.stride & 0xffff) << 8 |
                   (uint32_t)(p.skip_null_keys ? 1u : 0u) << 24 | (uint32_t)(tp->collapse ? 1u : 0u) << 25;
  uint32_t pin_mult = tp->mult;
  uint32_t pin_rows = (uint32_t)((p.rows_op >= 0 ? p.ops[p.rows_op].off : -1) + 1);
  uint64_t pin_vdata[NC], pin_vmask[NC], pin_prog[NC];
  uint32_t pin_v[NC];
  int64_t pin_voff[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const ValCol& v = p.vals[c];
    pin_vdata[c] = (uint64_t)v.data;
    pin_vmask[c] = (uint64_t)v.mask;
    pin_voff[c] = v.bit_off;
    pin_v[c] = (uint32_t)(v.width & 15) | (uint32_t)(v.vclass & 7) << 4 | (uint32_t)(v.stride & 0xffff) << 8;
    pin_prog[c] = tp->prog[c];
    asm volatile("" : "+v"(pin_vdata[c]), "+v"(pin_vmask[c]), "+v"(pin_voff[c]), "+v"(pin_v[c]), "+v"(pin_prog[c]));
  }
  asm volatile("" : "+v"(pin_kdata), "+v"(pin_k), "+v"(pin_mult), "+v"(pin_rows));
  __syncthreads();

  constexpr int kRows = 4;
  const int64_t step = kLdsBlock;
  const unsigned lane = threadIdx.x & 63u;
  const uint32_t copy = (lane % kTinyCopies) * stride;
  const uint32_t dum = dum_base + lane * 8u;
  auto sread64 = [](uint64_t v) {
    return (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v) |
           (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32)) << 32;
  };
  // the raw bits of the keys and values of the rows starting at `from`
  // (rows past the end load row 0); nothing is computed from them here, so
  // the loads stay in flight through the current rows' updates
  auto fetch = [&](int64_t from, uint64_t (&key)[kRows], uint64_t (&x)[NC][kRows]) {
    const uint32_t kp = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_k);
    bool pr[kRows];
#pragma unroll
    for (int u = 0; u < kRows; ++u) pr[u] = from + threadIdx.x + u * step < end;
    load_rows_raw(reinterpret_cast<const uint8_t*>(sread64(pin_kdata)), KW ? KW : (int)(kp & 15),
                  KW ? KW : (int)((kp >> 8) & 0xffff), from + threadIdx.x, step, pr, key);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const uint32_t vp = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_v[c]);
      if (c == 0 && TV != TV_ANY) {
        constexpr int w = TV == TV_I32 || TV == TV_F32 ? 4 : 8;
        load_rows_raw(reinterpret_cast<const uint8_t*>(sread64(pin_vdata[c])), w, w, from + threadIdx.x, step, pr,
                      x[c]);
      } else {
        load_rows_raw(reinterpret_cast<const uint8_t*>(sread64(pin_vdata[c])), (int)(vp & 15),
                      (int)((vp >> 8) & 0xffff), from + threadIdx.x, step, pr, x[c]);
      }
    }
  };
  uint64_t nkey[kRows], nx[NC][kRows];
  fetch(begin, nkey, nx);
  // the loop runs alike on every lane of the workgroup (the collapse's
  // cross-lane reductions need the whole wave); lanes past the end hold no
  // row. The next rows' loads are issued before this iteration's updates
  // (past the end: loads of row 0, unused).
  for (int64_t first = begin; first < end; first += (int64_t)kRows * kLdsBlock) {
    const int64_t base = first + threadIdx.x;
    const uint32_t kp = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_k);
    const int kw = KW ? KW : (int)(kp & 15), kvc = (int)((kp >> 4) & 7);
    const bool collapse = (kp >> 25) & 1u;
    const uint32_t mult = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_mult);
    const int rows_off = (int)(uint32_t)__builtin_amdgcn_readfirstlane((int)pin_rows) - 1;
    bool present[kRows];
#pragma unroll
    for (int u = 0; u < kRows; ++u) present[u] = base + u * step < end;
    uint64_t key[kRows];
    uint64_t x[NC][kRows];
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      key[u] = nkey[u];
#pragma unroll
      for (int c = 0; c < NC; ++c) x[c][u] = nx[c][u];
    }
    fetch(first + (int64_t)kRows * kLdsBlock, nkey, nx);
    // the values' canonical form (integers widened, floats as raw bits)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c == 0 && TV != TV_ANY) {
        if constexpr (TV == TV_I32) {
#pragma unroll
          for (int u = 0; u < kRows; ++u) x[c][u] = (uint64_t)(int64_t)(int32_t)(uint32_t)x[c][u];
        }
      } else {
        const uint32_t vp = (uint32_t)__builtin_amdgcn_readfirstlane((int)pin_v[c]);
#pragma unroll
        for (int u = 0; u < kRows; ++u) x[c][u] = widen_raw(x[c][u], (int)(vp & 15), (int)((vp >> 4) & 7));
      }
    }
    if ((kp >> 24) & 1u) {
#pragma unroll
      for (int u = 0; u < kRows; ++u)
        if (present[u]) present[u] = key_row_valid(p, base + u * step);
    }
    uint32_t cidx[kRows], loff[kRows];
    bool hit[kRows];
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      const uint64_t kk = (uint64_t)widen_int(key[u], kw, kvc);
      cidx[u] = lut[tiny_hash(kk, mult)];
      hit[u] = present[u] && cidx[u] != kTinyNone && ckeys[cidx[u] & (kTinyKeys - 1)] == kk && kk != kEmptyIntKey;
      cidx[u] &= kTinyKeys - 1;
      loff[u] = tiny_off + cidx[u] * (kTinyCopies * stride) + copy;
    }
#pragma unroll
    for (int u = 0; u < kRows; ++u)
      if (present[u] && !hit[u]) tiny_miss_row(p, t, base + u * step);
    // the row counts
    if (rows_off >= 0) {
      if (!collapse) {
#pragma unroll
        for (int u = 0; u < kRows; ++u) lds_add(reinterpret_cast<uint32_t*>(lds + (hit[u] ? loff[u] + rows_off : dum)), 1u);
      } else {
        bool h[kRows], ok[kRows];
        uint64_t unused[kRows];
        uint32_t n[kRows];
#pragma unroll
        for (int u = 0; u < kRows; ++u) {
          h[u] = hit[u];
          ok[u] = true;
          unused[u] = 0;
          n[u] = 0;
        }
        Program none{};
        tiny_collapse(none, cidx, h, ok, unused, n, true);
#pragma unroll
        for (int u = 0; u < kRows; ++u)
          if (h[u]) lds_add(reinterpret_cast<uint32_t*>(lds + loff[u] + rows_off), n[u]);
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const Program g = unpack_program(sread64(pin_prog[c]));
      const uint64_t vmask = sread64(pin_vmask[c]);
      const int64_t voff = (int64_t)sread64((uint64_t)pin_voff[c]);
      bool ok[kRows];
#pragma unroll
      for (int u = 0; u < kRows; ++u)
        ok[u] = hit[u] && bit_valid(reinterpret_cast<const uint32_t*>(vmask), voff + base + u * step);
      if (!collapse || g.mm_kind >= 0) {
        // (the collapse combines SUM and null counts; with MIN / MAX the
        // column takes the per-lane updates)
        if (g.nulls_off >= 0) {
#pragma unroll
          for (int u = 0; u < kRows; ++u)
            lds_add(reinterpret_cast<uint32_t*>(lds + (hit[u] && !ok[u] ? loff[u] + g.nulls_off : dum)), 1u);
        }
        uint32_t lo[kRows];
#pragma unroll
        for (int u = 0; u < kRows; ++u) lo[u] = hit[u] ? loff[u] : dum;
        if (c == 0 && TV != TV_ANY) tiny_column_fixed<TV>(lds, g, lo, x[c], ok, dum);
        else lds_column_program(lds, g, lo, x[c], ok, dum);
      } else {
        bool h[kRows];
        uint32_t n[kRows];
        uint64_t xs[kRows];
#pragma unroll
        for (int u = 0; u < kRows; ++u) {
          h[u] = hit[u];
          n[u] = 0;
          xs[u] = x[c][u];
        }
        tiny_collapse(g, cidx, h, ok, xs, n, false);
#pragma unroll
        for (int u = 0; u < kRows; ++u) {
          if (!h[u]) continue;
          if (g.nulls_off >= 0 && n[u]) lds_add(reinterpret_cast<uint32_t*>(lds + loff[u] + g.nulls_off), n[u]);
          if (g.sum_off >= 0 && ok[u]) {
            if (g.vclass == VC_F32 || g.vclass == VC_F64)
              lds_add(reinterpret_cast<double*>(lds + loff[u] + g.sum_off), __longlong_as_double((long long)xs[u]));
            else
              lds_add(reinterpret_cast<unsigned long long*>(lds + loff[u] + g.sum_off), (unsigned long long)xs[u]);
          }
        }
      }
    }
  }
  __syncthreads();
  // copies -> the candidates' records of the table
  for (int i = threadIdx.x; i < K * kTinyCopies; i += kLdsBlock) {
    const int g = i / kTinyCopies;
    if (coff[g] == kRecNone) continue;
    merge_record(p, lds + coff[g], lds + tiny_off + (uint32_t)i * stride);
  }
  __syncthreads();
}

}  // namespace mgb
