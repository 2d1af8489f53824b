// Device-side building blocks of the groupby kernels: MurmurHash3_x86_32
// (cudf and Spark variants), hash_combine, typed element loads, the
// order-preserving float <-> unsigned transform used for MIN / MAX, and the
// accumulator update ops.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mgb {

// ----------------------------------------------------------------- layout

constexpr int kMaxKeys = 8;     // key columns handled per launch
constexpr int kMaxVals = 16;    // value columns per aggregate launch
constexpr int kMaxOps = 48;     // accumulator ops per aggregate launch
constexpr int kMaxColOps = 8;   // accumulator ops of one value column (distinct kinds: at most 7)
constexpr int kProgramCols = 4; // value columns the LDS row loop handles unrolled
constexpr int kBlock = 256;     // threads per workgroup (4 wave64)

// Header of an empty record. Single integer keys store the key itself in the
// header (sign-extended), so the one 8-byte key equal to this value is routed
// to a dedicated record placed after the hashed region.
constexpr uint64_t kEmptyIntKey = 0x8000000000000000ull;
// Header of an empty record when the header holds (row hash << 32 | row).
constexpr uint64_t kEmptyRowRef = ~0ull;

enum KeyMode : int32_t { KEY_INT = 0, KEY_ROWREF = 1 };

// Value classes: how raw element bits are widened.
enum ValClass : int32_t { VC_SIGNED = 0, VC_UNSIGNED = 1, VC_BOOL = 2, VC_F32 = 3, VC_F64 = 4 };

// Accumulator ops. Each one updates one field of a group's record.
enum AccOp : int32_t {
  OP_ROWS = 0,    // u32  += 1 for every row of the group
  OP_COUNT,       // u32  += 1 for every valid value
  OP_SUM_I64,     // i64  += x          (integer sources, wraps)
  OP_SUM_F64,     // f64  += x          (float sources; f32 is widened)
  OP_SUMSQ_I64,   // i64  += x*x
  OP_SUMSQ_F64,   // f64  += x*x
  OP_PROD_I64,    // i64  *= x          (CAS loop)
  OP_PROD_F64,    // f64  *= x          (CAS loop)
  OP_MIN_I32, OP_MAX_I32, OP_MIN_U32, OP_MAX_U32,
  OP_MIN_I64, OP_MAX_I64, OP_MIN_U64, OP_MAX_U64,
  OP_MIN_F32, OP_MAX_F32,  // u32 order-preserving codes
  OP_MIN_F64, OP_MAX_F64,  // u64 order-preserving codes
  OP_NULLS,       // u32  += 1 for every null value (the row count minus it is COUNT_VALID)
  // second-pass fields, updated by a pass over the rows after the first pass
  // has completed the group's SUM / MIN / MAX:
  OP_M2,          // f64  += (x - mean)^2
  OP_ARGMIN_IDX,  // u32  = min row index whose value equals the group's MIN
  OP_ARGMAX_IDX,  // u32  = min row index whose value equals the group's MAX
  OP_NUM
};

struct KeyCol {
  const uint8_t* data;      // element 0 of the view (offset applied)
  const uint32_t* mask;     // nullptr when no nulls
  int64_t bit_off;          // mask bit of row 0
  int32_t type_id;
  int32_t width;            // bytes
  int32_t vclass;
  int32_t stride;           // bytes from one row to the next (width for a plain column)
};

struct ValCol {
  const uint8_t* data;
  const uint32_t* mask;
  int64_t bit_off;
  int32_t width;
  int32_t vclass;
  int32_t op_begin;         // ops[op_begin, op_end) read this column
  int32_t op_end;
  int32_t flag_bit;         // validity from bit `flag_bit` of the row's flags word, or -1
  int32_t nulls_off;        // OP_NULLS field of this column, or -1
  int32_t stride;           // bytes from one row to the next (width for a plain column)
  // the ops as the LDS path applies them (column_program): SUM, MIN and MAX
  // fields (-1: none), the MIN op kind naming the class of both, and 1 when
  // the column has other ops (PRODUCT, SUM_OF_SQUARES), which take the
  // general op-by-op path
  int32_t sum_off;
  int32_t min_off;
  int32_t max_off;
  int32_t mm_kind;
  int32_t lds_generic;
};

struct OpDesc {
  int32_t kind;
  int32_t off;              // byte offset inside the record
};

// ----------------------------------------------------------------- hashing

__device__ __host__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

__device__ __host__ __forceinline__ uint32_t fmix32(uint32_t h)
{
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

__device__ __host__ __forceinline__ uint32_t mix_k1(uint32_t k1)
{
  k1 *= 0xcc9e2d51u;
  k1 = rotl32(k1, 15);
  return k1 * 0x1b873593u;
}

__device__ __host__ __forceinline__ uint32_t mix_block(uint32_t h, uint32_t k1)
{
  h ^= mix_k1(k1);
  h = rotl32(h, 13);
  return h * 5u + 0xe6546b64u;
}

// MurmurHash3_x86_32 of an element of `width` bytes whose little-endian bits
// are `bits` (cudf variant: 1-3 byte tails are mixed as one block).
__device__ __host__ __forceinline__ uint32_t murmur3_bits(uint64_t bits, int width, uint32_t seed)
{
  uint32_t h = seed;
  if (width >= 4) h = mix_block(h, (uint32_t)bits);
  if (width == 8) h = mix_block(h, (uint32_t)(bits >> 32));
  if (width < 4) {
    uint32_t k1 = (uint32_t)bits & (width == 1 ? 0xffu : 0xffffu);
    h ^= mix_k1(k1);
  }
  h ^= (uint32_t)width;
  return fmix32(h);
}

// Spark variant for 4- and 8-byte inputs (narrower types are widened to 4
// bytes before hashing, so the signed-tail rule never applies here).
__device__ __host__ __forceinline__ uint32_t spark_murmur3_bits(uint64_t bits, int width, uint32_t seed)
{
  uint32_t h = mix_block(seed, (uint32_t)bits);
  if (width == 8) h = mix_block(h, (uint32_t)(bits >> 32));
  h ^= (uint32_t)width;
  return fmix32(h);
}

__device__ __host__ __forceinline__ uint32_t hash_combine(uint32_t lhs, uint32_t rhs)
{
  return lhs ^ (rhs + 0x9e3779b9u + (lhs << 6) + (lhs >> 2));
}

// ----------------------------------------------------------------- loads

// Global-memory pointers. Column and output pointers read from memory (not
// from kernel arguments) are flat pointers to the compiler, and accesses
// through them flat accesses that also hold the LDS counter; loads and
// stores of column data go through to_global(...) instead.
template <typename T>
using global_t = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ global_t<T>* to_global(T* p)
{
  return (global_t<T>*)p;
}

__device__ __forceinline__ uint64_t load_raw(const uint8_t* base, int width, int64_t row)
{
  const global_t<const uint8_t>* g = to_global(base);
  switch (width) {
    case 1: return g[row];
    case 2: return reinterpret_cast<const global_t<uint16_t>*>(g)[row];
    case 4: return reinterpret_cast<const global_t<uint32_t>*>(g)[row];
    default: return reinterpret_cast<const global_t<uint64_t>*>(g)[row];
  }
}

// Element `row` of a column whose rows are `stride` bytes apart.
__device__ __forceinline__ uint64_t load_at(const uint8_t* base, int width, int stride, int64_t row)
{
  const global_t<const uint8_t>* q = to_global(base) + row * (int64_t)stride;
  switch (width) {
    case 1: return *q;
    case 2: return *reinterpret_cast<const global_t<uint16_t>*>(q);
    case 4: return *reinterpret_cast<const global_t<uint32_t>*>(q);
    default: return *reinterpret_cast<const global_t<uint64_t>*>(q);
  }
}

__device__ __forceinline__ bool bit_valid(const uint32_t* mask, int64_t bit)
{
  return mask == nullptr || ((to_global(mask)[bit >> 5] >> (bit & 31)) & 1u);
}

// Raw bits widened to a 64-bit integer (sign- or zero-extended, bool -> 0/1).
__device__ __forceinline__ int64_t widen_int(uint64_t raw, int width, int vclass)
{
  if (vclass == VC_BOOL) return raw != 0;
  if (vclass == VC_SIGNED) {
    switch (width) {
      case 1: return (int8_t)raw;
      case 2: return (int16_t)raw;
      case 4: return (int32_t)raw;
      default: return (int64_t)raw;
    }
  }
  return (int64_t)raw;
}

__device__ __forceinline__ double widen_f64(uint64_t raw, int width, int vclass)
{
  if (vclass == VC_F32) return (double)__uint_as_float((uint32_t)raw);
  if (vclass == VC_F64) return __longlong_as_double((long long)raw);
  if (vclass == VC_UNSIGNED && width == 8) return (double)raw;
  return (double)widen_int(raw, width, vclass);
}

// Floats canonicalised for hashing and equality: -0.0 -> +0.0, NaN -> qNaN
// (cpp/include/cudf/hashing/detail/hash_functions.cuh:29-47).
__device__ __host__ __forceinline__ uint64_t normalize_float_bits(uint64_t raw, int vclass)
{
  if (vclass == VC_F32) {
    uint32_t b = (uint32_t)raw;
    if ((b & 0x7fffffffu) == 0) return 0;
    if ((b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu)) return 0x7fc00000u;
    return b;
  }
  if (vclass == VC_F64) {
    if ((raw & 0x7fffffffffffffffull) == 0) return 0;
    if ((raw & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (raw & 0x000fffffffffffffull))
      return 0x7ff8000000000000ull;
    return raw;
  }
  if (vclass == VC_BOOL) return raw != 0;
  return raw;
}

// ------------------------------------------------------- ordered float codes
// Monotone map float -> unsigned: negative floats have all bits flipped,
// non-negative floats get the sign bit set. NaN never reaches it: MIN maps
// NaN to (identity - 1) and MAX to (identity + 1), so a NaN wins only when
// every valid value of the group is NaN.

__device__ __host__ __forceinline__ uint32_t f32_code(uint32_t b) { return (b & 0x80000000u) ? ~b : (b | 0x80000000u); }
__device__ __host__ __forceinline__ uint32_t f32_decode(uint32_t c) { return (c & 0x80000000u) ? (c & 0x7fffffffu) : ~c; }
__device__ __host__ __forceinline__ unsigned long long f64_code(unsigned long long b)
{
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __host__ __forceinline__ uint64_t f64_decode(uint64_t c)
{
  return (c & 0x8000000000000000ull) ? (c & 0x7fffffffffffffffull) : ~c;
}

constexpr uint32_t kF32MinIdentity = 0xffffffffu, kF32MinNaN = 0xfffffffeu;
constexpr uint32_t kF32MaxIdentity = 0u, kF32MaxNaN = 1u;
constexpr unsigned long long kF64MinIdentity = ~0ull, kF64MinNaN = ~0ull - 1;
constexpr unsigned long long kF64MaxIdentity = 0ull, kF64MaxNaN = 1ull;

// ----------------------------------------------------------------- atomics

template <typename T>
__device__ __forceinline__ void atom_add(T* p, T v)
{
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void atom_min(T* p, T v)
{
  __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void atom_max(T* p, T v)
{
  __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long atom_cas(unsigned long long* p, unsigned long long expected,
                                                       unsigned long long desired)
{
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expected;
}

// Applies one accumulator op for a valid value (raw bits of a column of the
// given width / class). OP_ROWS and OP_COUNT ignore the value.
__device__ __forceinline__ void apply_op(uint8_t* rec, int kind, int off, uint64_t raw, int width, int vclass)
{
  uint8_t* f = rec + off;
  switch (kind) {
    case OP_ROWS:
    case OP_COUNT: atom_add(reinterpret_cast<uint32_t*>(f), 1u); break;
    case OP_SUM_I64:
      atom_add(reinterpret_cast<unsigned long long*>(f), (unsigned long long)widen_int(raw, width, vclass));
      break;
    case OP_SUM_F64: atom_add(reinterpret_cast<double*>(f), widen_f64(raw, width, vclass)); break;
    case OP_SUMSQ_I64: {
      unsigned long long x = (unsigned long long)widen_int(raw, width, vclass);
      atom_add(reinterpret_cast<unsigned long long*>(f), x * x);
      break;
    }
    case OP_SUMSQ_F64: {
      double x = widen_f64(raw, width, vclass);
      if (vclass == VC_F32) {
        float xf = (float)x;
        x = (double)(xf * xf);
      } else {
        x = x * x;
      }
      atom_add(reinterpret_cast<double*>(f), x);
      break;
    }
    case OP_PROD_I64: {
      unsigned long long* p = reinterpret_cast<unsigned long long*>(f);
      unsigned long long x = (unsigned long long)widen_int(raw, width, vclass);
      unsigned long long old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (;;) {
        unsigned long long seen = atom_cas(p, old, old * x);
        if (seen == old) break;
        old = seen;
      }
      break;
    }
    case OP_PROD_F64: {
      unsigned long long* p = reinterpret_cast<unsigned long long*>(f);
      double x = widen_f64(raw, width, vclass);
      unsigned long long old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (;;) {
        double cur = __longlong_as_double((long long)old);
        double nxt = (vclass == VC_F32) ? (double)((float)cur * (float)x) : cur * x;
        unsigned long long seen = atom_cas(p, old, (unsigned long long)__double_as_longlong(nxt));
        if (seen == old) break;
        old = seen;
      }
      break;
    }
    case OP_MIN_I32: atom_min(reinterpret_cast<int32_t*>(f), (int32_t)widen_int(raw, width, vclass)); break;
    case OP_MAX_I32: atom_max(reinterpret_cast<int32_t*>(f), (int32_t)widen_int(raw, width, vclass)); break;
    case OP_MIN_U32: atom_min(reinterpret_cast<uint32_t*>(f), (uint32_t)widen_int(raw, width, vclass)); break;
    case OP_MAX_U32: atom_max(reinterpret_cast<uint32_t*>(f), (uint32_t)widen_int(raw, width, vclass)); break;
    case OP_MIN_I64: atom_min(reinterpret_cast<long long*>(f), (long long)widen_int(raw, width, vclass)); break;
    case OP_MAX_I64: atom_max(reinterpret_cast<long long*>(f), (long long)widen_int(raw, width, vclass)); break;
    case OP_MIN_U64: atom_min(reinterpret_cast<unsigned long long*>(f), (unsigned long long)raw); break;
    case OP_MAX_U64: atom_max(reinterpret_cast<unsigned long long*>(f), (unsigned long long)raw); break;
    case OP_MIN_F32: {
      uint32_t b = (uint32_t)raw;
      bool nan = (b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu);
      atom_min(reinterpret_cast<uint32_t*>(f), nan ? kF32MinNaN : f32_code(b));
      break;
    }
    case OP_MAX_F32: {
      uint32_t b = (uint32_t)raw;
      bool nan = (b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu);
      atom_max(reinterpret_cast<uint32_t*>(f), nan ? kF32MaxNaN : f32_code(b));
      break;
    }
    case OP_MIN_F64: {
      bool nan = (raw & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (raw & 0x000fffffffffffffull);
      atom_min(reinterpret_cast<unsigned long long*>(f), nan ? kF64MinNaN : f64_code(raw));
      break;
    }
    case OP_MAX_F64: {
      bool nan = (raw & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (raw & 0x000fffffffffffffull);
      atom_max(reinterpret_cast<unsigned long long*>(f), nan ? kF64MaxNaN : f64_code(raw));
      break;
    }
    default: break;
  }
}

// The field value a MIN / MAX op of `kind` would take from this one value
// (apply_op's operand): the group's field equals it exactly when the value is
// the group's minimum / maximum in that op's order (NaN below every number for
// MAX and above every number for MIN, -0.0 below +0.0).
__device__ __forceinline__ uint64_t minmax_code(int kind, uint64_t raw, int width, int vclass)
{
  switch (kind) {
    case OP_MIN_I32: case OP_MAX_I32: return (uint32_t)(int32_t)widen_int(raw, width, vclass);
    case OP_MIN_U32: case OP_MAX_U32: return (uint32_t)widen_int(raw, width, vclass);
    case OP_MIN_I64: case OP_MAX_I64: return (uint64_t)widen_int(raw, width, vclass);
    case OP_MIN_U64: case OP_MAX_U64: return raw;
    case OP_MIN_F32: case OP_MAX_F32: {
      uint32_t b = (uint32_t)raw;
      bool nan = (b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu);
      return nan ? (kind == OP_MIN_F32 ? kF32MinNaN : kF32MaxNaN) : f32_code(b);
    }
    case OP_MIN_F64: case OP_MAX_F64: {
      bool nan = (raw & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (raw & 0x000fffffffffffffull);
      return nan ? (kind == OP_MIN_F64 ? kF64MinNaN : kF64MaxNaN) : f64_code(raw);
    }
    default: return 0ull;
  }
}

// Identity value of an op's field, written by the table fill.
__device__ __host__ __forceinline__ uint64_t op_identity(int kind)
{
  switch (kind) {
    case OP_PROD_I64: return 1ull;
    case OP_PROD_F64: return 0x3ff0000000000000ull;  // 1.0
    case OP_MIN_I32: return 0x7fffffffull;
    case OP_MAX_I32: return 0x80000000ull;
    case OP_MIN_U32: return 0xffffffffull;
    case OP_MAX_U32: return 0ull;
    case OP_MIN_I64: return 0x7fffffffffffffffull;
    case OP_MAX_I64: return 0x8000000000000000ull;
    case OP_MIN_U64: return ~0ull;
    case OP_MAX_U64: return 0ull;
    case OP_MIN_F32: return kF32MinIdentity;
    case OP_MAX_F32: return kF32MaxIdentity;
    case OP_MIN_F64: return kF64MinIdentity;
    case OP_MAX_F64: return kF64MaxIdentity;
    case OP_ARGMIN_IDX: case OP_ARGMAX_IDX: return 0xffffffffull;
    default: return 0ull;
  }
}

__device__ __host__ __forceinline__ int op_field_bytes(int kind)
{
  switch (kind) {
    case OP_ROWS: case OP_COUNT: case OP_MIN_I32: case OP_MAX_I32: case OP_MIN_U32: case OP_MAX_U32:
    case OP_MIN_F32: case OP_MAX_F32: case OP_NULLS: case OP_ARGMIN_IDX: case OP_ARGMAX_IDX: return 4;
    default: return 8;
  }
}

}  // namespace mgb
