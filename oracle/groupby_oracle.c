/*
 * groupby_oracle.c -- CPU restatement of libcudf 24.04's hash groupby-aggregate
 * path and of the MurmurHash3_x86_32 row hashers it uses.
 *
 * THIS IS TEST INFRASTRUCTURE. It is the checker that the HIP product path is
 * compared against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * What it restates (citations relative to the cudf source tree):
 *   - MurmurHash3_x86_32 (cudf variant):
 *       cpp/include/cudf/hashing/detail/murmurhash3_x86_32.cuh:41-130
 *   - Spark variant (signed tail bytes, small ints widened to 4 bytes,
 *     seed chaining over columns, nulls keep the running hash):
 *       cpp/src/hash/spark_murmurhash3_x86_32.cu:42-300
 *   - hash_combine: cpp/include/cudf/hashing/detail/hashing.hpp:90-93
 *   - NaN / -0.0 normalisation: cpp/include/cudf/hashing/detail/hash_functions.cuh:29-47
 *   - row hash = fold of hash_combine over key columns starting at the seed,
 *     null element -> 0xFFFFFFFF:
 *       cpp/include/cudf/table/experimental/row_operators.cuh:1799-1883
 *   - row equality with NaN == NaN and null == null:
 *       cpp/include/cudf/table/experimental/row_operators.cuh:1285-1364
 *   - the hash groupby driver: key-row validity (bitmask_and of the key
 *     columns when nulls are excluded), insert-or-find over row indices, one
 *     update per flattened aggregation into sparse targets indexed by the
 *     group's representative row, compaction of the populated keys and the
 *     MEAN = SUM / COUNT_VALID finalisation:
 *       cpp/src/groupby/hash/groupby.cu:351-383 (flatten), 424-454 (sparse
 *       targets), 460-504 (single pass), 510-520 (populated keys),
 *       287-306 (MEAN), 548-620 (driver)
 *   - element update rules and identities:
 *       cpp/include/cudf/detail/aggregation/aggregation.cuh:131-434
 *       cpp/include/cudf/detail/utilities/device_operators.cuh:63-200
 *       cpp/include/cudf/fixed_point/temporary.hpp:62-72 (min / max ops)
 *   - target types: cpp/include/cudf/detail/aggregation/aggregation.hpp:1168-1424
 *   - the sort-path aggregations (MEDIAN, QUANTILE, NUNIQUE, NTH_ELEMENT,
 *     COLLECT_LIST, COLLECT_SET) over each group's rows in row order:
 *       cpp/src/groupby/sort/aggregate.cpp, group_quantiles.cu,
 *       group_nunique.cu, group_nth_element.cu, group_collect.cu,
 *       cpp/src/lists/stream_compaction/distinct.cu
 *
 * The rows are processed in order, so float sums are the naive left-to-right
 * sums and float MIN/MAX follow the CAS-loop operator `old < x ? old : x`
 * applied in row order.
 *
 * Pinning: the murmur core is checked against SMHasher's published
 * verification value for MurmurHash3_x86_32 (0xB0F57EE3) and against the
 * Spark golden vectors in cpp/tests/hashing/spark_murmurhash3_x86_32_test.cpp;
 * hash_combine against the list-column vectors of
 * cpp/tests/hashing/murmurhash3_x86_32_test.cpp:123-160; the groupby against
 * the literal fixtures of the cpp/tests/groupby test files, brute force and pandas;
 * COLLECT_LIST / COLLECT_SET against the fixtures of
 * cpp/tests/groupby/collect_list_tests.cpp and collect_set_tests.cpp and
 * against set and list identities on random inputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* cudf::type_id numbering (cpp/include/cudf/types.hpp:201-233). */
enum {
  T_EMPTY = 0, T_INT8, T_INT16, T_INT32, T_INT64, T_UINT8, T_UINT16, T_UINT32, T_UINT64,
  T_FLOAT32, T_FLOAT64, T_BOOL8, T_TS_D, T_TS_S, T_TS_MS, T_TS_US, T_TS_NS,
  T_DUR_D, T_DUR_S, T_DUR_MS, T_DUR_US, T_DUR_NS
};

/* cudf::aggregation::Kind numbering (cpp/include/cudf/aggregation.hpp:85-122). */
enum {
  K_SUM = 0, K_PRODUCT = 1, K_MIN = 2, K_MAX = 3, K_COUNT_VALID = 4, K_COUNT_ALL = 5,
  K_SUM_OF_SQUARES = 8, K_MEAN = 9, K_M2 = 10, K_VARIANCE = 11, K_STD = 12,
  K_MEDIAN = 13, K_QUANTILE = 14, K_ARGMAX = 15, K_ARGMIN = 16, K_NUNIQUE = 17, K_NTH_ELEMENT = 18,
  K_COLLECT_LIST = 21, K_COLLECT_SET = 22
};
enum { T_LIST = 24 };

typedef struct {
  int32_t type_id;
  const void* data;
  const uint32_t* mask; /* LSB-first words, bit = 1 valid; NULL = all valid */
} mgo_col;

static int type_size(int t)
{
  switch (t) {
    case T_INT8: case T_UINT8: case T_BOOL8: return 1;
    case T_INT16: case T_UINT16: return 2;
    case T_INT32: case T_UINT32: case T_FLOAT32: case T_TS_D: case T_DUR_D: return 4;
    case T_INT64: case T_UINT64: case T_FLOAT64: case T_TS_S: case T_TS_MS: case T_TS_US:
    case T_TS_NS: case T_DUR_S: case T_DUR_MS: case T_DUR_US: case T_DUR_NS: return 8;
    default: return 0;
  }
}
static int is_float(int t) { return t == T_FLOAT32 || t == T_FLOAT64; }
static int is_signed_int(int t) { return t == T_INT8 || t == T_INT16 || t == T_INT32 || t == T_INT64; }
static int is_unsigned_int(int t) { return t == T_UINT8 || t == T_UINT16 || t == T_UINT32 || t == T_UINT64; }
static int is_integral(int t) { return is_signed_int(t) || is_unsigned_int(t) || t == T_BOOL8; }
static int is_timestamp(int t) { return t >= T_TS_D && t <= T_TS_NS; }
static int is_duration(int t) { return t >= T_DUR_D && t <= T_DUR_NS; }

static int bit_is_set(const uint32_t* m, int64_t i) { return m == NULL || ((m[i >> 5] >> (i & 31)) & 1u); }

/* ---------------------------------------------------------------- murmur3 */

static uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
static uint32_t fmix32(uint32_t h)
{
  h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
  return h;
}
static uint32_t mix_k1(uint32_t k1) { k1 *= 0xcc9e2d51u; k1 = rotl32(k1, 15); return k1 * 0x1b873593u; }

/* cudf / SMHasher MurmurHash3_x86_32 over bytes (murmurhash3_x86_32.cuh:93-121). */
uint32_t mgo_murmur3_x86_32(const void* key, int32_t len, uint32_t seed)
{
  const uint8_t* d = (const uint8_t*)key;
  int32_t nblocks = len / 4;
  uint32_t h = seed;
  for (int32_t i = 0; i < nblocks; ++i) {
    uint32_t k1 = (uint32_t)d[4 * i] | ((uint32_t)d[4 * i + 1] << 8) |
                  ((uint32_t)d[4 * i + 2] << 16) | ((uint32_t)d[4 * i + 3] << 24);
    h ^= mix_k1(k1);
    h = rotl32(h, 13);
    h = h * 5u + 0xe6546b64u;
  }
  const uint8_t* tail = d + 4 * nblocks;
  uint32_t k1 = 0;
  switch (len & 3) {
    case 3: k1 ^= (uint32_t)tail[2] << 16; /* fall through */
    case 2: k1 ^= (uint32_t)tail[1] << 8;  /* fall through */
    case 1: k1 ^= tail[0]; h ^= mix_k1(k1);
  }
  h ^= (uint32_t)len;
  return fmix32(h);
}

/* Spark variant: tail bytes are sign-extended and mixed one at a time
 * (spark_murmurhash3_x86_32.cu:78-96). */
uint32_t mgo_spark_murmur3_x86_32(const void* key, int32_t len, uint32_t seed)
{
  const uint8_t* d = (const uint8_t*)key;
  int32_t nblocks = len / 4;
  uint32_t h = seed;
  for (int32_t i = 0; i < nblocks; ++i) {
    uint32_t k1 = (uint32_t)d[4 * i] | ((uint32_t)d[4 * i + 1] << 8) |
                  ((uint32_t)d[4 * i + 2] << 16) | ((uint32_t)d[4 * i + 3] << 24);
    h ^= mix_k1(k1);
    h = rotl32(h, 13);
    h = h * 5u + 0xe6546b64u;
  }
  for (int32_t i = 4 * nblocks; i < len; ++i) {
    uint32_t k1 = (uint32_t)(int32_t)(int8_t)d[i];
    h ^= mix_k1(k1);
    h = rotl32(h, 13);
    h = h * 5u + 0xe6546b64u;
  }
  h ^= (uint32_t)len;
  return fmix32(h);
}

/* Boost-style combine (hashing.hpp:90-93). */
uint32_t mgo_hash_combine(uint32_t lhs, uint32_t rhs)
{
  return lhs ^ (rhs + 0x9e3779b9u + (lhs << 6) + (lhs >> 2));
}

/* Element hash of the cudf variant: floats normalised (-0 -> +0, NaN -> qNaN),
 * bool hashed as one byte, every other type over its storage bytes. */
static uint32_t element_hash(int t, const void* base, int64_t i, uint32_t seed)
{
  int sz = type_size(t);
  const uint8_t* p = (const uint8_t*)base + (size_t)i * sz;
  if (t == T_FLOAT32) {
    float f; memcpy(&f, p, 4);
    if (f == 0.0f) f = 0.0f;
    if (isnan(f)) f = NAN;
    return mgo_murmur3_x86_32(&f, 4, seed);
  }
  if (t == T_FLOAT64) {
    double f; memcpy(&f, p, 8);
    if (f == 0.0) f = 0.0;
    if (isnan(f)) f = (double)NAN;
    return mgo_murmur3_x86_32(&f, 8, seed);
  }
  if (t == T_BOOL8) { uint8_t b = p[0] != 0; return mgo_murmur3_x86_32(&b, 1, seed); }
  return mgo_murmur3_x86_32(p, sz, seed);
}

static uint32_t row_hash(const mgo_col* cols, int32_t ncols, int64_t i, uint32_t seed, int check_nulls)
{
  uint32_t h = seed;
  for (int32_t c = 0; c < ncols; ++c) {
    uint32_t e = (check_nulls && !bit_is_set(cols[c].mask, i)) ? 0xFFFFFFFFu
                                                                : element_hash(cols[c].type_id, cols[c].data, i, seed);
    h = mgo_hash_combine(h, e);
  }
  return h;
}

/* cudf::hashing::murmurhash3_x86_32(table, seed) for flat columns. */
int mgo_murmurhash3_x86_32_table(const mgo_col* cols, int32_t ncols, int64_t nrows, uint32_t seed, uint32_t* out)
{
  int check_nulls = 0;
  for (int32_t c = 0; c < ncols; ++c) {
    if (type_size(cols[c].type_id) == 0) return 1;
    if (cols[c].mask) check_nulls = 1;
  }
  for (int64_t i = 0; i < nrows; ++i) out[i] = row_hash(cols, ncols, i, seed, check_nulls);
  return 0;
}

/* cudf::hashing::spark_murmurhash3_x86_32(table, seed) for flat columns. */
int mgo_spark_murmurhash3_x86_32_table(const mgo_col* cols, int32_t ncols, int64_t nrows, uint32_t seed, int32_t* out)
{
  for (int32_t c = 0; c < ncols; ++c)
    if (type_size(cols[c].type_id) == 0) return 1;
  for (int64_t i = 0; i < nrows; ++i) {
    uint32_t h = seed;
    for (int32_t c = 0; c < ncols; ++c) {
      if (!bit_is_set(cols[c].mask, i)) continue; /* null keeps the running hash */
      int t = cols[c].type_id, sz = type_size(t);
      const uint8_t* p = (const uint8_t*)cols[c].data + (size_t)i * sz;
      uint32_t w;
      switch (t) {
        case T_BOOL8: w = p[0] != 0; h = mgo_spark_murmur3_x86_32(&w, 4, h); break;  /* bool reads as 0/1 */
        case T_UINT8: w = p[0]; h = mgo_spark_murmur3_x86_32(&w, 4, h); break;
        case T_INT8: w = (uint32_t)(int32_t)(int8_t)p[0]; h = mgo_spark_murmur3_x86_32(&w, 4, h); break;
        case T_INT16: { int16_t s; memcpy(&s, p, 2); w = (uint32_t)(int32_t)s; h = mgo_spark_murmur3_x86_32(&w, 4, h); break; }
        case T_UINT16: { uint16_t s; memcpy(&s, p, 2); w = s; h = mgo_spark_murmur3_x86_32(&w, 4, h); break; }
        case T_FLOAT32: { float f; memcpy(&f, p, 4); if (isnan(f)) f = NAN; h = mgo_spark_murmur3_x86_32(&f, 4, h); break; }
        case T_FLOAT64: { double f; memcpy(&f, p, 8); if (isnan(f)) f = (double)NAN; h = mgo_spark_murmur3_x86_32(&f, 8, h); break; }
        default: h = mgo_spark_murmur3_x86_32(p, sz, h);
      }
    }
    out[i] = (int32_t)h;
  }
  return 0;
}

/* ---------------------------------------------------------------- groupby */

static double load_f(int t, const void* base, int64_t i)
{
  if (t == T_FLOAT32) return (double)((const float*)base)[i];
  return ((const double*)base)[i];
}
/* Integer load widened to int64 (uint64 values are carried in the same bits). */
static int64_t load_i(int t, const void* base, int64_t i)
{
  switch (type_size(t)) {
    case 1: return is_signed_int(t) ? (int64_t)((const int8_t*)base)[i] : (int64_t)((const uint8_t*)base)[i];
    case 2: return is_signed_int(t) ? (int64_t)((const int16_t*)base)[i] : (int64_t)((const uint16_t*)base)[i];
    case 4: return (is_signed_int(t) || is_timestamp(t) || is_duration(t)) ? (int64_t)((const int32_t*)base)[i]
                                                                          : (int64_t)((const uint32_t*)base)[i];
    default: return ((const int64_t*)base)[i];
  }
}

static int elements_equal(int t, const void* base, int64_t a, int64_t b)
{
  if (is_float(t)) {
    double x = load_f(t, base, a), y = load_f(t, base, b);
    return (isnan(x) && isnan(y)) ? 1 : x == y;
  }
  int sz = type_size(t);
  if (t == T_BOOL8) return (((const uint8_t*)base)[a] != 0) == (((const uint8_t*)base)[b] != 0);
  return memcmp((const uint8_t*)base + (size_t)a * sz, (const uint8_t*)base + (size_t)b * sz, sz) == 0;
}

static int rows_equal(const mgo_col* k, int32_t nk, int64_t a, int64_t b)
{
  for (int32_t c = 0; c < nk; ++c) {
    int va = bit_is_set(k[c].mask, a), vb = bit_is_set(k[c].mask, b);
    if (va != vb) return 0;
    if (!va) continue; /* null == null (null_equality::EQUAL) */
    if (!elements_equal(k[c].type_id, k[c].data, a, b)) return 0;
  }
  return 1;
}

/* Target type of (source type, kind); -1 when the pair is invalid
 * (aggregation.hpp:1168-1424). */
int mgo_target_type(int t, int kind)
{
  if (type_size(t) == 0) return -1;
  switch (kind) {
    case K_MIN: case K_MAX: return t;
    case K_COUNT_VALID: case K_COUNT_ALL: return T_INT32;
    case K_ARGMIN: case K_ARGMAX: return T_INT32;
    /* sort-path aggregations (aggregation.hpp:1290-1330): quantiles of numeric
     * sources are double, NUNIQUE is size_type, NTH_ELEMENT keeps the type */
    case K_MEDIAN: case K_QUANTILE:
      if (!is_integral(t) && !is_float(t)) return -1;
      return T_FLOAT64;
    case K_NUNIQUE: return T_INT32;
    case K_NTH_ELEMENT: return t;
    /* a list of the source type (aggregation.hpp:1335-1350) */
    case K_COLLECT_LIST: case K_COLLECT_SET: return T_LIST;
    case K_SUM:
      if (is_integral(t)) return T_INT64;
      if (is_float(t) || is_duration(t)) return t;
      return -1;
    case K_PRODUCT: case K_SUM_OF_SQUARES:
      if (is_integral(t)) return T_INT64;
      if (is_float(t)) return t;
      return -1;
    case K_MEAN:
      if (is_duration(t)) return t;
      if (is_timestamp(t)) return -1;
      return T_FLOAT64;
    case K_M2: case K_VARIANCE: case K_STD:
      /* the target is double for every source (aggregation.hpp:1272-1289), but
       * the hash path's functor accepts numeric sources only
       * (multi_pass_kernels.cuh:59-71: other types reach CUDF_UNREACHABLE) */
      if (!is_integral(t) && !is_float(t)) return -1;
      return T_FLOAT64;
    default: return -1;
  }
}

/* Per-group scratch for one request. Integers accumulate in int64 (uint64 in
 * the same bits, wrapping like the device's two's complement atomics), floats
 * in the source type for SUM/PRODUCT and MIN/MAX, MEAN's sum in its SUM target. */
typedef struct {
  int64_t* i;
  double* f;
  float* f32;
  int64_t* cnt;
  double* m2;
  uint8_t* valid;
} acc_t;

/*
 * keys/values: host columns; requests: (value column index, kind, ddof).
 * Outputs: out_rep_rows[g] = representative (first) row of group g, in order of
 * first appearance; out_data[j] / out_valid[j] = request j's result per group
 * (capacity num_rows each). Returns 0, or 1 for an invalid type/aggregation pair.
 */
/* ------------------------------------------------ sort-path aggregations */

/* Value order of cudf's sort of a column (ascending, NaN after every number;
 * -0.0 and +0.0 equal): the context of the qsort comparator below. */
static const mgo_col* g_sort_col;
static int value_cmp(int64_t a, int64_t b)
{
  int t = g_sort_col->type_id;
  if (is_float(t)) {
    double x = load_f(t, g_sort_col->data, a), y = load_f(t, g_sort_col->data, b);
    int nx = isnan(x), ny = isnan(y);
    if (nx || ny) return nx - ny;
    return x < y ? -1 : x > y ? 1 : 0;
  }
  if (is_unsigned_int(t)) {
    uint64_t x = (uint64_t)load_i(t, g_sort_col->data, a), y = (uint64_t)load_i(t, g_sort_col->data, b);
    return x < y ? -1 : x > y ? 1 : 0;
  }
  int64_t x = load_i(t, g_sort_col->data, a), y = load_i(t, g_sort_col->data, b);
  return x < y ? -1 : x > y ? 1 : 0;
}
static int row_cmp(const void* pa, const void* pb)
{
  int64_t a = *(const int64_t*)pa, b = *(const int64_t*)pb;
  int c = value_cmp(a, b);
  return c ? c : (a < b ? -1 : a > b);
}
static double value_f64(const mgo_col* v, int64_t r)
{
  return is_float(v->type_id) ? load_f(v->type_id, v->data, r)
         : is_unsigned_int(v->type_id) ? (double)(uint64_t)load_i(v->type_id, v->data, r)
                                       : (double)load_i(v->type_id, v->data, r);
}

/* One quantile of the sorted values rows[0..n) (quantiles_util.hpp:90-105
 * quantile_index, :161-181 select_quantile_data, :40-80 interpolation). */
static double select_quantile(const mgo_col* v, const int64_t* rows, int64_t n, double q, int interp)
{
  if (q < 0.0) q = 0.0;
  if (q > 1.0) q = 1.0;
  double val = q * (double)(n - 1);
  int64_t lo = (int64_t)floor(val), hi = (int64_t)ceil(val), ne = (int64_t)nearbyint(val);
  double frac = val - (double)lo;
  switch (interp) {
    case 1: return value_f64(v, rows[lo]);   /* LOWER */
    case 2: return value_f64(v, rows[hi]);   /* HIGHER */
    case 4: return value_f64(v, rows[ne]);   /* NEAREST */
    case 3:                                  /* MIDPOINT */
      if (v->type_id == T_INT64) {
        int64_t a = load_i(T_INT64, v->data, rows[lo]), b = load_i(T_INT64, v->data, rows[hi]);
        int64_t half = a / 2 + b / 2, rest = a % 2 + b % 2;
        return (double)half + (double)rest * 0.5;
      }
      return value_f64(v, rows[lo]) / 2 + value_f64(v, rows[hi]) / 2;
    default: {                               /* LINEAR */
      double a = value_f64(v, rows[lo]), b = value_f64(v, rows[hi]);
      return (1.0 - frac) * a + frac * b;
    }
  }
}

/* MEDIAN, QUANTILE, NUNIQUE, NTH_ELEMENT of every group: grp_rows[grp_off[g]..
 * grp_off[g+1]) are group g's rows in row order. */
static void sort_path_agg(const mgo_col* v, int kind, int ddof, int nulls, int interp, double q, int64_t ng,
                          const int64_t* grp_off, const int64_t* grp_rows, void* out, uint8_t* valid)
{
  int64_t max_n = 1;
  for (int64_t g = 0; g < ng; ++g)
    if (grp_off[g + 1] - grp_off[g] > max_n) max_n = grp_off[g + 1] - grp_off[g];
  int64_t* buf = (int64_t*)malloc(sizeof(int64_t) * (size_t)max_n);
  int w = type_size(v->type_id);
  for (int64_t g = 0; g < ng; ++g) {
    const int64_t* rows = grp_rows + grp_off[g];
    int64_t size = grp_off[g + 1] - grp_off[g], n = 0, has_null = 0;
    if (kind == K_NTH_ELEMENT) {
      /* the n-th row (n < 0: from the end) of the group, of all its rows
       * (INCLUDE) or of its valid values (EXCLUDE) (group_nth_element.cu:60-120) */
      for (int64_t i = 0; i < size; ++i)
        if (nulls == 1 || bit_is_set(v->mask, rows[i])) buf[n++] = rows[i];
      int64_t idx = ddof < 0 ? n + ddof : ddof;
      int ok = idx >= 0 && idx < n && bit_is_set(v->mask, buf[idx]);
      valid[g] = (uint8_t)ok;
      memset((uint8_t*)out + g * w, 0, (size_t)w);
      if (ok) memcpy((uint8_t*)out + g * w, (const uint8_t*)v->data + buf[idx] * w, (size_t)w);
      continue;
    }
    for (int64_t i = 0; i < size; ++i) {
      if (bit_is_set(v->mask, rows[i])) buf[n++] = rows[i];
      else has_null = 1;
    }
    g_sort_col = v;
    qsort(buf, (size_t)n, sizeof(int64_t), row_cmp);
    if (kind == K_NUNIQUE) {
      /* distinct values in sorted order, NaN == NaN; nulls one more value
       * under INCLUDE (group_nunique.cu:60-70) */
      int32_t c = 0;
      for (int64_t i = 0; i < n; ++i) c += (i == 0 || value_cmp(buf[i - 1], buf[i]) != 0);
      if (nulls == 1 && has_null) c += 1;
      ((int32_t*)out)[g] = c;
      valid[g] = 1;
      continue;
    }
    /* MEDIAN (0.5, LINEAR) and QUANTILE: null without values (group_quantiles.cu:66-74) */
    valid[g] = (uint8_t)(n > 0);
    ((double*)out)[g] = n > 0 ? select_quantile(v, buf, n, kind == K_MEDIAN ? 0.5 : q, kind == K_MEDIAN ? 0 : interp)
                              : 0.0;
  }
  free(buf);
}

/* Groups of the key rows in order of first appearance: group_of_row[i] (-1
 * for a row dropped for a null key), out_rep_rows[g] = first row of group g;
 * returns the group count. Open addressing over row indices, linear probing. */
static int64_t form_groups(const mgo_col* keys, int32_t n_keys, int64_t num_rows, int32_t include_null_keys,
                           int64_t* out_rep_rows, int64_t* group_of_row)
{
  int keys_have_nulls = 0;
  for (int32_t c = 0; c < n_keys; ++c) keys_have_nulls |= keys[c].mask != NULL;
  int skip_null_rows = keys_have_nulls && !include_null_keys;

  int64_t cap = 16;
  while (cap < 2 * num_rows) cap <<= 1;
  int64_t* slots = (int64_t*)malloc(sizeof(int64_t) * cap);
  for (int64_t s = 0; s < cap; ++s) slots[s] = -1;

  int64_t ng = 0;
  for (int64_t i = 0; i < num_rows; ++i) {
    group_of_row[i] = -1;
    int valid = 1;
    if (skip_null_rows)
      for (int32_t c = 0; c < n_keys; ++c) valid &= bit_is_set(keys[c].mask, i);
    if (!valid) continue;
    uint32_t h = row_hash(keys, n_keys, i, 0u, keys_have_nulls);
    int64_t s = h & (cap - 1);
    for (;;) {
      if (slots[s] < 0) { slots[s] = i; out_rep_rows[ng] = i; group_of_row[i] = ng++; break; }
      if (rows_equal(keys, n_keys, slots[s], i)) { group_of_row[i] = group_of_row[slots[s]]; break; }
      s = (s + 1) & (cap - 1);
    }
  }
  free(slots);
  return ng;
}

/* Each group's rows in row order: grp_rows[grp_off[g]..grp_off[g+1]). */
static void group_rows(const int64_t* group_of_row, int64_t num_rows, int64_t ng, int64_t* grp_off,
                       int64_t* grp_rows)
{
  for (int64_t g = 0; g <= ng; ++g) grp_off[g] = 0;
  for (int64_t r = 0; r < num_rows; ++r)
    if (group_of_row[r] >= 0) grp_off[group_of_row[r] + 1]++;
  for (int64_t g = 0; g < ng; ++g) grp_off[g + 1] += grp_off[g];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ng ? ng : 1));
  for (int64_t g = 0; g < ng; ++g) fill[g] = grp_off[g];
  for (int64_t r = 0; r < num_rows; ++r)
    if (group_of_row[r] >= 0) grp_rows[fill[group_of_row[r]]++] = r;
  free(fill);
}

int mgo_groupby_aggregate_ex(const mgo_col* keys, int32_t n_keys, int64_t num_rows, int32_t include_null_keys,
                             const mgo_col* values, int32_t n_values, const int32_t* req_col,
                             const int32_t* req_kind, const int32_t* req_ddof, const int32_t* req_nulls,
                             const int32_t* req_interp, const double* req_q, int32_t n_req,
                             int64_t* out_rep_rows, int64_t* out_num_groups, void* const* out_data,
                             uint8_t* const* out_valid)
{
  (void)n_values;
  for (int32_t j = 0; j < n_req; ++j)
    if (mgo_target_type(values[req_col[j]].type_id, req_kind[j]) < 0) return 1;

  int64_t* group_of_row = (int64_t*)malloc(sizeof(int64_t) * (num_rows ? num_rows : 1));
  int64_t ng = form_groups(keys, n_keys, num_rows, include_null_keys, out_rep_rows, group_of_row);
  *out_num_groups = ng;

  size_t gsz = (size_t)(ng ? ng : 1);
  acc_t a;
  a.i = (int64_t*)malloc(sizeof(int64_t) * gsz);
  a.f = (double*)malloc(sizeof(double) * gsz);
  a.f32 = (float*)malloc(sizeof(float) * gsz);
  a.cnt = (int64_t*)malloc(sizeof(int64_t) * gsz);
  a.m2 = (double*)malloc(sizeof(double) * gsz);
  a.valid = (uint8_t*)malloc(gsz);

  /* each group's rows in row order (for the sort-path aggregations) */
  int64_t* grp_off = (int64_t*)calloc((size_t)ng + 1, sizeof(int64_t));
  int64_t* grp_rows = (int64_t*)malloc(sizeof(int64_t) * (size_t)(num_rows ? num_rows : 1));
  group_rows(group_of_row, num_rows, ng, grp_off, grp_rows);

  for (int32_t j = 0; j < n_req; ++j) {
    const mgo_col* v = &values[req_col[j]];
    int t = v->type_id, kind = req_kind[j];
    int tt = mgo_target_type(t, kind);
    int fl = is_float(t);
    if (kind == K_MEDIAN || kind == K_QUANTILE || kind == K_NUNIQUE || kind == K_NTH_ELEMENT) {
      sort_path_agg(v, kind, req_ddof ? req_ddof[j] : 0, req_nulls ? req_nulls[j] : 0,
                    req_interp ? req_interp[j] : 0, req_q ? req_q[j] : 0.5, ng, grp_off, grp_rows, out_data[j],
                    out_valid[j]);
      continue;
    }
    for (int64_t g = 0; g < ng; ++g) {
      a.cnt[g] = 0; a.valid[g] = 0; a.m2[g] = 0.0;
      /* identities (device_operators.cuh:63-200) */
      if (kind == K_PRODUCT) { a.i[g] = 1; a.f[g] = 1.0; a.f32[g] = 1.0f; }
      else if (kind == K_MIN) { a.i[g] = 0; a.f[g] = INFINITY; a.f32[g] = INFINITY; }
      else if (kind == K_MAX) { a.i[g] = 0; a.f[g] = -INFINITY; a.f32[g] = -INFINITY; }
      else if (kind == K_ARGMIN || kind == K_ARGMAX) { a.i[g] = -1; }
      else { a.i[g] = 0; a.f[g] = 0.0; a.f32[g] = 0.0f; }
    }
    for (int64_t r = 0; r < num_rows; ++r) {
      int64_t g = group_of_row[r];
      if (g < 0) continue;
      int vv = bit_is_set(v->mask, r);
      if (kind == K_COUNT_ALL) { a.cnt[g]++; continue; }
      if (!vv) continue; /* null source rows are skipped (aggregation.cuh:144,193,242,408) */
      a.cnt[g]++;
      switch (kind) {
        case K_COUNT_VALID: break;
        case K_SUM: case K_MEAN: case K_VARIANCE: case K_STD: case K_M2:
          if (t == T_FLOAT32) a.f32[g] += (float)load_f(t, v->data, r);
          else if (fl) a.f[g] += load_f(t, v->data, r);
          else a.i[g] = (int64_t)((uint64_t)a.i[g] + (uint64_t)load_i(t, v->data, r));
          break;
        case K_SUM_OF_SQUARES:
          if (t == T_FLOAT32) { float x = (float)load_f(t, v->data, r); a.f32[g] += x * x; }
          else if (fl) { double x = load_f(t, v->data, r); a.f[g] += x * x; }
          else { uint64_t x = (uint64_t)load_i(t, v->data, r); a.i[g] = (int64_t)((uint64_t)a.i[g] + x * x); }
          break;
        case K_PRODUCT:
          if (t == T_FLOAT32) a.f32[g] *= (float)load_f(t, v->data, r);
          else if (fl) a.f[g] *= load_f(t, v->data, r);
          else a.i[g] = (int64_t)((uint64_t)a.i[g] * (uint64_t)load_i(t, v->data, r));
          break;
        case K_MIN: case K_MAX: {
          int is_min = kind == K_MIN;
          if (fl) {
            /* The reference folds old < x ? old : x (device_operators.cuh:117-123,
             * fixed_point/temporary.hpp:63-72) in the order the atomics land, so a
             * NaN stands until the next number arrives and tied zeros keep the
             * last one. This oracle fixes the outcome of the order that brings the
             * NaN rows first: NaN only when every valid value of the group is NaN,
             * and -0.0 below +0.0. a.valid[g] marks a number seen. */
            double x = load_f(t, v->data, r);
            double o = t == T_FLOAT32 ? (double)a.f32[g] : a.f[g];
            double res;
            if (isnan(x)) res = a.valid[g] ? o : x;
            else if (!a.valid[g]) { res = x; a.valid[g] = 1; }
            else if (x == o) res = (is_min ? signbit(x) : !signbit(x)) ? x : o;
            else res = is_min ? (x < o ? x : o) : (x > o ? x : o);
            if (t == T_FLOAT32) a.f32[g] = (float)res;
            else a.f[g] = res;
          } else {
            int64_t x = load_i(t, v->data, r);
            if (a.cnt[g] == 1) a.i[g] = x;
            else if (is_unsigned_int(t)) {
              uint64_t o = (uint64_t)a.i[g], ux = (uint64_t)x;
              a.i[g] = (int64_t)(is_min ? (o < ux ? o : ux) : (o > ux ? o : ux));
            } else {
              int64_t o = a.i[g];
              a.i[g] = is_min ? (o < x ? o : x) : (o > x ? o : x);
            }
          }
          break;
        }
        case K_ARGMIN: case K_ARGMAX: {
          /* row index of the min / max valid element. The reference's CAS loop
           * (aggregation.cuh:436-490) keeps whichever tied row arrived first and,
           * for floats, lets a NaN win only when it arrives first; this oracle
           * fixes one of those outcomes: ties keep the smaller row index, NaN
           * counts only when every valid value of the group is NaN, and -0.0
           * orders below +0.0 (the order of the library's MIN / MAX). */
          int64_t cur = a.i[g];
          if (cur < 0) { a.i[g] = r; break; }
          int better;
          if (fl) {
            double x = load_f(t, v->data, r), o = load_f(t, v->data, cur);
            if (isnan(x)) better = 0;
            else if (isnan(o)) better = 1;
            else if (x == o) better = kind == K_ARGMIN ? (signbit(x) && !signbit(o)) : (!signbit(x) && signbit(o));
            else better = kind == K_ARGMIN ? (x < o) : (x > o);
          } else if (is_unsigned_int(t)) {
            uint64_t x = (uint64_t)load_i(t, v->data, r), o = (uint64_t)load_i(t, v->data, cur);
            better = kind == K_ARGMIN ? (x < o) : (x > o);
          } else {
            int64_t x = load_i(t, v->data, r), o = load_i(t, v->data, cur);
            better = kind == K_ARGMIN ? (x < o) : (x > o);
          }
          if (better) a.i[g] = r;
          break;
        }
      }
    }
    /* second pass for M2 / VARIANCE / STD (multi_pass_kernels.cuh:35-112) */
    if (kind == K_M2 || kind == K_VARIANCE || kind == K_STD) {
      for (int64_t r = 0; r < num_rows; ++r) {
        int64_t g = group_of_row[r];
        if (g < 0 || !bit_is_set(v->mask, r)) continue;
        double sum = (t == T_FLOAT32) ? (double)a.f32[g] : fl ? a.f[g] : (double)a.i[g];
        double mean = sum / (double)a.cnt[g];
        double x = fl ? load_f(t, v->data, r) : (double)load_i(t, v->data, r);
        a.m2[g] += (x - mean) * (x - mean);
      }
    }
    /* finalise into the dense result */
    for (int64_t g = 0; g < ng; ++g) {
      int has = a.cnt[g] > 0;
      uint8_t ok = 1;
      switch (kind) {
        case K_COUNT_VALID: case K_COUNT_ALL: ((int32_t*)out_data[j])[g] = (int32_t)a.cnt[g]; break;
        case K_MEAN:
          ok = has;
          if (is_duration(t)) ((int64_t*)out_data[j])[g] = has ? a.i[g] / a.cnt[g] : 0;
          else {
            double sum = (t == T_FLOAT32) ? (double)a.f32[g] : fl ? a.f[g] : (double)a.i[g];
            ((double*)out_data[j])[g] = has ? sum / (double)a.cnt[g] : 0.0;
          }
          if (is_duration(t) && type_size(t) == 4) ((int32_t*)out_data[j])[g] = (int32_t)(has ? a.i[g] / a.cnt[g] : 0);
          break;
        case K_M2: case K_VARIANCE: case K_STD: {
          int ddof = req_ddof ? req_ddof[j] : 1;
          double m2 = a.m2[g], n = (double)a.cnt[g];
          double res = kind == K_M2 ? m2 : m2 / (n - ddof);
          if (kind == K_STD) res = sqrt(res);
          /* valid when the group has more valid values than ddof
           * (multi_pass_kernels.cuh:80) */
          ok = kind == K_M2 ? has : (has && a.cnt[g] > ddof);
          ((double*)out_data[j])[g] = ok ? res : 0.0;
          break;
        }
        case K_ARGMIN: case K_ARGMAX:
          ok = a.i[g] >= 0;
          ((int32_t*)out_data[j])[g] = (int32_t)a.i[g];
          break;
        default: {
          ok = has;
          if (tt == T_FLOAT32) ((float*)out_data[j])[g] = (kind == K_MIN || kind == K_MAX || kind == K_SUM ||
                                                          kind == K_PRODUCT || kind == K_SUM_OF_SQUARES)
                                                             ? a.f32[g] : 0.0f;
          else if (tt == T_FLOAT64) ((double*)out_data[j])[g] = a.f[g];
          else {
            int sz = type_size(tt);
            int64_t x = a.i[g];
            void* o = out_data[j];
            if (sz == 1) ((int8_t*)o)[g] = (int8_t)x;
            else if (sz == 2) ((int16_t*)o)[g] = (int16_t)x;
            else if (sz == 4) ((int32_t*)o)[g] = (int32_t)x;
            else ((int64_t*)o)[g] = x;
          }
        }
      }
      if (kind == K_COUNT_VALID || kind == K_COUNT_ALL) ok = 1;
      out_valid[j][g] = ok;
    }
  }
  free(a.i); free(a.f); free(a.f32); free(a.cnt); free(a.m2); free(a.valid);
  free(grp_off); free(grp_rows);
  free(group_of_row);
  return 0;
}

/* The hash-path entry point: no sort-path parameters (defaults). */
int mgo_groupby_aggregate(const mgo_col* keys, int32_t n_keys, int64_t num_rows, int32_t include_null_keys,
                          const mgo_col* values, int32_t n_values, const int32_t* req_col,
                          const int32_t* req_kind, const int32_t* req_ddof, int32_t n_req,
                          int64_t* out_rep_rows, int64_t* out_num_groups, void* const* out_data,
                          uint8_t* const* out_valid)
{
  return mgo_groupby_aggregate_ex(keys, n_keys, num_rows, include_null_keys, values, n_values, req_col, req_kind,
                                  req_ddof, NULL, NULL, NULL, n_req, out_rep_rows, out_num_groups, out_data,
                                  out_valid);
}

/* ------------------------------------------------ COLLECT_LIST / COLLECT_SET */

/*
 * COLLECT_LIST (kind 21) or COLLECT_SET (kind 22) of one value column, with
 * the groups of mgo_groupby_aggregate_ex (order of first appearance,
 * out_rep_rows / out_num_groups as there). Group g's list holds the source
 * rows out_rows[out_offsets[g] .. out_offsets[g+1]) (out_offsets: ng + 1
 * entries; out_rows: capacity num_rows): an element's value and validity are
 * those of its row.
 *
 * COLLECT_LIST: the group's rows in row order, null values kept
 * (null_policy 1, INCLUDE) or dropped (0, EXCLUDE)
 * (cpp/src/groupby/sort/group_collect.cu:36-115: the grouped values are the
 * stable sort by key, so row order within a group).
 * COLLECT_SET: that list, then its distinct elements
 * (cpp/src/groupby/sort/aggregate.cpp:426-447 -> lists::distinct,
 * cpp/src/lists/stream_compaction/distinct.cu:36-70: stable_distinct over
 * {list label, element} with KEEP_ANY): values equal under cudf's row
 * equality (-0.0 == +0.0; NaNs all equal under nans_equal 0 (ALL_EQUAL),
 * each its own element under 1 (UNEQUAL)), nulls one element under
 * nulls_equal 0 (EQUAL), each its own under 1 (UNEQUAL). Of the choices
 * KEEP_ANY leaves open, this oracle keeps the first row of each run of equal
 * elements, and the elements in order of first appearance; callers compare
 * sets after sorting each list, as cpp/tests/groupby/collect_set_tests.cpp:
 * 50-65 does.
 * Returns 0, or 1 for a kind other than 21 / 22 or a type without a width.
 */
int mgo_groupby_collect(const mgo_col* keys, int32_t n_keys, int64_t num_rows, int32_t include_null_keys,
                        const mgo_col* value, int32_t kind, int32_t null_policy, int32_t nulls_equal,
                        int32_t nans_equal, int64_t* out_rep_rows, int64_t* out_num_groups, int64_t* out_offsets,
                        int64_t* out_rows)
{
  if ((kind != K_COLLECT_LIST && kind != K_COLLECT_SET) || mgo_target_type(value->type_id, kind) < 0) return 1;
  int64_t* group_of_row = (int64_t*)malloc(sizeof(int64_t) * (num_rows ? num_rows : 1));
  int64_t ng = form_groups(keys, n_keys, num_rows, include_null_keys, out_rep_rows, group_of_row);
  *out_num_groups = ng;
  int64_t* grp_off = (int64_t*)calloc((size_t)ng + 1, sizeof(int64_t));
  int64_t* grp_rows = (int64_t*)malloc(sizeof(int64_t) * (size_t)(num_rows ? num_rows : 1));
  group_rows(group_of_row, num_rows, ng, grp_off, grp_rows);
  int64_t* buf = (int64_t*)malloc(sizeof(int64_t) * (size_t)(num_rows ? num_rows : 1));
  uint8_t* keep = (uint8_t*)malloc((size_t)(num_rows ? num_rows : 1));
  const int fl = is_float(value->type_id);

  int64_t e = 0;
  for (int64_t g = 0; g < ng; ++g) {
    const int64_t* rows = grp_rows + grp_off[g];
    const int64_t size = grp_off[g + 1] - grp_off[g];
    out_offsets[g] = e;
    if (kind == K_COLLECT_LIST) {
      for (int64_t i = 0; i < size; ++i)
        if (null_policy == 1 || bit_is_set(value->mask, rows[i])) out_rows[e++] = rows[i];
      continue;
    }
    /* valid values sorted into runs of equal values, the first row of each kept */
    int64_t n = 0, first_null = -1;
    for (int64_t i = 0; i < size; ++i) {
      keep[rows[i]] = 0;
      if (bit_is_set(value->mask, rows[i])) buf[n++] = rows[i];
      else if (null_policy == 1 && (nulls_equal == 1 || first_null < 0)) {
        keep[rows[i]] = 1;
        if (first_null < 0) first_null = rows[i];
      }
    }
    g_sort_col = value;
    qsort(buf, (size_t)n, sizeof(int64_t), row_cmp); /* value order, ties by row */
    for (int64_t i = 0; i < n; ++i) {
      int starts = i == 0 || value_cmp(buf[i - 1], buf[i]) != 0;
      if (!starts && fl && nans_equal == 1 && isnan(load_f(value->type_id, value->data, buf[i]))) starts = 1;
      keep[buf[i]] = (uint8_t)starts;
    }
    for (int64_t i = 0; i < size; ++i)
      if (keep[rows[i]]) out_rows[e++] = rows[i];
  }
  out_offsets[ng] = e;
  free(buf); free(keep); free(grp_off); free(grp_rows); free(group_of_row);
  return 0;
}
