// Kernel parameter blocks of the groupby kernels (passed by value).
#pragma once

#include "mgb_device.hpp"

namespace mgb {

constexpr uint32_t kNoRecord = 0xffffffffu;

// A set of rows to aggregate (the input columns, or the partitioned copy of
// them), the accumulator ops, and the global record table.
struct AggParams {
  int64_t num_rows;
  uint8_t* table;                 // (capacity + 1) records of `stride` bytes
  int64_t capacity;               // records in the hashed region
  unsigned long long* counters;   // [0] groups, [1] overflow flag
  int64_t max_groups;             // inserting more sets the overflow flag
  const uint8_t* flags;           // per-row validity bits of the partitioned copy, or nullptr
  int32_t flag_bytes;             // 1, 2 or 4
  int32_t flag_stride;            // bytes between the flags of consecutive rows
  int32_t stride;
  int32_t key_mode;
  int32_t num_keys;
  int32_t skip_null_keys;         // rows with a null key are dropped
  int32_t compare_key_nulls;      // null == null participates in row equality
  int32_t num_vals;
  int32_t num_ops;
  int32_t rows_op;                // index of the OP_ROWS op in ops[], or -1
  // partitioned records of at most 32 bytes: row r is the record at
  // rec_base + r * 8 * rec_words, read whole (the key, value and flags
  // columns point into it); 0: rows are read column by column
  int32_t rec_words = 0;
  int32_t program_cols = 0;       // 1: at most kProgramCols value columns, none with ops beyond SUM/MIN/MAX
  // record mode with program_cols: each value column's program in one word
  // (pack_program), so the row loop keeps a few registers per column
  uint64_t prog[kProgramCols] = {};

  const uint8_t* rec_base = nullptr;
  uint32_t* row_rec = nullptr;    // global build of the sort path: each row's record index (kNoRecord: none)
  KeyCol keys[kMaxKeys];
  ValCol vals[kMaxVals];
  OpDesc ops[kMaxOps + kMaxColOps];  // the last kMaxColOps unused: a column's window of ops is read whole
};

// How one requested aggregation is produced from record fields.
enum FinKind : int32_t {
  FIN_COUNT = 0,      // u32 field -> int32
  FIN_INT,            // i64 / i32 / u32 / u64 field -> integer of out_width bytes
  FIN_F64,            // f64 field -> double
  FIN_F64_TO_F32,     // f64 field -> float (f32 sums accumulated in f64)
  FIN_F32_CODE,       // ordered u32 code -> float
  FIN_F64_CODE,       // ordered u64 code -> double
  FIN_MEAN_I64,       // i64 sum / u32 count -> double
  FIN_MEAN_F64,       // f64 sum / u32 count -> double
  FIN_MEAN_F32,       // (float)f64 sum / u32 count -> double
  FIN_MEAN_DURATION,  // i64 sum / u32 count -> integer of out_width bytes (truncating)
  FIN_VAR,            // f64 sum of squared deviations / (count - ddof) -> double, null unless count > ddof
  FIN_STD,            // square root of FIN_VAR's value
  FIN_ARGIDX,         // u32 row index -> int32 (-1 where the group has no valid value)
  FIN_M2,             // f64 sum of squared deviations -> double, null where the group has no valid value
  FIN_NONE            // written by the sort-path kernels after finalize (MEDIAN, QUANTILE, NUNIQUE, NTH_ELEMENT)
};

struct OutSpec {
  void* out;              // result data (out_width bytes per group)
  uint8_t* valid;         // validity byte per group, nullptr = never null
  int32_t fin;
  int32_t out_width;
  int32_t field_off;      // primary field
  int32_t count_off;      // row-count field (COUNT_ALL) or count of valid values
  int32_t nulls_off;      // OP_NULLS field subtracted from count_off, or -1
  int32_t ddof;           // FIN_VAR / FIN_STD: delta degrees of freedom
};

constexpr int kMaxOuts = 64;    // outputs per call on the LDS paths / per finalize launch

// Where finalised groups are written (device-resident, see OutTable).
struct OutTable {
  void* key_out[kMaxKeys];        // key columns
  uint8_t* key_valid[kMaxKeys];   // key validity bytes (null keys grouped), or nullptr
  int32_t num_outs;
  int32_t pad;
  // null results counted as they are written: [num_outs] results, then the
  // keys; nullptr: not counted
  unsigned long long* nulls;
  OutSpec outs[kMaxOuts];
};

struct FinParams {
  const uint8_t* table;
  int64_t capacity;
  int32_t stride;
  int32_t key_mode;
  int32_t num_keys;
  int32_t pad;
  const int64_t* block_offsets;   // exclusive prefix of occupied records per block
  const unsigned long long* pos_base;  // added to every output row, or nullptr
  const unsigned long long* table_groups;  // the table's group counter (0: nothing to write), or nullptr
  const OutTable* out;            // device memory
  KeyCol keys[kMaxKeys];          // representative-row gather source
  uint32_t* rec_pos = nullptr;    // per record: its output row (the sort-path aggregations), or nullptr
};

// Hash partitioning of the input rows (histogram and scatter passes).
struct PartParams {
  AggParams src;                  // input rows; keys / vals / skip_null_keys
  int64_t chunk;                  // rows per workgroup
  int32_t log2p;                  // 2^log2p partitions
  int32_t nblocks;                // workgroups of both passes
  uint32_t* hist;                 // [P][nblocks] row counts
  const uint32_t* offsets;        // [P][nblocks] exclusive prefix of hist
  uint8_t* out;                   // partitioned rows: records of rec_bytes
  int32_t rec_bytes;              // multiple of 8
  int32_t flag_bytes;             // 0, 1, 2 or 4
  int32_t flags_off;              // offset of the validity bits in a record
  int32_t tile_rows;              // rows staged per tile (tiled scatter)
  uint64_t rec_magic;             // ceil(2^32 / (rec_bytes / 8)): word index -> record index by multiply-shift
  int32_t key_off[kMaxKeys];      // offset of each key column in a record
  int32_t val_off[kMaxVals];      // offset of each value column in a record
  int32_t rowid_off = -1;         // offset of the source row index (u32) in a record, or -1 (none kept)
  int32_t pad = 0;
};

// Partitioned records of the compact partitioned execution (a 4- or 8-byte
// integer key, or two such columns; at most kLeanCols value columns of 4 or 8
// bytes whose ops are SUM / MIN / MAX and counts): the key at byte 0, the row
// hash at byte `key bytes` with the validity bits of the nullable columns in
// its low byte, and each value column at the lowest free offset aligned to
// its width, widest first (lean_offsets); 16 or 32 bytes a record. The
// aggregation takes the LDS slot from the stored hash instead of hashing the
// key again. Widths and offsets are compile-time constants of the kernels.
constexpr int kLeanCols = 3;
constexpr int kLeanMaxLog2P = 13;  // partitions of the compact scatter: at most 2^13

// Two-column integer keys (column widths K0, K1 of 4 or 8 bytes): the
// record's key region holds both columns, the 8-byte one first (column 0
// when both are as wide), so that its first 8 bytes are one 8-byte value
// ("key part A") and the rest (none, 4 or 8 bytes) "key part B". The LDS
// record keeps part A in its header and, in a tail after the plan's fields,
// part B and the index of the group's first record (lean_tail_bytes).
struct LeanKeyOffsets {
  int off0, off1, bytes;  // byte offsets of key columns 0 and 1 (-1: absent), key bytes
};
__host__ __device__ constexpr LeanKeyOffsets lean_key_offsets(int k0, int k1)
{
  return k1 == 0 ? LeanKeyOffsets{0, -1, k0}
                 : (k0 >= k1 ? LeanKeyOffsets{0, k0, k0 + k1} : LeanKeyOffsets{k1, 0, k0 + k1});
}
// LDS tail of a two-column key's record: 8 bytes (the first record's index,
// and part B when it is 4 bytes) or 16 (part B of 8 bytes, then the index).
__host__ __device__ constexpr int lean_tail_bytes(int key_bytes) { return key_bytes <= 12 ? 8 : 16; }
// A tail word whose record index is this is not written yet.
constexpr uint32_t kLeanNotReady = 0xffffffffu;

struct LeanOffsets {
  int off[kLeanCols];  // byte offset of each value column (-1: absent, or no room)
  int bytes;           // bytes used
};

__host__ __device__ constexpr LeanOffsets lean_offsets(int kw, int w0, int w1, int w2)
{
  LeanOffsets r{{-1, -1, -1}, kw + 4};
  const int ws[kLeanCols] = {w0, w1, w2};
  uint32_t used = (1u << (kw + 4)) - 1u;
  for (int w = 8; w >= 1; w >>= 1) {
    for (int c = 0; c < kLeanCols; ++c) {
      if (ws[c] != w) continue;
      for (int o = 0; o + w <= 32; o += w) {
        const uint32_t bits = ((1u << w) - 1u) << o;
        if ((used & bits) == 0) {
          used |= bits;
          r.off[c] = o;
          if (o + w > r.bytes) r.bytes = o + w;
          break;
        }
      }
    }
  }
  return r;
}

// Hot keys of a compact call: keys seen kHotMin times or more in the
// estimate's sample (at most kHotMax of them), as their raw key column
// values (zero-extended). The estimate's fine histogram leaves their rows
// out, and the compact scatter aggregates those rows in LDS records of its
// own, folded into the global table when the chunk is done, instead of
// writing them as partitioned records: under skew (a few keys holding much
// of the rows) the records and their aggregation shrink accordingly.
constexpr int kHotMax = 1024;
constexpr uint32_t kHotMin = 8;
constexpr int kHotSlots = 4096;  // LDS lookup slots (a power of two; few occupied, so a miss ends at its home slot)
struct HotKey {
  uint64_t c0, c1;  // key column 0 and 1 (0 for a single column)
};
struct HotSet {
  uint32_t n;
  uint32_t pad[3];
  HotKey keys[kHotMax];
};

// LDS layout of the compact tiled scatter: per partition its cursor and its
// tile count (scanned in place into its tile start; one more word for the
// tile's total), 16 scan words, then per staged row its destination, and the
// stage of records. Offsets in bytes.
// With hot keys (hot_stride: their LDS record stride, else 0) the hot keys,
// their lookup slots, a counter block (count, overflow, 2 words for the
// fold, kHotMax bits of reached records, 64 dummy 8-byte words) and
// kHotMax + 1 records follow.
struct LeanScatterLds {
  size_t dst, stage, hot_keys, hot_idx, hot_ctr, hot_rec, total;
};
__host__ __device__ constexpr size_t lean_align16(size_t x) { return (x + 15) & ~(size_t)15; }
__host__ __device__ constexpr LeanScatterLds lean_scatter_lds(int P, int T, int rec_bytes, int hot_stride = 0)
{
  const size_t d = (size_t)(2 * P + 1 + 16) * 4;
  const size_t st = lean_align16(d + (size_t)T * 4);
  const size_t hk = st + (size_t)T * rec_bytes;
  if (hot_stride == 0) return LeanScatterLds{d, st, hk, hk, hk, hk, hk};
  const size_t hi = hk + (size_t)kHotMax * 16;
  const size_t hc = hi + (size_t)kHotSlots * 4;
  const size_t hr = hc + 16 + kHotMax / 8 + 64 * 8;
  return LeanScatterLds{d, st, hk, hi, hc, hr, hr + (size_t)(kHotMax + 1) * hot_stride};
}

// A value column's updates in the compact aggregation: LDS record field
// offsets (-1: none) of its SUM, MIN, MAX and null count, its class and its
// validity bit in the hash word (-1: never null).
enum LeanClass : int32_t { LEAN_SIGNED = 0, LEAN_UNSIGNED = 1, LEAN_FLOAT = 2 };
struct LeanCol {
  int32_t sum_off;
  int32_t min_off;
  int32_t max_off;
  int32_t nulls_off;
  int32_t cls;       // LeanClass: f32 / f64 by the column's width
  int32_t flag_bit;
};

struct LdsParams;

// Arguments of the compact scatter. Masks point at the word holding row 0's
// bit (the columns' bit offsets are multiples of 32).
struct LeanScatterArgs {
  const uint8_t* key;               // key column (column 0 of two), 4 or 8 bytes a row
  const uint32_t* key_mask;         // rows whose key is null are dropped; nullptr: none dropped
  const uint8_t* key1;              // key column 1 of a two-column key
  const uint32_t* key1_mask;
  const uint8_t* val[kLeanCols];    // value columns
  const uint32_t* mask[kLeanCols];  // their null masks, nullptr: never null
  const uint32_t* offsets;          // [P][nblocks]: first destination row of partition p for chunk b
  uint8_t* out;                     // the records
  int64_t rows;
  int64_t chunk;                    // rows per workgroup
  int32_t log2p;
  int32_t nblocks;
  int32_t tile_rows;                // rows staged per tile: a multiple of 1024, at most 4096
                                    // (partitions: at most 2^kLeanMaxLog2P)
  int32_t flag_bits;                // bits 4c..4c+3: column c's validity bit in the hash word (15: never null)
  int32_t direct;                   // rows written straight to their records (no staging; tile_rows: rows
                                    // per step, no stage in LDS)
  // hot keys (hot == nullptr: none)
  const HotSet* hot;
  const LdsParams* lp;              // the global table (device memory), for the hot records' fold
  const uint64_t* tmpl;             // identity record (device memory)
  int32_t hot_words;                // 8-byte words of a hot record
  int32_t rows_off;                 // row-count field of a record, or -1
  int32_t key_signed;               // a single 4-byte key sign-extended into the header
  int32_t rowref;                   // two-column keys: headers are record indices
  int64_t rep_base;                 // two-column keys: record index of hot key 0's record
  LeanCol cols[kLeanCols];           // the value columns' updates of a hot record (flag_bit: -1)
};

// Arguments of the compact aggregation (one work item per workgroup).
struct LeanAggArgs {
  const uint8_t* recs;
  const uint32_t* offsets;          // as LdsParams::offsets
  const uint32_t* item_offsets;     // as LdsParams::item_offsets
  const unsigned long long* total_rows;
  const uint64_t* tmpl;             // identity record (device memory)
  const LdsParams* lp;              // every parameter (device memory): spilled rows, writing or folding the groups
  int64_t max_item_rows;
  int32_t nblocks;
  int32_t log2p;
  int32_t slots;
  int32_t cap;
  int32_t words;
  int32_t rows_off;                 // row-count field, or -1
  int32_t key_signed;               // 4-byte keys: sign-extended into the header
  int32_t tail_off;                 // two-column keys: byte offset of the LDS record's tail
  LeanCol cols[kLeanCols];
};

// Second pass over the rows (VARIANCE / STD / ARGMIN / ARGMAX): every row
// finds its group's record in a completed table (LDS or global) and updates
// the second-pass fields from the first pass's SUM / MIN / MAX fields.
constexpr int kMaxPass2 = 16;

struct Pass2Op {
  const uint8_t* data;  // the value column (an input column, or its field of the partitioned records)
  const uint32_t* mask; // its null mask, or nullptr
  int64_t bit_off;
  int32_t width;
  int32_t vclass;
  int32_t stride;       // bytes between consecutive rows' values
  int32_t flag_bit;     // its validity bit in the row's flags word, or -1
  int32_t kind;         // OP_M2, OP_ARGMIN_IDX or OP_ARGMAX_IDX
  int32_t val;          // the value column's index (ops of one column are adjacent)
  int32_t off;          // the updated field
  int32_t ref_off;      // OP_M2: the column's SUM field; ARG*: its MIN / MAX field
  int32_t ref_kind;     // op kind of the ref field
  int32_t count_off;    // the row-count field (OP_M2)
  int32_t nulls_off;    // the column's OP_NULLS field, or -1 (OP_M2)
  int32_t pad;
};

// The ops of a second pass, read by the kernels from device memory (so that
// their runtime indexing reads memory instead of copying a parameter block
// into private memory).
struct Pass2Ops {
  AggParams rows;            // the rows and the global table (LDS kernels read them from here, not from their
                             // parameter block)
  int32_t num_ops;
  int32_t rowid_stride;      // bytes between the row indices of consecutive rows
  const uint8_t* rowid;      // source row index (u32) of each row (partitioned records), or nullptr: the row itself
  uint32_t* item_global;     // partitioned: per work item, 1 when its groups went to the global table
  // per value column (the first kProgramCols): its M2, ARGMIN and ARGMAX
  // field offsets + 1 in bits 0-9, 10-19 and 20-29 (0: none), for the record
  // loop beside rows.prog
  uint32_t p2prog[kProgramCols];
  Pass2Op ops[kMaxPass2];
};

// LDS-resident aggregation of one work item (a partition of the partitioned
// copy, or a contiguous slice of the input rows).
struct LdsParams {
  AggParams src;                  // rows, ops; table/capacity/counters: the global table (direct mode)
  const uint32_t* offsets;        // partitioned mode: [P][nblocks] prefix; partition p starts at offsets[p*nblocks]
  const unsigned long long* total_rows;  // partitioned mode: rows in the partitioned copy
  int64_t rows_per_item;          // direct mode
  const uint64_t* tmpl;           // identity record (device memory)
  const OutTable* out;            // partitioned mode: outputs
  unsigned long long* out_counters;  // [0] groups written, [1] LDS overflow flag
  int32_t nblocks;                // partition pass workgroups
  int32_t log2p;
  int32_t slots;                  // LDS records (plus one for the 8-byte empty key)
  int32_t cap;                    // groups an item may hold
  int32_t words;                  // record words
  int32_t direct;                 // 1: merge into the global table, 0: write outputs
  const uint32_t* item_offsets;   // partitioned mode: [P+1] first work item of each partition
  int64_t max_item_rows;          // partitioned mode: rows per work item at most
  const Pass2Ops* p2;             // the second pass (device memory), or nullptr
};

}  // namespace mgb
