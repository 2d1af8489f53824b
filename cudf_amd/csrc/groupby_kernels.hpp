// Kernel parameter blocks and launch helpers shared by the groupby driver
// (groupby.hip) and the batched / served executors.
#pragma once

#include "mgb_device.hpp"

namespace mgb {

// Everything the aggregate kernel needs, passed by value as a kernel argument.
struct AggParams {
  int64_t num_rows;
  uint8_t* table;                 // (capacity + 1) records of `stride` bytes
  int64_t capacity;               // records in the hashed region
  unsigned long long* counters;   // [0] groups, [1] overflow flag
  int64_t max_groups;             // inserting more sets the overflow flag
  int32_t stride;
  int32_t key_mode;
  int32_t num_keys;
  int32_t skip_null_keys;         // rows with a null key are dropped
  int32_t compare_key_nulls;      // null == null participates in row equality
  int32_t num_vals;
  int32_t num_ops;
  int32_t rows_op;                // index of the OP_ROWS op in ops[], or -1
  KeyCol keys[kMaxKeys];
  ValCol vals[kMaxVals];
  OpDesc ops[kMaxOps];
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
  FIN_MEAN_DURATION   // i64 sum / u32 count -> integer of out_width bytes (truncating)
};

struct OutSpec {
  void* out;              // result data (out_width bytes per group)
  uint8_t* valid;         // validity byte per group (workspace), nullptr = never null
  int32_t fin;
  int32_t out_width;
  int32_t field_off;      // primary field
  int32_t count_off;      // count field deciding validity / the MEAN divisor
};

constexpr int kMaxOuts = 32;    // outputs per finalize launch

struct FinParams {
  const uint8_t* table;
  int64_t capacity;
  int32_t stride;
  int32_t key_mode;
  int32_t num_keys;
  int32_t num_outs;
  const int64_t* block_offsets;   // exclusive prefix of occupied records per block
  KeyCol keys[kMaxKeys];          // generic mode: gather source
  void* key_out[kMaxKeys];        // key columns out
  uint8_t* key_valid[kMaxKeys];   // generic mode with null keys: validity bytes
  OutSpec outs[kMaxOuts];
};

}  // namespace mgb
