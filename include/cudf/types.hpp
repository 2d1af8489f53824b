// Core types of the cudf C++ API used by cudf::groupby (subset of
// cpp/include/cudf/types.hpp in cudf 24.04; same names and values).
#pragma once

#include <cstddef>
#include <cstdint>

namespace cudf {

using size_type    = int32_t;   // cudf/types.hpp:93
using bitmask_type = uint32_t;  // cudf/types.hpp:94
using hash_value_type = uint32_t;

class column;
class column_view;
class table;
class table_view;

enum class order : bool { ASCENDING, DESCENDING };
enum class null_policy : bool { EXCLUDE, INCLUDE };
enum class interpolation : int32_t { LINEAR, LOWER, HIGHER, MIDPOINT, NEAREST };
enum class nan_policy : bool { NAN_IS_NULL, NAN_IS_VALID };
enum class null_equality : bool { EQUAL, UNEQUAL };
enum class nan_equality { ALL_EQUAL, UNEQUAL };
enum class null_order : bool { AFTER, BEFORE };
enum class sorted : bool { NO, YES };
enum class mask_state : int32_t { UNALLOCATED, UNINITIALIZED, ALL_VALID, ALL_NULL };

// Numbering of cpp/include/cudf/types.hpp:201-233.
enum class type_id : int32_t {
  EMPTY,
  INT8,
  INT16,
  INT32,
  INT64,
  UINT8,
  UINT16,
  UINT32,
  UINT64,
  FLOAT32,
  FLOAT64,
  BOOL8,
  TIMESTAMP_DAYS,
  TIMESTAMP_SECONDS,
  TIMESTAMP_MILLISECONDS,
  TIMESTAMP_MICROSECONDS,
  TIMESTAMP_NANOSECONDS,
  DURATION_DAYS,
  DURATION_SECONDS,
  DURATION_MILLISECONDS,
  DURATION_MICROSECONDS,
  DURATION_NANOSECONDS,
  DICTIONARY32,
  STRING,
  LIST,
  DECIMAL32,
  DECIMAL64,
  DECIMAL128,
  STRUCT,
  NUM_TYPE_IDS
};

class data_type {
 public:
  data_type() = default;
  explicit constexpr data_type(type_id id) : _id{id} {}
  data_type(type_id id, int32_t scale) : _id{id}, _fixed_point_scale{scale} {}
  [[nodiscard]] constexpr type_id id() const noexcept { return _id; }
  [[nodiscard]] constexpr int32_t scale() const noexcept { return _fixed_point_scale; }

 private:
  type_id _id{type_id::EMPTY};
  int32_t _fixed_point_scale{};
};

constexpr bool operator==(data_type const& lhs, data_type const& rhs)
{
  return lhs.id() == rhs.id() && lhs.scale() == rhs.scale();
}
constexpr bool operator!=(data_type const& lhs, data_type const& rhs) { return !(lhs == rhs); }

// Bytes per element of a fixed-width type (0 for everything else).
std::size_t size_of(data_type t);

}  // namespace cudf
