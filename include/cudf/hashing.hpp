// Row hashing of cudf::hashing (subset of cpp/include/cudf/hashing.hpp):
// the MurmurHash3_x86_32 row hasher that the groupby table uses, and the
// Spark-compatible variant.
#pragma once

#include <cudf/column/column.hpp>
#include <cudf/table/table_view.hpp>

#include <rmm/cuda_stream_view.hpp>
#include <rmm/mr/device/per_device_resource.hpp>

#include <memory>

namespace cudf {

rmm::cuda_stream_view const get_default_stream();

namespace hashing {

static constexpr uint32_t DEFAULT_HASH_SEED = 0;  // cpp/include/cudf/hashing.hpp:51

// cpp/include/cudf/hashing.hpp:90-94
std::unique_ptr<column> murmurhash3_x86_32(
  table_view const& input, uint32_t seed = DEFAULT_HASH_SEED,
  rmm::cuda_stream_view stream = cudf::get_default_stream(),
  rmm::mr::device_memory_resource* mr = rmm::mr::get_current_device_resource());

// cpp/include/cudf/hashing.hpp:130-134
std::unique_ptr<column> spark_murmurhash3_x86_32(
  table_view const& input, uint32_t seed = DEFAULT_HASH_SEED,
  rmm::cuda_stream_view stream = cudf::get_default_stream(),
  rmm::mr::device_memory_resource* mr = rmm::mr::get_current_device_resource());

}  // namespace hashing
}  // namespace cudf
