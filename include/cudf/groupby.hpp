// cudf::groupby::groupby -- the drop-in for cpp/include/cudf/groupby.hpp
// (cudf 24.04): same class, request/result structs and `aggregate`
// signatures. Every call groups through a hash table; the aggregations cudf
// serves only from its sort-based path (MEDIAN, QUANTILE, NUNIQUE,
// NTH_ELEMENT, M2) are computed from sorts of each value column's rows by
// group; list-valued and nested aggregations raise cudf::logic_error.
#pragma once

#include <cudf/aggregation.hpp>
#include <cudf/column/column.hpp>
#include <cudf/table/table.hpp>
#include <cudf/table/table_view.hpp>
#include <cudf/types.hpp>
#include <cudf/utilities/span.hpp>

#include <rmm/cuda_stream_view.hpp>
#include <rmm/mr/device/per_device_resource.hpp>

#include <memory>
#include <utility>
#include <vector>

namespace cudf {

// The stream used when none is passed (cudf::get_default_stream()).
rmm::cuda_stream_view const get_default_stream();

namespace groupby {

// cpp/include/cudf/groupby.hpp:59-62
struct aggregation_request {
  column_view values;
  std::vector<std::unique_ptr<groupby_aggregation>> aggregations;
};

// cpp/include/cudf/groupby.hpp:86-89
struct aggregation_result {
  std::vector<std::unique_ptr<column>> results{};
};

// cpp/include/cudf/groupby.hpp:94-197
class groupby {
 public:
  groupby() = delete;
  ~groupby();
  groupby(groupby const&) = delete;
  groupby(groupby&&) = delete;
  groupby& operator=(groupby const&) = delete;
  groupby& operator=(groupby&&) = delete;

  explicit groupby(table_view const& keys, null_policy null_handling = null_policy::EXCLUDE,
                   sorted keys_are_sorted = sorted::NO, std::vector<order> const& column_order = {},
                   std::vector<null_order> const& null_precedence = {});

  std::pair<std::unique_ptr<table>, std::vector<aggregation_result>> aggregate(
    host_span<aggregation_request const> requests,
    rmm::mr::device_memory_resource* mr = rmm::mr::get_current_device_resource());

  std::pair<std::unique_ptr<table>, std::vector<aggregation_result>> aggregate(
    host_span<aggregation_request const> requests, rmm::cuda_stream_view stream,
    rmm::mr::device_memory_resource* mr = rmm::mr::get_current_device_resource());

 private:
  table_view _keys;
  null_policy _include_null_keys{null_policy::EXCLUDE};
  sorted _keys_are_sorted{sorted::NO};
  std::vector<order> _column_order{};
  std::vector<null_order> _null_precedence{};
};

}  // namespace groupby
}  // namespace cudf
