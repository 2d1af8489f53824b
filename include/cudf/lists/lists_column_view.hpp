// View of a list column: child 0 holds size() + 1 INT32 offsets, child 1 the
// values of every list back to back (subset of
// cpp/include/cudf/lists/lists_column_view.hpp:41-120).
#pragma once

#include <cudf/column/column_view.hpp>
#include <cudf/utilities/error.hpp>

namespace cudf {

class lists_column_view : private column_view {
 public:
  static constexpr size_type offsets_column_index{0};
  static constexpr size_type child_column_index{1};

  explicit lists_column_view(column_view const& lists_column) : column_view(lists_column)
  {
    CUDF_EXPECTS(type().id() == type_id::LIST, "lists_column_view only supports lists");
  }

  using column_view::has_nulls;
  using column_view::is_empty;
  using column_view::null_count;
  using column_view::null_mask;
  using column_view::offset;
  using column_view::size;

  [[nodiscard]] column_view parent() const { return static_cast<column_view>(*this); }
  [[nodiscard]] column_view offsets() const
  {
    CUDF_EXPECTS(num_children() == 2, "lists column has no children");
    return column_view::child(offsets_column_index);
  }
  [[nodiscard]] column_view child() const
  {
    CUDF_EXPECTS(num_children() == 2, "lists column has no children");
    return column_view::child(child_column_index);
  }
};

}  // namespace cudf
