// Non-owning set of equal-length columns (subset of
// cpp/include/cudf/table/table_view.hpp).
#pragma once

#include <cudf/column/column_view.hpp>
#include <cudf/utilities/error.hpp>

#include <algorithm>
#include <vector>

namespace cudf {

class table_view {
 public:
  using iterator       = std::vector<column_view>::iterator;
  using const_iterator = std::vector<column_view>::const_iterator;

  table_view() = default;
  table_view(std::vector<column_view> const& cols) : _columns{cols}
  {
    if (!_columns.empty()) {
      _num_rows = _columns.front().size();
      CUDF_EXPECTS(std::all_of(_columns.begin(), _columns.end(),
                               [this](column_view const& c) { return c.size() == _num_rows; }),
                   "Column size mismatch.");
    }
  }
  [[nodiscard]] const_iterator begin() const noexcept { return _columns.begin(); }
  [[nodiscard]] const_iterator end() const noexcept { return _columns.end(); }
  [[nodiscard]] column_view const& column(size_type i) const { return _columns.at(static_cast<std::size_t>(i)); }
  [[nodiscard]] size_type num_columns() const noexcept { return static_cast<size_type>(_columns.size()); }
  [[nodiscard]] size_type num_rows() const noexcept { return _num_rows; }
  [[nodiscard]] bool is_empty() const noexcept { return num_columns() == 0; }

 private:
  std::vector<column_view> _columns{};
  size_type _num_rows{0};
};

// True if any column of the table has nulls (cpp/include/cudf/table/table_view.hpp:315).
inline bool has_nulls(table_view const& t)
{
  return std::any_of(t.begin(), t.end(), [](column_view const& c) { return c.has_nulls(); });
}

}  // namespace cudf
