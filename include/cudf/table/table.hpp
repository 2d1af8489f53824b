// Owning set of columns (subset of cpp/include/cudf/table/table.hpp).
#pragma once

#include <cudf/column/column.hpp>
#include <cudf/table/table_view.hpp>

#include <memory>
#include <vector>

namespace cudf {

class table {
 public:
  table() = default;
  ~table() = default;
  table(table&&) = default;
  table& operator=(table const&) = delete;
  table& operator=(table&&) = delete;
  explicit table(std::vector<std::unique_ptr<column>>&& columns) : _columns{std::move(columns)}
  {
    if (!_columns.empty()) _num_rows = _columns.front()->size();
  }
  [[nodiscard]] size_type num_columns() const noexcept { return static_cast<size_type>(_columns.size()); }
  [[nodiscard]] size_type num_rows() const noexcept { return _num_rows; }
  [[nodiscard]] table_view view() const
  {
    std::vector<column_view> views;
    views.reserve(_columns.size());
    for (auto const& c : _columns) views.push_back(c->view());
    return table_view{views};
  }
  operator table_view() const { return view(); }
  std::vector<std::unique_ptr<column>> release()
  {
    _num_rows = 0;
    return std::move(_columns);
  }
  column& get_column(size_type i) { return *_columns.at(static_cast<std::size_t>(i)); }
  [[nodiscard]] column const& get_column(size_type i) const { return *_columns.at(static_cast<std::size_t>(i)); }

 private:
  std::vector<std::unique_ptr<column>> _columns{};
  size_type _num_rows{0};
};

}  // namespace cudf
