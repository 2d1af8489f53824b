// Owning device column, with the child columns of a list column (subset of
// cpp/include/cudf/column/column.hpp).
#pragma once

#include <cudf/column/column_view.hpp>
#include <cudf/types.hpp>

#include <rmm/device_buffer.hpp>

#include <memory>
#include <vector>

namespace cudf {

class column {
 public:
  column() = default;
  ~column() = default;
  column(column const&) = delete;
  column& operator=(column const&) = delete;
  column(column&&) noexcept = default;
  column& operator=(column&&) noexcept = default;

  column(data_type dtype, size_type size, rmm::device_buffer&& data, rmm::device_buffer&& null_mask,
         size_type null_count, std::vector<std::unique_ptr<column>>&& children = {})
    : _type{dtype}, _size{size}, _data{std::move(data)}, _null_mask{std::move(null_mask)},
      _null_count{_null_mask.is_empty() ? 0 : null_count}, _children{std::move(children)}
  {
  }

  [[nodiscard]] data_type type() const noexcept { return _type; }
  [[nodiscard]] size_type size() const noexcept { return _size; }
  [[nodiscard]] size_type null_count() const noexcept { return _null_count; }
  [[nodiscard]] bool nullable() const noexcept { return !_null_mask.is_empty(); }
  [[nodiscard]] bool has_nulls() const noexcept { return _null_count > 0; }

  [[nodiscard]] size_type num_children() const noexcept { return static_cast<size_type>(_children.size()); }
  column& child(size_type i) noexcept { return *_children[i]; }
  [[nodiscard]] column const& child(size_type i) const noexcept { return *_children[i]; }

  [[nodiscard]] column_view view() const
  {
    std::vector<column_view> kids;
    kids.reserve(_children.size());
    for (auto const& c : _children) kids.push_back(c->view());
    return column_view{_type, _size, _data.data(),
                       nullable() ? static_cast<bitmask_type const*>(_null_mask.data()) : nullptr,
                       _null_count, 0, kids};
  }
  operator column_view() const { return view(); }

 private:
  data_type _type{type_id::EMPTY};
  size_type _size{0};
  rmm::device_buffer _data{};
  rmm::device_buffer _null_mask{};
  size_type _null_count{0};
  std::vector<std::unique_ptr<column>> _children{};
};

}  // namespace cudf
