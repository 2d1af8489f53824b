// Non-owning view of a device column: fixed-width, or a list column whose
// children are its offsets and values (subset of
// cpp/include/cudf/column/column_view.hpp:53-225).
#pragma once

#include <cudf/types.hpp>
#include <cudf/utilities/error.hpp>

#include <vector>

namespace cudf {

class column_view {
 public:
  column_view() = default;
  column_view(data_type type, size_type size, void const* data, bitmask_type const* null_mask = nullptr,
              size_type null_count = 0, size_type offset = 0, std::vector<column_view> const& children = {})
    : _type{type}, _size{size}, _data{data}, _null_mask{null_mask}, _null_count{null_count}, _offset{offset},
      _children{children}
  {
    CUDF_EXPECTS(size >= 0, "Column size cannot be negative.");
    if (null_mask == nullptr) _null_count = 0;
  }

  template <typename T = void>
  [[nodiscard]] T const* head() const noexcept
  {
    return static_cast<T const*>(_data);
  }
  template <typename T>
  [[nodiscard]] T const* data() const noexcept
  {
    return head<T>() + _offset;
  }
  [[nodiscard]] size_type size() const noexcept { return _size; }
  [[nodiscard]] bool is_empty() const noexcept { return _size == 0; }
  [[nodiscard]] data_type type() const noexcept { return _type; }
  [[nodiscard]] bool nullable() const noexcept { return _null_mask != nullptr; }
  [[nodiscard]] size_type null_count() const noexcept { return _null_count; }
  [[nodiscard]] bool has_nulls() const noexcept { return _null_count > 0; }
  [[nodiscard]] bitmask_type const* null_mask() const noexcept { return _null_mask; }
  [[nodiscard]] size_type offset() const noexcept { return _offset; }
  [[nodiscard]] size_type num_children() const noexcept { return static_cast<size_type>(_children.size()); }
  [[nodiscard]] column_view child(size_type i) const { return _children.at(i); }
  [[nodiscard]] auto child_begin() const noexcept { return _children.cbegin(); }
  [[nodiscard]] auto child_end() const noexcept { return _children.cend(); }

 private:
  data_type _type{type_id::EMPTY};
  size_type _size{0};
  void const* _data{nullptr};
  bitmask_type const* _null_mask{nullptr};
  size_type _null_count{0};
  size_type _offset{0};
  std::vector<column_view> _children{};
};

}  // namespace cudf
