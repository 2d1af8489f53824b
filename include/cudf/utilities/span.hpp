// cudf::host_span: a non-owning view of contiguous host objects
// (cpp/include/cudf/utilities/span.hpp).
#pragma once

#include <cstddef>
#include <vector>

namespace cudf {

template <typename T>
class host_span {
 public:
  using value_type = T;
  constexpr host_span() noexcept = default;
  constexpr host_span(T* data, std::size_t size) noexcept : _data{data}, _size{size} {}
  template <typename C, typename = decltype(std::declval<C&>().data())>
  constexpr host_span(C& in) : _data{in.data()}, _size{in.size()}
  {
  }
  template <typename C, typename = decltype(std::declval<C const&>().data())>
  constexpr host_span(C const& in) : _data{in.data()}, _size{in.size()}
  {
  }
  [[nodiscard]] constexpr T* begin() const noexcept { return _data; }
  [[nodiscard]] constexpr T* end() const noexcept { return _data + _size; }
  [[nodiscard]] constexpr T* data() const noexcept { return _data; }
  [[nodiscard]] constexpr std::size_t size() const noexcept { return _size; }
  [[nodiscard]] constexpr bool empty() const noexcept { return _size == 0; }
  constexpr T& operator[](std::size_t i) const { return _data[i]; }

 private:
  T* _data{nullptr};
  std::size_t _size{0};
};

}  // namespace cudf
