// Owning, untyped device allocation (rmm::device_buffer) tied to the resource
// and stream it came from.
#pragma once

#include <rmm/cuda_stream_view.hpp>
#include <rmm/mr/device/per_device_resource.hpp>

#include <cstddef>
#include <utility>

namespace rmm {

class device_buffer {
 public:
  device_buffer() = default;
  device_buffer(std::size_t size, cuda_stream_view stream,
                mr::device_memory_resource* mr = mr::get_current_device_resource())
    : size_{size}, stream_{stream}, mr_{mr}
  {
    data_ = size ? mr_->allocate(size, stream) : nullptr;
  }
  // Takes ownership of `data`, which `mr` allocated on `stream`.
  device_buffer(void* data, std::size_t size, cuda_stream_view stream, mr::device_memory_resource* mr) noexcept
    : data_{data}, size_{size}, stream_{stream}, mr_{mr}
  {
  }
  device_buffer(device_buffer const&) = delete;
  device_buffer& operator=(device_buffer const&) = delete;
  device_buffer(device_buffer&& o) noexcept { swap(o); }
  device_buffer& operator=(device_buffer&& o) noexcept
  {
    if (this != &o) {
      reset();
      swap(o);
    }
    return *this;
  }
  ~device_buffer() { reset(); }

  [[nodiscard]] void* data() noexcept { return data_; }
  [[nodiscard]] void const* data() const noexcept { return data_; }
  [[nodiscard]] std::size_t size() const noexcept { return size_; }
  [[nodiscard]] bool is_empty() const noexcept { return size_ == 0; }
  [[nodiscard]] cuda_stream_view stream() const noexcept { return stream_; }
  [[nodiscard]] mr::device_memory_resource* memory_resource() const noexcept { return mr_; }

 private:
  void reset() noexcept
  {
    if (data_ && mr_) mr_->deallocate(data_, size_, stream_);
    data_ = nullptr;
    size_ = 0;
  }
  void swap(device_buffer& o) noexcept
  {
    std::swap(data_, o.data_);
    std::swap(size_, o.size_);
    std::swap(stream_, o.stream_);
    std::swap(mr_, o.mr_);
  }
  void* data_{nullptr};
  std::size_t size_{0};
  cuda_stream_view stream_{};
  mr::device_memory_resource* mr_{mr::get_current_device_resource()};
};

}  // namespace rmm
