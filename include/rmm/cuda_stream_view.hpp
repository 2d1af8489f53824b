// HIP-native stand-in for rmm::cuda_stream_view: a non-owning hipStream_t
// handle, so that cudf::groupby::groupby::aggregate keeps its signature
// (cpp/include/cudf/groupby.hpp:194-197 in cudf 24.04).
#pragma once

#include <hip/hip_runtime_api.h>

namespace rmm {

class cuda_stream_view {
 public:
  constexpr cuda_stream_view() = default;
  constexpr cuda_stream_view(hipStream_t stream) noexcept : stream_{stream} {}
  [[nodiscard]] constexpr hipStream_t value() const noexcept { return stream_; }
  constexpr operator hipStream_t() const noexcept { return stream_; }
  [[nodiscard]] bool is_default() const noexcept { return stream_ == nullptr; }
  void synchronize() const { (void)hipStreamSynchronize(stream_); }

 private:
  hipStream_t stream_{nullptr};
};

static constexpr cuda_stream_view cuda_stream_default{};

}  // namespace rmm
