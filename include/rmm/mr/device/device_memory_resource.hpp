// HIP-native stand-in for rmm::mr::device_memory_resource (the `mr` argument
// of cudf's APIs) with a hipMalloc-backed default resource.
#pragma once

#include <rmm/cuda_stream_view.hpp>

#include <cstddef>
#include <cstdint>
#include <new>
#include <string>

namespace rmm {

class bad_alloc : public std::bad_alloc {
 public:
  explicit bad_alloc(std::string msg) : msg_{std::move(msg)} {}
  [[nodiscard]] const char* what() const noexcept override { return msg_.c_str(); }

 private:
  std::string msg_;
};

namespace mr {

class device_memory_resource {
 public:
  virtual ~device_memory_resource() = default;
  void* allocate(std::size_t bytes, cuda_stream_view stream = cuda_stream_view{})
  {
    return do_allocate(bytes, stream);
  }
  void deallocate(void* p, std::size_t bytes, cuda_stream_view stream = cuda_stream_view{})
  {
    do_deallocate(p, bytes, stream);
  }
  [[nodiscard]] bool is_equal(device_memory_resource const& other) const noexcept { return do_is_equal(other); }

 private:
  virtual void* do_allocate(std::size_t bytes, cuda_stream_view stream) = 0;
  virtual void do_deallocate(void* p, std::size_t bytes, cuda_stream_view stream) = 0;
  [[nodiscard]] virtual bool do_is_equal(device_memory_resource const& other) const noexcept
  {
    return this == &other;
  }
};

// Device memory from hipMallocAsync / hipFreeAsync on the given stream (the
// device's default memory pool, which keeps freed memory for reuse instead
// of returning it to the device at every synchronisation).
class cuda_memory_resource final : public device_memory_resource {
 public:
  cuda_memory_resource()
  {
    int dev = 0;
    hipMemPool_t pool = nullptr;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
      uint64_t keep = ~0ull;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
  }

 private:
  void* do_allocate(std::size_t bytes, cuda_stream_view stream) override
  {
    void* p = nullptr;
    if (bytes == 0) return nullptr;
    if (hipMallocAsync(&p, bytes, stream.value()) != hipSuccess || p == nullptr)
      throw rmm::bad_alloc("hipMallocAsync failed");
    return p;
  }
  void do_deallocate(void* p, std::size_t, cuda_stream_view stream) override
  {
    if (p) (void)hipFreeAsync(p, stream.value());
  }
};

}  // namespace mr
}  // namespace rmm
