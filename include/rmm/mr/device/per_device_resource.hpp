// rmm::mr::get_current_device_resource / set_current_device_resource.
#pragma once

#include <rmm/mr/device/device_memory_resource.hpp>

namespace rmm::mr {

namespace detail {
inline device_memory_resource*& current_resource_slot()
{
  static cuda_memory_resource default_resource;
  static device_memory_resource* current = &default_resource;
  return current;
}
}  // namespace detail

inline device_memory_resource* get_current_device_resource() { return detail::current_resource_slot(); }

inline device_memory_resource* set_current_device_resource(device_memory_resource* mr)
{
  auto*& slot = detail::current_resource_slot();
  auto* old = slot;
  if (mr) slot = mr;
  return old;
}

}  // namespace rmm::mr
