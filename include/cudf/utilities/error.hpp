// Exceptions and checking macros of cpp/include/cudf/utilities/error.hpp.
#pragma once

#include <stdexcept>
#include <string>

namespace cudf {

struct logic_error : public std::logic_error {
  logic_error(char const* const message) : std::logic_error(message) {}
  logic_error(std::string const& message) : std::logic_error(message) {}
};

struct cuda_error : public std::runtime_error {
  cuda_error(std::string const& message) : std::runtime_error(message) {}
};

struct fatal_cuda_error : public cuda_error {
  using cuda_error::cuda_error;
};

struct data_type_error : public std::invalid_argument {
  data_type_error(char const* const message) : std::invalid_argument(message) {}
  data_type_error(std::string const& message) : std::invalid_argument(message) {}
};

}  // namespace cudf

#define CUDF_STRINGIFY_DETAIL(x) #x
#define CUDF_STRINGIFY(x) CUDF_STRINGIFY_DETAIL(x)

#define CUDF_EXPECTS(cond, reason)                                                              \
  do {                                                                                          \
    if (!(cond))                                                                                \
      throw cudf::logic_error(std::string("CUDF failure at: " __FILE__ ":" CUDF_STRINGIFY(__LINE__) ": ") + \
                              (reason));                                                        \
  } while (0)

#define CUDF_FAIL(reason) \
  throw cudf::logic_error(std::string("CUDF failure at: " __FILE__ ":" CUDF_STRINGIFY(__LINE__) ": ") + (reason))
