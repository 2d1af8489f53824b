// The compact execution's kernels for two key columns of 8 and 8 bytes.
#include <hip/hip_runtime.h>

#include "lean_kernels.hpp"

namespace mgb {

const LeanFns kLeanFns5[15] = {MGB_LEAN_SHAPES(8, 8)};

}  // namespace mgb
