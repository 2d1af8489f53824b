// The compact execution's kernels for two key columns of 4 and 8 bytes.
#include <hip/hip_runtime.h>

#include "lean_kernels.hpp"

namespace mgb {

const LeanFns kLeanFns3[15] = {MGB_LEAN_SHAPES(4, 8)};

}  // namespace mgb
