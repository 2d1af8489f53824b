// The compact execution's kernels for two key columns of 8 and 4 bytes.
#include <hip/hip_runtime.h>

#include "lean_kernels.hpp"

namespace mgb {

const LeanFns kLeanFns4[15] = {MGB_LEAN_SHAPES(8, 4)};

}  // namespace mgb
