// The compact execution's kernels for two key columns of 4 and 4 bytes.
#include <hip/hip_runtime.h>

#include "lean_kernels.hpp"

namespace mgb {

const LeanFns kLeanFns2[15] = {MGB_LEAN_SHAPES(4, 4)};

}  // namespace mgb
