// The compact execution's kernels for a single 8-byte key column.
#include <hip/hip_runtime.h>

#include "lean_kernels.hpp"

namespace mgb {

const LeanFns kLeanFns1[15] = {MGB_LEAN_SHAPES(8, 0)};

}  // namespace mgb
