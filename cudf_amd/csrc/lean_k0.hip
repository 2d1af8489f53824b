// The compact execution's kernels for a single 4-byte key column.
#include <hip/hip_runtime.h>

#include "lean_kernels.hpp"

namespace mgb {

const LeanFns kLeanFns0[15] = {MGB_LEAN_SHAPES(4, 0)};

}  // namespace mgb
