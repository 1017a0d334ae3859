// gradient.hip -- shape gradient of v^T A(X) u (matrix-free adjoint kernel).
#include <hip/hip_runtime.h>

#include "core.hpp"

namespace bempp_amd {

void shape_gradient(const Assembler&, const FunctionSpace&, const FunctionSpace&, const void*, const void*, void*) {
  throw Error("shape gradient is not available in this build");
}

}  // namespace bempp_amd
