// Gather GEMM kernel instantiations for f16 operands.
#include "cs_gemm.h"

namespace cs {

hipError_t launch_gemm_rows_f16(int out_dtype, const GemmRowsArgs& a, hipStream_t s) {
  if (out_dtype == kF16) return launch_gm<_Float16, _Float16>(a, s);
  return hipErrorInvalidValue;
}

}  // namespace cs
