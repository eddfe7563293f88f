// Row-gather kernel instantiations for bf16 inputs.
#include "cs_rowgather.h"

namespace cs {

hipError_t launch_rowgather_bf16(int out_dtype, int block_n, const RowGatherArgs& a, hipStream_t s) {
  if (out_dtype == kBF16) return launch_rg<bf16, bf16>(a, block_n, s);
  if (out_dtype == kF32) return launch_rg<bf16, float>(a, block_n, s);
  return hipErrorInvalidValue;
}

}  // namespace cs
