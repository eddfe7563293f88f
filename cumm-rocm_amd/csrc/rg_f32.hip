// Row-gather kernel instantiations for f32 inputs.
#include "cs_rowgather.h"

namespace cs {

hipError_t launch_rowgather_f32(int out_dtype, int block_n, const RowGatherArgs& a, hipStream_t s) {
  if (out_dtype == kF32) return launch_rg<float, float>(a, block_n, s);
  return hipErrorInvalidValue;
}

}  // namespace cs
