// Row gather / scatter-add kernels: the glue of the gather-GEMM-scatter
// ("native") sparse-conv algorithm next to matmul2.
//
//   gather:       output[i] = input[indices[i]]          (any dtype, as bytes)
//   scatter-add:  output[indices[i]] += input[i]         (float32; indices of
//                 one call must be distinct, as in the reference, which adds
//                 without atomics)
//
// Reference: cumm/gemm/gather.py:29-618 (GatherKernelV2 / Gather / GatherAll,
// ScatterKernel / Scatter / ScatterAll).  The reference instantiates one kernel
// per (tile_m, tile_k_bytes, bytes_per_access, num_threads); here one kernel
// serves every published parameter set: a wave copies whole rows with 16-byte
// (or 4-byte) accesses, rows spread over the grid, indices checked against
// the source rows (out-of-range rows read zero / are skipped instead of
// faulting).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace cs {
namespace {

template <int VB>  // bytes per access: 16 or 4
__global__ __launch_bounds__(256) void gather_rows_kernel(const char* __restrict__ in, char* __restrict__ out,
                                                          const int32_t* __restrict__ idx, int64_t m,
                                                          int64_t n_in, int64_t row_bytes, int64_t in_ld,
                                                          int64_t out_ld) {
  const int64_t per_row = row_bytes / VB;
  const int64_t total = m * per_row;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / per_row, c = e - i * per_row;
    const int32_t r = idx[i];
    char* dst = out + i * out_ld + c * VB;
    if constexpr (VB == 16) {
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (r >= 0 && r < n_in) v = *reinterpret_cast<const uint4*>(in + (int64_t)r * in_ld + c * VB);
      *reinterpret_cast<uint4*>(dst) = v;
    } else {
      uint32_t v = 0u;
      if (r >= 0 && r < n_in) v = *reinterpret_cast<const uint32_t*>(in + (int64_t)r * in_ld + c * VB);
      *reinterpret_cast<uint32_t*>(dst) = v;
    }
  }
}

template <int VEC>  // floats per access: 4 or 1
__global__ __launch_bounds__(256) void scatter_add_f32_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                              const int32_t* __restrict__ idx, int64_t m,
                                                              int64_t n_out, int64_t cols, int64_t in_ld,
                                                              int64_t out_ld) {
  const int64_t per_row = cols / VEC;
  const int64_t total = m * per_row;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / per_row, c = e - i * per_row;
    const int32_t r = idx[i];
    if (r < 0 || r >= n_out) continue;
    float* dst = out + (int64_t)r * out_ld + c * VEC;
    const float* src = in + i * in_ld + c * VEC;
    if constexpr (VEC == 4) {
      float4 a = *reinterpret_cast<const float4*>(src);
      float4 b = *reinterpret_cast<float4*>(dst);
      b.x += a.x;
      b.y += a.y;
      b.z += a.z;
      b.w += a.w;
      *reinterpret_cast<float4*>(dst) = b;
    } else {
      *dst += *src;
    }
  }
}

unsigned grid_for(int64_t total) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
}

}  // namespace

hipError_t launch_gather_rows(const void* in, void* out, const int32_t* idx, int64_t m, int64_t n_in,
                              int64_t row_bytes, int64_t in_ld_bytes, int64_t out_ld_bytes, hipStream_t s) {
  if (m == 0 || row_bytes == 0) return hipSuccess;
  const bool v16 = row_bytes % 16 == 0 && in_ld_bytes % 16 == 0 && out_ld_bytes % 16 == 0 &&
                   (reinterpret_cast<uintptr_t>(in) & 15u) == 0 && (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
  if (v16) {
    hipLaunchKernelGGL(gather_rows_kernel<16>, dim3(grid_for(m * row_bytes / 16)), dim3(256), 0, s,
                       static_cast<const char*>(in), static_cast<char*>(out), idx, m, n_in, row_bytes, in_ld_bytes,
                       out_ld_bytes);
  } else {
    if (row_bytes % 4 != 0 || in_ld_bytes % 4 != 0 || out_ld_bytes % 4 != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gather_rows_kernel<4>, dim3(grid_for(m * row_bytes / 4)), dim3(256), 0, s,
                       static_cast<const char*>(in), static_cast<char*>(out), idx, m, n_in, row_bytes, in_ld_bytes,
                       out_ld_bytes);
  }
  return hipGetLastError();
}

hipError_t launch_scatter_add_f32(const float* in, float* out, const int32_t* idx, int64_t m, int64_t n_out,
                                  int64_t cols, int64_t in_ld, int64_t out_ld, hipStream_t s) {
  if (m == 0 || cols == 0) return hipSuccess;
  const bool v4 = cols % 4 == 0 && in_ld % 4 == 0 && out_ld % 4 == 0 &&
                  (reinterpret_cast<uintptr_t>(in) & 15u) == 0 && (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
  if (v4)
    hipLaunchKernelGGL(scatter_add_f32_kernel<4>, dim3(grid_for(m * cols / 4)), dim3(256), 0, s, in, out, idx, m,
                       n_out, cols, in_ld, out_ld);
  else
    hipLaunchKernelGGL(scatter_add_f32_kernel<1>, dim3(grid_for(m * cols)), dim3(256), 0, s, in, out, idx, m, n_out,
                       cols, in_ld, out_ld);
  return hipGetLastError();
}

}  // namespace cs
