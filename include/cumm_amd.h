/*
 * C ABI of libcumm_amd.so: the MI355X (gfx950) implementation of cumm's
 * gather-GEMM and mask-sparse implicit-GEMM path.
 *
 * Every struct is plain data: raw device pointers, shapes, strides, dtype codes
 * and flags.  No torch or tensorview types cross this boundary.  Calls are
 * synchronous on the host and asynchronous on the device (they enqueue on
 * `stream`, 0 = the null stream), exactly like the pybind entry points they
 * replace.  Tensors are borrowed; outputs are written in place.
 *
 * Return codes: 0 = ok, 1 = invalid argument (the reference raises
 * std::invalid_argument -> ValueError), 2 = runtime error (std::runtime_error ->
 * RuntimeError).  cs_last_error() returns the thread-local message of the last
 * failing call on this thread.
 *
 * Reference interfaces replaced (paths relative to the cumm source tree):
 *   cs_tensor              tv::Tensor borrowed view      include/tensorview/tensor.h:860-935
 *   cs_gemm_algo_desp      tv::gemm::GemmAlgoDesp        include/tensorview/gemm/core/params.h:24-164
 *   cs_conv_algo_desp      tv::gemm::ConvAlgoDesp        include/tensorview/gemm/core/params.h:166-327
 *   cs_gemm_params         tv::gemm::GemmParams          include/tensorview/gemm/core/params.h:329-390
 *   cs_conv_params         tv::gemm::ConvParams          include/tensorview/gemm/core/params.h:392-428
 */
#ifndef CUMM_AMD_H_
#define CUMM_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_ABI_VERSION 1
#define CS_MAX_DIM 4
#define CS_ALGO_NAME_LEN 32

/* Error codes. */
#define CS_OK 0
#define CS_ERR_INVALID_ARG 1
#define CS_ERR_RUNTIME 2

/* A borrowed strided tensor.  data == NULL means "empty" (tv::Tensor()).
 * dtype uses tv::DType codes (include/tensorview/dtypes.h:39-60).
 * stride is in elements.  device: -1 host, >= 0 GPU ordinal. */
typedef struct cs_tensor {
  void* data;
  int32_t dtype;
  int32_t ndim;
  int64_t shape[CS_MAX_DIM];
  int64_t stride[CS_MAX_DIM];
  int32_t device;
  int32_t reserved;
} cs_tensor;

/* tv::gemm::GemmAlgoDesp (params.h:24-49); min_arch is a (major, minor) pair. */
typedef struct cs_gemm_algo_desp {
  int32_t dtype_a, dtype_b, dtype_c;
  int32_t trans_a, trans_b, trans_c;
  int32_t tile_shape[3];
  int32_t warp_tile_shape[3];
  int32_t num_stage;
  int32_t dacc, dcomp;
  char algo[CS_ALGO_NAME_LEN];
  int32_t tensorop[3];
  int32_t split_k_serial, split_k_parallel;
  int32_t shuffle_type;
  int32_t element_per_access_a, element_per_access_b, element_per_access_c;
  int32_t access_per_vector;
  int32_t is_nvrtc;
  int32_t min_arch[2];
} cs_gemm_algo_desp;

/* tv::gemm::ConvAlgoDesp (params.h:166-190). */
typedef struct cs_conv_algo_desp {
  cs_gemm_algo_desp gemm;
  int32_t ndim, op_type, iter_algo;
  int32_t layout_i, layout_w, layout_o;
  int32_t interleave_i, interleave_w, interleave_o;
  int32_t mask_sparse, increment_k_first, dynamic_mask, is_int8_inference;
} cs_conv_algo_desp;

/* tv::gemm::GemmParams (params.h:329-350). */
typedef struct cs_gemm_params {
  cs_gemm_algo_desp algo_desp;
  cs_tensor a, b, c, d;
  cs_tensor workspace;
  cs_tensor a_inds, b_inds, c_inds;
  int32_t split_k_slices;
  float alpha, beta, act_alpha, act_beta;
  int32_t act_type;
  uint64_t stream;
} cs_gemm_params;

/* tv::gemm::ConvParams (params.h:392-421).  padding/stride/dilation are unused
 * by the mask-sparse path and not carried. */
typedef struct cs_conv_params {
  cs_conv_algo_desp conv_algo_desp;
  cs_tensor input, weight, output, output_add, bias, scale;
  cs_tensor workspace, mask, mask_argsort, indices, mask_output;
  int32_t split_k_slices;
  float alpha, beta, act_alpha, act_beta;
  int32_t act_type;
  int32_t mask_width;
  uint32_t mask_filter;
  int32_t reverse_mask, verbose, mask_int_count;
  uint64_t stream;
} cs_conv_params;

/* ABI version of the loaded library (CS_ABI_VERSION). */
int cs_abi_version(void);

/* ConvMainUnitTest.implicit_gemm2(params)  -- cumm/conv/main.py:483-745 with the
 * host preamble of cumm/conv/nvrtc_code.py:20-196.  Mask-sparse forward (op 0),
 * dgrad (op 1) and wgrad (op 2).  An empty workspace is allocated stream-ordered
 * and released after the launches. */
int cs_implicit_gemm2(const cs_conv_params* params);

/* GemmMainUnitTest.matmul2(params)  -- cumm/gemm/main.py:1176-1335 with the host
 * preamble of cumm/gemm/nvrtc_code.py:34-145.  NoShuffle, ShuffleAC (gather A /
 * scatter C) and ShuffleAB (gather A and B, sum over rows). */
int cs_matmul2(const cs_gemm_params* params);

/* ConvAlgoDesp::query_conv_workspace_size (params.h:286-306) evaluated on the
 * shapes of `params`; writes the byte count the call needs. */
int cs_conv_workspace_size(const cs_conv_params* params, int64_t* nbytes);

/* GemmAlgoDesp::query_workspace_size (params.h:116-132) on the shapes of
 * `params`. */
int cs_gemm_workspace_size(const cs_gemm_params* params, int64_t* nbytes);

/* ConvMainUnitTest.get_all_conv_algo_desp()  -- cumm/conv/main.py:983-1050.
 * Copies up to `cap` entries into `out` (may be NULL) and returns the total. */
int cs_get_all_conv_algo_desp(cs_conv_algo_desp* out, int cap);

/* GemmMainUnitTest.get_all_algo_desp()  -- cumm/gemm/main.py:914-965. */
int cs_get_all_algo_desp(cs_gemm_algo_desp* out, int cap);

/* GemmAlgoDesp::__repr__ / ConvAlgoDesp::__repr__ (params.h:56-78, 191-224):
 * writes the kernel name into buf (NUL-terminated, truncated to cap). */
int cs_gemm_algo_name(const cs_gemm_algo_desp* desp, char* buf, int cap);
int cs_conv_algo_name(const cs_conv_algo_desp* desp, char* buf, int cap);

/* GemmMainUnitTest.device_synchronize / stream_synchronize
 * (cumm/gemm/main.py:1050-1070). */
int cs_device_synchronize(void);
int cs_stream_synchronize(uint64_t stream);

/* Message of the last failing call on this thread ("" if none). */
const char* cs_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CUMM_AMD_H_ */
