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

/* cs_implicit_gemm2 without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The call's
 * workspace is a file-scope device array of CS_NOALLOC_SCRATCH_BYTES held by
 * each GPU the library runs on until the process ends; params->workspace is
 * ignored.  A call whose workspace exceeds the array fails with
 * CS_ERR_INVALID_ARG (wgrad first lowers its split pieces per offset to fit,
 * which changes the FP32 summation only for an offset with more than
 * pieces * 1024 pairs).  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other threads) race. */
#define CS_NOALLOC_SCRATCH_BYTES (64ll << 20)
int cs_implicit_gemm2_noalloc(const cs_conv_params* params);

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

/* ---- Pair tables, masks and argsort on the device.
 * These replace what the caller of implicit_gemm2 builds before each layer:
 * spconv's get_indice_pairs / generate_subm_conv_inds + sort_1d_by_key_split
 * (called at cumm/conv/main_real_spconv.py:91-145), over tensorview's GPU
 * linear-probing hash table (include/tensorview/hash/linear.cu.h:111-430).
 * Coordinates are int32 rows (b, z, y, x); spatial shapes are (z, y, x).
 * Outputs: pair tables [kv][N] int32 (-1 = none) with offsets
 * f = (kz*K1 + ky)*K2 + kx, masks [N][W] uint32 (W = div_up(kv, 32), offset f
 * in bit f % 32 of word f / 32) sorted stably by the words read as one number
 * (last word most significant), with their argsort.  All buffers are device
 * memory; work is queued on `stream`.  Return 0 / 1 (bad argument) / 2
 * (device error), message in cs_pairs_last_error().
 * The *_geom entry points take the layer's kernel geometry (kv <= 1024):
 * submanifold neighbours coord + dilation*(k - K/2) (odd K), strided outputs
 * o*stride - padding + k over an output shape (S + 2*padding - K)/stride + 1.
 * The entry points without _geom are 3x3x3 (strided: stride 2, padding 1). */
typedef struct cs_conv_geom {
  int32_t ksize[3];
  int32_t stride[3];
  int32_t padding[3];
  int32_t dilation[3];
} cs_conv_geom;
int cs_pairs_subm_geom_workspace(int64_t n, const cs_conv_geom* geom, int64_t* nbytes);
int cs_pairs_subm_geom(const int32_t* coords, int64_t n, const int32_t* spatial_shape, const cs_conv_geom* geom,
                       int32_t* pair_fwd, int32_t* pair_bwd, uint32_t* mask_sorted, int32_t* argsort,
                       void* workspace, int64_t ws_bytes, uint64_t stream);
int cs_pairs_strided_out_geom_workspace(int64_t n, const cs_conv_geom* geom, int32_t rule, int64_t* nbytes);
int cs_pairs_strided_out_geom(const int32_t* coords, int64_t n, const int32_t* in_shape, const cs_conv_geom* geom,
                              int32_t rule, int32_t* out_coords, int64_t cap, int64_t* n_out, void* workspace,
                              int64_t ws_bytes, uint64_t stream);
int cs_pairs_strided_geom_workspace(int64_t n_in, int64_t n_out, const cs_conv_geom* geom, int64_t* nbytes);
int cs_pairs_strided_geom(const int32_t* in_coords, int64_t n_in, const int32_t* in_shape, const cs_conv_geom* geom,
                          const int32_t* out_coords, int64_t n_out, int32_t* pair_fwd, int32_t* pair_bwd,
                          uint32_t* mask_fwd_sorted, int32_t* argsort_fwd, uint32_t* mask_bwd_sorted,
                          int32_t* argsort_bwd, void* workspace, int64_t ws_bytes, uint64_t stream);
int cs_pairs_subm_workspace(int64_t n, int64_t* nbytes);
int cs_pairs_subm(const int32_t* coords, int64_t n, const int32_t* spatial_shape, int32_t* pair_fwd,
                  int32_t* pair_bwd, uint32_t* mask_sorted, int32_t* argsort, void* workspace,
                  int64_t ws_bytes, uint64_t stream);
/* Output coordinates of a strided layer, sorted by key: rule 0 =
 * unique(floor(c / stride)), rule 1 = every output whose window reaches an
 * input.  Synchronises `stream` to return the count in *n_out (<= cap rows
 * written). */
int cs_pairs_strided_out_workspace(int64_t n, int32_t rule, int64_t* nbytes);
int cs_pairs_strided_out(const int32_t* coords, int64_t n, const int32_t* in_shape, int32_t rule,
                         int32_t* out_coords, int64_t cap, int64_t* n_out, void* workspace, int64_t ws_bytes,
                         uint64_t stream);
/* Strided pair tables: pair_fwd [kv][n_out] (input rows), pair_bwd [kv][n_in]
 * (output rows), forward masks over the outputs and backward masks over the
 * inputs, each sorted with its argsort.  Forward-only callers (inference)
 * pass pair_bwd, mask_bwd_sorted and argsort_bwd all NULL; cs_pairs_subm
 * likewise accepts pair_bwd = NULL.  out_coords may be any list of distinct
 * outputs in any order (rows outside the output grid
 * (S + 2*padding - k) / stride + 1 get no pairs); input and submanifold
 * coordinates must be distinct. */
int cs_pairs_strided_workspace(int64_t n_in, int64_t n_out, int64_t* nbytes);
int cs_pairs_strided(const int32_t* in_coords, int64_t n_in, const int32_t* in_shape, const int32_t* out_coords,
                     int64_t n_out, int32_t* pair_fwd, int32_t* pair_bwd, uint32_t* mask_fwd_sorted,
                     int32_t* argsort_fwd, uint32_t* mask_bwd_sorted, int32_t* argsort_bwd, void* workspace,
                     int64_t ws_bytes, uint64_t stream);
/* Per-row masks in row order, for callers that pass unsorted masks to
 * implicit_gemm2 (no counterpart in the reference, whose table generation
 * always sorts the masks; for inference stacks, where the tile grouping does
 * not change any row's result and skipping the sort saves its time): mask[i]
 * word w bit b <=> pair[32w + b][i] != -1 ([n][div_up(kv, 32)] words), and
 * identity[i] = i (the argsort to pass with them) when identity is not NULL.
 * Queued on `stream`.  The table builders leave the forward masks to this when
 * their mask_sorted / argsort (mask_fwd_sorted / argsort_fwd) are both NULL. */
int cs_pairs_row_masks(const int32_t* pair, int64_t n, int32_t kv, uint32_t* mask, int32_t* identity,
                       uint64_t stream);
const char* cs_pairs_last_error(void);

/* ---- Standalone row gather / scatter-add, the glue around matmul2 of
 * spconv's gather-GEMM-scatter ("native") algorithm.
 *   GatherAll.get_all_gather_params / ScatterAll.get_all_scatter_params
 *     (cumm/gemm/gather.py:445-455, 552-570): writes up to `cap` entries of
 *     (tile_m, tile_k_bytes, bytes_per_access, num_threads) into out[4*i..],
 *     returns the total.
 *   GatherAll.supported (gather.py:459-468): (channel_size * itemsize) %
 *     bytes_per_access == 0.  ScatterAll.supported_scatter (gather.py:574-589)
 *     additionally needs a published float32 parameter set.
 *   GatherAll.gather (gather.py:482-495 -> Gather.gather 224-257):
 *     output[i] = input[indices[i]] for i < indices.dim(0); any dtype.
 *   ScatterAll.scatter (gather.py:603-618 -> Scatter.scatter 356-389):
 *     output[indices[i]] += input[i]; float32; the indices of one call must be
 *     distinct (the reference adds without atomics).
 * Tile parameters must be a published set ("can't find a sutable gather" /
 * "...scatter" otherwise, code 2); the kernel's own launch shape is chosen for
 * the row width.  Rows whose index lies outside the source (gather: reads
 * zeros) or target (scatter: skipped) are not read or written. */
int cs_get_all_gather_params(int32_t* out, int cap);
int cs_get_all_scatter_params(int32_t* out, int cap);
int cs_gather_supported(int32_t bytes_per_access, int32_t channel_size, int32_t dtype);
int cs_scatter_supported(int64_t tile_m, int64_t tile_k_bytes, int64_t bytes_per_access, int64_t num_threads,
                         int32_t channel_size, int32_t dtype);
int cs_gather(const cs_tensor* output, const cs_tensor* input, const cs_tensor* indices, int64_t tile_m,
              int64_t tile_k_bytes, int64_t bytes_per_access, int64_t num_threads, uint64_t stream);
int cs_scatter(const cs_tensor* output, const cs_tensor* input, const cs_tensor* indices, int64_t tile_m,
               int64_t tile_k_bytes, int64_t bytes_per_access, int64_t num_threads, uint64_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CUMM_AMD_H_ */
