// Host-visible launch interface of the gfx950 kernels (no torch types).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cs {

// Element type tags used by the launchers; values are tv::DType codes.
enum DT : int {
  kF32 = 0, kI32 = 1, kI8 = 3, kF64 = 4, kF16 = 7, kU32 = 10, kBF16 = 12, kE4M3 = 14
};

// Row-gather implicit GEMM (sparse conv forward / dgrad, gather-A scatter-C GEMM).
//
// Sorted row r (0 <= r < m) of the output tile space accumulates, for every
// kernel offset f enabled in its 16-row tile mask,
//   acc[r][n] += sum_k A[g(r,f)][k] * B_f[n][k]
// with g(r,f) = pairs[f*pair_ld + col(r)] (or col(r) when pairs is null),
// col(r) = pair_col ? pair_col[r] : r, and rows with g = -1 contributing zero.
// B_f[n][k] = b[f*b_fs + n*b_ns + k].  The result row is written to
// out_row ? out_row[r] : r as act(alpha * scale[n] * acc + beta * src).
struct RowGatherArgs {
  const void* a; int64_t lda;
  const void* b; int64_t b_fs; int64_t b_ns;
  // b_kmajor: B_f[n][k] = b[k*b_ks + f*b_fs + n] instead (dgrad's KRSC weights);
  // kernels that need the n-major form build it in b_scratch ([kv][nout][kred])
  int32_t b_kmajor; int64_t b_ks; void* b_scratch;
  uint32_t a_bytes, b_bytes;  // extents of a and b (< 2^31): buffer-load range
  void* c; int64_t ldc;
  const void* d; int64_t ldd; int32_t d_is_bias;
  const int32_t* pairs; int64_t pair_ld;
  const int32_t* pair_col;
  const int32_t* out_row;
  const uint32_t* mask;      // sorted per-row masks ([m][mw] words); null -> every offset of the filter
  uint32_t* mask_out;        // per-16-row tile masks ([tiles][mw], forward), may be null
  int32_t mw;                // mask words per row: 1 (kv <= 32) or div_up(kv, 32) <= kMaxDynWords (dynamic masks)
  const float* scale;        // per-output-column multiplier, may be null
  // int8 inference (int8 a, b, c): c = act(scale[n] * acc + bias_f[n] + beta * d[o][n])
  const float* bias_f;
  int32_t m, kred, nout, kv;
  uint32_t mask_filter;
  int32_t reverse_mask;      // dgrad of a submanifold layer reuses forward masks
  int32_t skip_empty;        // dgrad: tiles whose (mask & filter) is empty are not written
  float alpha, beta, act_alpha, act_beta;
  int32_t act_type;
};

// Row-pair outer-product reduction (sparse conv wgrad, ShuffleAB GEMM).
//   out[f][mm][nn] = sum over offset f's pairs (i, j) of A[i][mm] * B[j][nn]
// Conv form (fixed_total < 0): sorted row r, original row i = a_rows[r]
// (argsort); i holds a pair for f when j = pairs[f*pair_ld + i] >= 0 and bit f
// of the mask word of r's group (mask[r / mask_width], all ones when mask is
// null) and of mask_filter are set.  Kernels build rowmask[i], then per
// offset the list of (i, j) in ascending i (list_a / list_b, total[f]).
// List form (fixed_total >= 0, kv = 1): list_a / list_b are given (a_inds,
// b_inds) with fixed_total entries.
// Offset f's list is cut into nb_f = min(max_blocks, div_up(total_f, min_rows))
// contiguous pieces; each block writes an FP32 partial for its piece and a
// fixed-order tree reduces the pieces (deterministic: the partition depends
// only on the inputs; no atomics).
struct WgradArgs {
  const void* a; int64_t lda;
  const void* b; int64_t ldb;
  uint32_t a_bytes, b_bytes;  // extents of a and b (< 2^31): buffer-load range
  const int32_t* a_rows;      // conv form: mask argsort
  const int32_t* pairs; int64_t pair_ld;
  const uint32_t* mask; int32_t mask_width; int32_t num_mask;
  int32_t mw;                 // mask words per group / row (1, or div_up(kv, 32) for dynamic masks)
  int32_t nrows, M, N, kv;
  uint32_t mask_filter;
  int32_t max_blocks;        // pieces per offset, upper bound, <= kWgradMaxBlocks
  int32_t min_rows;          // pairs per piece, lower bound
  int32_t fixed_total;       // list form: entries of the given lists; else -1
  uint32_t* rowmask;         // [nrows][mw]
  int32_t* bcount;           // [kv][div_up(nrows, 4096)]
  int32_t* list_a;           // [kv][list_ld] (list form: a_inds, read only)
  int32_t* list_b;           // (list form: b_inds, read only)
  int64_t list_ld;
  int32_t* total;            // [kv]
  float* partial;            // [kv][max_blocks][M][N] FP32
  int32_t xcd_grid;          // set by launch_wgrad: partial products on a grid grouped per XCD (kv <= 64)
};

// out[mm*o_ms + f*o_fs + nn] = act(alpha * tree_sum_b partial[f][b][mm][nn] + beta * src)
// over the nb_f pieces of offset f (from total / fixed_total, as in the wgrad kernel);
// src = src_is_bias ? src[f*N + nn] : src[mm*s_ms + f*s_fs + nn]  (src may alias out)
// Integer form (int_compute, int8 operands): the partials hold int32 sums,
// added in wrapping int32 arithmetic, and the epilogue is int_epilogue with
// ialpha / ibeta / iact_alpha.
struct ReduceArgs {
  const float* partial; const int32_t* total; int32_t fixed_total, max_blocks, min_rows, kv, M, N;
  void* out; int64_t o_ms, o_fs;
  const void* src; int64_t s_ms, s_fs; int32_t src_is_bias;
  uint32_t mask_filter;
  float alpha, beta, act_alpha;
  int32_t act_type;
  int32_t int_compute, ialpha, ibeta, iact_alpha;
};

// Gather GEMM (matmul2 NoShuffle / ShuffleAC):
//   C[c_rows ? c_rows[i] : i][n] = epilogue(sum_k A[a_rows ? a_rows[i] : i][k] * B(n, k))
// for 0 <= i < m, 0 <= n < nn, with B(n, k) = b[n*ldb + k] (n-major) or
// b[k*ldb + n] (b_kmajor).  Float epilogue: act(alpha * acc + beta * src);
// integer (int8 operands, int32 accumulation): int_epilogue, then int8
// saturation or the int32 value.  src = d_is_bias ? d[n] : d[o*ldd + n].
struct GemmRowsArgs {
  const void* a; int64_t lda; uint32_t a_bytes;
  const int32_t* a_rows;
  const int32_t* c_rows;
  const void* b; int64_t ldb; uint32_t b_bytes; int32_t b_kmajor;
  void* c; int64_t ldc;
  const void* d; int64_t ldd; int32_t d_is_bias;
  int32_t m, nn, k;
  float alpha, beta, act_alpha;
  int32_t act_type;
  int32_t ialpha, ibeta, iact_alpha;
};
hipError_t launch_gemm_rows(int in_dtype, int out_dtype, const GemmRowsArgs& args, hipStream_t s);
bool gemm_rows_supported(int in_dtype, int out_dtype);
hipError_t launch_gemm_rows_f32(int out_dtype, const GemmRowsArgs& args, hipStream_t s);
hipError_t launch_gemm_rows_f16(int out_dtype, const GemmRowsArgs& args, hipStream_t s);
hipError_t launch_gemm_rows_bf16(int out_dtype, const GemmRowsArgs& args, hipStream_t s);
hipError_t launch_gemm_rows_i8(int out_dtype, const GemmRowsArgs& args, hipStream_t s);

// block_n: output columns per wave, 64 or 128.
hipError_t launch_rowgather_f32(int out_dtype, int block_n, const RowGatherArgs& args, hipStream_t s);
hipError_t launch_rowgather_f16(int out_dtype, int block_n, const RowGatherArgs& args, hipStream_t s);
hipError_t launch_rowgather_bf16(int out_dtype, int block_n, const RowGatherArgs& args, hipStream_t s);
hipError_t launch_rowgather_e4m3(int out_dtype, int block_n, const RowGatherArgs& args, hipStream_t s);
hipError_t launch_rowgather_i8(int out_dtype, int block_n, const RowGatherArgs& args, hipStream_t s);
hipError_t launch_rowgather(int in_dtype, int out_dtype, int block_n, const RowGatherArgs& args, hipStream_t s);
bool rowgather_supported(int in_dtype, int out_dtype);
int rowgather_block_n(int nout);
// plan + main wgrad kernels
hipError_t launch_wgrad(int in_dtype, const WgradArgs& args, hipStream_t s);
hipError_t launch_reduce(int out_dtype, const ReduceArgs& args, hipStream_t s);
// out[f][c][r] = in[r][f][c] for a [R, F, Cc] tensor of `elem_bytes`-byte elements.
hipError_t launch_transpose(const void* in, void* out, int64_t R, int64_t F, int64_t Cc,
                            int64_t in_row_stride, int elem_bytes, hipStream_t s);
// 2-D copy with strides (elements), used to make operands dense.
hipError_t launch_zero(void* p, int64_t nbytes, hipStream_t s);
// out[i] = in[idx[i]] for m rows of row_bytes bytes (idx outside [0, n_in): zeros)
hipError_t launch_gather_rows(const void* in, void* out, const int32_t* idx, int64_t m, int64_t n_in,
                              int64_t row_bytes, int64_t in_ld_bytes, int64_t out_ld_bytes, hipStream_t s);
// out[idx[i]] += in[i] for m float rows of `cols` (idx outside [0, n_out): skipped)
hipError_t launch_scatter_add_f32(const float* in, float* out, const int32_t* idx, int64_t m, int64_t n_out,
                                  int64_t cols, int64_t in_ld, int64_t out_ld, hipStream_t s);

// The main kernel of this thread's last launch (a verbose caller reports its
// register count).  CS_LAUNCH records it and launches.
const void*& last_kernel();
#define CS_LAUNCH(K, ...)                                          \
  do {                                                             \
    ::cs::last_kernel() = reinterpret_cast<const void*>(&K);       \
    hipLaunchKernelGGL(K, __VA_ARGS__);                            \
  } while (0)

constexpr int kRowsPerWave = 64;     // four 16-row mask tiles per wave
constexpr int kWavesPerBlock = 4;
constexpr int kMaskTile = 16;        // rows per kernel-offset mask tile
constexpr int kMaxMaskWords = 4;     // dynamic masks: words per window of offsets
constexpr int kMaxDynWords = 32;     // dynamic masks: words per row, kv <= 1024
constexpr int kWgradRows = 64;       // rows staged per wgrad step
constexpr int kWgradMaxBlocks = 128; // pieces per offset, upper bound
constexpr int kWgradTile = 64;       // wgrad output tile (M and N)
constexpr int kWgradXcds = 8;        // XCDs of an MI355X: wgrad pieces are grouped per XCD

}  // namespace cs
