// C-ABI host layer of libcumm_amd.so: validation, kernel catalogue, workspace
// sizing and launch for ConvMainUnitTest.implicit_gemm2 and
// GemmMainUnitTest.matmul2.  Validation follows the reference host preambles
// (cumm/conv/nvrtc_code.py:20-196, cumm/gemm/nvrtc_code.py:34-145): the same
// checks raise the same error class (ValueError <-> code 1, RuntimeError <->
// code 2).  Messages carry file(line) like TV_THROW_* do
// (include/tensorview/core/common.h:102-182).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cumm_amd.h"
#include "cs_kernels.h"

namespace {

thread_local std::string g_last_error;

template <typename... Args>
std::string cat(Args&&... args) {
  std::ostringstream ss;
  (void)std::initializer_list<int>{((ss << args), 0)...};
  return ss.str();
}

#define CS_RT(cond, ...)                                                              \
  do {                                                                                \
    if (!(cond))                                                                      \
      throw std::runtime_error(cat(__FILE__, "(", __LINE__, ")\n", #cond, " assert faild. ", \
                                   __VA_ARGS__));                                     \
  } while (0)
#define CS_ARG(cond, ...)                                                             \
  do {                                                                                \
    if (!(cond))                                                                      \
      throw std::invalid_argument(cat(__FILE__, "(", __LINE__, ")\n", #cond, " assert faild. ", \
                                      __VA_ARGS__));                                  \
  } while (0)
#define CS_HIP(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw std::runtime_error(cat(__FILE__, "(", __LINE__, ")\nHIP error ", (int)e_, ": ", \
                                   hipGetErrorString(e_), " in ", #expr));           \
  } while (0)

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return CS_OK;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return CS_ERR_INVALID_ARG;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return CS_ERR_RUNTIME;
  } catch (...) {
    g_last_error = "unknown error";
    return CS_ERR_RUNTIME;
  }
}

// ------------------------------------------------------------------ dtypes
// tv::DType codes and short names (include/tensorview/dtypes.h:39-60, 217-267).
const char* dtype_short_str(int dt) {
  switch (dt) {
    case 5: return "b1";
    case 0: return "f32";
    case 3: return "s8";
    case 2: return "s16";
    case 1: return "s32";
    case 4: return "f64";
    case 8: return "s64";
    case 6: return "u8";
    case 9: return "u16";
    case 10: return "u32";
    case 11: return "u64";
    case 7: return "f16";
    case 14: return "e4m3";
    case 15: return "e5m2";
    case 13: return "tf32";
    case 12: return "bf16";
    default: return "unknown";
  }
}

int dtype_bytes(int dt) {
  switch (dt) {
    case 5: case 3: case 6: case 14: case 15: return 1;
    case 2: case 7: case 9: case 12: return 2;
    case 0: case 1: case 10: case 13: return 4;
    case 4: case 8: case 11: return 8;
    default: return 0;
  }
}

constexpr int kUnknown = -1;
const char* kAlgo = "CDNA4";

// ------------------------------------------------------------------ tensors
bool empty(const cs_tensor& t) { return t.data == nullptr; }

int64_t dim(const cs_tensor& t, int i) {
  CS_RT(i >= 0 && i < t.ndim, "dim index ", i, " out of range for ndim ", t.ndim);
  return t.shape[i];
}

int64_t numel(const cs_tensor& t) {
  if (empty(t)) return 0;
  int64_t n = 1;
  for (int i = 0; i < t.ndim; ++i) n *= t.shape[i];
  return n;
}

int64_t nbytes(const cs_tensor& t) { return numel(t) * dtype_bytes(t.dtype); }

// tv::check_shape: -1 matches any extent.
void check_shape(const cs_tensor& t, std::initializer_list<int64_t> shape, const char* name) {
  CS_RT(t.ndim == (int)shape.size(), name, " rank ", t.ndim, " != ", shape.size());
  int i = 0;
  for (int64_t s : shape) {
    CS_RT(s < 0 || t.shape[i] == s, name, " dim ", i, " is ", t.shape[i], ", expected ", s);
    ++i;
  }
}

// Rows of a 2-D (or flattened 3-D) operand must be contiguous along the last dim.
void check_rows(const cs_tensor& t, const char* name) {
  CS_ARG(t.ndim >= 1 && t.stride[t.ndim - 1] == 1, name, " must be contiguous in its last dim");
  for (int i = t.ndim - 2; i >= 0; --i)
    CS_ARG(t.stride[i] >= t.shape[i + 1] * t.stride[i + 1] || t.shape[i] <= 1, name,
           " has overlapping strides");
}

void check_dense(const cs_tensor& t, const char* name) {
  int64_t expect = 1;
  for (int i = t.ndim - 1; i >= 0; --i) {
    CS_ARG(t.shape[i] <= 1 || t.stride[i] == expect, name, " must be contiguous");
    expect *= t.shape[i];
  }
}

void check_device(const cs_tensor& t, const char* name) {
  CS_ARG(t.device >= 0, name, " must be a GPU tensor");
}

void check_index_dtype(const cs_tensor& t, const char* name) {
  CS_ARG(t.dtype == 1 || t.dtype == 10, name, " must be int32 or uint32, got ", dtype_short_str(t.dtype));
}

int64_t row_stride(const cs_tensor& t) { return t.ndim >= 2 ? t.stride[0] : t.shape[0]; }

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Bytes from the first to one past the last element of a strided tensor.
int64_t span_bytes(const cs_tensor& t) {
  if (empty(t) || numel(t) == 0) return 0;
  int64_t last = 0;
  for (int i = 0; i < t.ndim; ++i) last += (t.shape[i] - 1) * t.stride[i];
  return (last + 1) * dtype_bytes(t.dtype);
}

// The kernels address gathered operands with 32-bit byte offsets, the same
// int32 limit the reference asserts (cumm/conv/nvrtc_code.py:131-143,
// cumm/gemm/nvrtc_code.py:87-88, 104-107).
constexpr int64_t kInt32Max = 2147483647;
void check_int32_span(const cs_tensor& t, const char* name) {
  CS_RT(span_bytes(t) < kInt32Max, "your data exceed int32 range (", name, ").");
}

// ------------------------------------------------------------------ desp helpers
std::string gemm_repr(const cs_gemm_algo_desp& d) {
  // GemmAlgoDesp::__repr__  (include/tensorview/gemm/core/params.h:56-78)
  std::ostringstream ss;
  ss << d.algo << "_" << dtype_short_str(d.dtype_a) << dtype_short_str(d.dtype_b)
     << dtype_short_str(d.dtype_c) << dtype_short_str(d.dacc) << dtype_short_str(d.dcomp);
  ss << (d.trans_a == 1 ? "n" : "t") << (d.trans_b == 1 ? "n" : "t") << (d.trans_c == 1 ? "n" : "t");
  ss << "_m" << d.tile_shape[0] << "n" << d.tile_shape[1] << "k" << d.tile_shape[2];
  ss << "m" << d.warp_tile_shape[0] << "n" << d.warp_tile_shape[1] << "k" << d.warp_tile_shape[2];
  ss << "A" << d.access_per_vector;
  if (d.tensorop[0] != -1) ss << "T" << d.tensorop[0] << d.tensorop[1] << d.tensorop[2];
  ss << "_" << d.num_stage;
  ss << (d.split_k_serial == 1 ? 1 : 0) << (d.split_k_parallel == 1 ? 1 : 0);
  if (d.shuffle_type != 0) ss << "_S" << d.shuffle_type;
  return ss.str();
}

std::string conv_repr(const cs_conv_algo_desp& d) {
  // ConvAlgoDesp::__repr__  (include/tensorview/gemm/core/params.h:191-224)
  std::ostringstream ss;
  ss << gemm_repr(d.gemm);
  ss << "_C" << d.ndim << d.op_type << d.iter_algo;
  auto lay = [](int l, int il) {
    std::string s = l == 0 ? "F" : "L";
    if (il > 1) s += std::to_string(il);
    return s;
  };
  ss << lay(d.layout_i, d.interleave_i) << lay(d.layout_w, d.interleave_w) << lay(d.layout_o, d.interleave_o);
  if (d.mask_sparse) ss << "_" << (d.increment_k_first ? "SK" : "SF");
  if (d.dynamic_mask) ss << "D";
  if (d.is_int8_inference) ss << "_S8";
  return ss.str();
}

void check_valid(const cs_gemm_algo_desp& d) {
  // GemmAlgoDesp::check_valid  (params.h:84-108)
  CS_RT(d.trans_a != -1 && d.trans_b != -1 && d.trans_c != -1 && d.algo[0] != 0,
        "trans_a, trans_b, trans_c and algo must be set");
  for (int i = 0; i < 3; ++i)
    CS_RT(d.tile_shape[i] > 0 && d.warp_tile_shape[i] > 0, "tile_shape and warp_tile_shape must be set");
  std::string algo(d.algo);
  if (algo != "Simt" && algo != "SimtDP4A" && algo != "SimtDP2A")
    for (int i = 0; i < 3; ++i) CS_RT(d.tensorop[i] > 0, "tensorop must be set");
  CS_RT(d.dtype_a != kUnknown && d.dtype_b != kUnknown && d.dtype_c != kUnknown,
        "dacc and dcomp must be set to valid value");
  CS_RT(d.dacc != kUnknown && d.dcomp != kUnknown, "dacc and dcomp must be set to valid value");
  CS_RT(d.num_stage > 0, "num_stage must larger than zero");
}

bool gemm_supported(const cs_gemm_algo_desp& d, int64_t m, int64_t n, int64_t k) {
  // GemmAlgoDesp::supported  (params.h:133-149)
  bool res = true;
  int64_t lda = d.trans_a == 1 ? m : k;
  int64_t ldb = d.trans_b == 1 ? k : n;
  int64_t ldc = d.trans_c == 1 ? m : n;
  if (d.element_per_access_a > 0) res &= lda % d.element_per_access_a == 0;
  if (d.element_per_access_b > 0) res &= ldb % d.element_per_access_b == 0;
  if (d.element_per_access_c > 0) res &= ldc % d.element_per_access_c == 0;
  return res;
}

bool conv_supported(const cs_conv_algo_desp& d, int64_t m, int64_t n, int64_t k, int64_t C, int64_t K,
                    int mask_width) {
  // ConvAlgoDesp::supported  (params.h:266-285)
  bool res = gemm_supported(d.gemm, m, n, k);
  if (d.mask_sparse) {
    if (d.op_type == 0) {
      res &= C % d.gemm.element_per_access_a == 0;
    } else if (d.op_type == 1) {
      res &= K % d.gemm.element_per_access_a == 0;
    } else {
      res &= mask_width > 0 && mask_width % d.gemm.tile_shape[2] == 0;
      res &= K % d.gemm.element_per_access_a == 0;
      res &= C % d.gemm.element_per_access_b == 0;
    }
  }
  return res;
}

// ------------------------------------------------------------------ catalogue
int chunk_elems(int dt) { return dt == 0 ? 16 : (dt == 14 || dt == 3) ? 64 : 32; }
int vec_elems(int dt) { return 16 / dtype_bytes(dt); }
int mfma_k(int dt) { return dt == 0 ? 4 : dt == 3 ? 64 : 32; }

cs_gemm_algo_desp base_gemm_desp() {
  cs_gemm_algo_desp d;
  std::memset(&d, 0, sizeof(d));
  d.dtype_a = d.dtype_b = d.dtype_c = kUnknown;
  d.trans_a = d.trans_b = d.trans_c = -1;
  for (int i = 0; i < 3; ++i) d.tile_shape[i] = d.warp_tile_shape[i] = d.tensorop[i] = -1;
  d.num_stage = -1;
  d.dacc = d.dcomp = kUnknown;
  d.element_per_access_a = d.element_per_access_b = d.element_per_access_c = -1;
  d.access_per_vector = 1;
  std::strncpy(d.algo, kAlgo, CS_ALGO_NAME_LEN - 1);
  d.min_arch[0] = 9;
  d.min_arch[1] = 5;
  d.dacc = 0;
  d.dcomp = 0;
  d.num_stage = 1;
  return d;
}

// Row-gather kernels: tile = [rows per mask tile, output cols, reduction chunk].
cs_gemm_algo_desp rowgather_desp(int dta, int dtb, int dtc, int tile_m, int bn, int ta, int tb, int tc) {
  cs_gemm_algo_desp d = base_gemm_desp();
  d.dtype_a = dta;
  d.dtype_b = dtb;
  d.dtype_c = dtc;
  d.trans_a = ta;
  d.trans_b = tb;
  d.trans_c = tc;
  const int ch = chunk_elems(dta);
  d.tile_shape[0] = tile_m;
  d.tile_shape[1] = bn;
  d.tile_shape[2] = ch;
  d.warp_tile_shape[0] = cs::kRowsPerWave;
  d.warp_tile_shape[1] = bn;
  d.warp_tile_shape[2] = ch;
  d.tensorop[0] = 16;
  d.tensorop[1] = 16;
  d.tensorop[2] = mfma_k(dta);
  d.element_per_access_a = vec_elems(dta);
  d.element_per_access_b = vec_elems(dtb);
  d.element_per_access_c = 1;
  return d;
}

// Row-pair reduction kernels: tile = [M tile, N tile, rows per mask check].
cs_gemm_algo_desp wgrad_desp(int dta, int dtb, int dtc, int tile_k) {
  cs_gemm_algo_desp d = base_gemm_desp();
  d.dtype_a = dta;
  d.dtype_b = dtb;
  d.dtype_c = dtc;
  d.trans_a = 1;
  d.trans_b = 0;
  d.trans_c = 0;
  d.tile_shape[0] = cs::kWgradTile;
  d.tile_shape[1] = cs::kWgradTile;
  d.tile_shape[2] = tile_k;
  d.warp_tile_shape[0] = 32;
  d.warp_tile_shape[1] = 32;
  d.warp_tile_shape[2] = chunk_elems(dta);
  d.tensorop[0] = 16;
  d.tensorop[1] = 16;
  d.tensorop[2] = mfma_k(dta);
  d.element_per_access_a = vec_elems(dta);
  d.element_per_access_b = vec_elems(dtb);
  d.element_per_access_c = 1;
  d.split_k_parallel = 1;
  return d;
}

cs_conv_algo_desp conv_desp(int op, const cs_gemm_algo_desp& g) {
  cs_conv_algo_desp d;
  std::memset(&d, 0, sizeof(d));
  d.gemm = g;
  d.ndim = 3;
  d.op_type = op;
  d.iter_algo = 1;  // kOptimized
  d.layout_i = d.layout_w = d.layout_o = 1;  // kChannelLast
  d.interleave_i = d.interleave_w = d.interleave_o = 1;
  d.mask_sparse = 1;
  d.increment_k_first = 1;  // per offset, all reduction chunks
  d.dynamic_mask = 0;
  d.is_int8_inference = 0;
  return d;
}

struct ConvCombo { int in, w, out; };

const std::vector<cs_conv_algo_desp>& conv_catalogue() {
  static const std::vector<cs_conv_algo_desp> cat = [] {
    std::vector<cs_conv_algo_desp> v;
    const ConvCombo fwd_combos[] = {{0, 0, 0}, {7, 7, 7}, {12, 12, 12}, {14, 14, 7}, {14, 14, 12}};
    // tile_shape[1]: column slice of the weight-stationary kernel (16 / 32 / 64;
    // 64 switches to streamed weights past two slices), 128 = streamed weights
    for (int bn : {64, 32, 16, 128}) {
      for (auto c : fwd_combos) {
        // forward: a = input, b = weight, c = output; trans (F, T, F)
        v.push_back(conv_desp(0, rowgather_desp(c.in, c.w, c.out, cs::kMaskTile, bn, 0, 1, 0)));
      }
      for (auto c : fwd_combos) {
        // dgrad: a = dout (c.in here is dout's dtype), b = weight, c = din; trans (F, F, F)
        v.push_back(conv_desp(1, rowgather_desp(c.in, c.w, c.out, cs::kMaskTile, bn, 0, 0, 0)));
      }
    }
    const ConvCombo wg_combos[] = {{0, 0, 0}, {7, 7, 7}, {12, 12, 12}, {7, 7, 0}, {12, 12, 0}, {14, 14, 0}};
    for (auto c : wg_combos) {
      // wgrad: a = dout, b = input, c = dW; trans (T, F, F)
      v.push_back(conv_desp(2, wgrad_desp(c.in, c.w, c.out, cs::kMaskTile)));
    }
    // dynamic (multi-word) masks for kv > 32: forward / dgrad with streamed
    // weights (128) or the column-sliced kernels (64), and wgrad
    for (int bn : {64, 128}) {
      for (auto c : fwd_combos) {
        cs_conv_algo_desp f = conv_desp(0, rowgather_desp(c.in, c.w, c.out, cs::kMaskTile, bn, 0, 1, 0));
        f.dynamic_mask = 1;
        v.push_back(f);
        cs_conv_algo_desp d = conv_desp(1, rowgather_desp(c.in, c.w, c.out, cs::kMaskTile, bn, 0, 0, 0));
        d.dynamic_mask = 1;
        v.push_back(d);
      }
    }
    for (auto c : wg_combos) {
      cs_conv_algo_desp w = conv_desp(2, wgrad_desp(c.in, c.w, c.out, cs::kMaskTile));
      w.dynamic_mask = 1;
      v.push_back(w);
    }
    // int8 inference forward: int8 input / weight / output, int32 accumulation,
    // float epilogue act(scale * acc + bias + beta * output_add)
    for (int bn : {64, 32, 16, 128}) {
      cs_gemm_algo_desp g = rowgather_desp(3, 3, 3, cs::kMaskTile, bn, 0, 1, 0);
      g.dacc = 1;
      g.dcomp = 0;
      cs_conv_algo_desp d = conv_desp(0, g);
      d.is_int8_inference = 1;
      v.push_back(d);
    }
    return v;
  }();
  return cat;
}

const std::vector<cs_gemm_algo_desp>& gemm_catalogue() {
  static const std::vector<cs_gemm_algo_desp> cat = [] {
    std::vector<cs_gemm_algo_desp> v;
    const int dts[] = {0, 7, 12};
    const int rows = cs::kRowsPerWave;
    for (int bn : {64}) {
      for (int dt : dts) {
        for (int ta = 0; ta < 2; ++ta)
          for (int tb = 0; tb < 2; ++tb)
            for (int tc = 0; tc < 2; ++tc) v.push_back(rowgather_desp(dt, dt, dt, rows, bn, ta, tb, tc));
        for (int tb = 0; tb < 2; ++tb) {
          auto d = rowgather_desp(dt, dt, dt, rows, bn, 0, tb, 0);
          d.shuffle_type = 1;
          v.push_back(d);
        }
      }
    }
    for (int dt : dts) {
      auto d = wgrad_desp(dt, dt, dt, cs::kWgradRows);
      d.shuffle_type = 2;
      v.push_back(d);
    }
    // int8 operands with int32 accumulation and int32 epilogue arithmetic,
    // int8 (saturated) or int32 output: "s8,s8,s8,s32,s32" and
    // "s8,s8,s32,s32,s32" (cumm/gemm/main.py:348-366, 499-523)
    for (int dc : {3, 1}) {
      auto i8 = [&](cs_gemm_algo_desp d) {
        d.dacc = 1;
        d.dcomp = 1;
        return d;
      };
      for (int ta = 0; ta < 2; ++ta)
        for (int tb = 0; tb < 2; ++tb)
          for (int tc = 0; tc < 2; ++tc) v.push_back(i8(rowgather_desp(3, 3, dc, rows, 64, ta, tb, tc)));
      for (int tb = 0; tb < 2; ++tb) {
        auto d = i8(rowgather_desp(3, 3, dc, rows, 64, 0, tb, 0));
        d.shuffle_type = 1;
        v.push_back(d);
      }
      auto d = i8(wgrad_desp(3, 3, dc, cs::kWgradRows));
      d.shuffle_type = 2;
      v.push_back(d);
    }
    // "f16,f16,f16,f16,f16" (f16 accumulator and epilogue type in the
    // reference): the same kernels, which accumulate and compute in FP32 and
    // round once into the f16 output.  No ShuffleAB entry: its split-K
    // workspace, sized by the reference for f16 partial sums, cannot hold the
    // FP32 partials (the f16 ShuffleAB entry with FP32 accumulation serves it)
    {
      auto h = [&](cs_gemm_algo_desp d) {
        d.dacc = 7;
        d.dcomp = 7;
        return d;
      };
      for (int ta = 0; ta < 2; ++ta)
        for (int tb = 0; tb < 2; ++tb)
          for (int tc = 0; tc < 2; ++tc) v.push_back(h(rowgather_desp(7, 7, 7, rows, 64, ta, tb, tc)));
      for (int tb = 0; tb < 2; ++tb) {
        auto d = h(rowgather_desp(7, 7, 7, rows, 64, 0, tb, 0));
        d.shuffle_type = 1;
        v.push_back(d);
      }
    }
    return v;
  }();
  return cat;
}

// Kernel identity is the kernel name (ConvAlgoDesp::__repr__); the catalogue's
// names are built once.
const cs_conv_algo_desp* find_conv_kernel(const cs_conv_algo_desp& desp) {
  static const std::vector<std::string> names = [] {
    std::vector<std::string> v;
    for (const auto& d : conv_catalogue()) v.push_back(conv_repr(d));
    return v;
  }();
  const std::string want = conv_repr(desp);
  const auto& catalogue = conv_catalogue();
  for (size_t i = 0; i < catalogue.size(); ++i)
    if (names[i] == want) return &catalogue[i];
  return nullptr;
}

// ------------------------------------------------------------------ workspace
struct Workspace {
  void* ptr = nullptr;
  bool owned = false;
  hipStream_t stream = nullptr;
  ~Workspace() {
    if (owned && ptr) (void)hipFreeAsync(ptr, stream);
  }
};

void acquire_workspace(Workspace& ws, const cs_tensor& given, int64_t need, hipStream_t s) {
  ws.stream = s;
  if (need <= 0) return;
  if (!empty(given)) {
    CS_RT(nbytes(given) >= need, "workspace at least", need, "bytes.");
    ws.ptr = given.data;
    return;
  }
  CS_HIP(hipMallocAsync(&ws.ptr, (size_t)need, s));
  ws.owned = true;
}

int64_t align256(int64_t v) { return (v + 255) / 256 * 256; }

// Pieces per offset of the planned wgrad partition: the caller's split_k_slices
// when > 1, else a fixed default; depends only on the call's parameters.
int resolve_splits(int requested, int64_t M, int64_t N) {
  int s = requested > 1 ? std::min(requested, cs::kWgradMaxBlocks) : cs::kWgradMaxBlocks;
  // one offset's partials are addressed with 32-bit byte offsets
  const int64_t per = std::max<int64_t>(1, M * N * 4);
  s = (int)std::min<int64_t>(s, std::max<int64_t>(1, 2147483647LL / per));
  return s;
}

// Lower bound of pairs per wgrad piece: pieces of an offset scale with its
// pair count.  Fewer pieces mean fewer FP32 partials to write and reduce,
// more keep more blocks in flight.  Measured on MI355X: 1024 pairs (16 stages
// of 64 rows at 64 channels) for 64-wide output tiles; with 128 x 128 tiles a
// partial is four times larger and 2048 pairs were 14-20 % faster for every
// dtype at 200 000 voxels, 128 -> 128 (1024 and 4096 slower).
constexpr int kWgradMinRows = 1024;
int wgrad_min_rows(int64_t M, int64_t N) { return M > 64 && N > 64 ? 2 * kWgradMinRows : kWgradMinRows; }

// wgrad workspace: FP32 partials, then (conv form only) the effective row
// masks, per-block pair counts, the per-offset pair lists and their totals.
struct WgradLayout {
  int64_t partial = 0, rowmask = 0, bcount = 0, list_a = 0, list_b = 0, total = 0, size = 0;
};

WgradLayout wgrad_layout(int64_t kv, int64_t M, int64_t N, int64_t nrows, int splits, bool lists, int mw = 1) {
  WgradLayout w;
  w.partial = 0;
  int64_t off = align256(kv * splits * M * N * 4);
  if (lists) {
    const int64_t nblk = std::max<int64_t>(1, (nrows + 4095) / 4096);
    w.rowmask = off;
    off += align256(std::max<int64_t>(nrows, 1) * 4 * mw);
    w.bcount = off;
    off += align256(kv * nblk * 4);
    w.list_a = off;
    off += align256(kv * std::max<int64_t>(nrows, 1) * 4);
    w.list_b = off;
    off += align256(kv * std::max<int64_t>(nrows, 1) * 4);
    w.total = off;
    off += align256(kv * 4);
  }
  w.size = off;
  return w;
}

int popcount32(uint32_t v) { return __builtin_popcount(v); }

// ------------------------------------------------------------------ conv
struct ConvPlan {
  const cs_conv_algo_desp* kernel = nullptr;
  int op = 0;
  int64_t N = 0, C = 0, K = 0, kv = 0;
  int64_t m = 0, n = 0, k = 0;
  int splits = 1;
  int mw = 1;  // mask words per row (dynamic masks: div_up(kv, 32))
  int64_t workspace = 0;
};

ConvPlan plan_conv(const cs_conv_params& p) {
  // nvrtc_conv_template  (cumm/conv/nvrtc_code.py:20-196)
  const cs_conv_algo_desp& desp = p.conv_algo_desp;
  check_valid(desp.gemm);
  ConvPlan plan;
  plan.op = desp.op_type;
  CS_ARG(desp.op_type >= 0 && desp.op_type <= 2, "unknown op type ", desp.op_type);
  CS_ARG(desp.mask_sparse, "only mask-sparse (implicit gemm sparse) convolution is implemented");
  const int io_dim = 2;
  const int weight_ndim = 3;
  CS_RT(p.input.ndim == io_dim, "error");
  CS_RT(p.weight.ndim == weight_ndim, "error");
  CS_RT(p.output.ndim == io_dim, "error");
  // The reference's guard asserts a string literal and never fires
  // (cumm/conv/nvrtc_code.py:56-58); this build rejects the combination.
  if (!(desp.gemm.split_k_serial || desp.gemm.split_k_parallel) && p.split_k_slices > 1)
    CS_ARG(false, "algo don't support splitk but you provide split_k_slices > 1.", p.split_k_slices);
  int64_t N = dim(p.input, 0);
  const int64_t K = dim(p.weight, 0);
  const int64_t C = desp.layout_i == 0 ? dim(p.input, 1) : dim(p.input, io_dim - 1);
  const int64_t K2 = desp.layout_o == 0 ? dim(p.output, 1) : dim(p.output, io_dim - 1);
  CS_RT(K2 == K, "error");
  if (!desp.is_int8_inference) {
    CS_ARG(empty(p.output_add), "only int8 inference support output_add not empty ");
  } else {
    CS_RT(!empty(p.bias) && !empty(p.scale), "int8 inference must have both scale and bias");
  }
  if (!empty(p.output_add)) {
    // nvrtc_conv_template (cumm/conv/nvrtc_code.py:84-89)
    bool same = p.output_add.dtype == p.output.dtype && p.output_add.ndim == p.output.ndim;
    for (int i = 0; same && i < p.output.ndim; ++i) same = p.output_add.shape[i] == p.output.shape[i];
    CS_RT(same, "output and output_add must have same dtype and shape");
  }
  if (desp.op_type == 2) CS_RT(p.mask_width > 0 && p.mask_width % desp.gemm.tile_shape[2] == 0, "error");
  CS_RT(!empty(p.indices), "error");
  CS_RT(!empty(p.mask), "error");
  CS_RT(!empty(p.mask_argsort), "error");
  const int64_t kv = dim(p.weight, desp.layout_w == 0 ? 2 : 1);
  check_shape(p.indices, {kv, -1}, "indices");
  N = dim(p.indices, 1);
  if (desp.op_type == 2) {
    CS_RT(N == dim(p.output, 0), "error");
  } else if (desp.op_type == 0) {
    CS_RT(N == dim(p.output, 0), "error");
  } else {
    CS_RT(N == dim(p.input, 0), "error");
  }
  // implicit_gemm_mnk  (include/tensorview/gemm/core/utils.h:62-92)
  if (desp.op_type == 0) {
    plan.m = N; plan.n = K; plan.k = C * kv;
  } else if (desp.op_type == 1) {
    plan.m = N; plan.n = C; plan.k = K * kv;
  } else {
    plan.m = K; plan.n = C * kv; plan.k = N;
  }
  CS_RT(conv_supported(desp, plan.m, plan.n, plan.k, C, K, p.mask_width), "error");

  // kernel selection (conv_select_helper, cumm/conv/main.py:392-465): exact desp match
  plan.kernel = find_conv_kernel(desp);
  CS_ARG(plan.kernel != nullptr, "Can't Found Algorithm for params: ", conv_repr(desp));

  // Checks the generated kernels leave to the caller; here they guard memory safety.
  // Masks hold one 32-bit word per row, or div_up(kv, 32) words with dynamic
  // masks (MaskIGemmIteratorMaskLoaderDynamic, cumm/gemm/blockmma/mma.py:78-269).
  // As in the reference kernels (cumm/conv/kernel.py:829, 935) the word count
  // is div_up(kv, 32), derived from kv; ConvParams.mask_int_count is carried
  // by the reference but not read by its kernels, so here it must be left at
  // its default of 1 (or 0, a zero-filled struct) or agree with kv.
  if (desp.dynamic_mask) {
    CS_ARG(kv <= 32 * cs::kMaxDynWords, "dynamic masks support kernel volumes up to ",
           32 * cs::kMaxDynWords, ", got ", kv);
    plan.mw = (int)((kv + 31) / 32);
  } else {
    CS_ARG(kv <= 32, "kernel volume > 32 needs dynamic (multi-word) masks");
  }
  CS_ARG(p.mask_int_count <= 1 || p.mask_int_count == plan.mw, "mask_int_count ", p.mask_int_count,
         " does not match the kernel volume ", kv, " (", plan.mw, " mask words per row)");
  const int dt_in = p.input.dtype, dt_w = p.weight.dtype, dt_out = p.output.dtype;
  const int abc[3] = {desp.gemm.dtype_a, desp.gemm.dtype_b, desp.gemm.dtype_c};
  // conv_iwo_012_to_abc (params.h:225-237): i/w/o -> a/b/c
  const int iwo2abc[3][3] = {{0, 1, 2}, {2, 1, 0}, {1, 2, 0}};
  CS_ARG(dt_in == abc[iwo2abc[desp.op_type][0]] && dt_w == abc[iwo2abc[desp.op_type][1]] &&
             dt_out == abc[iwo2abc[desp.op_type][2]],
         "tensor dtypes (input ", dtype_short_str(dt_in), ", weight ", dtype_short_str(dt_w), ", output ",
         dtype_short_str(dt_out), ") do not match the algo ", conv_repr(desp));
  for (const cs_tensor* t : {&p.input, &p.weight, &p.output, &p.indices, &p.mask, &p.mask_argsort})
    check_device(*t, "tensor");
  for (const cs_tensor* t : {&p.input, &p.weight, &p.output}) check_int32_span(*t, "conv operand");
  check_rows(p.input, "input");
  check_rows(p.output, "output");
  check_dense(p.weight, "weight");
  check_dense(p.indices, "indices");
  check_index_dtype(p.indices, "indices");
  check_index_dtype(p.mask, "mask");
  check_index_dtype(p.mask_argsort, "mask_argsort");
  CS_RT(numel(p.mask_argsort) >= N, "mask_argsort must hold ", N, " entries");
  const int tile_m = desp.gemm.tile_shape[0];
  const int64_t mw = plan.mw;
  if (desp.op_type == 2) {
    CS_RT(numel(p.mask) >= (N + p.mask_width - 1) / p.mask_width * mw,
          "mask must hold div_up(N, mask_width) * ", mw, " entries");
  } else {
    CS_RT(numel(p.mask) >= N * mw, "mask must hold ", N * mw, " entries");
  }
  if (desp.op_type == 0 && !empty(p.mask_output))
    CS_RT(numel(p.mask_output) >= (N + tile_m - 1) / tile_m * mw, "mask_output must hold div_up(N, ", tile_m,
          ") * ", mw, " entries");
  const int dt_written = desp.op_type == 0 ? dt_out : desp.op_type == 1 ? dt_in : dt_w;
  if (!empty(p.bias)) {
    check_device(p.bias, "bias");
    // one value per GEMM column: K (forward), C (dgrad), kv * C (wgrad, the
    // [K, kv * C] view of the KRSC gradient; cumm/conv/kernel.py:457-460)
    CS_ARG(numel(p.bias) >= (desp.op_type == 0 ? K : desp.op_type == 1 ? C : kv * C), "bias too short");
    if (desp.is_int8_inference)  // bias is in the compute type (float32)
      CS_ARG(p.bias.dtype == desp.gemm.dcomp, "int8 inference bias must be float32");
    else
      CS_ARG(p.bias.dtype == dt_written, "bias dtype must equal the dtype of the written tensor");
    check_dense(p.bias, "bias");
  }
  if (!empty(p.output_add)) {
    check_device(p.output_add, "output_add");
    check_rows(p.output_add, "output_add");
  }
  // 16-byte operand loads: row strides and bases of the gathered operands.
  const int vec_a = vec_elems(desp.gemm.dtype_a), vec_b = vec_elems(desp.gemm.dtype_b);
  if (desp.op_type == 0) {
    CS_ARG(row_stride(p.input) % vec_a == 0 && aligned16(p.input.data) && aligned16(p.weight.data),
           "input rows and weight must be 16-byte aligned");
  } else if (desp.op_type == 1) {
    CS_ARG(row_stride(p.output) % vec_a == 0 && aligned16(p.output.data), "output rows must be 16-byte aligned");
  } else {
    CS_ARG(row_stride(p.output) % vec_a == 0 && aligned16(p.output.data) && row_stride(p.input) % vec_b == 0 &&
               aligned16(p.input.data),
           "input and output rows must be 16-byte aligned");
  }
  if (!empty(p.scale)) {
    CS_ARG(p.scale.dtype == 0 && numel(p.scale) >= (desp.op_type == 1 ? C : K),
           "scale must be float32 per output channel");
    check_device(p.scale, "scale");
    check_dense(p.scale, "scale");
  }

  plan.N = N; plan.C = C; plan.K = K; plan.kv = kv;
  if (desp.op_type == 1) {
    plan.workspace = align256(kv * C * K * dtype_bytes(dt_w));
  } else if (desp.op_type == 2) {
    CS_RT(K * C * 4 < kInt32Max, "your data exceed int32 range (wgrad output).");
    plan.splits = resolve_splits(p.split_k_slices, K, C);
    plan.workspace = wgrad_layout(kv, K, C, N, plan.splits, true, plan.mw).size;
  }
  return plan;
}

void run_conv(const cs_conv_params& p) {
  ConvPlan plan = plan_conv(p);
  const cs_conv_algo_desp& desp = *plan.kernel;
  hipStream_t s = reinterpret_cast<hipStream_t>(p.stream);
  Workspace ws;
  acquire_workspace(ws, p.workspace, plan.workspace, s);
  if (plan.N == 0 && plan.op != 2) return;  // wgrad still writes dW = act(beta * dW)
  const bool has_bias = !empty(p.bias);
  if (plan.op == 0 || plan.op == 1) {
    cs::RowGatherArgs a;
    std::memset(&a, 0, sizeof(a));
    const cs_tensor& src = plan.op == 0 ? p.input : p.output;   // gathered operand
    const cs_tensor& dst = plan.op == 0 ? p.output : p.input;   // scattered result
    a.a = src.data;
    a.lda = row_stride(src);
    a.a_bytes = (uint32_t)span_bytes(src);
    if (plan.op == 0) {
      a.b = p.weight.data;        // W[K][kv][C]: B_f[n=k][c]
      a.b_fs = plan.C;
      a.b_ns = plan.kv * plan.C;
      a.b_bytes = (uint32_t)span_bytes(p.weight);
    } else {
      // B_f[n=c][k] = W[k][f][c]: read k-major from the weights themselves; a
      // kernel that needs it n-major transposes into the workspace
      a.b = p.weight.data;
      a.b_kmajor = 1;
      a.b_ks = plan.kv * plan.C;
      a.b_fs = plan.C;
      a.b_ns = 1;
      a.b_scratch = ws.ptr;
      a.b_bytes = (uint32_t)span_bytes(p.weight);
    }
    a.c = dst.data;
    a.ldc = row_stride(dst);
    a.d = has_bias ? p.bias.data : dst.data;
    a.ldd = has_bias ? 0 : row_stride(dst);
    a.d_is_bias = has_bias ? 1 : 0;
    a.pairs = reinterpret_cast<const int32_t*>(p.indices.data);
    a.pair_ld = p.indices.stride[0];
    a.pair_col = reinterpret_cast<const int32_t*>(p.mask_argsort.data);
    a.out_row = a.pair_col;
    a.mask = reinterpret_cast<const uint32_t*>(p.mask.data);
    a.mask_out = plan.op == 0 && !empty(p.mask_output) ? reinterpret_cast<uint32_t*>(p.mask_output.data) : nullptr;
    a.mw = plan.mw;
    a.scale = empty(p.scale) ? nullptr : reinterpret_cast<const float*>(p.scale.data);
    a.m = (int32_t)plan.N;
    a.kred = (int32_t)(plan.op == 0 ? plan.C : plan.K);
    a.nout = (int32_t)(plan.op == 0 ? plan.K : plan.C);
    a.kv = (int32_t)plan.kv;
    a.mask_filter = p.mask_filter;
    a.reverse_mask = plan.op == 1 && p.reverse_mask ? 1 : 0;
    a.skip_empty = plan.op == 1 ? 1 : 0;
    a.alpha = p.alpha;
    a.beta = p.beta;
    a.act_alpha = p.act_alpha;
    a.act_beta = p.act_beta;
    a.act_type = p.act_type;
    if (desp.is_int8_inference) {
      // source = output_add, else the output itself (cumm/conv/main.py:604-610)
      const cs_tensor& srct = empty(p.output_add) ? p.output : p.output_add;
      a.bias_f = reinterpret_cast<const float*>(p.bias.data);
      a.d = srct.data;
      a.ldd = row_stride(srct);
      a.d_is_bias = 0;
    }
    CS_HIP(cs::launch_rowgather(src.dtype, dst.dtype, desp.gemm.tile_shape[1], a, s));
  } else {
    cs::WgradArgs w;
    std::memset(&w, 0, sizeof(w));
    w.a = p.output.data;
    w.lda = row_stride(p.output);
    w.b = p.input.data;
    w.ldb = row_stride(p.input);
    w.a_bytes = (uint32_t)span_bytes(p.output);
    w.b_bytes = (uint32_t)span_bytes(p.input);
    w.a_rows = reinterpret_cast<const int32_t*>(p.mask_argsort.data);
    w.pairs = reinterpret_cast<const int32_t*>(p.indices.data);
    w.pair_ld = p.indices.stride[0];
    w.mask = reinterpret_cast<const uint32_t*>(p.mask.data);
    w.mask_width = p.mask_width;
    w.num_mask = (int32_t)((plan.N + p.mask_width - 1) / p.mask_width);
    w.nrows = (int32_t)plan.N;
    w.M = (int32_t)plan.K;
    w.N = (int32_t)plan.C;
    w.kv = (int32_t)plan.kv;
    // dynamic masks filter offset f by bit f % 32 (mma.py:611-614)
    w.mask_filter = plan.mw > 1 ? p.mask_filter : p.mask_filter & (plan.kv >= 32 ? 0xffffffffu : ((1u << plan.kv) - 1u));
    w.mw = plan.mw;
    w.max_blocks = plan.splits;
    w.min_rows = wgrad_min_rows(plan.K, plan.C);
    w.fixed_total = -1;
    const WgradLayout lay = wgrad_layout(plan.kv, plan.K, plan.C, plan.N, plan.splits, true, plan.mw);
    char* base = static_cast<char*>(ws.ptr);
    w.partial = reinterpret_cast<float*>(base + lay.partial);
    w.rowmask = reinterpret_cast<uint32_t*>(base + lay.rowmask);
    w.bcount = reinterpret_cast<int32_t*>(base + lay.bcount);
    w.list_a = reinterpret_cast<int32_t*>(base + lay.list_a);
    w.list_b = reinterpret_cast<int32_t*>(base + lay.list_b);
    w.list_ld = plan.N;
    w.total = reinterpret_cast<int32_t*>(base + lay.total);
    CS_HIP(cs::launch_wgrad(p.output.dtype, w, s));
    cs::ReduceArgs r;
    std::memset(&r, 0, sizeof(r));
    r.partial = w.partial;
    r.total = w.total;
    r.fixed_total = -1;
    r.max_blocks = w.max_blocks;
    r.min_rows = w.min_rows;
    r.kv = (int32_t)plan.kv;
    r.M = (int32_t)plan.K;
    r.N = (int32_t)plan.C;
    r.out = p.weight.data;
    r.o_ms = p.weight.stride[0];
    r.o_fs = p.weight.stride[1];
    r.src = has_bias ? p.bias.data : p.weight.data;
    r.s_ms = r.o_ms;
    r.s_fs = r.o_fs;
    r.src_is_bias = has_bias ? 1 : 0;
    r.mask_filter = w.mask_filter;
    r.alpha = p.alpha;
    r.beta = p.beta;
    r.act_alpha = p.act_alpha;
    r.act_type = p.act_type;
    CS_HIP(cs::launch_reduce(p.weight.dtype, r, s));
  }
}

// ------------------------------------------------------------------ gemm
struct GemmPlan {
  const cs_gemm_algo_desp* kernel = nullptr;
  int shuffle = 0;
  bool ta = false, tb = false;
  cs_tensor a, b, c, d;
  int64_t m = 0, n = 0, k = 0;
  int splits = 1;
  bool b_kmajor = false;
  int64_t ws_a = 0, ws_b = 0, workspace = 0;
};

GemmPlan plan_gemm(const cs_gemm_params& p) {
  // GemmParams::check_valid  (params.h:351-364) + nvrtc_gemm_template (nvrtc_code.py:34-145)
  const cs_gemm_algo_desp& desp = p.algo_desp;
  check_valid(desp);
  if (desp.dtype_a == 3) {
    // int32 epilogue arithmetic: alpha, beta and act_alpha become int32 by
    // truncation, and sigmoid has no integer form
    CS_ARG(p.act_type != 2, "sigmoid is not available for int8 GEMMs (int32 compute)");
    for (float v : {p.alpha, p.beta, p.act_alpha})
      CS_ARG(std::isfinite(v) && std::fabs(v) < 2147483648.0f, "alpha, beta and act_alpha of an int8 GEMM must fit int32");
  }
  CS_RT(!empty(p.a) && !empty(p.b) && !empty(p.c), "a,b,c must not empty");
  if (desp.shuffle_type == 1) CS_RT(!empty(p.c_inds), "a_inds,c_inds tensor must not empty");
  if (desp.shuffle_type == 2) CS_RT(!empty(p.a_inds) && !empty(p.b_inds), "a_inds,b_inds tensor must not empty");
  GemmPlan g;
  g.shuffle = desp.shuffle_type;
  g.a = p.a;
  g.b = p.b;
  g.c = p.c;
  g.d = empty(p.d) ? p.c : p.d;
  check_shape(p.a, {-1, -1}, "a");
  check_shape(p.b, {-1, -1}, "b");
  check_shape(p.c, {-1, -1}, "c");
  CS_RT(p.a.device == p.b.device && p.b.device == p.c.device, "a, b, c must be on the same device");
  bool trans_a = desp.trans_a == 1, trans_b = desp.trans_b == 1;
  const bool trans_c = desp.trans_c == 1;
  if (!(desp.split_k_serial || desp.split_k_parallel) && p.split_k_slices > 1)
    CS_ARG(false, "algo don't support splitk but you provide split_k_slices > 1.", p.split_k_slices);
  int64_t k2 = 0;
  if (desp.shuffle_type == 1) {
    CS_ARG(!trans_c, "trans_c is not supported with ShuffleAC");
    CS_RT(!trans_a, "a of shuffle AB must be row major");
    g.m = !empty(p.a_inds) ? dim(p.a_inds, 0) : dim(p.a, 0);
    g.k = dim(p.a, 1);
    k2 = dim(p.b, trans_b ? 1 : 0);
    g.n = dim(p.b, trans_b ? 0 : 1);
    check_shape(p.c, {-1, g.n}, "c");
  } else if (desp.shuffle_type == 2) {
    CS_ARG(!trans_c, "trans_c is not supported with ShuffleAB");
    CS_RT(trans_a && !trans_b, "shuffle AB must be nt, i.e. backward weight");
    g.m = dim(p.a, 1);
    g.k = dim(p.a_inds, 0);
    k2 = dim(p.b_inds, 0);
    g.n = dim(p.b, 1);
    check_shape(p.c, {g.m, g.n}, "c");
  } else {
    // trans_c computes the transposed problem with A and B swapped and stores
    // it column-major; the stored result is the same row-major C = op(A) op(B)
    // (nvrtc_code.py:62-67 with the {n, m} check of the swapped problem).
    g.m = dim(p.a, trans_a ? 1 : 0);
    g.k = dim(p.a, trans_a ? 0 : 1);
    k2 = dim(p.b, trans_b ? 1 : 0);
    g.n = dim(p.b, trans_b ? 0 : 1);
    check_shape(p.c, {g.m, g.n}, "c");
  }
  g.ta = trans_a;
  g.tb = trans_b;
  CS_ARG(gemm_supported(desp, g.m, g.n, g.k), "this m, n, k isn't supported due to misaligned contiguous dim.");
  CS_ARG(g.k == k2, "error");
  if (g.d.ndim == 1) CS_RT(dim(g.d, 0) == g.n, "d must be a valid bias");
  CS_ARG(p.c.dtype == g.d.dtype, "d dtype must equal to c");

  static const std::vector<std::string> names = [] {
    std::vector<std::string> v;
    for (const auto& d : gemm_catalogue()) v.push_back(gemm_repr(d));
    return v;
  }();
  const auto& catalogue = gemm_catalogue();
  const std::string want = gemm_repr(desp);
  for (size_t i = 0; i < catalogue.size(); ++i)
    if (names[i] == want) { g.kernel = &catalogue[i]; break; }
  CS_ARG(g.kernel != nullptr, "Can't Found Algorithm for params: ", want);
  CS_ARG(p.a.dtype == desp.dtype_a && p.b.dtype == desp.dtype_b && p.c.dtype == desp.dtype_c,
         "tensor dtypes do not match the algo ", want);
  for (const cs_tensor* t : {&p.a, &p.b, &p.c}) {
    check_device(*t, "tensor");
    check_rows(*t, "gemm operand");
    check_int32_span(*t, "gemm operand");
  }
  if (g.d.ndim == 2) check_rows(g.d, "d");
  if (!empty(p.a_inds)) check_index_dtype(p.a_inds, "a_inds");
  if (!empty(p.b_inds)) check_index_dtype(p.b_inds, "b_inds");
  if (!empty(p.c_inds)) {
    check_index_dtype(p.c_inds, "c_inds");
    CS_RT(numel(p.c_inds) >= g.m, "c_inds must hold m entries");
  }
  const int es = dtype_bytes(p.a.dtype);
  const int vec = vec_elems(p.a.dtype);
  CS_ARG(aligned16(p.a.data) && aligned16(p.b.data), "a and b must be 16-byte aligned");
  if (g.shuffle == 2) {
    CS_ARG(g.m % vec == 0 && g.n % vec == 0 && row_stride(p.a) % vec == 0 && row_stride(p.b) % vec == 0,
           "m, n and row strides must be multiples of ", vec);
    const int64_t tiles = ((g.m + cs::kWgradTile - 1) / cs::kWgradTile) * ((g.n + cs::kWgradTile - 1) / cs::kWgradTile);
    (void)tiles;
    CS_RT(g.m * g.n * 4 < kInt32Max, "your data exceed int32 range (c).");
    g.splits = resolve_splits(p.split_k_slices, g.m, g.n);
    g.workspace = wgrad_layout(1, g.m, g.n, g.k, g.splits, false).size;
  } else {
    CS_ARG(g.k % vec == 0, "k must be a multiple of ", vec);
    if (!trans_a) CS_ARG(row_stride(p.a) % vec == 0, "lda must be a multiple of ", vec);
    if (trans_b) CS_ARG(row_stride(p.b) % vec == 0, "ldb must be a multiple of ", vec);
    g.ws_a = trans_a ? align256(g.m * g.k * es) : 0;
    // a row-major [k, n] b is read k-major in place when its rows are whole
    // 16-byte units; otherwise it is transposed into the workspace
    g.b_kmajor = !trans_b && g.n % vec == 0 && row_stride(p.b) % vec == 0;
    g.ws_b = trans_b || g.b_kmajor ? 0 : align256(g.k * g.n * es);
    g.workspace = g.ws_a + g.ws_b;
  }
  return g;
}

void run_gemm(const cs_gemm_params& p) {
  GemmPlan g = plan_gemm(p);
  hipStream_t s = reinterpret_cast<hipStream_t>(p.stream);
  Workspace ws;
  acquire_workspace(ws, p.workspace, g.workspace, s);
  if (g.m == 0 || g.n == 0) return;
  const bool bias = g.d.ndim == 1;
  if (g.shuffle == 2) {
    cs::WgradArgs w;
    std::memset(&w, 0, sizeof(w));
    w.a = p.a.data;
    w.lda = row_stride(p.a);
    w.b = p.b.data;
    w.ldb = row_stride(p.b);
    w.a_bytes = (uint32_t)span_bytes(p.a);
    w.b_bytes = (uint32_t)span_bytes(p.b);
    // list form: the index tensors are the pair list (every row counts)
    w.list_a = const_cast<int32_t*>(reinterpret_cast<const int32_t*>(p.a_inds.data));
    w.list_b = const_cast<int32_t*>(reinterpret_cast<const int32_t*>(p.b_inds.data));
    w.list_ld = 0;
    w.fixed_total = (int32_t)g.k;
    w.nrows = (int32_t)g.k;
    w.M = (int32_t)g.m;
    w.N = (int32_t)g.n;
    w.kv = 1;
    w.mask_filter = 1u;
    w.mw = 1;
    w.max_blocks = g.splits;
    w.min_rows = kWgradMinRows;
    const WgradLayout lay = wgrad_layout(1, g.m, g.n, g.k, g.splits, false);
    char* base = static_cast<char*>(ws.ptr);
    w.partial = reinterpret_cast<float*>(base + lay.partial);
    CS_HIP(cs::launch_wgrad(p.a.dtype, w, s));
    cs::ReduceArgs r;
    std::memset(&r, 0, sizeof(r));
    r.partial = w.partial;
    r.total = nullptr;
    r.fixed_total = w.fixed_total;
    r.max_blocks = w.max_blocks;
    r.min_rows = w.min_rows;
    r.kv = 1;
    r.M = (int32_t)g.m;
    r.N = (int32_t)g.n;
    r.out = p.c.data;
    r.o_ms = row_stride(p.c);
    r.o_fs = 0;
    r.src = g.d.data;
    r.s_ms = bias ? 0 : row_stride(g.d);
    r.s_fs = 0;
    r.src_is_bias = bias ? 1 : 0;
    r.mask_filter = 1u;
    r.alpha = p.alpha;
    r.beta = p.beta;
    r.act_alpha = p.act_alpha;
    r.act_type = p.act_type;
    r.int_compute = p.a.dtype == 3 ? 1 : 0;
    r.ialpha = (int32_t)p.alpha;
    r.ibeta = (int32_t)p.beta;
    r.iact_alpha = (int32_t)p.act_alpha;
    CS_HIP(cs::launch_reduce(p.c.dtype, r, s));
    return;
  }
  const int es = dtype_bytes(p.a.dtype);
  const void* aptr = p.a.data;
  int64_t lda = row_stride(p.a);
  if (g.ta) {  // a is [k][m] -> [m][k]
    CS_HIP(cs::launch_transpose(p.a.data, ws.ptr, g.k, 1, g.m, row_stride(p.a), es, s));
    aptr = ws.ptr;
    lda = g.k;
  }
  const void* bptr = p.b.data;
  int64_t ldb = row_stride(p.b);
  if (!g.tb && !g.b_kmajor) {  // b is [k][n] -> [n][k]
    void* bt = static_cast<char*>(ws.ptr) + g.ws_a;
    CS_HIP(cs::launch_transpose(p.b.data, bt, g.k, 1, g.n, row_stride(p.b), es, s));
    bptr = bt;
    ldb = g.k;
  }
  cs::GemmRowsArgs a;
  std::memset(&a, 0, sizeof(a));
  a.a = aptr;
  a.lda = lda;
  a.a_bytes = (uint32_t)(g.ta ? g.m * g.k * es : span_bytes(p.a));
  a.b = bptr;
  a.ldb = ldb;
  a.b_kmajor = g.b_kmajor ? 1 : 0;
  a.b_bytes = (uint32_t)(g.tb || g.b_kmajor ? span_bytes(p.b) : g.k * g.n * es);
  a.c = p.c.data;
  a.ldc = row_stride(p.c);
  a.d = g.d.data;
  a.ldd = bias ? 0 : row_stride(g.d);
  a.d_is_bias = bias ? 1 : 0;
  if (g.shuffle == 1) {
    a.a_rows = empty(p.a_inds) ? nullptr : reinterpret_cast<const int32_t*>(p.a_inds.data);
    a.c_rows = reinterpret_cast<const int32_t*>(p.c_inds.data);
  }
  a.m = (int32_t)g.m;
  a.nn = (int32_t)g.n;
  a.k = (int32_t)g.k;
  a.alpha = p.alpha;
  a.beta = p.beta;
  a.act_alpha = p.act_alpha;
  a.act_type = p.act_type;
  a.ialpha = (int32_t)p.alpha;
  a.ibeta = (int32_t)p.beta;
  a.iact_alpha = (int32_t)p.act_alpha;
  CS_HIP(cs::launch_gemm_rows(p.a.dtype, p.c.dtype, a, s));
}

// ------------------------------------------------------------------ gather / scatter
// Published parameter sets of GatherAll / ScatterAll (cumm/gemm/gather.py:392-420,
// 500-530), in the reference's enumeration order: tile k bytes, tile m, bytes
// per access, threads, keeping sets whose threads * access fit the tile.
using GsParam = std::array<int32_t, 4>;  // tile_m, tile_k_bytes, bytes_per_access, num_threads

std::vector<GsParam> gs_params(bool scatter) {
  std::vector<GsParam> out;
  const std::vector<int> tk_gather = {4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048};
  const std::vector<int> tk_scatter = {64, 128, 256, 512, 1024, 2048};
  const std::vector<int> th_gather = {64, 128, 256};
  const std::vector<int> th_scatter = {128, 256};
  for (int tk : scatter ? tk_scatter : tk_gather)
    for (int tm : {1, 2, 4, 8})
      for (int bpa : {4, 8, 16})
        for (int nt : scatter ? th_scatter : th_gather) {
          if (nt * bpa > tk * tm) continue;
          out.push_back({tm, tk, bpa, nt});
        }
  return out;
}

const std::vector<GsParam>& gather_params() {
  static const std::vector<GsParam> v = gs_params(false);
  return v;
}

const std::vector<GsParam>& scatter_params() {
  static const std::vector<GsParam> v = gs_params(true);
  return v;
}

bool has_param(const std::vector<GsParam>& v, int64_t tm, int64_t tk, int64_t bpa, int64_t nt) {
  for (const auto& q : v)
    if (q[0] == tm && q[1] == tk && q[2] == bpa && q[3] == nt) return true;
  return false;
}

// Gather::supported / Scatter::supported (cumm/gemm/gather.py:210-219, 345-353)
bool gs_supported(int64_t bpa, int64_t channel_size, int dtype) {
  const int isz = dtype_bytes(dtype);
  return isz > 0 && bpa > 0 && (channel_size * isz) % bpa == 0;
}

// Shared preamble of Gather::gather / Scatter::scatter (cumm/gemm/gather.py:236-243, 368-375)
void gs_check(const cs_tensor& out, const cs_tensor& in, const cs_tensor& idx, int64_t bpa) {
  CS_RT(in.ndim == 2 && out.ndim == 2 && in.shape[1] == out.shape[1], "error");
  CS_RT(in.dtype == out.dtype, "error");
  CS_RT(gs_supported(bpa, in.shape[1], in.dtype), "not supported. use supported() to check first");
  CS_ARG(idx.ndim == 1, "indices must be 1-D");
  check_index_dtype(idx, "indices");
  check_device(in, "input");
  check_device(out, "output");
  check_device(idx, "indices");
  check_rows(in, "input");
  check_rows(out, "output");
  CS_ARG(idx.shape[0] == 0 || idx.stride[0] == 1, "indices must be contiguous");
}

void run_gather(const cs_tensor& out, const cs_tensor& in, const cs_tensor& idx, int64_t tm, int64_t tk,
                int64_t bpa, int64_t nt, uint64_t stream) {
  CS_RT(has_param(gather_params(), tm, tk, bpa, nt), "can't find a sutable gather");
  gs_check(out, in, idx, bpa);
  const int64_t m = idx.shape[0];
  CS_ARG(out.shape[0] >= m, "output has ", out.shape[0], " rows, indices ", m);
  const int64_t isz = dtype_bytes(in.dtype);
  CS_HIP(cs::launch_gather_rows(in.data, out.data, reinterpret_cast<const int32_t*>(idx.data), m, in.shape[0],
                                in.shape[1] * isz, row_stride(in) * isz, row_stride(out) * isz,
                                reinterpret_cast<hipStream_t>(stream)));
}

void run_scatter(const cs_tensor& out, const cs_tensor& in, const cs_tensor& idx, int64_t tm, int64_t tk,
                 int64_t bpa, int64_t nt, uint64_t stream) {
  // the scatter catalogue holds float32 kernels only (cumm/gemm/gather.py:515)
  CS_RT(in.dtype == 0 && has_param(scatter_params(), tm, tk, bpa, nt), "can't find a sutable scatter");
  gs_check(out, in, idx, bpa);
  const int64_t m = idx.shape[0];
  CS_ARG(in.shape[0] >= m, "input has ", in.shape[0], " rows, indices ", m);
  CS_HIP(cs::launch_scatter_add_f32(static_cast<const float*>(in.data), static_cast<float*>(out.data),
                                    reinterpret_cast<const int32_t*>(idx.data), m, out.shape[0], in.shape[1],
                                    row_stride(in), row_stride(out), reinterpret_cast<hipStream_t>(stream)));
}

int copy_params(const std::vector<GsParam>& v, int32_t* out, int cap) {
  if (out != nullptr)
    for (int i = 0; i < cap && i < (int)v.size(); ++i)
      for (int j = 0; j < 4; ++j) out[4 * i + j] = v[i][j];
  return (int)v.size();
}

}  // namespace

// ====================================================================== allocation-free conv
// Used by cs_implicit_gemm2_noalloc (cs_noalloc.hip), which owns the memory.
namespace cs_detail {
// ConvParams.verbose (cumm/conv/main.py:704-729): time the call's launches on
// its stream (synchronising, as the reference's timer does) and print the
// kernel name, the main kernel's register count and the time in ms.
template <typename F>
void run_verbose(const cs_conv_params& p, F&& body) {
  hipStream_t s = reinterpret_cast<hipStream_t>(p.stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  struct Free {
    hipEvent_t a, b;
    ~Free() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } fr{e0, e1};
  cs::last_kernel() = nullptr;
  CS_HIP(hipEventRecord(e0, s));
  body();
  CS_HIP(hipEventRecord(e1, s));
  CS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CS_HIP(hipEventElapsedTime(&ms, e0, e1));
  int regs = -1;
  hipFuncAttributes attr;
  if (cs::last_kernel() != nullptr && hipFuncGetAttributes(&attr, cs::last_kernel()) == hipSuccess) regs = attr.numRegs;
  std::printf("%s kernel num regs: %d time: %f\n", conv_repr(p.conv_algo_desp).c_str(), regs, (double)ms);
  std::fflush(stdout);
}


__attribute__((visibility("hidden"))) int fail_runtime(const char* msg) {
  g_last_error = msg;
  return CS_ERR_RUNTIME;
}

__attribute__((visibility("hidden"))) int implicit_gemm2_in(const cs_conv_params* params, void* mem, int64_t cap,
                                                            int device) {
  return guarded([&] {
    CS_ARG(params != nullptr, "params is null");
    cs_conv_params q = *params;
    ConvPlan plan = plan_conv(q);
    if (plan.op == 2 && plan.workspace > cap) {
      // fewer split pieces per offset: the partition only changes for an
      // offset with more than pieces * wgrad_min_rows(K, C) pairs
      const int64_t other = wgrad_layout(plan.kv, plan.K, plan.C, plan.N, 0, true, plan.mw).size;
      const int64_t per = plan.kv * plan.K * plan.C * 4;
      const int64_t fit = (cap - other - 256) / std::max<int64_t>(per, 1);
      CS_ARG(fit >= 2, "the call needs a workspace of ", plan.workspace, " bytes (at least ",
             other + 2 * per + 256, " with split pieces capped), beyond the ", cap,
             " bytes of the allocation-free entry point's scratch");
      q.split_k_slices = (int)std::min<int64_t>(plan.splits, fit);
      plan = plan_conv(q);
    }
    CS_ARG(plan.workspace <= cap, "the call needs a workspace of ", plan.workspace, " bytes, beyond the ", cap,
           " bytes of the allocation-free entry point's scratch");
    std::memset(&q.workspace, 0, sizeof(q.workspace));
    q.workspace.data = mem;
    q.workspace.dtype = 6;  // u8
    q.workspace.ndim = 1;
    q.workspace.shape[0] = cap;
    q.workspace.stride[0] = 1;
    q.workspace.device = device;
    if (q.verbose) run_verbose(q, [&] { run_conv(q); });
    else run_conv(q);
  });
}

}  // namespace cs_detail

// ====================================================================== C ABI
extern "C" {

int cs_abi_version(void) { return CS_ABI_VERSION; }

int cs_implicit_gemm2(const cs_conv_params* params) {
  return guarded([&] {
    CS_ARG(params != nullptr, "params is null");
    if (params->verbose) cs_detail::run_verbose(*params, [&] { run_conv(*params); });
    else run_conv(*params);
  });
}

int cs_matmul2(const cs_gemm_params* params) {
  return guarded([&] {
    CS_ARG(params != nullptr, "params is null");
    run_gemm(*params);
  });
}

int cs_conv_workspace_size(const cs_conv_params* params, int64_t* out) {
  return guarded([&] {
    CS_ARG(params != nullptr && out != nullptr, "null argument");
    *out = plan_conv(*params).workspace;
  });
}

int cs_gemm_workspace_size(const cs_gemm_params* params, int64_t* out) {
  return guarded([&] {
    CS_ARG(params != nullptr && out != nullptr, "null argument");
    *out = plan_gemm(*params).workspace;
  });
}

int cs_get_all_conv_algo_desp(cs_conv_algo_desp* out, int cap) {
  const auto& catalogue = conv_catalogue();
  if (out != nullptr)
    for (int i = 0; i < cap && i < (int)catalogue.size(); ++i) out[i] = catalogue[i];
  return (int)catalogue.size();
}

int cs_get_all_algo_desp(cs_gemm_algo_desp* out, int cap) {
  const auto& catalogue = gemm_catalogue();
  if (out != nullptr)
    for (int i = 0; i < cap && i < (int)catalogue.size(); ++i) out[i] = catalogue[i];
  return (int)catalogue.size();
}

static int copy_name(const std::string& s, char* buf, int cap) {
  if (buf != nullptr && cap > 0) {
    const int n = std::min<int>(cap - 1, (int)s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return (int)s.size();
}

int cs_gemm_algo_name(const cs_gemm_algo_desp* desp, char* buf, int cap) {
  int n = -1;
  int rc = guarded([&] {
    CS_ARG(desp != nullptr, "null desp");
    check_valid(*desp);
    n = copy_name(gemm_repr(*desp), buf, cap);
  });
  return rc == CS_OK ? n : -rc;
}

int cs_conv_algo_name(const cs_conv_algo_desp* desp, char* buf, int cap) {
  int n = -1;
  int rc = guarded([&] {
    CS_ARG(desp != nullptr, "null desp");
    check_valid(desp->gemm);
    n = copy_name(conv_repr(*desp), buf, cap);
  });
  return rc == CS_OK ? n : -rc;
}

int cs_device_synchronize(void) {
  return guarded([] { CS_HIP(hipDeviceSynchronize()); });
}

int cs_stream_synchronize(uint64_t stream) {
  return guarded([&] { CS_HIP(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream))); });
}

const char* cs_last_error(void) { return g_last_error.c_str(); }

int cs_get_all_gather_params(int32_t* out, int cap) { return copy_params(gather_params(), out, cap); }

int cs_get_all_scatter_params(int32_t* out, int cap) { return copy_params(scatter_params(), out, cap); }

int cs_gather_supported(int32_t bytes_per_access, int32_t channel_size, int32_t dtype) {
  return gs_supported(bytes_per_access, channel_size, dtype) ? 1 : 0;
}

int cs_scatter_supported(int64_t tile_m, int64_t tile_k_bytes, int64_t bytes_per_access, int64_t num_threads,
                         int32_t channel_size, int32_t dtype) {
  if (dtype != 0 || !has_param(scatter_params(), tile_m, tile_k_bytes, bytes_per_access, num_threads)) return 0;
  return gs_supported(bytes_per_access, channel_size, dtype) ? 1 : 0;
}

int cs_gather(const cs_tensor* output, const cs_tensor* input, const cs_tensor* indices, int64_t tile_m,
              int64_t tile_k_bytes, int64_t bytes_per_access, int64_t num_threads, uint64_t stream) {
  return guarded([&] {
    CS_ARG(output != nullptr && input != nullptr && indices != nullptr, "null argument");
    run_gather(*output, *input, *indices, tile_m, tile_k_bytes, bytes_per_access, num_threads, stream);
  });
}

int cs_scatter(const cs_tensor* output, const cs_tensor* input, const cs_tensor* indices, int64_t tile_m,
               int64_t tile_k_bytes, int64_t bytes_per_access, int64_t num_threads, uint64_t stream) {
  return guarded([&] {
    CS_ARG(output != nullptr && input != nullptr && indices != nullptr, "null argument");
    run_scatter(*output, *input, *indices, tile_m, tile_k_bytes, bytes_per_access, num_threads, stream);
  });
}

}  // extern "C"
