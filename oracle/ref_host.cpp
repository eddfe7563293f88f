// C-ABI shim over the reference's header-only host code (tv::gemm desp and
// shape helpers), compiled straight from the reference tree:
//   include/tensorview/gemm/core/params.h  (GemmAlgoDesp / ConvAlgoDesp)
//   include/tensorview/gemm/core/utils.h   (implicit_gemm_mnk, tile counts)
// TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): tests compare the build's
// kernel names, `supported`, workspace formulas and mnk against these.
#include <tensorview/gemm/core/params.h>
#include <tensorview/gemm/core/utils.h>

#include <cstring>
#include <string>

#include "cumm_amd.h"

namespace {

void fill_gemm(tv::gemm::GemmAlgoDesp& d, const cs_gemm_algo_desp& c) {
  d.dtype_a = c.dtype_a;
  d.dtype_b = c.dtype_b;
  d.dtype_c = c.dtype_c;
  d.trans_a_ = c.trans_a;
  d.trans_b_ = c.trans_b;
  d.trans_c_ = c.trans_c;
  for (int i = 0; i < 3; ++i) {
    d.tile_shape[i] = c.tile_shape[i];
    d.warp_tile_shape[i] = c.warp_tile_shape[i];
    d.tensorop[i] = c.tensorop[i];
  }
  d.num_stage = c.num_stage;
  d.dacc = c.dacc;
  d.dcomp = c.dcomp;
  d.algo = std::string(c.algo);
  d.split_k_serial_ = c.split_k_serial;
  d.split_k_parallel_ = c.split_k_parallel;
  d.shuffle_type = static_cast<tv::gemm::ShuffleStrideType>(c.shuffle_type);
  d.element_per_access_a = c.element_per_access_a;
  d.element_per_access_b = c.element_per_access_b;
  d.element_per_access_c = c.element_per_access_c;
  d.access_per_vector = c.access_per_vector;
}

tv::gemm::ConvAlgoDesp make_conv(const cs_conv_algo_desp& c) {
  tv::gemm::ConvAlgoDesp d(c.ndim, static_cast<tv::gemm::ConvOpType>(c.op_type));
  fill_gemm(d, c.gemm);
  d.iter_algo = static_cast<tv::gemm::ConvIterAlgo>(c.iter_algo);
  d.layout_i = static_cast<tv::gemm::ConvLayoutType>(c.layout_i);
  d.layout_w = static_cast<tv::gemm::ConvLayoutType>(c.layout_w);
  d.layout_o = static_cast<tv::gemm::ConvLayoutType>(c.layout_o);
  d.interleave_i = c.interleave_i;
  d.interleave_w = c.interleave_w;
  d.interleave_o = c.interleave_o;
  d.mask_sparse = c.mask_sparse;
  d.increment_k_first = c.increment_k_first;
  d.dynamic_mask = c.dynamic_mask;
  d.is_int8_inference = c.is_int8_inference;
  return d;
}

int put(const std::string& s, char* buf, int cap) {
  const int n = (int)std::min<size_t>(s.size(), (size_t)cap - 1);
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
  return (int)s.size();
}

}  // namespace

extern "C" {

int ref_gemm_repr(const cs_gemm_algo_desp* c, char* buf, int cap) {
  tv::gemm::GemmAlgoDesp d;
  fill_gemm(d, *c);
  return put(d.__repr__(), buf, cap);
}

int ref_conv_repr(const cs_conv_algo_desp* c, char* buf, int cap) {
  auto d = make_conv(*c);
  return put(d.__repr__(), buf, cap);
}

int ref_gemm_supported(const cs_gemm_algo_desp* c, int m, int n, int k) {
  tv::gemm::GemmAlgoDesp d;
  fill_gemm(d, *c);
  return d.supported(m, n, k) ? 1 : 0;
}

int ref_conv_supported(const cs_conv_algo_desp* c, int m, int n, int k, int C, int K, int mask_width) {
  auto d = make_conv(*c);
  return d.supported(m, n, k, C, K, mask_width) ? 1 : 0;
}

int ref_gemm_workspace(const cs_gemm_algo_desp* c, int m, int n, int k, int split) {
  tv::gemm::GemmAlgoDesp d;
  fill_gemm(d, *c);
  return d.query_workspace_size(m, n, k, split);
}

int ref_conv_workspace(const cs_conv_algo_desp* c, int m, int n, int k, int split, int kv) {
  auto d = make_conv(*c);
  return d.query_conv_workspace_size(m, n, k, split, kv);
}

void ref_implicit_gemm_mnk(int op, int N, int C, int K, int kv, int in_prod, int out_prod, int mask_sparse,
                           int* out) {
  auto r = tv::gemm::implicit_gemm_mnk(static_cast<tv::gemm::ConvOpType>(op), N, C, K, kv, in_prod, out_prod,
                                       mask_sparse != 0);
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
}

void ref_spconv_tile_count(int m, int n, int k, int tm, int tn, int split, int kv, int op, int* out) {
  auto r = tv::gemm::get_spconv_logical_tile_count(m, n, k, tm, tn, split, kv,
                                                   static_cast<tv::gemm::ConvOpType>(op));
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
}

}  // extern "C"
