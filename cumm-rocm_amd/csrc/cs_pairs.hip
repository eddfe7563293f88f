// Device-side pair tables, per-row masks and mask argsort for 3x3x3 sparse
// conv layers: what the caller of implicit_gemm2 prepares before each layer.
//
// In the reference that caller is spconv (get_indice_pairs /
// generate_subm_conv_inds and sort_1d_by_key_split, called from
// cumm/conv/main_real_spconv.py:91-145), built on tensorview's linear-probing
// GPU hash table (include/tensorview/hash/linear.cu.h:111-430,
// hash.cu.h:43-95).  The formats are the ones implicit_gemm2 consumes:
//   pair[kv][N_out] int32 (-1 = no neighbour), offsets f = kz*9 + ky*3 + kx,
//   neighbour of a submanifold row = coord + (k - 1), of a strided output
//   o*2 - 1 + k;  mask[r] bit f <=> pair[f][r] != -1, sorted stably by value
//   with its argsort.  Coordinates are (b, z, y, x) int32 rows; keys are
//   ((b*S0 + z)*S1 + y)*S2 + x as uint64.
//
// Kernels: hash insert (open addressing, linear probing, 64-bit CAS), one
// thread per (offset, row) lookup writing pair_fwd / pair_bwd coalesced along
// rows, a per-row mask kernel, and rocPRIM radix sorts (stable LSD) for the
// mask argsort and for the unique output coordinates of strided layers.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <algorithm>

#include "cumm_amd.h"

namespace cs {
namespace {

constexpr uint64_t kEmptyKey = ~0ull;
constexpr int kKv = 27;

struct Shape3 {
  int64_t s0, s1, s2;
};

__device__ __forceinline__ uint64_t mix64(uint64_t k) {
  // splitmix64 finaliser
  k ^= k >> 30;
  k *= 0xbf58476d1ce4e5b9ull;
  k ^= k >> 27;
  k *= 0x94d049bb133111ebull;
  k ^= k >> 31;
  return k;
}

__device__ __forceinline__ uint64_t key_of(int64_t b, int64_t z, int64_t y, int64_t x, Shape3 s) {
  return (uint64_t)(((b * s.s0 + z) * s.s1 + y) * s.s2 + x);
}

__global__ void fill_u64(uint64_t* p, int64_t n, uint64_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void iota_i32(int32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int32_t)i;
}

__global__ void hash_insert(const int32_t* __restrict__ coords, int64_t n, Shape3 s, uint64_t* tkeys,
                            int32_t* tvals, uint64_t tmask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t* c = coords + 4 * i;
  const uint64_t key = key_of(c[0], c[1], c[2], c[3], s);
  uint64_t h = mix64(key) & tmask;
  for (uint64_t probe = 0; probe <= tmask; ++probe) {
    const unsigned long long prev =
        atomicCAS(reinterpret_cast<unsigned long long*>(tkeys + h), (unsigned long long)kEmptyKey,
                  (unsigned long long)key);
    if (prev == kEmptyKey || prev == key) {
      tvals[h] = (int32_t)i;  // duplicate coordinates keep one of their rows
      return;
    }
    h = (h + 1) & tmask;
  }
}

__device__ __forceinline__ int32_t hash_find(uint64_t key, const uint64_t* __restrict__ tkeys,
                                             const int32_t* __restrict__ tvals, uint64_t tmask) {
  uint64_t h = mix64(key) & tmask;
  for (uint64_t probe = 0; probe <= tmask; ++probe) {
    const uint64_t k = tkeys[h];
    if (k == key) return tvals[h];
    if (k == kEmptyKey) return -1;
    h = (h + 1) & tmask;
  }
  return -1;
}

// Submanifold: thread (f, i) looks up coord_i + (k_f - 1).  pair_bwd[f] is
// pair_fwd[26 - f] (the neighbour relation is symmetric).
__global__ void subm_lookup(const int32_t* __restrict__ coords, int64_t n, Shape3 s, const uint64_t* tkeys,
                            const int32_t* tvals, uint64_t tmask, int32_t* pair_fwd, int32_t* pair_bwd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= n) return;
  const int32_t* c = coords + 4 * i;
  const int kz = f / 9, ky = (f / 3) % 3, kx = f % 3;
  const int64_t z = c[1] + kz - 1, y = c[2] + ky - 1, x = c[3] + kx - 1;
  int32_t j = -1;
  if (z >= 0 && z < s.s0 && y >= 0 && y < s.s1 && x >= 0 && x < s.s2) j = hash_find(key_of(c[0], z, y, x, s), tkeys, tvals, tmask);
  pair_fwd[(int64_t)f * n + i] = j;
  if (pair_bwd) pair_bwd[(int64_t)(kKv - 1 - f) * n + i] = j;
}

// Strided (stride 2, padding 1): thread (f, o) looks up o*2 - 1 + k_f among
// the inputs; the inverse entry goes to pair_bwd[f][i] (when pair_bwd is set).
__global__ void strided_lookup(const int32_t* __restrict__ out_coords, int64_t n_out, int64_t n_in, Shape3 in_s,
                               const uint64_t* tkeys, const int32_t* tvals, uint64_t tmask, int32_t* pair_fwd,
                               int32_t* pair_bwd) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (o >= n_out) return;
  const int32_t* c = out_coords + 4 * o;
  const int kz = f / 9, ky = (f / 3) % 3, kx = f % 3;
  const int64_t z = 2 * (int64_t)c[1] - 1 + kz, y = 2 * (int64_t)c[2] - 1 + ky, x = 2 * (int64_t)c[3] - 1 + kx;
  int32_t i = -1;
  if (z >= 0 && z < in_s.s0 && y >= 0 && y < in_s.s1 && x >= 0 && x < in_s.s2)
    i = hash_find(key_of(c[0], z, y, x, in_s), tkeys, tvals, tmask);
  pair_fwd[(int64_t)f * n_out + o] = i;
  if (pair_bwd != nullptr && i >= 0 && i < n_in) pair_bwd[(int64_t)f * n_in + i] = (int32_t)o;
}

__global__ void row_masks(const int32_t* __restrict__ pair, int64_t n, uint32_t* mask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t m = 0;
#pragma unroll
  for (int f = 0; f < kKv; ++f) m |= (uint32_t)(pair[(int64_t)f * n + i] >= 0) << f;
  mask[i] = m;
}

// Output cell keys of a strided layer: rule 0 ("down") = floor(c / 2) of each
// input; rule 1 ("reach") = every output whose window o*2 - 1 + k holds the
// input (27 candidates per input, kEmptyKey where none).
__global__ void strided_cand_keys(const int32_t* __restrict__ coords, int64_t n, Shape3 out_s, int rule,
                                  uint64_t* keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t* c = coords + 4 * i;
  if (rule == 0) {
    const int64_t z = std::min<int64_t>(c[1] / 2, out_s.s0 - 1), y = std::min<int64_t>(c[2] / 2, out_s.s1 - 1),
                  x = std::min<int64_t>(c[3] / 2, out_s.s2 - 1);
    keys[i] = key_of(c[0], z, y, x, out_s);
    return;
  }
  for (int f = 0; f < kKv; ++f) {
    const int kz = f / 9, ky = (f / 3) % 3, kx = f % 3;
    const int64_t nz = c[1] + 1 - kz, ny = c[2] + 1 - ky, nx = c[3] + 1 - kx;
    uint64_t k = kEmptyKey;
    if (((nz | ny | nx) & 1) == 0 && nz >= 0 && ny >= 0 && nx >= 0) {
      const int64_t z = nz / 2, y = ny / 2, x = nx / 2;
      if (z < out_s.s0 && y < out_s.s1 && x < out_s.s2) k = key_of(c[0], z, y, x, out_s);
    }
    keys[(int64_t)f * n + i] = k;
  }
}

__global__ void decode_keys(const uint64_t* __restrict__ keys, const int64_t* n_sel, int64_t cap, Shape3 s,
                            int32_t* coords) {
  const int64_t n = cap < *n_sel ? cap : *n_sel;
  // grid-stride: the launch caps the grid at 65535 blocks
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    const int64_t x = (int64_t)(k % (uint64_t)s.s2);
    k /= (uint64_t)s.s2;
    const int64_t y = (int64_t)(k % (uint64_t)s.s1);
    k /= (uint64_t)s.s1;
    const int64_t z = (int64_t)(k % (uint64_t)s.s0);
    const int64_t b = (int64_t)(k / (uint64_t)s.s0);
    int32_t* c = coords + 4 * i;
    c[0] = (int32_t)b;
    c[1] = (int32_t)z;
    c[2] = (int32_t)y;
    c[3] = (int32_t)x;
  }
}

// ------------------------------------------------------------ workspace carving
int64_t a256(int64_t v) { return (v + 255) / 256 * 256; }

uint64_t table_cap(int64_t n) {
  uint64_t c = 64;
  while ((int64_t)c < 2 * std::max<int64_t>(n, 1)) c <<= 1;
  return c;
}

struct Carver {
  char* base;
  int64_t off = 0;
  template <typename T> T* take(int64_t count) {
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += a256(count * (int64_t)sizeof(T));
    return p;
  }
};

int64_t sort_pairs_bytes(int64_t n) {
  size_t t = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, t, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (const int32_t*)nullptr, (int32_t*)nullptr, (int)std::max<int64_t>(n, 1),
                                     0, kKv, (hipStream_t)0);
  return (int64_t)t;
}

int64_t sort_keys_bytes(int64_t n) {
  size_t t = 0;
  (void)hipcub::DeviceRadixSort::SortKeys(nullptr, t, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                    (int)std::max<int64_t>(n, 1), 0, 64, (hipStream_t)0);
  size_t u = 0;
  (void)hipcub::DeviceSelect::Unique(nullptr, u, (const uint64_t*)nullptr, (uint64_t*)nullptr, (int64_t*)nullptr,
                               (int)std::max<int64_t>(n, 1), (hipStream_t)0);
  return (int64_t)std::max(t, u);
}

thread_local char g_err[256];
int fail(const char* msg, hipError_t e = hipSuccess) {
  if (e != hipSuccess)
    snprintf(g_err, sizeof(g_err), "%s: %s", msg, hipGetErrorString(e));
  else
    snprintf(g_err, sizeof(g_err), "%s", msg);
  return e != hipSuccess ? 2 : 1;
}

#define PCHK(x)                                   \
  do {                                            \
    hipError_t e_ = (x);                          \
    if (e_ != hipSuccess) return fail(#x, e_);    \
  } while (0)

dim3 grid1(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65535))); }

// masks of pair[kv][n] -> stably sorted masks + argsort
int sort_masks(const int32_t* pair, int64_t n, uint32_t* mask_sorted, int32_t* argsort, Carver& cv,
               hipStream_t s) {
  uint32_t* masks = cv.take<uint32_t>(n);
  int32_t* iota = cv.take<int32_t>(n);
  const int64_t tb = sort_pairs_bytes(n);
  void* tmp = cv.take<char>(tb);
  if (!cv.base || n == 0) return 0;
  hipLaunchKernelGGL(row_masks, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pair, n, masks);
  PCHK(hipGetLastError());
  hipLaunchKernelGGL(iota_i32, grid1(n), dim3(256), 0, s, iota, n);
  PCHK(hipGetLastError());
  size_t t = (size_t)tb;
  PCHK(hipcub::DeviceRadixSort::SortPairs(tmp, t, masks, mask_sorted, iota, argsort, (int)n, 0, kKv, s));
  return 0;
}

int subm_impl(const int32_t* coords, int64_t n, const int32_t* shape, int32_t* pair_fwd, int32_t* pair_bwd,
              uint32_t* mask_sorted, int32_t* argsort, Carver& cv, hipStream_t s) {
  const Shape3 sh{shape[0], shape[1], shape[2]};
  const uint64_t cap = table_cap(n);
  uint64_t* tkeys = cv.take<uint64_t>((int64_t)cap);
  int32_t* tvals = cv.take<int32_t>((int64_t)cap);
  if (!cv.base) return sort_masks(nullptr, n, nullptr, nullptr, cv, s);
  if (n == 0) return 0;
  hipLaunchKernelGGL(fill_u64, grid1((int64_t)cap), dim3(256), 0, s, tkeys, (int64_t)cap, kEmptyKey);
  PCHK(hipGetLastError());
  hipLaunchKernelGGL(hash_insert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coords, n, sh, tkeys, tvals,
                     cap - 1);
  PCHK(hipGetLastError());
  hipLaunchKernelGGL(subm_lookup, dim3((unsigned)((n + 255) / 256), kKv), dim3(256), 0, s, coords, n, sh, tkeys,
                     tvals, cap - 1, pair_fwd, pair_bwd);
  PCHK(hipGetLastError());
  return sort_masks(pair_fwd, n, mask_sorted, argsort, cv, s);
}

Shape3 out_shape_of(const int32_t* in_shape) {
  // (S + 2*padding - k) / stride + 1 with k = 3, padding = 1, stride = 2
  return Shape3{(in_shape[0] - 1) / 2 + 1, (in_shape[1] - 1) / 2 + 1, (in_shape[2] - 1) / 2 + 1};
}

int strided_out_impl(const int32_t* coords, int64_t n, const int32_t* in_shape, int rule, int32_t* out_coords,
                     int64_t cap, int64_t* n_out, Carver& cv, hipStream_t s) {
  const Shape3 os = out_shape_of(in_shape);
  const int64_t nc = rule == 0 ? n : kKv * n;
  uint64_t* k0 = cv.take<uint64_t>(nc);
  uint64_t* k1 = cv.take<uint64_t>(nc);
  uint64_t* k2 = cv.take<uint64_t>(nc);
  int64_t* nsel = cv.take<int64_t>(1);
  const int64_t tb = sort_keys_bytes(nc);
  void* tmp = cv.take<char>(tb);
  if (!cv.base) return 0;
  *n_out = 0;
  if (n == 0) return 0;
  hipLaunchKernelGGL(strided_cand_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coords, n, os, rule, k0);
  PCHK(hipGetLastError());
  size_t t = (size_t)tb;
  PCHK(hipcub::DeviceRadixSort::SortKeys(tmp, t, k0, k1, (int)nc, 0, 64, s));
  t = (size_t)tb;
  PCHK(hipcub::DeviceSelect::Unique(tmp, t, k1, k2, nsel, (int)nc, s));
  int64_t cnt = 0;
  PCHK(hipMemcpyAsync(&cnt, nsel, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  PCHK(hipStreamSynchronize(s));
  if (rule == 1 && cnt > 0) {
    // the empty-candidate marker sorts last
    uint64_t last = 0;
    PCHK(hipMemcpy(&last, k2 + cnt - 1, sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (last == kEmptyKey) --cnt;
  }
  *n_out = cnt;
  if (cnt > cap) return fail("output coordinate buffer too small");
  // the device count is >= cnt (rule 1 drops the empty marker on the host);
  // decode_keys stops at min(cnt, *nsel) = cnt, queued on the stream
  hipLaunchKernelGGL(decode_keys, grid1(cnt), dim3(256), 0, s, k2, nsel, cnt, os, out_coords);
  PCHK(hipGetLastError());
  return 0;
}

int strided_impl(const int32_t* in_coords, int64_t n_in, const int32_t* in_shape, const int32_t* out_coords,
                 int64_t n_out, int32_t* pair_fwd, int32_t* pair_bwd, uint32_t* mfs, int32_t* afs, uint32_t* mbs,
                 int32_t* abs_, Carver& cv, hipStream_t s) {
  const Shape3 is{in_shape[0], in_shape[1], in_shape[2]};
  const uint64_t cap = table_cap(n_in);
  uint64_t* tkeys = cv.take<uint64_t>((int64_t)cap);
  int32_t* tvals = cv.take<int32_t>((int64_t)cap);
  if (!cv.base) {
    sort_masks(nullptr, std::max(n_in, n_out), nullptr, nullptr, cv, s);
    return 0;
  }
  if (n_in > 0) {
    hipLaunchKernelGGL(fill_u64, grid1((int64_t)cap), dim3(256), 0, s, tkeys, (int64_t)cap, kEmptyKey);
    PCHK(hipGetLastError());
    hipLaunchKernelGGL(hash_insert, dim3((unsigned)((n_in + 255) / 256)), dim3(256), 0, s, in_coords, n_in, is,
                       tkeys, tvals, cap - 1);
    PCHK(hipGetLastError());
    if (pair_bwd != nullptr) {
      hipLaunchKernelGGL(fill_i32, grid1(kKv * n_in), dim3(256), 0, s, pair_bwd, kKv * n_in, -1);
      PCHK(hipGetLastError());
    }
  }
  if (n_out > 0) {
    hipLaunchKernelGGL(strided_lookup, dim3((unsigned)((n_out + 255) / 256), kKv), dim3(256), 0, s, out_coords,
                       n_out, n_in, is, tkeys, tvals, cap - 1, pair_fwd, pair_bwd);
    PCHK(hipGetLastError());
  }
  const int64_t mark = cv.off;
  int r = sort_masks(pair_fwd, n_out, mfs, afs, cv, s);
  if (r || pair_bwd == nullptr) return r;  // forward-only tables (inference)
  cv.off = mark;  // the backward sort reuses the scratch; both are ordered on one stream
  return sort_masks(pair_bwd, n_in, mbs, abs_, cv, s);
}

bool bad_shape(const int32_t* s) { return s == nullptr || s[0] <= 0 || s[1] <= 0 || s[2] <= 0; }

}  // namespace
}  // namespace cs

extern "C" {

const char* cs_pairs_last_error(void) { return cs::g_err; }

int cs_pairs_subm_workspace(int64_t n, int64_t* nbytes) {
  if (nbytes == nullptr || n < 0) return cs::fail("bad argument");
  cs::Carver cv{nullptr};
  const int32_t shape[3] = {1, 1, 1};
  cs::subm_impl(nullptr, n, shape, nullptr, nullptr, nullptr, nullptr, cv, 0);
  *nbytes = cv.off;
  return 0;
}

int cs_pairs_subm(const int32_t* coords, int64_t n, const int32_t* spatial_shape, int32_t* pair_fwd,
                  int32_t* pair_bwd, uint32_t* mask_sorted, int32_t* argsort, void* workspace, int64_t ws_bytes,
                  uint64_t stream) {
  int64_t need = 0;
  cs_pairs_subm_workspace(n, &need);
  if (n < 0 || n > 0x7fffffff || cs::bad_shape(spatial_shape)) return cs::fail("bad size or spatial shape");
  if (ws_bytes < need || (n > 0 && (!coords || !pair_fwd || !mask_sorted || !argsort || !workspace)))
    return cs::fail("null buffer or workspace too small");
  cs::Carver cv{static_cast<char*>(workspace)};
  return cs::subm_impl(coords, n, spatial_shape, pair_fwd, pair_bwd, mask_sorted, argsort, cv,
                       reinterpret_cast<hipStream_t>(stream));
}

int cs_pairs_strided_out_workspace(int64_t n, int32_t rule, int64_t* nbytes) {
  if (nbytes == nullptr || n < 0 || (rule != 0 && rule != 1)) return cs::fail("bad argument");
  cs::Carver cv{nullptr};
  const int32_t shape[3] = {1, 1, 1};
  int64_t dummy = 0;
  cs::strided_out_impl(nullptr, n, shape, rule, nullptr, 0, &dummy, cv, 0);
  *nbytes = cv.off;
  return 0;
}

int cs_pairs_strided_out(const int32_t* coords, int64_t n, const int32_t* in_shape, int32_t rule, int32_t* out_coords,
                         int64_t cap, int64_t* n_out, void* workspace, int64_t ws_bytes, uint64_t stream) {
  int64_t need = 0;
  if (cs_pairs_strided_out_workspace(n, rule, &need)) return 1;
  if (n > 0x7fffffff / 27 || cs::bad_shape(in_shape)) return cs::fail("bad size or spatial shape");
  if (ws_bytes < need || n_out == nullptr || (n > 0 && (!coords || !out_coords || !workspace)))
    return cs::fail("null buffer or workspace too small");
  cs::Carver cv{static_cast<char*>(workspace)};
  return cs::strided_out_impl(coords, n, in_shape, rule, out_coords, cap, n_out, cv,
                              reinterpret_cast<hipStream_t>(stream));
}

int cs_pairs_strided_workspace(int64_t n_in, int64_t n_out, int64_t* nbytes) {
  if (nbytes == nullptr || n_in < 0 || n_out < 0) return cs::fail("bad argument");
  cs::Carver cv{nullptr};
  const int32_t shape[3] = {1, 1, 1};
  cs::strided_impl(nullptr, n_in, shape, nullptr, n_out, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, cv, 0);
  *nbytes = cv.off;
  return 0;
}

int cs_pairs_strided(const int32_t* in_coords, int64_t n_in, const int32_t* in_shape, const int32_t* out_coords,
                     int64_t n_out, int32_t* pair_fwd, int32_t* pair_bwd, uint32_t* mask_fwd_sorted,
                     int32_t* argsort_fwd, uint32_t* mask_bwd_sorted, int32_t* argsort_bwd, void* workspace,
                     int64_t ws_bytes, uint64_t stream) {
  int64_t need = 0;
  cs_pairs_strided_workspace(n_in, n_out, &need);
  if (n_in > 0x7fffffff / 27 || n_out > 0x7fffffff / 27 || cs::bad_shape(in_shape))
    return cs::fail("bad size or spatial shape");
  // the backward side (pair_bwd, mask_bwd_sorted, argsort_bwd) is all set or all null
  const bool bwd = pair_bwd != nullptr;
  if (bwd != (mask_bwd_sorted != nullptr) || bwd != (argsort_bwd != nullptr))
    return cs::fail("backward tables must be all set or all null");
  if (ws_bytes < need || !workspace || (n_out > 0 && (!out_coords || !pair_fwd || !mask_fwd_sorted || !argsort_fwd)) ||
      (n_in > 0 && !in_coords))
    return cs::fail("null buffer or workspace too small");
  cs::Carver cv{static_cast<char*>(workspace)};
  return cs::strided_impl(in_coords, n_in, in_shape, out_coords, n_out, pair_fwd, pair_bwd, mask_fwd_sorted,
                          argsort_fwd, mask_bwd_sorted, argsort_bwd, cv, reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
