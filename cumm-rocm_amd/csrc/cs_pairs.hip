// Device-side pair tables, per-row masks and mask argsort for sparse conv
// layers of any kernel size up to kv = 1024 (3x3x3, 5x5x5, 7x7x7, 2x2x2
// stride-2 ...): what the caller of implicit_gemm2 prepares before each layer.
//
// In the reference that caller is spconv (get_indice_pairs /
// generate_subm_conv_inds and sort_1d_by_key_split, called from
// cumm/conv/main_real_spconv.py:91-145), built on tensorview's linear-probing
// GPU hash table (include/tensorview/hash/linear.cu.h:111-430,
// hash.cu.h:43-95).  The formats are the ones implicit_gemm2 consumes:
//   pair[kv][N_out] int32 (-1 = no neighbour), offsets
//   f = (kz*K1 + ky)*K2 + kx, neighbour of a submanifold row =
//   coord + dilation*(k - K/2), of a strided output o*stride - padding + k;
//   mask[r] bit f % 32 of word f / 32 <=> pair[f][r] != -1 ([N][W] words,
//   W = div_up(kv, 32)), sorted stably by the words read as one number with
//   the last word most significant, with its argsort.  Coordinates are
//   (b, z, y, x) int32 rows; keys are ((b*S0 + z)*S1 + y)*S2 + x as uint64.
//
// Kernels: hash insert (open addressing, linear probing, 64-bit CAS), one
// thread per (offset, row) lookup writing pair_fwd / pair_bwd coalesced along
// rows (submanifold: a lookup per output row; strided: the outputs are
// hashed and each input finds the output it reaches per offset), a per-row
// mask kernel, and rocPRIM radix sorts (stable LSD, one pass
// per mask word, least significant word first) for the mask argsort and for
// the unique output coordinates of strided layers.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <algorithm>

#include "cs_device.h"
#include "cumm_amd.h"

namespace cs {
namespace {

constexpr uint64_t kEmptyKey = ~0ull;
constexpr int kMaxKv = 1024;

struct Shape3 {
  int64_t s0, s1, s2;
};

// Kernel geometry of one layer (per dimension z, y, x).
struct Geom {
  int k0, k1, k2;   // kernel size
  int st0, st1, st2;  // stride (strided layers)
  int p0, p1, p2;   // padding (strided layers)
  int d0, d1, d2;   // dilation (submanifold layers)
  __host__ __device__ int kv() const { return k0 * k1 * k2; }
};

__device__ __forceinline__ void offset_of(const Geom& g, int f, int& kz, int& ky, int& kx) {
  kz = f / (g.k1 * g.k2);
  ky = (f / g.k2) % g.k1;
  kx = f % g.k2;
}

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

// bounded: rows outside the shape are left out (their keys would alias
// other cells)
__global__ void hash_insert(const int32_t* __restrict__ coords, int64_t n, Shape3 s, uint64_t* tkeys,
                            int32_t* tvals, uint64_t tmask, int bounded) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t* c = coords + 4 * i;
  if (bounded && (c[1] < 0 || c[1] >= s.s0 || c[2] < 0 || c[2] >= s.s1 || c[3] < 0 || c[3] >= s.s2)) return;
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

// Submanifold: thread (f, i) looks up coord_i + dilation*(k_f - K/2).
// pair_bwd[f] is pair_fwd[kv - 1 - f] (odd kernel sizes: the neighbour
// relation is symmetric).
__global__ void subm_lookup(const int32_t* __restrict__ coords, int64_t n, Shape3 s, Geom g, const uint64_t* tkeys,
                            const int32_t* tvals, uint64_t tmask, int32_t* pair_fwd, int32_t* pair_bwd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= n) return;
  const int32_t* c = coords + 4 * i;
  int kz, ky, kx;
  offset_of(g, f, kz, ky, kx);
  const int64_t z = c[1] + (int64_t)g.d0 * (kz - g.k0 / 2), y = c[2] + (int64_t)g.d1 * (ky - g.k1 / 2),
                x = c[3] + (int64_t)g.d2 * (kx - g.k2 / 2);
  int32_t j = -1;
  if (z >= 0 && z < s.s0 && y >= 0 && y < s.s1 && x >= 0 && x < s.s2) j = hash_find(key_of(c[0], z, y, x, s), tkeys, tvals, tmask);
  pair_fwd[(int64_t)f * n + i] = j;
  if (pair_bwd) pair_bwd[(int64_t)(g.kv() - 1 - f) * n + i] = j;
}

// Forward-only submanifold tables with half the lookups.  For a centred odd
// kernel the neighbour relation is symmetric: coord_i + d_f = coord_j <=>
// coord_j + d_{kv-1-f} = coord_i.  Thread (f, i), f <= kv/2, looks up
// coord_i + d_f and writes pair[f][i] = j and, for f < kv/2 and j found,
// pair[kv-1-f][j] = i (one writer per entry for distinct coordinates); the
// rows f > kv/2 hold -1 beforehand, which rows nobody reaches keep.
__global__ void subm_lookup_half(const int32_t* __restrict__ coords, int64_t n, Shape3 s, Geom g,
                                 const uint64_t* tkeys, const int32_t* tvals, uint64_t tmask, int32_t* pair_fwd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= n) return;
  const int32_t* c = coords + 4 * i;
  int kz, ky, kx;
  offset_of(g, f, kz, ky, kx);
  const int64_t z = c[1] + (int64_t)g.d0 * (kz - g.k0 / 2), y = c[2] + (int64_t)g.d1 * (ky - g.k1 / 2),
                x = c[3] + (int64_t)g.d2 * (kx - g.k2 / 2);
  int32_t j = -1;
  if (z >= 0 && z < s.s0 && y >= 0 && y < s.s1 && x >= 0 && x < s.s2) j = hash_find(key_of(c[0], z, y, x, s), tkeys, tvals, tmask);
  pair_fwd[(int64_t)f * n + i] = j;
  const int kv = g.kv();
  if (f < kv / 2 && j >= 0 && j < n) pair_fwd[(int64_t)(kv - 1 - f) * n + j] = (int32_t)i;
}

// Strided, from the input side: thread (f, i) finds the one output o with
// o*stride - padding + k_f = coord_i (where the stride divides and o lies in
// the output grid) and looks o up among the outputs: pair_fwd[f][o] = i and
// pair_bwd[f][i] = o, or -1.  An input reaches at most one output per offset,
// so each pair_fwd entry has at most one writer (distinct coordinates), and
// the entries no input reaches keep the -1 written before.  About
// kv / stride^3 hash probes per input, where a lookup per output row would
// make kv of them.  Output coordinates outside the output grid get no pairs.
__global__ void strided_scatter(const int32_t* __restrict__ in_coords, int64_t n_in, Shape3 in_s, Shape3 out_s,
                                Geom g, const uint64_t* tkeys, const int32_t* tvals, uint64_t tmask, int64_t n_out,
                                int32_t* pair_fwd, int32_t* pair_bwd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= n_in) return;
  const int32_t* c = in_coords + 4 * i;
  int kz, ky, kx;
  offset_of(g, f, kz, ky, kx);
  const int64_t nz = (int64_t)c[1] + g.p0 - kz, ny = (int64_t)c[2] + g.p1 - ky, nx = (int64_t)c[3] + g.p2 - kx;
  int32_t o = -1;
  const bool inside = c[1] >= 0 && c[1] < in_s.s0 && c[2] >= 0 && c[2] < in_s.s1 && c[3] >= 0 && c[3] < in_s.s2;
  if (inside && nz >= 0 && ny >= 0 && nx >= 0 && nz % g.st0 == 0 && ny % g.st1 == 0 && nx % g.st2 == 0) {
    const int64_t z = nz / g.st0, y = ny / g.st1, x = nx / g.st2;
    if (z < out_s.s0 && y < out_s.s1 && x < out_s.s2) o = hash_find(key_of(c[0], z, y, x, out_s), tkeys, tvals, tmask);
  }
  if (o >= n_out) o = -1;
  if (o >= 0) pair_fwd[(int64_t)f * n_out + o] = (int32_t)i;
  if (pair_bwd != nullptr) pair_bwd[(int64_t)f * n_in + i] = o;
}

// mask[i][w] bit b <=> pair[32w + b][i] >= 0 (W words per row)
__global__ void row_masks(const int32_t* __restrict__ pair, int64_t n, int kv, uint32_t* mask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int W = (kv + 31) / 32;
  if (W == 1 && (int64_t)kv * n * 4 < ((int64_t)1 << 31)) {
    // one word: the row's kv entries load together as unconditional buffer
    // loads (offsets f >= kv are out of range, read 0 and are masked by
    // f < kv), so the row costs one round trip instead of one per offset
    const __amdgpu_buffer_rsrc_t rp = make_rsrc(pair, (uint32_t)((int64_t)kv * n * 4));
    int32_t v[32];
#pragma unroll
    for (int f = 0; f < 32; ++f)
      v[f] = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(
          rp, f < kv ? (uint32_t)(((int64_t)f * n + i) * 4) : kBufOOB, 0, 0);
    uint32_t m = 0;
#pragma unroll
    for (int f = 0; f < 32; ++f) m |= (uint32_t)(f < kv && v[f] >= 0) << f;
    mask[i] = m;
    return;
  }
  for (int w = 0; w < W; ++w) {
    uint32_t m = 0;
    const int f1 = min(kv, 32 * w + 32);
    for (int f = 32 * w; f < f1; ++f) m |= (uint32_t)(pair[(int64_t)f * n + i] >= 0) << (f - 32 * w);
    mask[i * W + w] = m;
  }
}

// key[i] = word w of row perm[i] (one LSD pass of the multi-word mask sort)
__global__ void gather_word(const uint32_t* __restrict__ mask, const int32_t* __restrict__ perm, int64_t n, int W,
                            int w, uint32_t* key) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    key[i] = mask[(int64_t)perm[i] * W + w];
}

// sorted[r] = mask[perm[r]] (W words) and argsort[r] = perm[r]
__global__ void gather_rows(const uint32_t* __restrict__ mask, const int32_t* __restrict__ perm, int64_t n, int W,
                            uint32_t* sorted, int32_t* argsort) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t i = perm[r];
    for (int w = 0; w < W; ++w) sorted[r * W + w] = mask[(int64_t)i * W + w];
    argsort[r] = i;
  }
}

// Output cell keys of a strided layer: rule 0 ("down") = floor(c / stride) of
// each input (clamped to the output shape); rule 1 ("reach") = every output
// whose window o*stride - padding + k holds the input (kv candidates per
// input, kEmptyKey where none).
__global__ void strided_cand_keys(const int32_t* __restrict__ coords, int64_t n, Shape3 out_s, Geom g, int rule,
                                  uint64_t* keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t* c = coords + 4 * i;
  if (rule == 0) {
    const int64_t z = std::min<int64_t>(c[1] / g.st0, out_s.s0 - 1), y = std::min<int64_t>(c[2] / g.st1, out_s.s1 - 1),
                  x = std::min<int64_t>(c[3] / g.st2, out_s.s2 - 1);
    keys[i] = key_of(c[0], z, y, x, out_s);
    return;
  }
  const int kv = g.kv();
  for (int f = 0; f < kv; ++f) {
    int kz, ky, kx;
    offset_of(g, f, kz, ky, kx);
    const int64_t nz = (int64_t)c[1] + g.p0 - kz, ny = (int64_t)c[2] + g.p1 - ky, nx = (int64_t)c[3] + g.p2 - kx;
    uint64_t k = kEmptyKey;
    if (nz >= 0 && ny >= 0 && nx >= 0 && nz % g.st0 == 0 && ny % g.st1 == 0 && nx % g.st2 == 0) {
      const int64_t z = nz / g.st0, y = ny / g.st1, x = nx / g.st2;
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
                                     0, 32, (hipStream_t)0);
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

// masks of pair[kv][n] -> stably sorted masks ([n][W] words) + argsort.  One
// word: one key/value radix sort over the kv mask bits.  Several words: LSD
// passes from word 0 to word W - 1, each a stable sort of the permutation by
// that word, so the final order is by the words as one number with the last
// word most significant.
int sort_masks(const int32_t* pair, int64_t n, int kv, uint32_t* mask_sorted, int32_t* argsort, Carver& cv,
               hipStream_t s) {
  const int W = (kv + 31) / 32;
  uint32_t* masks = cv.take<uint32_t>(n * W);
  int32_t* iota = cv.take<int32_t>(n);
  uint32_t* key_in = W > 1 ? cv.take<uint32_t>(n) : nullptr;
  uint32_t* key_out = W > 1 ? cv.take<uint32_t>(n) : nullptr;
  int32_t* perm2 = W > 1 ? cv.take<int32_t>(n) : nullptr;
  const int64_t tb = sort_pairs_bytes(n);
  void* tmp = cv.take<char>(tb);
  if (!cv.base || n == 0) return 0;
  hipLaunchKernelGGL(row_masks, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pair, n, kv, masks);
  PCHK(hipGetLastError());
  hipLaunchKernelGGL(iota_i32, grid1(n), dim3(256), 0, s, iota, n);
  PCHK(hipGetLastError());
  size_t t = (size_t)tb;
  if (W == 1) {
    PCHK(hipcub::DeviceRadixSort::SortPairs(tmp, t, masks, mask_sorted, iota, argsort, (int)n, 0, kv, s));
    return 0;
  }
  int32_t* pa = iota;
  int32_t* pb = perm2;
  for (int w = 0; w < W; ++w) {
    hipLaunchKernelGGL(gather_word, grid1(n), dim3(256), 0, s, masks, pa, n, W, w, key_in);
    PCHK(hipGetLastError());
    const int end_bit = w == W - 1 ? kv - 32 * w : 32;
    t = (size_t)tb;
    PCHK(hipcub::DeviceRadixSort::SortPairs(tmp, t, key_in, key_out, pa, pb, (int)n, 0, end_bit, s));
    std::swap(pa, pb);
  }
  hipLaunchKernelGGL(gather_rows, grid1(n), dim3(256), 0, s, masks, pa, n, W, mask_sorted, argsort);
  PCHK(hipGetLastError());
  return 0;
}

int subm_impl(const int32_t* coords, int64_t n, const int32_t* shape, const Geom& g, int32_t* pair_fwd,
              int32_t* pair_bwd, uint32_t* mask_sorted, int32_t* argsort, Carver& cv, hipStream_t s) {
  const Shape3 sh{shape[0], shape[1], shape[2]};
  const uint64_t cap = table_cap(n);
  uint64_t* tkeys = cv.take<uint64_t>((int64_t)cap);
  int32_t* tvals = cv.take<int32_t>((int64_t)cap);
  if (!cv.base) return sort_masks(nullptr, n, g.kv(), nullptr, nullptr, cv, s);
  if (n == 0) return 0;
  hipLaunchKernelGGL(fill_u64, grid1((int64_t)cap), dim3(256), 0, s, tkeys, (int64_t)cap, kEmptyKey);
  PCHK(hipGetLastError());
  hipLaunchKernelGGL(hash_insert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coords, n, sh, tkeys, tvals,
                     cap - 1, 0);
  PCHK(hipGetLastError());
  if (pair_bwd == nullptr) {
    // forward-only: half the lookups, the mirrored offsets by symmetry
    const int ctr = g.kv() / 2;
    const int64_t upper = (int64_t)(g.kv() - 1 - ctr) * n;
    if (upper > 0) {
      hipLaunchKernelGGL(fill_i32, grid1(upper), dim3(256), 0, s, pair_fwd + (int64_t)(ctr + 1) * n, upper, -1);
      PCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(subm_lookup_half, dim3((unsigned)((n + 255) / 256), (unsigned)(ctr + 1)), dim3(256), 0, s,
                       coords, n, sh, g, tkeys, tvals, cap - 1, pair_fwd);
  } else {
    hipLaunchKernelGGL(subm_lookup, dim3((unsigned)((n + 255) / 256), (unsigned)g.kv()), dim3(256), 0, s, coords,
                       n, sh, g, tkeys, tvals, cap - 1, pair_fwd, pair_bwd);
  }
  PCHK(hipGetLastError());
  if (mask_sorted == nullptr) return 0;  // pair tables only (row masks from cs_pairs_row_masks)
  return sort_masks(pair_fwd, n, g.kv(), mask_sorted, argsort, cv, s);
}

// (S + 2*padding - k) / stride + 1 per dimension
Shape3 out_shape_of(const int32_t* in_shape, const Geom& g) {
  return Shape3{(in_shape[0] + 2 * g.p0 - g.k0) / g.st0 + 1, (in_shape[1] + 2 * g.p1 - g.k1) / g.st1 + 1,
                (in_shape[2] + 2 * g.p2 - g.k2) / g.st2 + 1};
}

int strided_out_impl(const int32_t* coords, int64_t n, const int32_t* in_shape, const Geom& g, int rule,
                     int32_t* out_coords, int64_t cap, int64_t* n_out, Carver& cv, hipStream_t s) {
  const Shape3 os = out_shape_of(in_shape, g);
  const int64_t nc = rule == 0 ? n : (int64_t)g.kv() * n;
  uint64_t* k0 = cv.take<uint64_t>(nc);
  uint64_t* k1 = cv.take<uint64_t>(nc);
  uint64_t* k2 = cv.take<uint64_t>(nc);
  int64_t* nsel = cv.take<int64_t>(1);
  const int64_t tb = sort_keys_bytes(nc);
  void* tmp = cv.take<char>(tb);
  if (!cv.base) return 0;
  *n_out = 0;
  if (n == 0) return 0;
  hipLaunchKernelGGL(strided_cand_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coords, n, os, g, rule,
                     k0);
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

int strided_impl(const int32_t* in_coords, int64_t n_in, const int32_t* in_shape, const Geom& g,
                 const int32_t* out_coords, int64_t n_out, int32_t* pair_fwd, int32_t* pair_bwd, uint32_t* mfs,
                 int32_t* afs, uint32_t* mbs, int32_t* abs_, Carver& cv, hipStream_t s) {
  const int kv = g.kv();
  const Shape3 is{in_shape[0], in_shape[1], in_shape[2]};
  // the hash holds the outputs (strided_scatter); sized for either side
  const uint64_t cap = table_cap(std::max(n_in, n_out));
  uint64_t* tkeys = cv.take<uint64_t>((int64_t)cap);
  int32_t* tvals = cv.take<int32_t>((int64_t)cap);
  if (!cv.base) {
    sort_masks(nullptr, std::max(n_in, n_out), kv, nullptr, nullptr, cv, s);
    return 0;
  }
  const Shape3 os = out_shape_of(in_shape, g);
  if (n_out > 0) {
    hipLaunchKernelGGL(fill_u64, grid1((int64_t)cap), dim3(256), 0, s, tkeys, (int64_t)cap, kEmptyKey);
    PCHK(hipGetLastError());
    hipLaunchKernelGGL(hash_insert, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0, s, out_coords, n_out, os,
                       tkeys, tvals, cap - 1, 1);
    PCHK(hipGetLastError());
    hipLaunchKernelGGL(fill_i32, grid1(kv * n_out), dim3(256), 0, s, pair_fwd, kv * n_out, -1);
    PCHK(hipGetLastError());
  }
  if (n_in > 0 && n_out > 0) {
    hipLaunchKernelGGL(strided_scatter, dim3((unsigned)((n_in + 255) / 256), (unsigned)kv), dim3(256), 0, s,
                       in_coords, n_in, is, os, g, tkeys, tvals, cap - 1, n_out, pair_fwd, pair_bwd);
    PCHK(hipGetLastError());
  } else if (n_in > 0 && pair_bwd != nullptr) {
    hipLaunchKernelGGL(fill_i32, grid1(kv * n_in), dim3(256), 0, s, pair_bwd, kv * n_in, -1);
    PCHK(hipGetLastError());
  }
  const int64_t mark = cv.off;
  // mfs null: pair tables only (row masks from cs_pairs_row_masks)
  int r = mfs != nullptr ? sort_masks(pair_fwd, n_out, kv, mfs, afs, cv, s) : 0;
  if (r || pair_bwd == nullptr) return r;  // forward-only tables (inference)
  cv.off = mark;  // the backward sort reuses the scratch; both are ordered on one stream
  return sort_masks(pair_bwd, n_in, kv, mbs, abs_, cv, s);
}

bool bad_shape(const int32_t* s) { return s == nullptr || s[0] <= 0 || s[1] <= 0 || s[2] <= 0; }

Geom geom_of(const cs_conv_geom* g) {
  Geom r;
  r.k0 = g->ksize[0], r.k1 = g->ksize[1], r.k2 = g->ksize[2];
  r.st0 = g->stride[0], r.st1 = g->stride[1], r.st2 = g->stride[2];
  r.p0 = g->padding[0], r.p1 = g->padding[1], r.p2 = g->padding[2];
  r.d0 = g->dilation[0], r.d1 = g->dilation[1], r.d2 = g->dilation[2];
  return r;
}

// 3x3x3; strided layers stride 2, padding 1
cs_conv_geom default_geom() {
  cs_conv_geom g;
  for (int i = 0; i < 3; ++i) {
    g.ksize[i] = 3;
    g.stride[i] = 2;
    g.padding[i] = 1;
    g.dilation[i] = 1;
  }
  return g;
}

// nullptr when the geometry is usable, else the reason
const char* bad_geom(const cs_conv_geom* g, bool subm) {
  if (g == nullptr) return "null geometry";
  int64_t kv = 1;
  for (int i = 0; i < 3; ++i) {
    if (g->ksize[i] < 1) return "kernel sizes must be positive";
    kv *= g->ksize[i];
    if (subm) {
      if (g->ksize[i] % 2 == 0) return "submanifold kernel sizes must be odd";
      if (g->dilation[i] < 1) return "dilation must be positive";
    } else {
      if (g->stride[i] < 1 || g->padding[i] < 0) return "stride must be positive and padding non-negative";
    }
  }
  if (kv > kMaxKv) return "kernel volume above 1024";
  return nullptr;
}

}  // namespace
}  // namespace cs

extern "C" {

const char* cs_pairs_last_error(void) { return cs::g_err; }

int cs_pairs_subm_geom_workspace(int64_t n, const cs_conv_geom* geom, int64_t* nbytes) {
  if (nbytes == nullptr || n < 0) return cs::fail("bad argument");
  if (const char* why = cs::bad_geom(geom, true)) return cs::fail(why);
  cs::Carver cv{nullptr};
  const int32_t shape[3] = {1, 1, 1};
  cs::subm_impl(nullptr, n, shape, cs::geom_of(geom), nullptr, nullptr, nullptr, nullptr, cv, 0);
  *nbytes = cv.off;
  return 0;
}

int cs_pairs_subm_geom(const int32_t* coords, int64_t n, const int32_t* spatial_shape, const cs_conv_geom* geom,
                       int32_t* pair_fwd, int32_t* pair_bwd, uint32_t* mask_sorted, int32_t* argsort, void* workspace,
                       int64_t ws_bytes, uint64_t stream) {
  int64_t need = 0;
  if (cs_pairs_subm_geom_workspace(n, geom, &need)) return 1;
  if (n > 0x7fffffff || cs::bad_shape(spatial_shape)) return cs::fail("bad size or spatial shape");
  // mask_sorted and argsort are both set (sorted masks) or both null (pair tables only)
  if ((mask_sorted == nullptr) != (argsort == nullptr)) return cs::fail("mask_sorted and argsort must be both set or both null");
  if (ws_bytes < need || (n > 0 && (!coords || !pair_fwd || !workspace)))
    return cs::fail("null buffer or workspace too small");
  cs::Carver cv{static_cast<char*>(workspace)};
  return cs::subm_impl(coords, n, spatial_shape, cs::geom_of(geom), pair_fwd, pair_bwd, mask_sorted, argsort, cv,
                       reinterpret_cast<hipStream_t>(stream));
}

int cs_pairs_subm_workspace(int64_t n, int64_t* nbytes) {
  const cs_conv_geom g = cs::default_geom();
  return cs_pairs_subm_geom_workspace(n, &g, nbytes);
}

int cs_pairs_subm(const int32_t* coords, int64_t n, const int32_t* spatial_shape, int32_t* pair_fwd,
                  int32_t* pair_bwd, uint32_t* mask_sorted, int32_t* argsort, void* workspace, int64_t ws_bytes,
                  uint64_t stream) {
  const cs_conv_geom g = cs::default_geom();
  return cs_pairs_subm_geom(coords, n, spatial_shape, &g, pair_fwd, pair_bwd, mask_sorted, argsort, workspace,
                            ws_bytes, stream);
}

int cs_pairs_strided_out_geom_workspace(int64_t n, const cs_conv_geom* geom, int32_t rule, int64_t* nbytes) {
  if (nbytes == nullptr || n < 0 || (rule != 0 && rule != 1)) return cs::fail("bad argument");
  if (const char* why = cs::bad_geom(geom, false)) return cs::fail(why);
  cs::Carver cv{nullptr};
  const int32_t shape[3] = {1, 1, 1};
  int64_t dummy = 0;
  cs::strided_out_impl(nullptr, n, shape, cs::geom_of(geom), rule, nullptr, 0, &dummy, cv, 0);
  *nbytes = cv.off;
  return 0;
}

int cs_pairs_strided_out_geom(const int32_t* coords, int64_t n, const int32_t* in_shape, const cs_conv_geom* geom,
                              int32_t rule, int32_t* out_coords, int64_t cap, int64_t* n_out, void* workspace,
                              int64_t ws_bytes, uint64_t stream) {
  int64_t need = 0;
  if (cs_pairs_strided_out_geom_workspace(n, geom, rule, &need)) return 1;
  const cs::Geom g = cs::geom_of(geom);
  if (n > 0x7fffffff / g.kv() || cs::bad_shape(in_shape)) return cs::fail("bad size or spatial shape");
  const cs::Shape3 os = cs::out_shape_of(in_shape, g);
  if (os.s0 < 1 || os.s1 < 1 || os.s2 < 1) return cs::fail("kernel larger than the padded input");
  if (ws_bytes < need || n_out == nullptr || (n > 0 && (!coords || !out_coords || !workspace)))
    return cs::fail("null buffer or workspace too small");
  cs::Carver cv{static_cast<char*>(workspace)};
  return cs::strided_out_impl(coords, n, in_shape, g, rule, out_coords, cap, n_out, cv,
                              reinterpret_cast<hipStream_t>(stream));
}

int cs_pairs_strided_out_workspace(int64_t n, int32_t rule, int64_t* nbytes) {
  const cs_conv_geom g = cs::default_geom();
  return cs_pairs_strided_out_geom_workspace(n, &g, rule, nbytes);
}

int cs_pairs_strided_out(const int32_t* coords, int64_t n, const int32_t* in_shape, int32_t rule, int32_t* out_coords,
                         int64_t cap, int64_t* n_out, void* workspace, int64_t ws_bytes, uint64_t stream) {
  const cs_conv_geom g = cs::default_geom();
  return cs_pairs_strided_out_geom(coords, n, in_shape, &g, rule, out_coords, cap, n_out, workspace, ws_bytes,
                                   stream);
}

int cs_pairs_strided_geom_workspace(int64_t n_in, int64_t n_out, const cs_conv_geom* geom, int64_t* nbytes) {
  if (nbytes == nullptr || n_in < 0 || n_out < 0) return cs::fail("bad argument");
  if (const char* why = cs::bad_geom(geom, false)) return cs::fail(why);
  cs::Carver cv{nullptr};
  const int32_t shape[3] = {1, 1, 1};
  cs::strided_impl(nullptr, n_in, shape, cs::geom_of(geom), nullptr, n_out, nullptr, nullptr, nullptr, nullptr,
                   nullptr, nullptr, cv, 0);
  *nbytes = cv.off;
  return 0;
}

int cs_pairs_strided_geom(const int32_t* in_coords, int64_t n_in, const int32_t* in_shape, const cs_conv_geom* geom,
                          const int32_t* out_coords, int64_t n_out, int32_t* pair_fwd, int32_t* pair_bwd,
                          uint32_t* mask_fwd_sorted, int32_t* argsort_fwd, uint32_t* mask_bwd_sorted,
                          int32_t* argsort_bwd, void* workspace, int64_t ws_bytes, uint64_t stream) {
  int64_t need = 0;
  if (cs_pairs_strided_geom_workspace(n_in, n_out, geom, &need)) return 1;
  const cs::Geom g = cs::geom_of(geom);
  if (n_in > 0x7fffffff / g.kv() || n_out > 0x7fffffff / g.kv() || cs::bad_shape(in_shape))
    return cs::fail("bad size or spatial shape");
  // the backward side (pair_bwd, mask_bwd_sorted, argsort_bwd) is all set or all null
  const bool bwd = pair_bwd != nullptr;
  if (bwd != (mask_bwd_sorted != nullptr) || bwd != (argsort_bwd != nullptr))
    return cs::fail("backward tables must be all set or all null");
  // the forward masks are sorted (both set) or left to cs_pairs_row_masks (both null)
  if ((mask_fwd_sorted == nullptr) != (argsort_fwd == nullptr))
    return cs::fail("mask_fwd_sorted and argsort_fwd must be both set or both null");
  if (ws_bytes < need || !workspace || (n_out > 0 && (!out_coords || !pair_fwd)) || (n_in > 0 && !in_coords))
    return cs::fail("null buffer or workspace too small");
  cs::Carver cv{static_cast<char*>(workspace)};
  return cs::strided_impl(in_coords, n_in, in_shape, g, out_coords, n_out, pair_fwd, pair_bwd, mask_fwd_sorted,
                          argsort_fwd, mask_bwd_sorted, argsort_bwd, cv, reinterpret_cast<hipStream_t>(stream));
}

int cs_pairs_row_masks(const int32_t* pair, int64_t n, int32_t kv, uint32_t* mask, int32_t* identity,
                       uint64_t stream) {
  if (n < 0 || kv < 1 || kv > cs::kMaxKv || (n > 0 && (pair == nullptr || mask == nullptr)))
    return cs::fail("bad argument");
  if (n == 0) return 0;
  if (n > 0x7fffffff / kv) return cs::fail("bad size");
  const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(cs::row_masks, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pair, n, kv, mask);
  const hipError_t em = hipGetLastError();
  if (em != hipSuccess) return cs::fail("row_masks launch failed", em);
  if (identity != nullptr) {
    hipLaunchKernelGGL(cs::iota_i32, cs::grid1(n), dim3(256), 0, s, identity, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cs::fail("iota launch failed", e);
  }
  return 0;
}

int cs_pairs_strided_workspace(int64_t n_in, int64_t n_out, int64_t* nbytes) {
  const cs_conv_geom g = cs::default_geom();
  return cs_pairs_strided_geom_workspace(n_in, n_out, &g, nbytes);
}

int cs_pairs_strided(const int32_t* in_coords, int64_t n_in, const int32_t* in_shape, const int32_t* out_coords,
                     int64_t n_out, int32_t* pair_fwd, int32_t* pair_bwd, uint32_t* mask_fwd_sorted,
                     int32_t* argsort_fwd, uint32_t* mask_bwd_sorted, int32_t* argsort_bwd, void* workspace,
                     int64_t ws_bytes, uint64_t stream) {
  const cs_conv_geom g = cs::default_geom();
  return cs_pairs_strided_geom(in_coords, n_in, in_shape, &g, out_coords, n_out, pair_fwd, pair_bwd, mask_fwd_sorted,
                               argsort_fwd, mask_bwd_sorted, argsort_bwd, workspace, ws_bytes, stream);
}

}  // extern "C"
