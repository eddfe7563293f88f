// Device-side building blocks shared by the gfx950 kernels: element types,
// 16-byte fragment loads, the per-dtype MFMA step and the epilogue op.
//
// Fragment convention (all kernels): a "chunk" is one 16-byte load per lane.
// In a 16x16 MFMA tile, lane l works on row/col (l & 15) and lane group
// g = l >> 4 (0..3) supplies reduction elements [g*E, g*E+E) of the chunk,
// E = 16 / sizeof(T).  A and B use the same element order, so the chunk's
// reduction order is a permutation shared by both operands and the sum is
// unchanged.  One chunk covers 4*E reduction elements:
//   f32  : 16 elements -> 4 x v_mfma_f32_16x16x4_f32
//   f16  : 32 elements -> 1 x v_mfma_f32_16x16x32_f16
//   bf16 : 32 elements -> 1 x v_mfma_f32_16x16x32_bf16
//   e4m3 : 64 elements -> 2 x v_mfma_f32_16x16x32_fp8_fp8
//   int8 : 64 elements -> 1 x v_mfma_i32_16x16x64_i8 (int32 accumulators,
//          kept as bit patterns in the same f32x4 registers)
// The 16x16 accumulator holds, in lane l, column (l & 15) and rows
// 4*(l >> 4) + i, i = 0..3.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cs {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));

struct bf16 { uint16_t bits; };
struct e4m3 { uint8_t bits; };

// ---------------------------------------------------------------- loads
__device__ __forceinline__ u32x4 ld16(const void* p) {
  return *reinterpret_cast<const u32x4*>(p);
}
__device__ __forceinline__ u32x4 zero16() { return u32x4{0u, 0u, 0u, 0u}; }

// Raw buffer resource over [base, base + nbytes) (gfx9 dword3: 32-bit data
// format, range-checked).  A load at an offset >= nbytes returns zeros without
// touching memory, which replaces a branch or a select for missing rows.
constexpr uint32_t kBufOOB = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)nbytes, 0x00020000);
}
__device__ __forceinline__ u32x4 bld16(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  return __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0);
}

// ---------------------------------------------------------------- conversions
__device__ __forceinline__ float to_f32(float v) { return v; }
__device__ __forceinline__ float to_f32(_Float16 v) { return (float)v; }
__device__ __forceinline__ float to_f32(bf16 v) {
  return __uint_as_float(((uint32_t)v.bits) << 16);
}
__device__ __forceinline__ float to_f32(int8_t v) { return (float)v; }

template <typename OutT> __device__ __forceinline__ OutT from_f32(float v);
template <> __device__ __forceinline__ float from_f32<float>(float v) { return v; }
template <> __device__ __forceinline__ _Float16 from_f32<_Float16>(float v) {
  return (_Float16)v;  // v_cvt_f16_f32: round to nearest even
}
template <> __device__ __forceinline__ bf16 from_f32<bf16>(float v) {
  // IEEE-754 binary32 -> bfloat16, round to nearest even; NaN stays quiet NaN.
  uint32_t u = __float_as_uint(v);
  bf16 r;
  if ((u & 0x7fffffffu) > 0x7f800000u) {
    r.bits = (uint16_t)((u >> 16) | 0x40u);
  } else {
    u += 0x7fffu + ((u >> 16) & 1u);
    r.bits = (uint16_t)(u >> 16);
  }
  return r;
}
template <> __device__ __forceinline__ int8_t from_f32<int8_t>(float v) {
  // clamp to [-128, 127], then round to nearest even (NaN -> -128)
  v = fminf(fmaxf(v, -128.f), 127.f);
  return (int8_t)(int)__builtin_rintf(v);
}

// ---------------------------------------------------------------- MFMA step
template <typename T> struct Mma;

template <> struct Mma<float> {
  static constexpr int kElems = 4;   // per lane per chunk
  static constexpr int kChunk = 16;  // reduction elements per chunk
  __device__ __forceinline__ static void step(f32x4& acc, u32x4 a, u32x4 b) {
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a.x), __uint_as_float(b.x), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a.y), __uint_as_float(b.y), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a.z), __uint_as_float(b.z), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a.w), __uint_as_float(b.w), acc, 0, 0, 0);
  }
};

template <> struct Mma<_Float16> {
  static constexpr int kElems = 8;
  static constexpr int kChunk = 32;
  __device__ __forceinline__ static void step(f32x4& acc, u32x4 a, u32x4 b) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                 __builtin_bit_cast(f16x8, b), acc, 0, 0, 0);
  }
};

template <> struct Mma<bf16> {
  static constexpr int kElems = 8;
  static constexpr int kChunk = 32;
  __device__ __forceinline__ static void step(f32x4& acc, u32x4 a, u32x4 b) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                  __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
  }
};

template <> struct Mma<e4m3> {
  static constexpr int kElems = 16;
  static constexpr int kChunk = 64;
  __device__ __forceinline__ static void step(f32x4& acc, u32x4 a, u32x4 b) {
    long a0 = (long)(((uint64_t)a.y << 32) | a.x), a1 = (long)(((uint64_t)a.w << 32) | a.z);
    long b0 = (long)(((uint64_t)b.y << 32) | b.x), b1 = (long)(((uint64_t)b.w << 32) | b.z);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a1, b1, acc, 0, 0, 0);
  }
};

template <> struct Mma<int8_t> {
  static constexpr int kElems = 16;
  static constexpr int kChunk = 64;
  __device__ __forceinline__ static void step(f32x4& acc, u32x4 a, u32x4 b) {
    const i32x4 r = __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(i32x4, a), __builtin_bit_cast(i32x4, b),
                                                          __builtin_bit_cast(i32x4, acc), 0, 0, 0);
    acc = __builtin_bit_cast(f32x4, r);
  }
};

// Slot of 16-byte chunk c of row n in an LDS tile of CHUNKS chunks per row
// (CHUNKS a power of two) read one chunk per row by 16 consecutive rows at a
// time (MFMA operand reads): slot = c ^ chunk_swz<CHUNKS>(n).  LDS has 64
// banks of 4 bytes, so 16 / CHUNKS rows share one bank span; the swizzle gives
// those rows distinct slots and the 16 reads land on distinct banks.
template <int CHUNKS> __device__ __forceinline__ int chunk_swz(int n) {
  constexpr int kRowsPerSpan = CHUNKS >= 16 ? 1 : 16 / CHUNKS;
  constexpr int kMask = (CHUNKS < 16 ? CHUNKS : 16) - 1;
  return (n / kRowsPerSpan) & kMask;
}

template <typename T> struct IsInt8 { static constexpr bool value = false; };
template <> struct IsInt8<int8_t> { static constexpr bool value = true; };

// ---------------------------------------------------------------- kernel-offset masks
// Offset f of a kv-offset kernel is bit (f & 31) of word (f >> 5); kv <= 32
// uses one word, larger kernels (dynamic masks) MW words per row.
// Bits of word w that are offsets below kv.
__device__ __forceinline__ uint32_t kv_word(int kv, int w) {
  const int b = kv - 32 * w;
  return b >= 32 ? 0xffffffffu : (b > 0 ? ((1u << b) - 1u) : 0u);
}
// Bit f of a multi-word mask (f uniform; no dynamic register indexing).
template <int MW>
__device__ __forceinline__ uint32_t mask_bit(const uint32_t (&m)[MW], int f) {
  uint32_t v = m[0];
#pragma unroll
  for (int w = 1; w < MW; ++w) v = (f >> 5) == w ? m[w] : v;
  return (v >> (f & 31)) & 1u;
}
// Lowest set offset above f (f = -1: the lowest), -1 if none.
template <int MW>
__device__ __forceinline__ int next_offset(const uint32_t (&m)[MW], int f) {
  const int g = f + 1;
#pragma unroll
  for (int w = 0; w < MW; ++w) {
    if (g < 32 * (w + 1)) {
      const int b = g > 32 * w ? g - 32 * w : 0;
      const uint32_t rest = m[w] & (0xffffffffu << b);
      if (rest) return 32 * w + __builtin_ctz(rest);
    }
  }
  return -1;
}
// Reversed kernel order (dgrad of a submanifold layer): bit f of the result is
// bit kv - 1 - f of m.  Words past the mask are zero.
template <int MW>
__device__ __forceinline__ void reverse_kv(uint32_t (&m)[MW], int kv) {
  if constexpr (MW == 1) {
    m[0] = __builtin_bitreverse32(m[0]) >> (32 - kv);
  } else {
    uint32_t r[MW];
#pragma unroll
    for (int w = 0; w < MW; ++w) r[w] = __builtin_bitreverse32(m[MW - 1 - w]);
    const int sh = 32 * MW - kv, q = sh >> 5, b = sh & 31;
#pragma unroll
    for (int w = 0; w < MW; ++w) {
      uint32_t lo = 0u, hi = 0u;
#pragma unroll
      for (int j = 0; j < MW; ++j) {
        lo = j == w + q ? r[j] : lo;
        hi = j == w + q + 1 ? r[j] : hi;
      }
      m[w] = b ? ((lo >> b) | (hi << (32 - b))) : lo;
    }
  }
}

// Word w, in kernel order, of a kv-offset mask whose original words are
// word(q) (zero outside [0, div_up(kv, 32)), bits at or past kv clear):
// forward, word(w); reversed (dgrad of a submanifold layer), bit f is
// original bit kv - 1 - f.  For kv <= 32 this is reverse_kv of one word.
template <typename F>
__device__ __forceinline__ uint32_t kernel_word(const F& word, int w, int kv, bool rev) {
  if (!rev) return word(w);
  const int hi = kv - 1 - 32 * w;  // original bit of kernel bit 32w
  if (hi < 0) return 0u;
  const int lo = hi - 31;          // original bit of kernel bit 32w + 31
  uint32_t x;                      // bit j: original bit lo + j (zero below bit 0)
  if (lo >= 0) {
    const int q = lo >> 5, sh = lo & 31;
    x = word(q) >> sh;
    if (sh) x |= word(q + 1) << (32 - sh);
  } else {
    x = word(0) << (-lo);
  }
  return __builtin_bitreverse32(x);
}

// ---------------------------------------------------------------- epilogue
// D = act(alpha * acc + beta * source); act codes follow tv::gemm::Activation.
enum : int { kActNone = 0, kActReLU = 1, kActSigmoid = 2, kActLeakyReLU = 3 };

__device__ __forceinline__ float apply_act(float v, int act, float act_alpha) {
  switch (act) {
    case kActReLU: return v > 0.f ? v : 0.f;
    case kActSigmoid: return 1.f / (1.f + expf(-v));
    case kActLeakyReLU: return v >= 0.f ? v : v * act_alpha;
    default: return v;
  }
}

// Integer GEMM epilogue (LinearCombination with int32 accumulation and
// compute, cumm/gemm/output_op/linear.py:112-140 with dtype_comp s32):
// x = alpha * acc + beta * src in wrapping int32 arithmetic, then ReLU /
// LeakyReLU (x * act_alpha) on int32; alpha, beta and act_alpha are the
// caller's floats converted to int32 by truncation.  Sigmoid has no integer
// form and is refused by the host.
template <typename Args>
__device__ __forceinline__ int32_t int_epilogue(const Args& p, int32_t acc, int32_t src) {
  uint32_t x = (uint32_t)p.ialpha * (uint32_t)acc;
  if (p.ibeta != 0) x += (uint32_t)p.ibeta * (uint32_t)src;
  int32_t v = (int32_t)x;
  if (p.act_type == kActReLU) v = v > 0 ? v : 0;
  else if (p.act_type == kActLeakyReLU) v = v >= 0 ? v : (int32_t)((uint32_t)v * (uint32_t)p.iact_alpha);
  return v;
}
// int32 -> output element: int8 saturates to [-128, 127] (the saturating
// int32 -> int8 array conversion, include/tensorview/gemm/math/
// numeric_convert.h:1091-1150); int32 is stored as is.
template <typename OutT> __device__ __forceinline__ OutT int_out(int32_t v);
template <> __device__ __forceinline__ int8_t int_out<int8_t>(int32_t v) {
  return (int8_t)(v < -128 ? -128 : (v > 127 ? 127 : v));
}
template <> __device__ __forceinline__ int32_t int_out<int32_t>(int32_t v) { return v; }
__device__ __forceinline__ int32_t to_i32(int8_t v) { return v; }
__device__ __forceinline__ int32_t to_i32(int32_t v) { return v; }

// int8 inference (Int8Inference output op, cumm/gemm/output_op/linear.py:198-392,
// bias not pre-scaled): v = act(scale[n] * acc + bias[n] + beta * source[o][n]),
// fused multiply-adds in float; source is output_add (or the output) row o.
// The int8 conversion (clamp, round to nearest even) is from_f32<int8_t>.
// sc / bi are scale[n] / bias[n], which the kernels keep in registers.
template <typename Args>
__device__ __forceinline__ float i8_epilogue(const Args& p, int32_t acc, float sc, float bi, int n, int64_t o) {
  float v = __builtin_fmaf(sc, (float)acc, bi);
  if (p.beta != 0.f) v = __builtin_fmaf(p.beta, to_f32(reinterpret_cast<const int8_t*>(p.d)[o * p.ldd + n]), v);
  return apply_act(v, p.act_type, p.act_alpha);
}

}  // namespace cs
