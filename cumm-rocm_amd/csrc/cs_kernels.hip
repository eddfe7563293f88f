// gfx950 kernels for the gather-GEMM / mask-sparse implicit-GEMM path.
//
// rowgather_kernel  : (cs_rowgather.h, instantiated per input dtype in
//                     rg_<dtype>.hip) forward / dgrad, ShuffleAC / dense GEMM.
// wgrad_kernel      : sparse-conv wgrad / ShuffleAB GEMM.  A block stages 64
//                     gathered rows of both operands into LDS transposed
//                     (channel-major), so MFMA operands are contiguous along
//                     the reduction (row) dimension.  FP32 partials per split.
// splitk_reduce     : fixed pairwise tree over the split partials + epilogue.
#include "cs_device.h"
#include "cs_kernels.h"

namespace cs {

// ============================================================ wgrad
template <typename T> struct TrStore;
template <> struct TrStore<float> {
  __device__ __forceinline__ static void put(char* lds, int pitch, int ch, int row, u32x4 v) {
#pragma unroll
    for (int e = 0; e < 4; ++e)
      *reinterpret_cast<uint32_t*>(lds + (ch + e) * pitch + row * 4) = v[e];
  }
};
template <> struct TrStore<_Float16> {
  __device__ __forceinline__ static void put(char* lds, int pitch, int ch, int row, u32x4 v) {
#pragma unroll
    for (int e = 0; e < 8; ++e)
      *reinterpret_cast<uint16_t*>(lds + (ch + e) * pitch + row * 2) =
          (uint16_t)((v[e >> 1] >> (16 * (e & 1))) & 0xffffu);
  }
};
template <> struct TrStore<bf16> : TrStore<_Float16> {};
template <> struct TrStore<e4m3> {
  __device__ __forceinline__ static void put(char* lds, int pitch, int ch, int row, u32x4 v) {
#pragma unroll
    for (int e = 0; e < 16; ++e)
      *reinterpret_cast<uint8_t*>(lds + (ch + e) * pitch + row) =
          (uint8_t)((v[e >> 2] >> (8 * (e & 3))) & 0xffu);
  }
};

template <typename T>
__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs p) {
  using MM = Mma<T>;
  constexpr int E = MM::kElems;
  constexpr int CH = MM::kChunk;
  constexpr int R = kWgradRows;
  constexpr int TS = kWgradTile;
  constexpr int PITCH = R * (int)sizeof(T) + 16;
  __shared__ __attribute__((aligned(16))) char lds_a[TS * PITCH];
  __shared__ __attribute__((aligned(16))) char lds_b[TS * PITCH];

  const int f = blockIdx.y;
  if (!((p.mask_filter >> f) & 1u)) return;
  const int tiles_n = (p.N + TS - 1) / TS;
  const int m0 = (blockIdx.x / tiles_n) * TS;
  const int nn0 = (blockIdx.x % tiles_n) * TS;
  const int split = blockIdx.z;
  const int rbeg = split * p.rows_per_split;
  const int rend = min(p.nrows, rbeg + p.rows_per_split);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int srow = tid & 63, sq = tid >> 6;
  constexpr int CHUNKS = TS / E;  // 16-byte chunks per 64 channels

  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  const T* A = reinterpret_cast<const T*>(p.a);
  const T* B = reinterpret_cast<const T*>(p.b);
  for (int r0 = rbeg; r0 < rend; r0 += R) {
    // 16-row groups allowed by the reduced forward mask (block-uniform).
    uint32_t gvalid = 0;
#pragma unroll
    for (int gi = 0; gi < 4; ++gi) {
      const int rg = r0 + 16 * gi;
      bool v = rg < rend;
      if (v && p.mask != nullptr) {
        const int mi = rg / p.mask_width;
        v = mi < p.num_mask && ((p.mask[mi] >> f) & 1u);
      }
      gvalid |= (uint32_t)v << gi;
    }
    if (gvalid == 0u) continue;
    const int r = r0 + srow;
    const bool rv = r < rend && ((gvalid >> (srow >> 4)) & 1u);
    int ra = 0, rb = -1;
    if (rv) {
      ra = p.a_rows ? p.a_rows[r] : r;
      rb = p.pairs ? p.pairs[(int64_t)f * p.pair_ld + ra] : (p.b_rows ? p.b_rows[r] : r);
    }
    for (int q = sq; q < CHUNKS; q += 4) {
      const int ch = q * E;
      const u32x4 va = (rv && m0 + ch < p.M) ? ld16(A + (int64_t)ra * p.lda + m0 + ch) : zero16();
      const u32x4 vb = (rb >= 0 && nn0 + ch < p.N) ? ld16(B + (int64_t)rb * p.ldb + nn0 + ch) : zero16();
      TrStore<T>::put(lds_a, PITCH, ch, srow, va);
      TrStore<T>::put(lds_b, PITCH, ch, srow, vb);
    }
    __syncthreads();
#pragma unroll
    for (int kc = 0; kc < R; kc += CH) {
      u32x4 af[2], bfr[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        af[i] = *reinterpret_cast<const u32x4*>(lds_a + (wm + 16 * i + lr) * PITCH + (kc + lg * E) * (int)sizeof(T));
#pragma unroll
      for (int j = 0; j < 2; ++j)
        bfr[j] = *reinterpret_cast<const u32x4*>(lds_b + (wn + 16 * j + lr) * PITCH + (kc + lg * E) * (int)sizeof(T));
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) MM::step(acc[i][j], af[i], bfr[j]);
    }
    __syncthreads();
  }

  float* out = p.partial + ((int64_t)split * p.kv + f) * (int64_t)p.M * p.N;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int mm = m0 + wm + 16 * i + 4 * lg + e;
        const int nn = nn0 + wn + 16 * j + lr;
        if (mm < p.M && nn < p.N) out[(int64_t)mm * p.N + nn] = acc[i][j][e];
      }
}

// ============================================================ split-K tree reduce
template <typename OutT, int MAXS>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(ReduceArgs p) {
  const int64_t mn = (int64_t)p.M * p.N;
  const int64_t total = (int64_t)p.kv * mn;
  const int64_t sstride = total;
  OutT* out = reinterpret_cast<OutT*>(p.out);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int f = (int)(e / mn);
    if (!((p.mask_filter >> f) & 1u)) continue;
    const int64_t rem = e - (int64_t)f * mn;
    const int mm = (int)(rem / p.N);
    const int nn = (int)(rem - (int64_t)mm * p.N);
    float v[MAXS];
#pragma unroll
    for (int s = 0; s < MAXS; ++s) v[s] = s < p.splits ? p.partial[(int64_t)s * sstride + e] : 0.f;
    // fixed pairwise tree: ((v0+v1)+(v2+v3))+... independent of scheduling
#pragma unroll
    for (int w = MAXS / 2; w >= 1; w >>= 1)
#pragma unroll
      for (int i = 0; i < w; ++i) v[i] = v[2 * i] + v[2 * i + 1];
    float r = p.alpha * v[0];
    const int64_t oi = (int64_t)mm * p.o_ms + (int64_t)f * p.o_fs + nn;
    if (p.beta != 0.f) {
      const OutT* src = reinterpret_cast<const OutT*>(p.src);
      r += p.beta * to_f32(p.src_is_bias ? src[nn] : src[(int64_t)mm * p.s_ms + (int64_t)f * p.s_fs + nn]);
    }
    r = apply_act(r, p.act_type, p.act_alpha);
    out[oi] = from_f32<OutT>(r);
  }
}

// ============================================================ helpers
template <int EB>
__global__ void transpose_kernel(const char* in, char* out, int64_t R, int64_t F, int64_t Cc,
                                 int64_t in_row_stride) {
  // out[f][c][r] = in[r][f][c]
  const int64_t total = R * F * Cc;
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
       o += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = o % R;
    const int64_t fc = o / R;
    const int64_t c = fc % Cc;
    const int64_t f = fc / Cc;
    const int64_t i = r * in_row_stride + f * Cc + c;
#pragma unroll
    for (int b = 0; b < EB; ++b) out[o * EB + b] = in[i * EB + b];
  }
}

__global__ void zero_kernel(uint8_t* p, int64_t nbytes) {
  const int64_t nwords = nbytes >> 2;
  uint32_t* w = reinterpret_cast<uint32_t*>(p);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = 0u;
  const int64_t t = (nwords << 2) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nbytes) p[t] = 0u;
}

// ============================================================ launchers
int rowgather_block_n(int nout) { return nout <= 64 ? 64 : 128; }

bool rowgather_supported(int in_dtype, int out_dtype) {
  switch (in_dtype) {
    case kF32: return out_dtype == kF32;
    case kF16: return out_dtype == kF16 || out_dtype == kF32;
    case kBF16: return out_dtype == kBF16 || out_dtype == kF32;
    case kE4M3: return out_dtype == kF16 || out_dtype == kBF16 || out_dtype == kF32;
    default: return false;
  }
}

hipError_t launch_rowgather(int in_dtype, int out_dtype, int block_n, const RowGatherArgs& a, hipStream_t s) {
  switch (in_dtype) {
    case kF32: return launch_rowgather_f32(out_dtype, block_n, a, s);
    case kF16: return launch_rowgather_f16(out_dtype, block_n, a, s);
    case kBF16: return launch_rowgather_bf16(out_dtype, block_n, a, s);
    case kE4M3: return launch_rowgather_e4m3(out_dtype, block_n, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_wgrad(int in_dtype, const WgradArgs& a, hipStream_t s) {
  const int tiles = ((a.M + kWgradTile - 1) / kWgradTile) * ((a.N + kWgradTile - 1) / kWgradTile);
  dim3 grid(tiles, a.kv, a.splits);
  dim3 block(256, 1, 1);
  if (tiles == 0) return hipSuccess;
  switch (in_dtype) {
    case kF32: hipLaunchKernelGGL(wgrad_kernel<float>, grid, block, 0, s, a); break;
    case kF16: hipLaunchKernelGGL(wgrad_kernel<_Float16>, grid, block, 0, s, a); break;
    case kBF16: hipLaunchKernelGGL(wgrad_kernel<bf16>, grid, block, 0, s, a); break;
    case kE4M3: hipLaunchKernelGGL(wgrad_kernel<e4m3>, grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename OutT>
static hipError_t launch_reduce_t(const ReduceArgs& a, hipStream_t s) {
  const int64_t total = (int64_t)a.kv * a.M * a.N;
  int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
  if (blocks == 0) return hipSuccess;
  dim3 grid(blocks), block(256);
  if (a.splits <= 1)
    hipLaunchKernelGGL((splitk_reduce_kernel<OutT, 1>), grid, block, 0, s, a);
  else if (a.splits <= 8)
    hipLaunchKernelGGL((splitk_reduce_kernel<OutT, 8>), grid, block, 0, s, a);
  else if (a.splits <= 32)
    hipLaunchKernelGGL((splitk_reduce_kernel<OutT, 32>), grid, block, 0, s, a);
  else if (a.splits <= 64)
    hipLaunchKernelGGL((splitk_reduce_kernel<OutT, 64>), grid, block, 0, s, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_reduce(int out_dtype, const ReduceArgs& a, hipStream_t s) {
  switch (out_dtype) {
    case kF32: return launch_reduce_t<float>(a, s);
    case kF16: return launch_reduce_t<_Float16>(a, s);
    case kBF16: return launch_reduce_t<bf16>(a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_transpose(const void* in, void* out, int64_t R, int64_t F, int64_t Cc,
                            int64_t in_row_stride, int elem_bytes, hipStream_t s) {
  const int64_t total = R * F * Cc;
  if (total == 0) return hipSuccess;
  int blocks = (int)std::min<int64_t>((total + 255) / 256, 8192);
  dim3 grid(blocks), block(256);
  const char* i = reinterpret_cast<const char*>(in);
  char* o = reinterpret_cast<char*>(out);
  switch (elem_bytes) {
    case 1: hipLaunchKernelGGL(transpose_kernel<1>, grid, block, 0, s, i, o, R, F, Cc, in_row_stride); break;
    case 2: hipLaunchKernelGGL(transpose_kernel<2>, grid, block, 0, s, i, o, R, F, Cc, in_row_stride); break;
    case 4: hipLaunchKernelGGL(transpose_kernel<4>, grid, block, 0, s, i, o, R, F, Cc, in_row_stride); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_zero(void* p, int64_t nbytes, hipStream_t s) {
  if (nbytes <= 0) return hipSuccess;
  // the buffer base must be 4-byte aligned (device allocations are)
  const int64_t nwords = (nbytes + 3) / 4;
  int blocks = (int)std::min<int64_t>((nwords + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, s, reinterpret_cast<uint8_t*>(p), nbytes);
  return hipGetLastError();
}

}  // namespace cs
