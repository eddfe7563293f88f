// gfx950 kernels for the gather-GEMM / mask-sparse implicit-GEMM path.
//
// rowgather_kernel  : (cs_rowgather.h, instantiated per input dtype in
//                     rg_<dtype>.hip) forward / dgrad, ShuffleAC / dense GEMM.
// wgrad_kernel      : sparse-conv wgrad / ShuffleAB GEMM.  A block stages 64
//                     gathered rows of both operands into LDS transposed
//                     (channel-major), so MFMA operands are contiguous along
//                     the reduction (row) dimension.  FP32 partials per split.
// splitk_reduce     : fixed pairwise tree over the split partials + epilogue.
#include <climits>
#include <cstdint>

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

// pieces of offset f's group list: nb = min(max_blocks, div_up(count, min_groups))
__device__ __forceinline__ int wgrad_pieces(int count, int max_blocks, int min_groups) {
  const int nb = (count + min_groups - 1) / min_groups;
  return nb < max_blocks ? nb : max_blocks;
}

// One block per offset: list the 16-row groups whose reduced mask word has the
// offset, in row order.
__global__ __launch_bounds__(256) void wgrad_plan_kernel(WgradArgs p) {
  const int f = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ int wave_cnt[4];
  if (!((p.mask_filter >> f) & 1u)) {
    if (tid == 0) p.gcount[f] = 0;
    return;
  }
  int base = 0;
  for (int c = 0; c < p.ngroups; c += 256) {
    const int g = c + tid;
    bool act = g < p.ngroups;
    if (act && p.mask != nullptr) {
      const int mi = (g * 16) / p.mask_width;
      act = mi < p.num_mask && ((p.mask[mi] >> f) & 1u);
    }
    const uint64_t bal = __ballot(act);
    const int pre = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wave_cnt[wave] = __popcll(bal);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wave; ++w) off += wave_cnt[w];
    if (act) p.glist[(int64_t)f * p.ngroups + off + pre] = g;
    base += wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
    __syncthreads();
  }
  if (tid == 0) p.gcount[f] = base;
}

template <typename T>
__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs p) {
  using MM = Mma<T>;
  constexpr int E = MM::kElems;
  constexpr int CH = MM::kChunk;
  constexpr int R = kWgradRows;
  constexpr int TS = kWgradTile;
  constexpr int PITCH = R * (int)sizeof(T) + 16;
  constexpr int CHUNKS = TS / E;          // 16-byte chunks per 64 channels
  constexpr int CPT = (CHUNKS + 3) / 4;   // chunks per thread per operand
  __shared__ __attribute__((aligned(16))) char lds_a[TS * PITCH];
  __shared__ __attribute__((aligned(16))) char lds_b[TS * PITCH];
  __shared__ int s_ra[256], s_rb[256];

  const int f = blockIdx.y;
  if (!((p.mask_filter >> f) & 1u)) return;
  const int count = p.gcount[f];
  const int nb = wgrad_pieces(count, p.max_blocks, p.min_groups);
  const int piece = blockIdx.z;
  const int tiles_n = (p.N + TS - 1) / TS;
  const int m0 = (blockIdx.x / tiles_n) * TS;
  const int nn0 = (blockIdx.x % tiles_n) * TS;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int srow = tid & 63, sq = tid >> 6;
  if (piece >= nb) return;
  const int gpb = (count + nb - 1) / nb;
  const int g0 = piece * gpb;
  const int g1 = min(g0 + gpb, count);
  const int nrows_blk = (g1 - g0) * 16;
  const int* glist = p.glist + (int64_t)f * p.ngroups + g0;

  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  const T* A = reinterpret_cast<const T*>(p.a);
  const T* B = reinterpret_cast<const T*>(p.b);

  // rows of super-stage j (256 list positions): thread t resolves position j*256+t
  auto fetch_rows = [&](int j, int& ra, int& rb) {
    const int pos = j * 256 + tid;
    ra = -1;
    rb = -1;
    if (pos < nrows_blk) {
      const int r = glist[pos >> 4] * 16 + (pos & 15);
      if (r < p.nrows) {
        ra = p.a_rows ? p.a_rows[r] : r;
        rb = p.pairs ? p.pairs[(int64_t)f * p.pair_ld + ra] : (p.b_rows ? p.b_rows[r] : r);
      }
    }
  };
  // 16-byte chunks of stage st (64 rows of the current super-stage)
  auto load_data = [&](int st, u32x4 (&va)[CPT], u32x4 (&vb)[CPT]) {
    const int ra = s_ra[st * R + srow];
    const int rb = s_rb[st * R + srow];
    const bool ok = rb >= 0;
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      const int ch = (sq + 4 * i) * E;
      const bool oka = ok && ch < TS && m0 + ch < p.M;
      const bool okb = ok && ch < TS && nn0 + ch < p.N;
      const u32x4 x = ld16(oka ? A + (int64_t)ra * p.lda + m0 + ch : A);
      const u32x4 y = ld16(okb ? B + (int64_t)rb * p.ldb + nn0 + ch : B);
      va[i] = oka ? x : zero16();
      vb[i] = okb ? y : zero16();
    }
  };

  const int nsuper = (nrows_blk + 255) / 256;
  int cra, crb;
  fetch_rows(0, cra, crb);
  for (int j = 0; j < nsuper; ++j) {
    __syncthreads();  // previous super-stage done with s_ra / s_rb
    s_ra[tid] = cra;
    s_rb[tid] = crb;
    if (j + 1 < nsuper) fetch_rows(j + 1, cra, crb);
    __syncthreads();
    const int rows_left = nrows_blk - j * 256;
    const int nst = min(4, (rows_left + R - 1) / R);
    u32x4 va[CPT], vb[CPT];
    load_data(0, va, vb);
    for (int st = 0; st < nst; ++st) {
#pragma unroll
      for (int i = 0; i < CPT; ++i) {
        const int ch = (sq + 4 * i) * E;
        if (ch < TS) {
          TrStore<T>::put(lds_a, PITCH, ch, srow, va[i]);
          TrStore<T>::put(lds_b, PITCH, ch, srow, vb[i]);
        }
      }
      if (st + 1 < nst) load_data(st + 1, va, vb);
      __syncthreads();
#pragma unroll
      for (int kc = 0; kc < R; kc += CH) {
        u32x4 af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
          af[i] = *reinterpret_cast<const u32x4*>(lds_a + (wm + 16 * i + lr) * PITCH + (kc + lg * E) * (int)sizeof(T));
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
          bfr[jj] = *reinterpret_cast<const u32x4*>(lds_b + (wn + 16 * jj + lr) * PITCH + (kc + lg * E) * (int)sizeof(T));
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int jj = 0; jj < 2; ++jj) MM::step(acc[i][jj], af[i], bfr[jj]);
      }
      __syncthreads();
    }
  }

  float* out = p.partial + ((int64_t)f * p.max_blocks + piece) * (int64_t)p.M * p.N;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int mm = m0 + wm + 16 * i + 4 * lg + e;
        const int nn = nn0 + wn + 16 * jj + lr;
        if (mm < p.M && nn < p.N) out[(int64_t)mm * p.N + nn] = acc[i][jj][e];
      }
}

// ============================================================ split-K tree reduce
template <typename OutT>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(ReduceArgs p) {
  constexpr int MAXS = kWgradMaxBlocks;
  const int64_t mn = (int64_t)p.M * p.N;
  const int64_t total = (int64_t)p.kv * mn;
  OutT* out = reinterpret_cast<OutT*>(p.out);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int f = (int)(e / mn);
    if (!((p.mask_filter >> f) & 1u)) continue;
    const int64_t rem = e - (int64_t)f * mn;
    const int mm = (int)(rem / p.N);
    const int nn = (int)(rem - (int64_t)mm * p.N);
    const int nb = wgrad_pieces(p.gcount[f], p.max_blocks, p.min_groups);
    const float* src = p.partial + (int64_t)f * p.max_blocks * mn + rem;
    float v[MAXS];
#pragma unroll
    for (int s = 0; s < MAXS; ++s) v[s] = s < nb ? src[(int64_t)s * mn] : 0.f;
    // fixed pairwise tree: ((v0+v1)+(v2+v3))+... independent of scheduling
#pragma unroll
    for (int w = MAXS / 2; w >= 1; w >>= 1)
#pragma unroll
      for (int i = 0; i < w; ++i) v[i] = v[2 * i] + v[2 * i + 1];
    float r = p.alpha * v[0];
    const int64_t oi = (int64_t)mm * p.o_ms + (int64_t)f * p.o_fs + nn;
    if (p.beta != 0.f) {
      const OutT* sp = reinterpret_cast<const OutT*>(p.src);
      r += p.beta * to_f32(p.src_is_bias ? sp[nn] : sp[(int64_t)mm * p.s_ms + (int64_t)f * p.s_fs + nn]);
    }
    r = apply_act(r, p.act_type, p.act_alpha);
    out[oi] = from_f32<OutT>(r);
  }
}

// ============================================================ helpers
// out[f][c][r] = in[r*in_row_stride + f*Cc + c]: one 32x32 tile of one f per
// block, staged through LDS so both the read and the write are coalesced.
template <typename U>
__global__ __launch_bounds__(256) void transpose_kernel(const U* in, U* out, int R, int Cc,
                                                        int64_t in_row_stride) {
  __shared__ U tile[32][33];
  const int f = blockIdx.z;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
#pragma unroll
  for (int i = 0; i < 32; i += 8) {
    const int r = r0 + ty + i, c = c0 + tx;
    if (r < R && c < Cc) tile[ty + i][tx] = in[(int64_t)r * in_row_stride + (int64_t)f * Cc + c];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 32; i += 8) {
    const int c = c0 + ty + i, r = r0 + tx;
    if (r < R && c < Cc) out[((int64_t)f * Cc + c) * R + r] = tile[tx][ty + i];
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
  if (a.max_blocks < 1 || a.max_blocks > kWgradMaxBlocks || a.min_groups < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(wgrad_plan_kernel, dim3(a.kv), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int tiles = ((a.M + kWgradTile - 1) / kWgradTile) * ((a.N + kWgradTile - 1) / kWgradTile);
  dim3 grid(tiles, a.kv, a.max_blocks);
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
  if (a.max_blocks < 1 || a.max_blocks > kWgradMaxBlocks) return hipErrorInvalidValue;
  hipLaunchKernelGGL((splitk_reduce_kernel<OutT>), dim3(blocks), dim3(256), 0, s, a);
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
  if (R * F * Cc == 0) return hipSuccess;
  if (R > INT32_MAX || Cc > INT32_MAX || F > 65535) return hipErrorInvalidValue;
  dim3 grid((unsigned)((Cc + 31) / 32), (unsigned)((R + 31) / 32), (unsigned)F), block(256);
  switch (elem_bytes) {
    case 1:
      hipLaunchKernelGGL(transpose_kernel<uint8_t>, grid, block, 0, s, reinterpret_cast<const uint8_t*>(in),
                         reinterpret_cast<uint8_t*>(out), (int)R, (int)Cc, in_row_stride);
      break;
    case 2:
      hipLaunchKernelGGL(transpose_kernel<uint16_t>, grid, block, 0, s, reinterpret_cast<const uint16_t*>(in),
                         reinterpret_cast<uint16_t*>(out), (int)R, (int)Cc, in_row_stride);
      break;
    case 4:
      hipLaunchKernelGGL(transpose_kernel<uint32_t>, grid, block, 0, s, reinterpret_cast<const uint32_t*>(in),
                         reinterpret_cast<uint32_t*>(out), (int)R, (int)Cc, in_row_stride);
      break;
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
