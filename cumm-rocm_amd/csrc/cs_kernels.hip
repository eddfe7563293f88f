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
// offset, in row order.  Each thread reads PG consecutive mask words at once
// (one round trip per 256*PG groups), then a block-wide exclusive scan of the
// per-thread counts places its groups.
__global__ __launch_bounds__(256) void wgrad_plan_kernel(WgradArgs p) {
  constexpr int PG = 32;
  const int f = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ int wave_sum[4];
  if (!((p.mask_filter >> f) & 1u)) {
    if (tid == 0) p.gcount[f] = 0;
    return;
  }
  const __amdgpu_buffer_rsrc_t mrs = make_rsrc(p.mask, p.mask ? (uint32_t)p.num_mask * 4u : 0u);
  int base = 0;
  for (int c0 = 0; c0 < p.ngroups; c0 += 256 * PG) {
    const int g0 = c0 + tid * PG;
    // unconditional buffer loads (out-of-range words read as 0), all in flight at once
    uint32_t w[PG];
    if (p.mask) {
#pragma unroll
      for (int i = 0; i < PG; ++i) {
        const uint32_t wi = (uint32_t)(g0 + i) * 16u / (uint32_t)p.mask_width;
        w[i] = __builtin_amdgcn_raw_buffer_load_b32(mrs, g0 + i < p.ngroups ? wi * 4u : kBufOOB, 0, 0);
      }
    } else {
#pragma unroll
      for (int i = 0; i < PG; ++i) w[i] = g0 + i < p.ngroups ? 0xffffffffu : 0u;
    }
    uint32_t bits = 0u;
#pragma unroll
    for (int i = 0; i < PG; ++i) bits |= ((w[i] >> f) & 1u) << i;
    const int cnt = __popc(bits);
    // inclusive scan of cnt over the wave
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(incl, d);
      if (lane >= d) incl += v;
    }
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();
    int off = base + incl - cnt;
    for (int wv = 0; wv < wave; ++wv) off += wave_sum[wv];
    int* out = p.glist + (int64_t)f * p.ngroups + off;
    int k = 0;
    while (bits) {
      const int i = __builtin_ctz(bits);
      bits &= bits - 1u;
      out[k++] = g0 + i;
    }
    base += wave_sum[0] + wave_sum[1] + wave_sum[2] + wave_sum[3];
    __syncthreads();
  }
  if (tid == 0) p.gcount[f] = base;
}

// Rows staged per thread and chunk: 4 (256-row chunks) for 1- and 2-byte types,
// 2 for f32, whose 16-byte loads cover half as many channels.
template <typename T> struct WgRows { static constexpr int value = 4; };
template <> struct WgRows<float> { static constexpr int value = 2; };

template <typename T>
__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs p) {
  using MM = Mma<T>;
  constexpr int E = MM::kElems;
  constexpr int CH = MM::kChunk;
  constexpr int R = kWgradRows;
  constexpr int TS = kWgradTile;
  constexpr int PITCH = R * (int)sizeof(T) + 16;
  constexpr int CHUNKS = TS / E;          // 16-byte chunks per 64 channels
  constexpr int CPT = (CHUNKS + 3) / 4;   // chunks per thread per operand and row
  constexpr int RPT = WgRows<T>::value;   // rows per thread per chunk
  constexpr int CROWS = RPT * R;          // rows per chunk
  __shared__ __attribute__((aligned(16))) char lds_a[TS * PITCH];
  __shared__ __attribute__((aligned(16))) char lds_b[TS * PITCH];

  const int f = blockIdx.y;
  if (!((p.mask_filter >> f) & 1u)) return;
  const int count = p.gcount[f];
  const int nb = wgrad_pieces(count, p.max_blocks, p.min_groups);
  const int piece = blockIdx.z;
  if (piece >= nb) return;
  const int tiles_n = (p.N + TS - 1) / TS;
  const int m0 = (blockIdx.x / tiles_n) * TS;
  const int nn0 = (blockIdx.x % tiles_n) * TS;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int srow = tid & 63, sq = tid >> 6;
  const int gpb = (count + nb - 1) / nb;
  const int g0 = piece * gpb;
  const int g1 = min(g0 + gpb, count);
  const int nrows_blk = (g1 - g0) * 16;
  const int* glist = p.glist + (int64_t)f * p.ngroups + g0;
  const int nchunks = (nrows_blk + CROWS - 1) / CROWS;

  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  // Every index and data load is an unconditional buffer load: invalid lanes
  // use an out-of-range offset and read zeros, so no load sits behind a branch
  // and all of a chunk's loads are in flight together.
  const __amdgpu_buffer_rsrc_t rA = make_rsrc(p.a, p.a_bytes);
  const __amdgpu_buffer_rsrc_t rB = make_rsrc(p.b, p.b_bytes);
  const __amdgpu_buffer_rsrc_t rG = make_rsrc(glist, (uint32_t)(g1 - g0) * 4u);
  const __amdgpu_buffer_rsrc_t rAr = make_rsrc(p.a_rows, (uint32_t)p.nrows * 4u);
  const __amdgpu_buffer_rsrc_t rBr =
      p.pairs ? make_rsrc(p.pairs + (int64_t)f * p.pair_ld, (uint32_t)p.nrows * 4u)
              : make_rsrc(p.b_rows, (uint32_t)p.nrows * 4u);
  const bool by_pairs = p.pairs != nullptr;

  // Four-stage index pipeline per staged row: in the iteration of chunk j the
  // thread loads the data rows of chunk j, the b-row index of chunk j+1, the
  // a-row (argsort) index of chunk j+2 and the group id of chunk j+3, so each
  // chunk waits for one round trip.  A value loaded in one iteration is first
  // read in the next one.
  auto posof = [&](int j, int st) { return j * CROWS + st * R + srow; };
  auto ld_group = [&](int j, int st) -> int {
    const int pos = posof(j, st);
    return (int)__builtin_amdgcn_raw_buffer_load_b32(rG, pos < nrows_blk ? (uint32_t)(pos >> 4) * 4u : kBufOOB, 0, 0);
  };
  int graw2[RPT], araw1[RPT], r1[RPT], brw0[RPT], aeff0[RPT];
  bool v1[RPT];
  {
    int g0r[RPT], g1r[RPT];
#pragma unroll
    for (int st = 0; st < RPT; ++st) {
      g0r[st] = ld_group(0, st);
      g1r[st] = ld_group(1, st);
      graw2[st] = ld_group(2, st);
    }
    int r0[RPT], araw0[RPT];
    bool v0[RPT];
#pragma unroll
    for (int st = 0; st < RPT; ++st) {
      const int p0 = posof(0, st), p1 = posof(1, st);
      r0[st] = g0r[st] * 16 + (p0 & 15);
      r1[st] = g1r[st] * 16 + (p1 & 15);
      v0[st] = p0 < nrows_blk && r0[st] < p.nrows;
      v1[st] = p1 < nrows_blk && r1[st] < p.nrows;
      araw0[st] = (int)__builtin_amdgcn_raw_buffer_load_b32(rAr, v0[st] ? (uint32_t)r0[st] * 4u : kBufOOB, 0, 0);
      araw1[st] = (int)__builtin_amdgcn_raw_buffer_load_b32(rAr, v1[st] ? (uint32_t)r1[st] * 4u : kBufOOB, 0, 0);
    }
#pragma unroll
    for (int st = 0; st < RPT; ++st) {
      aeff0[st] = v0[st] ? araw0[st] : -1;
      const uint32_t off = by_pairs ? (uint32_t)aeff0[st] * 4u : (uint32_t)r0[st] * 4u;
      brw0[st] = (int)__builtin_amdgcn_raw_buffer_load_b32(rBr, aeff0[st] >= 0 ? off : kBufOOB, 0, 0);
    }
  }

  for (int j = 0; j < nchunks; ++j) {
    u32x4 va[RPT][CPT], vb[RPT][CPT];
#pragma unroll
    for (int st = 0; st < RPT; ++st) {
      const int rb = aeff0[st] >= 0 ? brw0[st] : -1;
      const bool ok = rb >= 0;
#pragma unroll
      for (int i = 0; i < CPT; ++i) {
        const int ch = (sq + 4 * i) * E;
        const bool oka = ok && ch < TS && m0 + ch < p.M;
        const bool okb = ok && ch < TS && nn0 + ch < p.N;
        va[st][i] = bld16(rA, oka ? (uint32_t)(((int64_t)aeff0[st] * p.lda + m0 + ch) * (int)sizeof(T)) : kBufOOB);
        vb[st][i] = bld16(rB, okb ? (uint32_t)(((int64_t)rb * p.ldb + nn0 + ch) * (int)sizeof(T)) : kBufOOB);
      }
    }
    int brw1[RPT], aeff1[RPT], araw2[RPT], r2[RPT], graw3[RPT];
    bool v2[RPT];
#pragma unroll
    for (int st = 0; st < RPT; ++st) {
      aeff1[st] = v1[st] ? araw1[st] : -1;
      const uint32_t off = by_pairs ? (uint32_t)aeff1[st] * 4u : (uint32_t)r1[st] * 4u;
      brw1[st] = (int)__builtin_amdgcn_raw_buffer_load_b32(rBr, aeff1[st] >= 0 ? off : kBufOOB, 0, 0);
      const int p2 = posof(j + 2, st);
      r2[st] = graw2[st] * 16 + (p2 & 15);
      v2[st] = p2 < nrows_blk && r2[st] < p.nrows;
      araw2[st] = (int)__builtin_amdgcn_raw_buffer_load_b32(rAr, v2[st] ? (uint32_t)r2[st] * 4u : kBufOOB, 0, 0);
      graw3[st] = ld_group(j + 3, st);
    }
#pragma unroll
    for (int st = 0; st < RPT; ++st) {
      __syncthreads();  // previous stage's MFMAs are done with the tiles
#pragma unroll
      for (int i = 0; i < CPT; ++i) {
        const int ch = (sq + 4 * i) * E;
        if (ch < TS) {
          TrStore<T>::put(lds_a, PITCH, ch, srow, va[st][i]);
          TrStore<T>::put(lds_b, PITCH, ch, srow, vb[st][i]);
        }
      }
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
    }
#pragma unroll
    for (int st = 0; st < RPT; ++st) {
      aeff0[st] = aeff1[st];
      brw0[st] = brw1[st];
      araw1[st] = araw2[st];
      r1[st] = r2[st];
      v1[st] = v2[st];
      graw2[st] = graw3[st];
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
  constexpr int L1 = 32;                     // first tree level
  constexpr int L2 = kWgradMaxBlocks / L1;   // second tree level
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
    // fixed two-level pairwise tree over the pieces, independent of scheduling
    float s2[L2];
#pragma unroll
    for (int c = 0; c < L2; ++c) {
      float v[L1];
#pragma unroll
      for (int s = 0; s < L1; ++s) {
        const int b = c * L1 + s;
        v[s] = b < nb ? src[(int64_t)b * mn] : 0.f;
      }
#pragma unroll
      for (int w = L1 / 2; w >= 1; w >>= 1)
#pragma unroll
        for (int i = 0; i < w; ++i) v[i] = v[2 * i] + v[2 * i + 1];
      s2[c] = v[0];
    }
#pragma unroll
    for (int w = L2 / 2; w >= 1; w >>= 1)
#pragma unroll
      for (int i = 0; i < w; ++i) s2[i] = s2[2 * i] + s2[2 * i + 1];
    float r = p.alpha * s2[0];
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
