// vct_mip.hip - anisotropic mip-chain builder (kernel b).
//
// Computes the same values as Renderer::filter (renderer.cpp:283-314)
// dispatching mipmap.comp (mipmap.comp:45-99) once per level: every texel of
// level k+1 in direction map d is the mean of four front-to-back alpha blends
// alpha_blend(f, b) = f + (1 - f.a) * b (mipmap.comp:40-43) of the 2x2x2 block
// below it, paired along the axis of d, then converted to the storage type.
//
// MI355X design: levels are Morton ordered, so the 2x2x2 block under output
// texel m is the eight consecutive source elements [8m, 8m+8).  A 512-lane
// workgroup owns 4096 consecutive source texels: every lane streams its
// eight texels with wide loads (64 B per lane for FP16, fully coalesced),
// produces one texel of the next level for all six maps (a shared level 0
// is read once for the six), and writes it with coalesced stores.  The
// following levels (64, 8 and 1 texel per map) are computed from the stored
// values kept in LDS, exactly as the reference's per-level dispatches read
// the level below, so the fused result is bit-identical to level-by-level.
// Every source texel is read once from HBM and every mip texel written once.
#include "vct_internal.h"

namespace vct {
namespace {

// Source texel i of mipmap.comp's voxel_offsets (mipmap.comp:10-20) sits at
// (x, y, z) = (1 - (i >> 2), 1 - ((i >> 1) & 1), 1 - (i & 1)) inside the
// block, i.e. at Morton offset x | y << 1 | z << 2.
constexpr int kMorton[8] = {7, 3, 5, 1, 6, 2, 4, 0};

// (front, back) pairs of each direction, summed in this order (mipmap.comp:58-98).
constexpr uint8_t kPairs[6][4][2] = {
    {{0, 4}, {1, 5}, {2, 6}, {3, 7}},  // -x
    {{4, 0}, {5, 1}, {6, 2}, {7, 3}},  // +x
    {{0, 2}, {1, 3}, {5, 7}, {4, 6}},  // -y
    {{2, 0}, {3, 1}, {7, 5}, {6, 4}},  // +y
    {{0, 1}, {2, 3}, {4, 5}, {6, 7}},  // -z
    {{1, 0}, {3, 2}, {5, 4}, {7, 6}},  // +z
};

template <int DT>
struct Codec;

template <>
struct Codec<VCT_RGBA8> {
  using T = uint32_t;
  static __device__ __forceinline__ float4 dec(T v) {
    return make_float4(unorm8_exact(v & 255u), unorm8_exact((v >> 8) & 255u), unorm8_exact((v >> 16) & 255u),
                       unorm8_exact(v >> 24));
  }
  static __device__ __forceinline__ T enc(float4 f) {
    return to_unorm8(f.x) | (to_unorm8(f.y) << 8) | (to_unorm8(f.z) << 16) | (to_unorm8(f.w) << 24);
  }
};
template <>
struct Codec<VCT_FP16> {
  using T = uint2;
  static __device__ __forceinline__ float4 dec(T v) {
    return make_float4(f16_to_f32(v.x & 0xffffu), f16_to_f32(v.x >> 16), f16_to_f32(v.y & 0xffffu),
                       f16_to_f32(v.y >> 16));
  }
  static __device__ __forceinline__ T enc(float4 f) {
    return make_uint2((uint32_t)f32_to_f16(f.x) | ((uint32_t)f32_to_f16(f.y) << 16),
                      (uint32_t)f32_to_f16(f.z) | ((uint32_t)f32_to_f16(f.w) << 16));
  }
};
template <>
struct Codec<VCT_BF16> {
  using T = uint2;
  static __device__ __forceinline__ float4 dec(T v) {
    return make_float4(bf16_to_f32(v.x & 0xffffu), bf16_to_f32(v.x >> 16), bf16_to_f32(v.y & 0xffffu),
                       bf16_to_f32(v.y >> 16));
  }
  static __device__ __forceinline__ T enc(float4 f) {
    return make_uint2((uint32_t)f32_to_bf16(f.x) | ((uint32_t)f32_to_bf16(f.y) << 16),
                      (uint32_t)f32_to_bf16(f.z) | ((uint32_t)f32_to_bf16(f.w) << 16));
  }
};
template <>
struct Codec<VCT_FP32> {
  using T = float4;
  static __device__ __forceinline__ float4 dec(T v) { return v; }
  static __device__ __forceinline__ T enc(float4 f) { return f; }
};

// f + (1 - f.a) * b, each product and sum rounded on its own (no fused
// multiply-add), as the oracle evaluates it.
__device__ __forceinline__ float4 alpha_blend(float4 f, float4 b) {
#pragma clang fp contract(off)
  float k = 1.0f - f.w;
  return make_float4(f.x + k * b.x, f.y + k * b.y, f.z + k * b.z, f.w + k * b.w);
}
__device__ __forceinline__ float4 add4(float4 a, float4 b) {
#pragma clang fp contract(off)
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// One mip texel of direction D from the eight source texels, in
// voxel_offsets order: ((b0 + b1) + b2) + b3, then / 4.
template <int D>
__device__ __forceinline__ float4 dir_texel(const float4 (&v)[8]) {
  float4 acc = alpha_blend(v[kPairs[D][0][0]], v[kPairs[D][0][1]]);
  acc = add4(acc, alpha_blend(v[kPairs[D][1][0]], v[kPairs[D][1][1]]));
  acc = add4(acc, alpha_blend(v[kPairs[D][2][0]], v[kPairs[D][2][1]]));
  acc = add4(acc, alpha_blend(v[kPairs[D][3][0]], v[kPairs[D][3][1]]));
  return make_float4(acc.x / 4.0f, acc.y / 4.0f, acc.z / 4.0f, acc.w / 4.0f);
}

__device__ __forceinline__ float4 dir_texel_rt(int d, const float4 (&v)[8]) {
  switch (d) {
    case 0: return dir_texel<0>(v);
    case 1: return dir_texel<1>(v);
    case 2: return dir_texel<2>(v);
    case 3: return dir_texel<3>(v);
    case 4: return dir_texel<4>(v);
    default: return dir_texel<5>(v);
  }
}

struct MipArgs {
  const void* src[6];  // source level per direction (all equal when shared)
  void* dst[4][6];     // up to four output levels per direction
  uint32_t n_src;      // texels in the source level
  int nout;            // output levels produced by this launch (1..4)
  int shared_src;      // 1: one source for all six maps (level 0)
  // level 0 -> 1 launch, paired storage: the marcher's x-paired copy of the
  // source level and of the first output level per map (null: not written)
  void* pair0;
  void* pair1[6];
  uint32_t res;        // source level edge
};

constexpr int kThreads = 512;  // one output texel of the first level per lane

// Eight consecutive stored elements as decoded texels in voxel_offsets order
// (raw: as stored, in Morton order x | y << 1 | z << 2).
template <int DT>
__device__ __forceinline__ void load_block(const typename Codec<DT>::T* p, float4 (&v)[8],
                                           typename Codec<DT>::T (&raw)[8]) {
  using T = typename Codec<DT>::T;
  if constexpr (sizeof(T) == 4) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
    const uint4 a = q[0], b = q[1];
    raw[0] = a.x; raw[1] = a.y; raw[2] = a.z; raw[3] = a.w;
    raw[4] = b.x; raw[5] = b.y; raw[6] = b.z; raw[7] = b.w;
  } else if constexpr (sizeof(T) == 8) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint4 a = q[k];
      raw[2 * k] = make_uint2(a.x, a.y);
      raw[2 * k + 1] = make_uint2(a.z, a.w);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) raw[k] = p[k];
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = Codec<DT>::dec(raw[kMorton[i]]);
}

template <int DT>
__global__ __launch_bounds__(kThreads) void mip_kernel(MipArgs a) {
  using C = Codec<DT>;
  using T = typename C::T;
  // stored values of the levels produced inside the workgroup
  __shared__ T s1[6][512];
  __shared__ T s2[6][64];
  __shared__ T s3[6][8];
  // x = 0 face of each source block (the paired copy, level 0 -> 1 launch)
  __shared__ T s0[sizeof(T) < 16 ? 512 : 1][4];

  const int dir_lo = a.shared_src ? 0 : (int)blockIdx.y;
  const int ndir = a.shared_src ? 6 : 1;
  const uint32_t per_blk = a.n_src < 4096u ? a.n_src : 4096u;
  const uint32_t n1 = per_blk >> 3;  // output texels of the first level in this block
  const uint32_t t = threadIdx.x;

  // ---- first level: straight from the source, coalesced ----------------------
  T raw[8];
  if (t < n1) {
    const uint32_t m = blockIdx.x * n1 + t;
    float4 v[8];
    if (a.shared_src) {
      load_block<DT>(reinterpret_cast<const T*>(a.src[0]) + 8ull * m, v, raw);
#pragma unroll
      for (int d = 0; d < 6; ++d) {
        const T e = C::enc(dir_texel_rt(d, v));
        reinterpret_cast<T*>(a.dst[0][d])[m] = e;
        s1[d][t] = e;
      }
      if constexpr (sizeof(T) < 16) {
        if (a.pair0) {  // the x = 0 face of this 2x2x2 block, for the -x neighbour's pairs
#pragma unroll
          for (int k = 0; k < 4; ++k) s0[t][k] = raw[2 * k];
        }
      }
    } else {
      const int d = dir_lo;
      load_block<DT>(reinterpret_cast<const T*>(a.src[d]) + 8ull * m, v, raw);
      const T e = C::enc(dir_texel_rt(d, v));
      reinterpret_cast<T*>(a.dst[0][d])[m] = e;
      s1[0][t] = e;
    }
  }
  // ---- the x-paired copy of the source and first output levels -----------------
  // Entry i of a paired map holds texel i and its +x neighbour (zero past the
  // last column).  Lane t holds the 2x2x2 source block t of the block's cube
  // (edge: 16 texels, or the whole level when smaller) and output texel t:
  // the +x neighbours inside the cube come from the neighbouring lane through
  // LDS, those of the cube's last block column from the first block of the
  // next cube in x, loaded again (its output texels recomputed from it).
  if constexpr (sizeof(T) < 16) {
    if (a.pair0) {
      using PR = decltype(pair_pack(raw[0], raw[0]));
      // 16-byte entries of whole waves go out through LDS (below)
      const bool staged = sizeof(PR) == 16 && n1 >= 64u;
      PR e0[8];  // (y, z) = (k & 1, k >> 1): entries 2k (x = 0) and 2k + 1 (x = 1)
      __syncthreads();
      if (t < n1) {
        const uint32_t edge = a.res < 16u ? a.res : 16u;  // cube edge in source texels
        const uint32_t nb = edge >> 1;                     // 2x2x2 blocks per cube edge
        const uint32_t bx = compact3(t), by = compact3(t >> 1), bz = compact3(t >> 2);
        T nb0[4];   // x = 0 face of the +x neighbour block
        T nb1[6];   // the +x neighbour output texel per map
        if (bx + 1u < nb) {
          const uint32_t tn = morton3(bx + 1u, by, bz);
#pragma unroll
          for (int k = 0; k < 4; ++k) nb0[k] = s0[tn][k];
#pragma unroll
          for (int d = 0; d < 6; ++d) nb1[d] = s1[d][tn];
        } else {
          const uint32_t c = blockIdx.x, cubes = a.res / edge;
          const uint32_t cx = compact3(c) + 1u, cy = compact3(c >> 1), cz = compact3(c >> 2);
          if (cx < cubes) {
            const uint32_t mn = morton3(cx, cy, cz) * n1 + morton3(0u, by, bz);
            float4 vn[8];
            T rn[8];
            load_block<DT>(reinterpret_cast<const T*>(a.src[0]) + 8ull * mn, vn, rn);
#pragma unroll
            for (int k = 0; k < 4; ++k) nb0[k] = rn[2 * k];
#pragma unroll
            for (int d = 0; d < 6; ++d) nb1[d] = C::enc(dir_texel_rt(d, vn));
          } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) memset(&nb0[k], 0, sizeof(T));
#pragma unroll
            for (int d = 0; d < 6; ++d) memset(&nb1[d], 0, sizeof(T));
          }
        }
        const uint32_t m = blockIdx.x * n1 + t;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          e0[2 * k] = pair_pack(raw[2 * k], raw[2 * k + 1]);
          e0[2 * k + 1] = pair_pack(raw[2 * k + 1], nb0[k]);
        }
#pragma unroll
        for (int d = 0; d < 6; ++d) reinterpret_cast<PR*>(a.pair1[d])[m] = pair_pack(s1[d][t], nb1[d]);
        if (!staged) {
          PR* p0 = reinterpret_cast<PR*>(a.pair0) + 8ull * m;
#pragma unroll
          for (int q = 0; q < 8; ++q) p0[q] = e0[q];
        }
      }
      if (staged) {
        // A lane's 8 entries are 128 contiguous bytes, so lane-by-lane
        // stores touch 64 cache lines each.  Once every wave has read its
        // neighbours from s0, each wave's 512 entries go out through its own
        // 2 KB of s0 in 4 rounds of 128: 16 lanes write 8 entries each, then
        // every lane stores 2 of them, 1 KB contiguous per store.  Entry q of
        // lane j sits in slot q ^ (j & 7) of its 128 bytes, so the 16 writing
        // lanes spread over the banks instead of sharing one 16-byte column.
        __syncthreads();
        if (t < n1) {  // n1 is a multiple of 64 here: whole waves
          const uint32_t ln = t & 63u;
          PR* buf = reinterpret_cast<PR*>(&s0[t - ln][0]);
          PR* w0 = reinterpret_cast<PR*>(a.pair0) + 8ull * (blockIdx.x * n1 + t - ln);
          const uint32_t jw = ln & 15u;  // writing lane within its round
          const uint32_t ra = ln ^ ((ln >> 3) & 7u), rb = (64u + ln) ^ ((ln >> 3) & 7u);
#pragma unroll
          for (uint32_t r = 0; r < 4; ++r) {
            if ((ln >> 4) == r) {
#pragma unroll
              for (uint32_t q = 0; q < 8; ++q) buf[jw * 8u + (q ^ (jw & 7u))] = e0[q];
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            w0[128u * r + ln] = buf[ra];
            w0[128u * r + 64u + ln] = buf[rb];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          }
        }
      }
    }
  }
  // ---- further levels from LDS --------------------------------------------------
  uint32_t n_prev = n1;
  for (int l = 1; l < a.nout; ++l) {
    __syncthreads();
    const uint32_t n = n_prev >> 3;
    if (t < n * (uint32_t)ndir) {
      const int dl = (int)(t / n);  // local direction slot
      const uint32_t q = t % n;
      const int d = dir_lo + dl;
      const T* src = l == 1 ? &s1[dl][8 * q] : (l == 2 ? &s2[dl][8 * q] : &s3[dl][8 * q]);
      float4 v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = C::dec(src[kMorton[i]]);
      const T e = C::enc(dir_texel_rt(d, v));
      reinterpret_cast<T*>(a.dst[l][d])[blockIdx.x * n + q] = e;
      if (l == 1) s2[dl][q] = e;
      else if (l == 2) s3[dl][q] = e;
    }
    n_prev = n;
  }
}

// paired: the level 0 -> 1 launch also writes the x-paired copy of levels 0
// and 1 (ctx->pair_off laid out beforehand).
template <int DT>
int launch_chain(vct_context* ctx, bool paired) {
  const int L = ctx->levels;
  int src_level = 0;
  char* base = reinterpret_cast<char*>(ctx->volume.p);
  while (src_level < L - 1) {
    const uint32_t s = ctx->res >> src_level;
    const uint32_t n_src = s * s * s;
    const uint32_t per_blk = n_src < 4096u ? n_src : 4096u;
    int nout = 0;
    for (uint32_t c = per_blk; c >= 8 && nout < 4 && src_level + nout < L - 1; c >>= 3) ++nout;
    if (nout == 0) break;
    MipArgs a;
    for (int d = 0; d < 6; ++d) a.src[d] = base + ctx->lvl_off[src_level][src_level == 0 ? 0 : d];
    for (int l = 0; l < 4; ++l)
      for (int d = 0; d < 6; ++d)
        a.dst[l][d] = (l < nout) ? (void*)(base + ctx->lvl_off[src_level + 1 + l][d]) : nullptr;
    a.n_src = n_src;
    a.nout = nout;
    a.shared_src = src_level == 0 ? 1 : 0;
    a.res = s;
    a.pair0 = nullptr;
    for (int d = 0; d < 6; ++d) a.pair1[d] = nullptr;
    if (paired && src_level == 0) {
      char* pb = reinterpret_cast<char*>(ctx->pairs.p);
      a.pair0 = pb + ctx->pair_off[0][0];
      for (int d = 0; d < 6; ++d) a.pair1[d] = pb + ctx->pair_off[1][d];
    }
    dim3 grid(n_src / per_blk, a.shared_src ? 1 : 6);
    hipLaunchKernelGGL(mip_kernel<DT>, grid, dim3(kThreads), 0, ctx->stream, a);
    VCT_HIP(hipGetLastError());
    src_level += nout;
  }
  return VCT_OK;
}

}  // namespace
}  // namespace vct

// The mip chain, then the cone marcher's x-paired copy of the finished volume:
// the first launch of the chain writes the copy of levels 0 and 1 from the
// blocks it holds, a copy pass the rest (1/17 of the texels).
int vct_launch_mip(vct_context* ctx) {
  int rc;
  const bool paired = ctx->dtype != VCT_FP32 && ctx->levels >= 2;
  if (paired && (rc = vct_pair_layout(ctx))) return rc;
  switch (ctx->dtype) {
    case VCT_RGBA8: rc = vct::launch_chain<VCT_RGBA8>(ctx, paired); break;
    case VCT_FP16: rc = vct::launch_chain<VCT_FP16>(ctx, paired); break;
    case VCT_BF16: rc = vct::launch_chain<VCT_BF16>(ctx, paired); break;
    case VCT_FP32: rc = vct::launch_chain<VCT_FP32>(ctx, false); break;
    default: return vct_set_error(VCT_ERR_INVALID_ARG, "bad dtype");
  }
  if (rc) return rc;
  if (!paired) return vct_ensure_pairs(ctx);
  if ((rc = vct_pairs_from(ctx, 2))) return rc;
  ctx->pairs_version = ctx->volume_version;
  return VCT_OK;
}
